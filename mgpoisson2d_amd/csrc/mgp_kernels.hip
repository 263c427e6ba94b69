// mgp_kernels.hip -- CDNA4 (gfx950) kernels of the semicoarsening multigrid
// hot path.  FP64 throughout; HBM-bandwidth bound (about 0.8 flop/byte), so
// the kernels are written for coalesced row access (i contiguous, 64-wide
// wavefronts along i) and few passes, not for MFMA.
//
// Every kernel operates on the padded level layout of mgp_common.h.
#include <hip/hip_runtime.h>

#include <string>

#include "mgp_host.h"
#include "mgp_kernels.h"
#include "mgp_reduce.h"
#include "mgp_stencil.h"

namespace mgp {

// ===========================================================================
// Ghost fills (arrayutils.f90:871-1492)
// ===========================================================================
__device__ void neum_x_ghost(Fld f, const Grid &g, const Geo &geo, const Bc &bc, int homog, int side, int j) {
    const bool lo = side == XLO;
    const int e = lo ? -1 : 1, i = lo ? 0 : g.nx - 1, ny = g.ny;
    const double half = 0.5, threehalves = 1.5;
    const double dx = lo ? -g.dx : g.dx, dy = g.dy;
    if (bc.mode[side] == MODE_LAH) {
        f(i + e, j) = f(i, j);
        return;
    }
    const int nu = bc.mode[side] == MODE_NONUNIFORM;
    const int P = g.pitch;
    const double Jxy = geo.Jxy[ix(P, i, j)], Jxx = geo.Jxx[ix(P, i, j)];
    // homogeneous: "(-cross)" is evaluated as (-0.0 - cross), identical for every cross
    const bool end_lo = homog && j == 0;
    const bool end_hi = j == ny - 1;
    double gv;
    if (homog) gv = -0.0;
    else gv = bcval(bc, side, nu ? j : 0);
    double cross;
    if (end_lo) {
        const double dpn = -(3.0 * f(i, j) - 4.0 * f(i, j + 1) + f(i, j + 2)) * half / dy;
        const double dpf = -(3.0 * f(i - e, j) - 4.0 * f(i - e, j + 1) + f(i - e, j + 2)) * half / dy;
        cross = half * (3.0 * dpn - dpf) * Jxy;
    } else if (end_hi) {
        const double dpn = (3.0 * f(i, j) - 4.0 * f(i, j - 1) + f(i, j - 2)) * half / dy;
        const double dpf = (3.0 * f(i - e, j) - 4.0 * f(i - e, j - 1) + f(i - e, j - 2)) * half / dy;
        cross = half * (3.0 * dpn - dpf) * Jxy;
    } else {
        const double dpn = (f(i, j + 1) - f(i, j - 1)) * half / dy;
        const double dpf = (f(i - e, j + 1) - f(i - e, j - 1)) * half / dy;
        cross = (threehalves * dpn - half * dpf) * Jxy;
    }
    f(i + e, j) = f(i, j) - (gv - cross) * dx / Jxx;
}

__device__ void neum_y_ghost(Fld f, const Grid &g, const Geo &geo, const Bc &bc, int homog, int side, int i) {
    const bool lo = side == YLO;
    const int e = lo ? -1 : 1, j = lo ? 0 : g.ny - 1, nx = g.nx;
    const double half = 0.5, threehalves = 1.5;
    const double dy = lo ? -g.dy : g.dy, dx = g.dx;
    if (bc.mode[side] == MODE_LAH) {
        f(i, j + e) = f(i, j);
        return;
    }
    const int nu = bc.mode[side] == MODE_NONUNIFORM;
    const int P = g.pitch;
    const double Jyx = geo.Jyx[ix(P, i, j)], Jyy = geo.Jyy[ix(P, i, j)];
    double gv;
    if (homog) gv = -0.0;
    else gv = bcval(bc, side, nu ? i : 0);
    double cross;
    // the y-side loops cover ilo..ihi in every branch, so only the ihi end
    // keeps its one-sided stencil (arrayutils.f90:1002-1012, 1022-1032, 1042-1052)
    if (i == nx - 1) {
        const double dpn = (3.0 * f(i, j) - 4.0 * f(i - 1, j) + f(i - 2, j)) * half / dx;
        const double dpf = (3.0 * f(i, j - e) - 4.0 * f(i - 1, j - e) + f(i - 2, j - e)) * half / dx;
        cross = half * (3.0 * dpn - dpf) * Jyx;
    } else {
        const double dpn = (f(i + 1, j) - f(i - 1, j)) * half / dx;
        const double dpf = (f(i + 1, j - e) - f(i - 1, j - e)) * half / dx;
        cross = (threehalves * dpn - half * dpf) * Jyx;
    }
    f(i, j + e) = f(i, j) - (gv - cross) * dy / Jyy;
}

// x sides iterate j = -1 .. ny (the extrapolating types fill whole ghost
// columns, corners included); the other types use j = 0 .. ny-1.
__device__ void fill_x_ghost(Fld f, const Grid &g, const Geo &geo, const Bc &bc, int homog, int do_neum,
                             int side, int j) {
    const int t = bc.type[side];
    const bool lo = side == XLO;
    const int nx = g.nx;
    const bool edge_row = (j < 0 || j >= g.ny);
    if (t == BC_EXTRAP0 || t == BC_EXTRAP1 || t == BC_EXTRAP2) {
        const int gi = lo ? -1 : nx, a = lo ? 0 : nx - 1, s = lo ? 1 : -1;
        const double p0 = f(a, j), p1 = f(a + s, j);
        if (t == BC_EXTRAP0) f(gi, j) = p0;
        else if (t == BC_EXTRAP1) f(gi, j) = 2.0 * p0 - p1;
        else f(gi, j) = 3.0 * (p0 - p1) + f(a + 2 * s, j);
        return;
    }
    if (edge_row) return;
    if (t == BC_NEUM) {
        if (do_neum) neum_x_ghost(f, g, geo, bc, homog, side, j);
    } else if (t == BC_DIRI) {
        const int e = lo ? -1 : 1, i = lo ? 0 : nx - 1;
        f(i + e, j) = diri_ghost(bc, side, homog, f(i, j), f(i - e, j), f(i - 2 * e, j), j);
    } else if (t == BC_PERIODIC) {
        f(lo ? -1 : nx, j) = f(lo ? nx - 1 : 0, j);
    }
}

__device__ void fill_y_ghost(Fld f, const Grid &g, const Geo &geo, const Bc &bc, int homog, int do_neum,
                             int side, int i) {
    const int t = bc.type[side];
    const bool lo = side == YLO;
    const int ny = g.ny;
    const bool edge_col = (i < 0 || i >= g.nx);
    if (t == BC_EXTRAP0 || t == BC_EXTRAP1 || t == BC_EXTRAP2) {
        const int gj = lo ? -1 : ny, a = lo ? 0 : ny - 1, s = lo ? 1 : -1;
        const double p0 = f(i, a), p1 = f(i, a + s);
        if (t == BC_EXTRAP0) f(i, gj) = p0;
        else if (t == BC_EXTRAP1) f(i, gj) = 2.0 * p0 - p1;
        else f(i, gj) = 3.0 * (p0 - p1) + f(i, a + 2 * s);
        return;
    }
    if (edge_col) return;
    if (t == BC_NEUM) {
        if (do_neum) neum_y_ghost(f, g, geo, bc, homog, side, i);
    } else if (t == BC_DIRI) {
        const int e = lo ? -1 : 1, j = lo ? 0 : ny - 1;
        f(i, j + e) = diri_ghost(bc, side, homog, f(i, j), f(i, j - e), f(i, j - 2 * e), i);
    } else if (t == BC_PERIODIC) {
        f(i, lo ? -1 : ny) = f(i, lo ? ny - 1 : 0);
    }
}

__device__ void fill_corners(Fld f, const Grid &g, const Bc &bc) {
    const int nx = g.nx, ny = g.ny;
    const bool px = bc.type[XLO] == BC_PERIODIC, py = bc.type[YLO] == BC_PERIODIC;
    if (!px && !py) {
        f(-1, -1) = BOGUS;
        f(-1, ny) = BOGUS;
        f(nx, -1) = BOGUS;
        f(nx, ny) = BOGUS;
    } else if (px && py) {
        f(-1, -1) = f(nx - 1, ny - 1);
        f(-1, ny) = f(nx - 1, 0);
        f(nx, -1) = f(0, ny - 1);
        f(nx, ny) = f(0, 0);
    } else if (px) {
        f(-1, -1) = f(nx - 1, -1);
        f(-1, ny) = f(nx - 1, ny);
        f(nx, -1) = f(0, -1);
        f(nx, ny) = f(0, ny);
    } else {
        f(-1, -1) = f(-1, ny - 1);
        f(nx, -1) = f(nx, ny - 1);
        f(-1, ny) = f(-1, 0);
        f(nx, ny) = f(nx, 0);
    }
}

__device__ void block_fill_ghosts(Fld f, const Grid &g, const Geo &geo, const Bc &bc, int homog, int do_neum) {
    const int tid = threadIdx.x, nt = blockDim.x;
    // the two x sides are independent of each other
    for (int k = tid; k < 2 * (g.ny + 2); k += nt) {
        const int side = k < g.ny + 2 ? XLO : XHI;
        const int j = (k % (g.ny + 2)) - 1;
        fill_x_ghost(f, g, geo, bc, homog, do_neum, side, j);
    }
    __syncthreads();
    for (int k = tid; k < 2 * (g.nx + 2); k += nt) {
        const int side = k < g.nx + 2 ? YLO : YHI;
        const int i = (k % (g.nx + 2)) - 1;
        fill_y_ghost(f, g, geo, bc, homog, do_neum, side, i);
    }
    __syncthreads();
    if (tid == 0) fill_corners(f, g, bc);
    __syncthreads();
}

__global__ void __launch_bounds__(256) k_fill_ghosts(Fld f, Grid g, Geo geo, Bc bc, int homog, int do_neum) {
    block_fill_ghosts(f, g, geo, bc, homog, do_neum);
}

// ===========================================================================
// Operator, residual, inverse diagonal (poisson2d.f90:16-388, 886-1353)
// ===========================================================================

// out = L[phi] (or rhs - L[phi] when rhs != null, or L[phi]/J when jscale).
// Ghosts of phi must already hold the fill_ghosts(do_neum = .false.) values.
__global__ void __launch_bounds__(256) k_laplacian(double *out, CFld phi, const double *rhs, Grid g, Geo geo,
                                                   Bc bc, int homog, int jscale) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= g.nx || j >= g.ny) return;
    const size_t c = ix(g.pitch, i, j);
    const bool interior = i > 0 && i < g.nx - 1 && j > 0 && j < g.ny - 1;
    double v = interior ? cell_lap_interior(phi, g, geo, i, j) : cell_lap(phi, g, geo, bc, homog, i, j);
    if (jscale) v = v / geo.J[c];
    out[c] = rhs ? rhs[c] - v : v;
}

// out = rhs - L[phi] on a homogeneous MG level without a prior ghost fill.
__global__ void __launch_bounds__(256) k_residual_synth(double *out, const double *phi, const double *rhs, Grid g,
                                                        Geo geo, Bc bc) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= g.nx || j >= g.ny) return;
    const size_t c = ix(g.pitch, i, j);
    const bool interior = i > 0 && i < g.nx - 1 && j > 0 && j < g.ny - 1;
    const double v = interior ? cell_lap_interior(CFld{phi, g.pitch}, g, geo, i, j)
                              : cell_lap(synth(phi, g, bc), g, geo, bc, 1, i, j);
    out[c] = rhs ? rhs[c] - v : v;
}

__global__ void __launch_bounds__(256) k_invdiags(double *out, Grid g, Geo geo, Bc bc, int homog) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= g.nx || j >= g.ny) return;
    Impulse imp{i, j};
    out[ix(g.pitch, i, j)] = 1.0 / cell_lap(imp, g, geo, bc, homog, i, j);
}

// DeferredPoisson2D::compute_invdiags (MGPoisson2D_DD.f90:16-39)
__global__ void __launch_bounds__(256) k_dc_invdiags(double *out, Grid g, Geo geo) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= g.nx || j >= g.ny) return;
    const int P = g.pitch;
    const size_t c = ix(P, i, j);
    const double idx2 = 1.0 / (g.dx * g.dx), idy2 = 1.0 / (g.dy * g.dy);
    const double d = (geo.Jxx[c + 1] + geo.Jxx[c]) * idx2 + (geo.Jyy[c + P] + geo.Jyy[c]) * idy2;
    out[c] = -1.0 / d;
}

// DeferredPoisson2D::compute_laplacian / compute_residual.  Ghosts must hold
// the fill_ghosts(do_neum = .true.) values; the cross part uses the order-0
// extrapolation the reference applies first.
__global__ void __launch_bounds__(256) k_dc_laplacian(double *out, CFld phi, const double *rhs, Grid g, Geo geo) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= g.nx || j >= g.ny) return;
    const size_t c = ix(g.pitch, i, j);
    double ax, ay;
    dc_five(phi, g, geo, i, j, ax, ay);
    const double v = dc_cross(phi, g, geo, i, j) + ax + ay;
    out[c] = rhs ? rhs[c] - v : v;
}

// modrhs = rhs - lagged cross terms (MGPoisson2D_DD.f90:227-250)
__global__ void __launch_bounds__(256) k_dc_modrhs(double *modrhs, const double *rhs, CFld phi, Grid g, Geo geo) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= g.nx || j >= g.ny) return;
    const size_t c = ix(g.pitch, i, j);
    modrhs[c] = rhs[c] - dc_lagged_cross(phi, g, geo, i, j);
}

// ===========================================================================
// Smoothers
// ===========================================================================

// Multi-colour GS, interior cells of colour (ci, cj) (i & 1 == ci, j & 1 == cj).
// Interior cells never read ghosts and never read a cell of their own colour.
__global__ void __launch_bounds__(256) k_gs_mc_interior(Fld phi, const double *rhs, const double *invd, Grid g,
                                                        Geo geo, double omega, int ci, int cj) {
    const int a = blockIdx.x * 64 + (threadIdx.x & 63);
    const int b = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i = (ci ? 1 : 2) + 2 * a, j = (cj ? 1 : 2) + 2 * b;
    if (i > g.nx - 2 || j > g.ny - 2) return;
    const size_t c = ix(g.pitch, i, j);
    const double lphi = cell_lap_interior(CFld{phi.p, phi.pitch}, g, geo, i, j);
    phi.p[c] = gs_update(phi.p[c], rhs[c], lphi, invd[c], omega);
}

// The four interior colours of a multi-colour sweep in one pass over LDS
// tiles of MC_TX x MC_TY interior cells.  Colour c (in the order of the
// per-colour launches: (i, j) parity (even, even), (odd, even), (even, odd),
// (odd, odd), even meaning 2, 4, ..) reads its neighbours as left by colours
// < c, so the tile also updates colour c on a halo of 3 - c cells (the
// neighbouring tiles' cells, recomputed from the same values): its window
// holds the tile and a halo of MC_H = 4 cells of phi.  Every update is the
// per-colour kernel's expression on the same operands, so the sweep gives its
// bits.  phi is read once per tile window and written once; the metrics, rhs
// and invdiags of the halo updates come from the L2.  Ring cells and ghosts
// stay as they are (the ring phase follows).
constexpr int MC_TX = 64, MC_TY = 16, MC_H = 4;
constexpr int MC_WX = MC_TX + 2 * MC_H, MC_WY = MC_TY + 2 * MC_H;

__global__ void __launch_bounds__(256) k_gs_mc_tile(Fld phi, const double *rhs, const double *invd, Grid g, Geo geo,
                                                    double omega) {
    __shared__ double w[MC_WY][MC_WX + 1];
    const int P = g.pitch, nx = g.nx, ny = g.ny;
    const int ti0 = 1 + blockIdx.x * MC_TX, tj0 = 1 + blockIdx.y * MC_TY;  // tile origin (first interior cell 1)
    for (int k = threadIdx.x; k < MC_WX * MC_WY; k += 256) {
        const int wi = k % MC_WX, wj = k / MC_WX;
        const int i = ti0 - MC_H + wi, j = tj0 - MC_H + wj;
        // only valid cells are ever read by an interior update
        w[wj][wi] = (i >= 0 && i < nx && j >= 0 && j < ny) ? phi.p[ix(P, i, j)] : 0.0;
    }
    __syncthreads();
    const double idx = 1.0 / g.dx, idy = 1.0 / g.dy;
    for (int colour = 0; colour < 4; ++colour) {
        const int pi = (colour & 1) ? 1 : 0, pj = (colour >> 1) ? 1 : 0, h = 3 - colour;
        // the colour's cells in the tile plus halo h, within the interior 1..n-2
        int x0 = ti0 - h, x1 = ti0 + MC_TX + h, y0 = tj0 - h, y1 = tj0 + MC_TY + h;
        x0 = x0 < 1 ? 1 : x0;
        y0 = y0 < 1 ? 1 : y0;
        x1 = x1 > nx - 1 ? nx - 1 : x1;
        y1 = y1 > ny - 1 ? ny - 1 : y1;
        x0 += (x0 & 1) != pi;
        y0 += (y0 & 1) != pj;
        const int nxr = x0 < x1 ? (x1 - 1 - x0) / 2 + 1 : 0, nyr = y0 < y1 ? (y1 - 1 - y0) / 2 + 1 : 0;
        for (int k = threadIdx.x; k < nxr * nyr; k += 256) {
            const int i = x0 + 2 * (k % nxr), j = y0 + 2 * (k / nxr);
            const int a = i - ti0 + MC_H, b = j - tj0 + MC_H;
            const size_t c = ix(P, i, j);
            const double lphi = interior_lap_values(geo, c, P, idx, idy, w[b][a], w[b][a + 1], w[b][a - 1],
                                                    w[b + 1][a], w[b - 1][a], w[b + 1][a + 1], w[b + 1][a - 1],
                                                    w[b - 1][a + 1], w[b - 1][a - 1]);
            w[b][a] = gs_update(w[b][a], rhs[c], lphi, invd[c], omega);
        }
        __syncthreads();
    }
    for (int k = threadIdx.x; k < MC_TX * MC_TY; k += 256) {
        const int i = ti0 + k % MC_TX, j = tj0 + k / MC_TX;
        if (i <= nx - 2 && j <= ny - 2) phi.p[ix(P, i, j)] = w[j - tj0 + MC_H][i - ti0 + MC_H];
    }
}

__device__ __forceinline__ void gs_cell(Fld phi, const double *rhs, const double *invd, const Grid &g,
                                        const Geo &geo, const Bc &bc, int homog, double omega, int i, int j) {
    const size_t c = ix(g.pitch, i, j);
    const double lphi = cell_lap(CFld{phi.p, phi.pitch}, g, geo, bc, homog, i, j);
    phi.p[c] = gs_update(phi.p[c], rhs[c], lphi, invd[c], omega);
}

// Ring phase of the multi-colour sweep, one workgroup: refresh ghosts
// (do_neum = .false.), ring cells in 4 colours, then the 4 corners.
__device__ void block_gs_mc_ring(Fld phi, const double *rhs, const double *invd, const Grid &g, const Geo &geo,
                                 const Bc &bc, int homog, double omega) {
    const int nx = g.nx, ny = g.ny;
    block_fill_ghosts(phi, g, geo, bc, homog, 0);
    const int nring = 2 * (nx - 2) + 2 * (ny - 2);
    for (int colour = 0; colour < 4; ++colour) {
        for (int k = threadIdx.x; k < nring; k += blockDim.x) {
            int i, j;
            if (k < nx - 2) { i = 1 + k; j = 0; }
            else if (k < 2 * (nx - 2)) { i = 1 + k - (nx - 2); j = ny - 1; }
            else if (k < 2 * (nx - 2) + (ny - 2)) { i = 0; j = 1 + k - 2 * (nx - 2); }
            else { i = nx - 1; j = 1 + k - 2 * (nx - 2) - (ny - 2); }
            if (((i & 1) | ((j & 1) << 1)) == colour) gs_cell(phi, rhs, invd, g, geo, bc, homog, omega, i, j);
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) {
        const int i = (threadIdx.x & 1) ? nx - 1 : 0, j = (threadIdx.x & 2) ? ny - 1 : 0;
        gs_cell(phi, rhs, invd, g, geo, bc, homog, omega, i, j);
    }
    __syncthreads();
}

__global__ void __launch_bounds__(1024) k_gs_mc_ring(Fld phi, const double *rhs, const double *invd, Grid g,
                                                     Geo geo, Bc bc, int homog, double omega) {
    block_gs_mc_ring(phi, rhs, invd, g, geo, bc, homog, omega);
}

// Multi-workgroup forms of the single-workgroup ghost fill and ring sweep
// above, for large grids: the same cells with the same arithmetic, each
// phase (x sides, y sides, corners; ring colours 0..3, corner cells) one
// launch, in the same order.
__global__ void __launch_bounds__(256) k_fill_x_ghosts(Fld f, Grid g, Geo geo, Bc bc, int homog, int do_neum) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= 2 * (g.ny + 2)) return;
    fill_x_ghost(f, g, geo, bc, homog, do_neum, k < g.ny + 2 ? XLO : XHI, (k % (g.ny + 2)) - 1);
}

__global__ void __launch_bounds__(256) k_fill_y_ghosts(Fld f, Grid g, Geo geo, Bc bc, int homog, int do_neum) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= 2 * (g.nx + 2)) return;
    fill_y_ghost(f, g, geo, bc, homog, do_neum, k < g.nx + 2 ? YLO : YHI, (k % (g.nx + 2)) - 1);
}

__global__ void k_fill_corner_ghosts(Fld f, Grid g, Bc bc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) fill_corners(f, g, bc);
}

__global__ void __launch_bounds__(256) k_gs_mc_ring_colour(Fld phi, const double *rhs, const double *invd, Grid g,
                                                           Geo geo, Bc bc, int homog, double omega, int colour) {
    const int nx = g.nx, ny = g.ny;
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= 2 * (nx - 2) + 2 * (ny - 2)) return;
    int i, j;
    if (k < nx - 2) { i = 1 + k; j = 0; }
    else if (k < 2 * (nx - 2)) { i = 1 + k - (nx - 2); j = ny - 1; }
    else if (k < 2 * (nx - 2) + (ny - 2)) { i = 0; j = 1 + k - 2 * (nx - 2); }
    else { i = nx - 1; j = 1 + k - 2 * (nx - 2) - (ny - 2); }
    if (((i & 1) | ((j & 1) << 1)) == colour) gs_cell(phi, rhs, invd, g, geo, bc, homog, omega, i, j);
}

__global__ void k_gs_mc_corner_cells(Fld phi, const double *rhs, const double *invd, Grid g, Geo geo, Bc bc,
                                     int homog, double omega) {
    if (threadIdx.x < 4) {
        const int i = (threadIdx.x & 1) ? g.nx - 1 : 0, j = (threadIdx.x & 2) ? g.ny - 1 : 0;
        gs_cell(phi, rhs, invd, g, geo, bc, homog, omega, i, j);
    }
}

// Exact lexicographic order (poisson2d.f90:1561-1959) by wavefronts t = i + 3j:
// every value a cell reads from an earlier row (reach up to (i+2, j-1)) lies
// on an earlier front and every later-row or later-column read on a later
// one.  One workgroup; ghosts refreshed once per sweep as in the reference.
__device__ void block_gs_lex(Fld phi, const double *rhs, const double *invd, const Grid &g, const Geo &geo,
                             const Bc &bc, int homog, double omega, int sweeps) {
    const int nx = g.nx, ny = g.ny;
    for (int s = 0; s < sweeps; ++s) {
        block_fill_ghosts(phi, g, geo, bc, homog, 0);
        const int tmax = (nx - 1) + 3 * (ny - 1);
        for (int t = 0; t <= tmax; ++t) {
            int jlo = t - (nx - 1);
            jlo = jlo <= 0 ? 0 : (jlo + 2) / 3;
            int jhi = t / 3;
            if (jhi > ny - 1) jhi = ny - 1;
            for (int j = jlo + (int)threadIdx.x; j <= jhi; j += blockDim.x)
                gs_cell(phi, rhs, invd, g, geo, bc, homog, omega, t - 3 * j, j);
            __syncthreads();
        }
    }
}

__global__ void __launch_bounds__(1024) k_gs_lex(Fld phi, const double *rhs, const double *invd, Grid g, Geo geo,
                                                 Bc bc, int homog, double omega, int sweeps) {
    block_gs_lex(phi, rhs, invd, g, geo, bc, homog, omega, sweeps);
}

// DC smoother point update on the 5-point part with modrhs.
__device__ __forceinline__ void dc_cell(Fld phi, const double *modrhs, const double *invd, const Grid &g,
                                        const Geo &geo, double omega, int i, int j) {
    const size_t c = ix(g.pitch, i, j);
    double ax, ay;
    dc_five(CFld{phi.p, phi.pitch}, g, geo, i, j, ax, ay);
    const double lphi = ax + ay;
    phi.p[c] = gs_update(phi.p[c], modrhs[c], lphi, invd[c], omega);
}

// Red-black colour of the DC smoother (MGPoisson2D_DD.f90:262-285): colour c
// updates cells with (i + j + c) even, i.e. imin = ilo + mod(c + j, 2).
__global__ void __launch_bounds__(256) k_dc_rb(Fld phi, const double *modrhs, const double *invd, Grid g, Geo geo,
                                               double omega, int colour) {
    const int a = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (j >= g.ny) return;
    const int i = ((colour + j) & 1) + 2 * a;
    if (i >= g.nx) return;
    dc_cell(phi, modrhs, invd, g, geo, omega, i, j);
}

// Lexicographic DC sweeps (MGPoisson2D_DD.f90:288-306) by wavefronts t = i + j,
// one workgroup, ghosts refreshed per sweep (do_neum = .true.).
__device__ void block_dc_lex(Fld phi, const double *modrhs, const double *invd, const Grid &g, const Geo &geo,
                             const Bc &bc, int homog, double omega, int sweeps) {
    const int nx = g.nx, ny = g.ny;
    for (int s = 0; s < sweeps; ++s) {
        block_fill_ghosts(phi, g, geo, bc, homog, 1);
        for (int t = 0; t <= nx - 1 + ny - 1; ++t) {
            const int jlo = t - (nx - 1) > 0 ? t - (nx - 1) : 0;
            const int jhi = t < ny - 1 ? t : ny - 1;
            for (int j = jlo + (int)threadIdx.x; j <= jhi; j += blockDim.x)
                dc_cell(phi, modrhs, invd, g, geo, omega, t - j, j);
            __syncthreads();
        }
    }
}

__global__ void __launch_bounds__(1024) k_dc_lex(Fld phi, const double *modrhs, const double *invd, Grid g,
                                                 Geo geo, Bc bc, int homog, double omega, int sweeps) {
    block_dc_lex(phi, modrhs, invd, g, geo, bc, homog, omega, sweeps);
}

// ===========================================================================
// Transfers (MGPoisson2D.f90:135-394)
// ===========================================================================

// Cell-centred average of rx x ry children, summed in the reference's loop
// order (fine rows outer, fine columns inner).  When invd != null also sets
// the next level's initial guess e = r * invdiag (MGPoisson2D.f90:805-806).
__global__ void __launch_bounds__(256) k_restrict_cc(double *crse, const double *fine, Grid cg, int fpitch, int rx,
                                                     int ry, double *e, const double *invd) {
    const int ci = blockIdx.x * 64 + (threadIdx.x & 63);
    const int cj = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (ci >= cg.nx || cj >= cg.ny) return;
    const double scale = 1.0 / (rx * ry);
    double s = 0.0;
    for (int b = 0; b < ry; ++b)
        for (int a = 0; a < rx; ++a) s = s + scale * fine[ix(fpitch, ci * rx + a, cj * ry + b)];
    const size_t c = ix(cg.pitch, ci, cj);
    crse[c] = s;
    if (e) e[c] = s * invd[c];
}

// Face-centred metric restriction: injection at coincident faces times the
// average along the cell direction (full_weighting = .false.).
__global__ void __launch_bounds__(256) k_restrict_xf(double *crse, const double *fine, Grid cg, int fpitch, int rx,
                                                     int ry) {
    const int ci = blockIdx.x * 64 + (threadIdx.x & 63);
    const int cj = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (ci > cg.nx || cj >= cg.ny) return;
    const double scale = 1.0 / ry;
    double s = 0.0;
    for (int b = 0; b < ry; ++b) s = s + scale * fine[ix(fpitch, ci * rx, cj * ry + b)];
    crse[ix(cg.pitch, ci, cj)] = s;
}

__global__ void __launch_bounds__(256) k_restrict_yf(double *crse, const double *fine, Grid cg, int fpitch, int rx,
                                                     int ry) {
    const int ci = blockIdx.x * 64 + (threadIdx.x & 63);
    const int cj = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (ci >= cg.nx || cj > cg.ny) return;
    const double scale = 1.0 / rx;
    double s = 0.0;
    for (int a = 0; a < rx; ++a) s = s + scale * fine[ix(fpitch, ci * rx + a, cj * ry)];
    crse[ix(cg.pitch, ci, cj)] = s;
}

// 2x2 prolongation and correction (MGPoisson2D.f90:304-369).  The coarse
// ghosts must already hold fill_ghosts(crse, homog, do_neum) values; for
// order 2 the second coarse array holds the order-3 extrapolated ghosts.
// Per fine cell the additions happen in the reference's pass order.
__global__ void __launch_bounds__(256) k_prolong2x2(Fld fine, CFld crse, CFld crse_x3, Grid cg, int order) {
    const int ci = blockIdx.x * 64 + (threadIdx.x & 63);
    const int cj = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (ci >= cg.nx || cj >= cg.ny) return;
    const int fi = 2 * ci, fj = 2 * cj;
    const double c = crse(ci, cj);
    double f00 = fine(fi, fj) + c, f10 = fine(fi + 1, fj) + c;
    double f01 = fine(fi, fj + 1) + c, f11 = fine(fi + 1, fj + 1) + c;
    if (order >= 1) {
        const double mx = 0.5 * (crse(ci + 1, cj) - crse(ci - 1, cj));
        const double my = 0.5 * (crse(ci, cj + 1) - crse(ci, cj - 1));
        f00 = f00 + 0.25 * (-mx - my);
        f10 = f10 + 0.25 * (mx - my);
        f01 = f01 + 0.25 * (-mx + my);
        f11 = f11 + 0.25 * (mx + my);
    }
    if (order >= 2) {
        const double mx = 0.25 * (crse(ci + 1, cj) - 2.0 * c + crse(ci - 1, cj));
        const double my = 0.25 * (crse(ci, cj + 1) - 2.0 * c + crse(ci, cj - 1));
        const double q = 0.125 * (mx + my);
        f00 = f00 + q;
        f10 = f10 + q;
        f01 = f01 + q;
        f11 = f11 + q;
        const double mc = (1.0 / 16.0) * (crse_x3(ci + 1, cj + 1) - crse_x3(ci + 1, cj - 1)
                                          - crse_x3(ci - 1, cj + 1) + crse_x3(ci - 1, cj - 1));
        f00 = f00 + 0.25 * mc;
        f10 = f10 - 0.25 * mc;
        f01 = f01 - 0.25 * mc;
        f11 = f11 + 0.25 * mc;
    }
    fine(fi, fj) = f00;
    fine(fi + 1, fj) = f10;
    fine(fi, fj + 1) = f01;
    fine(fi + 1, fj + 1) = f11;
}

// Order-3 extrapolated copy of a coarse field's ghosts (extrapolate_ghosts(3),
// arrayutils.f90:1551-1560): x ghosts over all rows first, then y ghosts over
// all columns, so corners come from the y pass.  One workgroup.
__global__ void __launch_bounds__(256) k_extrap3(Fld dst, CFld src, Grid g) {
    const int nx = g.nx, ny = g.ny;
    for (int k = threadIdx.x; k < (nx + 2) * (ny + 2); k += blockDim.x) {
        const int i = k % (nx + 2) - 1, j = k / (nx + 2) - 1;
        dst(i, j) = src(i, j);
    }
    __syncthreads();
    for (int j = (int)threadIdx.x - 1; j <= ny; j += blockDim.x) {
        dst(-1, j) = 4.0 * (dst(0, j) + dst(2, j)) - 6.0 * dst(1, j) - dst(3, j);
        dst(nx, j) = 4.0 * (dst(nx - 1, j) + dst(nx - 3, j)) - 6.0 * dst(nx - 2, j) - dst(nx - 4, j);
    }
    __syncthreads();
    for (int i = (int)threadIdx.x - 1; i <= nx; i += blockDim.x) {
        dst(i, -1) = 4.0 * (dst(i, 0) + dst(i, 2)) - 6.0 * dst(i, 1) - dst(i, 3);
        dst(i, ny) = 4.0 * (dst(i, ny - 1) + dst(i, ny - 3)) - 6.0 * dst(i, ny - 2) - dst(i, ny - 4);
    }
}

// extrapolate_ghosts (arrayutils.f90:1497-1566), in place, orders 0..3: the
// x-side ghost columns over every row of the array (ghost rows included),
// then the y-side ghost rows over every column, so the corners come from the
// y rule applied to the new x ghosts, as in the reference's array syntax.
__device__ __forceinline__ double extrap(int order, double p0, double p1, double p2, double p3) {
    switch (order) {
    case 0: return p0;
    case 1: return 2.0 * p0 - p1;
    case 2: return 3.0 * (p0 - p1) + p2;
    default: return 4.0 * (p0 + p2) - 6.0 * p1 - p3;
    }
}

__global__ void __launch_bounds__(256) k_extrapolate(Fld f, Grid g, int order) {
    const int nx = g.nx, ny = g.ny;
    for (int j = (int)threadIdx.x - 1; j <= ny; j += blockDim.x) {
        f(-1, j) = extrap(order, f(0, j), f(1, j), f(2, j), order == 3 ? f(3, j) : 0.0);
        f(nx, j) = extrap(order, f(nx - 1, j), f(nx - 2, j), f(nx - 3, j), order == 3 ? f(nx - 4, j) : 0.0);
    }
    __syncthreads();
    for (int i = (int)threadIdx.x - 1; i <= nx; i += blockDim.x) {
        f(i, -1) = extrap(order, f(i, 0), f(i, 1), f(i, 2), order == 3 ? f(i, 3) : 0.0);
        f(i, ny) = extrap(order, f(i, ny - 1), f(i, ny - 2), f(i, ny - 3), order == 3 ? f(i, ny - 4) : 0.0);
    }
}

// ===========================================================================
// Pointwise helpers
// ===========================================================================

// e = r * invd over the valid cells (MGPoisson2D.f90:600)
__global__ void __launch_bounds__(256) k_einit(double *e, const double *r, const double *invd, Grid g) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= g.nx || j >= g.ny) return;
    const size_t c = ix(g.pitch, i, j);
    e[c] = r[c] * invd[c];
}

// y = y + a*x over the cells and the ghost ring (phi%data = phi%data + e%data).
__global__ void __launch_bounds__(256) k_axpy_ghosted(double *y, const double *x, double a, Grid g) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63) - 1;
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6) - 1;
    if (i > g.nx || j > g.ny) return;
    const size_t c = ix(g.pitch, i, j);
    y[c] = a == 1.0 ? y[c] + x[c] : (a == -1.0 ? y[c] - x[c] : y[c] + a * x[c]);
}

// y = a*y + b*x + c*z over valid cells, pointer-generic linear combination.
__global__ void __launch_bounds__(256) k_lincomb(double *out, double a, const double *y, double b, const double *x,
                                                 Grid g, int ghosted) {
    const int o = ghosted ? 1 : 0;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63) - o;
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6) - o;
    if (i >= g.nx + o || j >= g.ny + o) return;
    const size_t c = ix(g.pitch, i, j);
    out[c] = (y ? a * y[c] : 0.0) + (x ? b * x[c] : 0.0);
}

// out = a * b (op 0) or a / b (op 1) over valid cells.
__global__ void __launch_bounds__(256) k_muldiv(double *out, const double *a, const double *b, Grid g, int op) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= g.nx || j >= g.ny) return;
    const size_t c = ix(g.pitch, i, j);
    out[c] = op == 0 ? a[c] * b[c] : a[c] / b[c];
}

// Fill valid (and optionally ghost) cells with a constant.
__global__ void __launch_bounds__(256) k_setval(double *y, double v, Grid g, int ghosted) {
    const int o = ghosted ? 1 : 0;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63) - o;
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6) - o;
    if (i >= g.nx + o || j >= g.ny + o) return;
    y[ix(g.pitch, i, j)] = v;
}

// ===========================================================================
// Reductions.  Two passes: per-block partial sums in a fixed order, then one
// block sums the partials.  Compensated (Neumaier) FP64 accumulation.
// ===========================================================================
// |x|^p for an integer p >= 1 as Fortran's real**integer evaluates it:
// binary powering from the lowest bit.
__device__ __forceinline__ double abs_pow(double x, int p) {
    double b = fabs(x), r = 1.0;
    while (true) {
        if (p & 1) r *= b;
        p /= 2;
        if (p == 0) break;
        b *= b;
    }
    return r;
}

// kind 0: sum w*a*b, 1: sum w*a*a, 2: max |a|, 3: sum w*a, 4: sum w*|a|^p
// (w = vol or vol*J).  Block b takes rows b, b + gridDim.x, ...; threads
// stride along the row.
__global__ void __launch_bounds__(256) k_reduce_partial(double *partial, const double *a, const double *b,
                                                        const double *J, double vol, Grid g, int kind, int p) {
    __shared__ double ss[256], sc[256];
    Acc acc{0.0, 0.0};
    double mx = 0.0;
    for (int j = blockIdx.x; j < g.ny; j += gridDim.x) {
        const size_t row = ix(g.pitch, 0, j);
#pragma unroll 4
        for (int i = threadIdx.x; i < g.nx; i += blockDim.x) {
            const size_t c = row + i;
            const double w = J ? vol * J[c] : vol;
            const double x = a[c];
            if (kind == 0) acc.add(w * (x * b[c]));
            else if (kind == 1) acc.add(w * (x * x));
            else if (kind == 2) mx = fmax(mx, fabs(x));
            else if (kind == 3) acc.add(w * x);
            else acc.add(w * abs_pow(x, p));
        }
    }
    const double v = block_combine(ss, sc, acc, mx, kind);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

__global__ void __launch_bounds__(256) k_reduce_final(double *out, const double *partial, int n, int kind) {
    __shared__ double ss[256], sc[256];
    Acc acc{0.0, 0.0};
    double mx = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        if (kind == 2) mx = fmax(mx, partial[k]);
        else acc.add(partial[k]);
    }
    const double v = block_combine(ss, sc, acc, mx, kind);
    if (threadIdx.x == 0) *out = v;
}

// Sequential J-weighted sums in the reference's loop order (one thread):
// out[0] = sum dvol*fine, out[1] = sum dvol, dvol = dxprod*J.
__global__ void k_seq_mean(double *out, const double *fine, const double *J, Grid g, double dxprod) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double sum = 0.0, vol = 0.0;
    for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
            const size_t c = ix(g.pitch, i, j);
            const double dvol = dxprod * J[c];
            sum = sum + dvol * fine[c];
            vol = vol + dvol;
        }
    out[0] = sum;
    out[1] = vol;
}

// fine = fine - sums[0]/sums[1] over the cells and the ghost ring.
__global__ void __launch_bounds__(256) k_sub_mean(double *fine, const double *sums, Grid g) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63) - 1;
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6) - 1;
    if (i > g.nx || j > g.ny) return;
    const size_t c = ix(g.pitch, i, j);
    fine[c] = fine[c] - sums[0] / sums[1];
}

// ===========================================================================
// Bottom solver: BiCGStab2 with restarts in one workgroup (poisson2d.f90
// solve_bicgstab2 :2253-2477 / MGPoisson2D_DD.f90 :339-566).  Scalars are
// reduced in the reference's sequential order by thread 0, so the bottom
// trajectory (restart and convergence decisions included) matches.
// ===========================================================================
using Bottom = BottomArgs;  // work arrays use the padded layout

__device__ double seq_dot(const Bottom &B, const double *a, const double *b) {
    // inner_prod (arrayutils.f90:637-661): per-row dot, scaled, accumulated
    __shared__ double res;
    if (threadIdx.x == 0) {
        const double vol = B.g.dx * B.g.dy;
        double ip = 0.0;
        for (int j = 0; j < B.g.ny; ++j) {
            double d = 0.0;
            for (int i = 0; i < B.g.nx; ++i) d = d + a[ix(B.g.pitch, i, j)] * b[ix(B.g.pitch, i, j)];
            ip = ip + vol * d;
        }
        res = ip;
    }
    __syncthreads();
    const double v = res;
    __syncthreads();
    return v;
}

__device__ double seq_norm(const Bottom &B, const double *a, int p) {
    __shared__ double res;
    if (threadIdx.x == 0) {
        const double vs = B.g.dx * B.g.dy;
        double pn = 0.0;
        for (int j = 0; j < B.g.ny; ++j)
            for (int i = 0; i < B.g.nx; ++i) {
                const double x = a[ix(B.g.pitch, i, j)];
                if (p == 0) pn = fmax(pn, fabs(x));
                else pn = pn + vs * (x * x);
            }
        res = p == 0 ? pn : sqrt(pn);
    }
    __syncthreads();
    const double v = res;
    __syncthreads();
    return v;
}

template <class F>
__device__ __forceinline__ void block_cells(const Grid &g, int ghosted, F fn) {
    const int o = ghosted ? 1 : 0;
    const int w = g.nx + 2 * o, h = g.ny + 2 * o;
    for (int k = threadIdx.x; k < w * h; k += blockDim.x) fn(k % w - o, k / w - o, ix(g.pitch, k % w - o, k / w - o));
    __syncthreads();
}

__device__ void block_residual(const Bottom &B, int homog, const double *rhs, double *phi, double *res) {
    Fld f{phi, B.g.pitch};
    if (B.dc) {
        block_fill_ghosts(f, B.g, B.geo, B.bc, homog, 1);
        block_cells(B.g, 0, [&](int i, int j, size_t c) {
            double ax, ay;
            dc_five(CFld{phi, B.g.pitch}, B.g, B.geo, i, j, ax, ay);
            res[c] = rhs[c] - (dc_cross(CFld{phi, B.g.pitch}, B.g, B.geo, i, j) + ax + ay);
        });
    } else {
        block_fill_ghosts(f, B.g, B.geo, B.bc, homog, 0);
        block_cells(B.g, 0, [&](int i, int j, size_t c) {
            res[c] = rhs[c] - cell_lap(CFld{phi, B.g.pitch}, B.g, B.geo, B.bc, homog, i, j);
        });
    }
}

__device__ void block_laplacian(const Bottom &B, double *phi, double *out) {
    Fld f{phi, B.g.pitch};
    if (B.dc) {
        block_fill_ghosts(f, B.g, B.geo, B.bc, 1, 1);
        block_cells(B.g, 0, [&](int i, int j, size_t c) {
            double ax, ay;
            dc_five(CFld{phi, B.g.pitch}, B.g, B.geo, i, j, ax, ay);
            out[c] = dc_cross(CFld{phi, B.g.pitch}, B.g, B.geo, i, j) + ax + ay;
        });
    } else {
        block_fill_ghosts(f, B.g, B.geo, B.bc, 1, 0);
        block_cells(B.g, 0, [&](int i, int j, size_t c) {
            out[c] = cell_lap(CFld{phi, B.g.pitch}, B.g, B.geo, B.bc, 1, i, j);
        });
    }
}

// One relax_gs call (tol < 0) on the bottom level, by one workgroup.
__device__ void block_relax(const Bottom &B, int homog, double omega, int sweeps, const double *rhs, double *phi) {
    if (sweeps == 0) return;
    Fld f{phi, B.g.pitch};
    if (B.dc) {
        block_cells(B.g, 0, [&](int i, int j, size_t c) {
            B.mod[c] = rhs[c] - dc_lagged_cross(CFld{phi, B.g.pitch}, B.g, B.geo, i, j);
        });
        if (B.order == 1) {  // red-black
            for (int s = 0; s < sweeps; ++s) {
                block_fill_ghosts(f, B.g, B.geo, B.bc, homog, 1);
                for (int colour = 0; colour < 2; ++colour)
                    block_cells(B.g, 0, [&](int i, int j, size_t) {
                        if (((i + j + colour) & 1) == 0) dc_cell(f, B.mod, B.invd, B.g, B.geo, omega, i, j);
                    });
            }
        } else {
            block_dc_lex(f, B.mod, B.invd, B.g, B.geo, B.bc, homog, omega, sweeps);
        }
    } else if (B.order == 2) {  // multi-colour
        for (int s = 0; s < sweeps; ++s) {
            for (int colour = 0; colour < 4; ++colour)
                block_cells(B.g, 0, [&](int i, int j, size_t) {
                    if (i > 0 && i < B.g.nx - 1 && j > 0 && j < B.g.ny - 1 && ((i & 1) | ((j & 1) << 1)) == colour)
                        gs_cell(f, rhs, B.invd, B.g, B.geo, B.bc, homog, omega, i, j);
                });
            block_gs_mc_ring(f, rhs, B.invd, B.g, B.geo, B.bc, homog, omega);
        }
    } else {
        block_gs_lex(f, rhs, B.invd, B.g, B.geo, B.bc, homog, omega, sweeps);
    }
}

__global__ void __launch_bounds__(256) k_bicgstab2(Bottom B) {
    const double reps = 1.0e-12, hang = 1.0e-7, small = 1.0e-30;
    const int norm_type = B.dc ? 0 : 2;
    const double relax_omega = B.dc ? 4.0 / 5.0 : 1.0 - 1.0 / 3.0;
    const int relax_iters = 2;
    const Grid &g = B.g;
    if (B.max_iters == 0) return;
    // e, p_tilde, s_tilde start at zero, ghosts included
    block_cells(g, 1, [&](int, int, size_t c) { B.e[c] = 0.0; B.pt[c] = 0.0; B.st[c] = 0.0; });
    int recount = 0, it = 0, restarts = 0, init = 1;
    block_residual(B, B.homog, B.rhs, B.phi, B.r);
    block_cells(g, 0, [&](int, int, size_t c) { B.rt[c] = B.r[c]; });
    double norm0 = seq_norm(B, B.r, norm_type);
    const double initial_norm = norm0, initial_rnorm = norm0;
    double norm1 = norm0;
    double rho1 = 0, rho2 = 0, alpha0 = 0, alpha1 = 0, omega0 = 0, omega1 = 0, beta1 = 0;
    while (it < B.max_iters && norm0 > B.tol * norm1 && norm1 > 0.0) {
        ++it;
        norm1 = norm0;
        alpha1 = alpha0;
        beta1 = 0.0;
        omega1 = omega0;
        rho2 = rho1;
        rho1 = seq_dot(B, B.rt, B.r);
        if (rho1 == 0.0) break;
        if (init) {
            block_cells(g, 0, [&](int, int, size_t c) { B.p[c] = B.r[c]; });
            init = 0;
        } else {
            beta1 = (rho1 / rho2) * (alpha1 / omega1);
            const double b1 = beta1, o1 = omega1;
            block_cells(g, 0, [&](int, int, size_t c) { B.p[c] = b1 * (B.p[c] - o1 * B.v[c]) + B.r[c]; });
        }
        block_cells(g, 0, [&](int, int, size_t c) { B.pt[c] = B.p[c] * B.invd[c]; });
        block_relax(B, 1, relax_omega, relax_iters, B.p, B.pt);
        block_laplacian(B, B.pt, B.v);
        const double m = seq_dot(B, B.rt, B.v);
        alpha0 = rho1 / m;
        if (fabs(m) > small * fabs(rho1)) {
            const double a0 = alpha0;
            block_cells(g, 0, [&](int, int, size_t c) { B.r[c] = B.r[c] - a0 * B.v[c]; });
            norm0 = seq_norm(B, B.r, norm_type);
            block_cells(g, 1, [&](int, int, size_t c) { B.e[c] = B.e[c] + a0 * B.pt[c]; });
        } else {
            block_cells(g, 0, [&](int, int, size_t c) { B.r[c] = 0.0; });
            norm0 = 0.0;
        }
        if (norm0 > B.tol * initial_norm && norm0 > reps * initial_rnorm) {
            block_cells(g, 0, [&](int, int, size_t c) { B.st[c] = B.r[c] * B.invd[c]; });
            block_relax(B, 1, relax_omega, relax_iters, B.r, B.st);
            block_laplacian(B, B.st, B.t);
            omega0 = seq_dot(B, B.t, B.r) / seq_dot(B, B.t, B.t);
            const double o0 = omega0;
            block_cells(g, 1, [&](int, int, size_t c) { B.e[c] = B.e[c] + o0 * B.st[c]; });
            block_cells(g, 0, [&](int, int, size_t c) { B.r[c] = B.r[c] - o0 * B.t[c]; });
            norm0 = seq_norm(B, B.r, norm_type);
        }
        if (norm0 <= B.tol * initial_norm || norm0 <= reps * initial_rnorm) break;
        if (omega0 == 0.0 || norm0 > (1.0 - hang) * norm1) {
            if (recount == 0) {
                recount = 1;
            } else {
                recount = 0;
                block_cells(g, 1, [&](int, int, size_t c) { B.phi[c] = B.phi[c] + B.e[c]; });
                // the reference exits here and adds e once more after the loop
                if (restarts == B.max_restarts) break;
                block_residual(B, B.homog, B.rhs, B.phi, B.r);
                norm0 = seq_norm(B, B.r, norm_type);
                rho1 = rho2 = 0.0;
                alpha0 = 0.0;
                omega0 = 0.0;
                block_cells(g, 0, [&](int, int, size_t c) { B.rt[c] = B.r[c]; });
                block_cells(g, 1, [&](int, int, size_t c) { B.e[c] = 0.0; });
                ++restarts;
                init = 1;
            }
        }
    }
    (void)beta1;
    block_cells(g, 1, [&](int, int, size_t c) { B.phi[c] = B.phi[c] + B.e[c]; });
    if (threadIdx.x == 0 && B.scal) {
        B.scal[0] = (double)it;
        B.scal[1] = norm0;
    }
}

// Whole relax_gs call on a small level in one workgroup (coarse levels,
// parity mode): the same block functions the bottom solver uses.
__global__ void __launch_bounds__(256) k_relax_block(Bottom B, double omega, int sweeps, int homog) {
    block_relax(B, homog, omega, sweeps, B.rhs, B.phi);
}

}  // namespace mgp

namespace mgp {

// y-direction prolongation for 1x2 semicoarsening (prolong_y_cell).
template <typename T>
__global__ void __launch_bounds__(256) k_prolong_y(T *ef, const T *ec, Grid gf, Grid gc, T slo, T shi, int off) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= gf.nx || j >= gf.ny) return;
    prolong_y_cell(ef, ec, gf, gc, slo, shi, i, j, off);
}

template <typename TO, typename TI>
__global__ void __launch_bounds__(256) k_convert(TO *out, const TI *in, Grid g, int ghosted) {
    const int o = ghosted ? 1 : 0;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63) - o;
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6) - o;
    if (i >= g.nx + o || j >= g.ny + o) return;
    const size_t c = ix(g.pitch, i, j);
    out[c] = (TO)in[c];
}

__global__ void __launch_bounds__(256) k_convert_flat(float *out, const double *in, size_t n) {
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k < n) out[k] = (float)in[k];
}

__global__ void __launch_bounds__(256) k_setval_f(float *y, float v, Grid g, int ghosted) {
    const int o = ghosted ? 1 : 0;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63) - o;
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6) - o;
    if (i >= g.nx + o || j >= g.ny + o) return;
    y[ix(g.pitch, i, j)] = v;
}

__global__ void __launch_bounds__(256) k_begin32(float *f32, float *e32, const double *f, Grid g) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63) - 1;
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6) - 1;
    if (i > g.nx || j > g.ny) return;
    const size_t c = ix(g.pitch, i, j);
    e32[c] = 0.0f;
    if (i >= 0 && i < g.nx && j >= 0 && j < g.ny) f32[c] = (float)f[c];
}

// out = a*x + b*y + c*z over valid cells (null operands skipped).
__global__ void __launch_bounds__(256) k_axpbypcz(double *out, double a, const double *x, double b, const double *y,
                                                  double c, const double *z, Grid g) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= g.nx || j >= g.ny) return;
    const size_t q = ix(g.pitch, i, j);
    double v = a * x[q];
    if (y) v += b * y[q];
    if (z) v += c * z[q];
    out[q] = v;
}

// ---------------------------------------------------------------------------
// BiCGStab scalars kept in device memory (so an iteration needs no host
// round trip and can be replayed as a graph).  Layout of sc[]: see SC_* in
// mgp_kernels.h.  Reductions deliver sum(vol * a * b).
// ---------------------------------------------------------------------------
__global__ void k_bicg_scalars(double *sc, int op) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (op == 0) {          // beta = (rho_new/rho)(alpha/omega); rho = rho_new
        sc[SC_BETA] = (sc[SC_RHO_NEW] / sc[SC_RHO]) * (sc[SC_ALPHA] / sc[SC_OMEGA]);
        sc[SC_RHO] = sc[SC_RHO_NEW];
    } else if (op == 1) {   // alpha = rho / (rhat . v)
        sc[SC_ALPHA] = sc[SC_RHO] / sc[SC_RV];
    } else if (op == 2) {   // omega = (t . s) / (t . t)
        sc[SC_OMEGA] = sc[SC_TT] > 0.0 ? sc[SC_TS] / sc[SC_TT] : 0.0;
    } else {                // restart
        sc[SC_RHO] = 1.0;
        sc[SC_ALPHA] = 1.0;
        sc[SC_OMEGA] = 1.0;
        sc[SC_BETA] = 0.0;
    }
}

// Vector updates with device scalars (valid cells):
//  0: p = r + beta (p - omega v)     (out=p, x=r, y=v)
//  1: s = r - alpha v                (out=s, x=r, y=v)
//  2: phi = phi + alpha phat         (out=phi, x=phat)
//  3: phi = phi + omega shat         (out=phi, x=shat)
//  4: r = s - omega t                (out=r, x=s, y=t)
__global__ void __launch_bounds__(256) k_bicg_update(double *out, const double *x, const double *y,
                                                     const double *sc, Grid g, int op) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= g.nx || j >= g.ny) return;
    const size_t c = ix(g.pitch, i, j);
    switch (op) {
    case 0: out[c] = x[c] + sc[SC_BETA] * (out[c] - sc[SC_OMEGA] * y[c]); break;
    case 1: out[c] = x[c] - sc[SC_ALPHA] * y[c]; break;
    case 2: out[c] = out[c] + sc[SC_ALPHA] * x[c]; break;
    case 3: out[c] = out[c] + sc[SC_OMEGA] * x[c]; break;
    default: out[c] = x[c] - sc[SC_OMEGA] * y[c]; break;
    }
}

}  // namespace mgp

// ===========================================================================
// Launchers
// ===========================================================================
namespace mgp {

static inline dim3 tiles(int nx, int ny) { return dim3((unsigned)((nx + 63) / 64), (unsigned)((ny + 3) / 4)); }

// Grids whose perimeter one workgroup walks too slowly: the phases of the
// ghost fill and of the ring sweep as multi-workgroup launches.
static bool wide_perimeter(const Grid &g) { return g.nx + g.ny >= 512; }

void launch_fill_ghosts(hipStream_t s, double *phi, const Grid &g, const Geo &geo, const Bc &bc, int homog,
                        int do_neum) {
    const Fld f{phi, g.pitch};
    if (!wide_perimeter(g)) {
        k_fill_ghosts<<<1, 256, 0, s>>>(f, g, geo, bc, homog, do_neum);
        return;
    }
    k_fill_x_ghosts<<<(2 * (g.ny + 2) + 255) / 256, 256, 0, s>>>(f, g, geo, bc, homog, do_neum);
    k_fill_y_ghosts<<<(2 * (g.nx + 2) + 255) / 256, 256, 0, s>>>(f, g, geo, bc, homog, do_neum);
    k_fill_corner_ghosts<<<1, 64, 0, s>>>(f, g, bc);
}
void launch_laplacian(hipStream_t s, double *out, const double *phi, const double *rhs, const Grid &g,
                      const Geo &geo, const Bc &bc, int homog, int jscale) {
    k_laplacian<<<tiles(g.nx, g.ny), 256, 0, s>>>(out, CFld{phi, g.pitch}, rhs, g, geo, bc, homog, jscale);
}
void launch_residual_synth(hipStream_t s, double *out, const double *phi, const double *rhs, const Grid &g,
                           const Geo &geo, const Bc &bc) {
    k_residual_synth<<<tiles(g.nx, g.ny), 256, 0, s>>>(out, phi, rhs, g, geo, bc);
}
void launch_invdiags(hipStream_t s, double *out, const Grid &g, const Geo &geo, const Bc &bc, int homog) {
    k_invdiags<<<tiles(g.nx, g.ny), 256, 0, s>>>(out, g, geo, bc, homog);
}
void launch_dc_invdiags(hipStream_t s, double *out, const Grid &g, const Geo &geo) {
    k_dc_invdiags<<<tiles(g.nx, g.ny), 256, 0, s>>>(out, g, geo);
}
void launch_dc_laplacian(hipStream_t s, double *out, const double *phi, const double *rhs, const Grid &g,
                         const Geo &geo) {
    k_dc_laplacian<<<tiles(g.nx, g.ny), 256, 0, s>>>(out, CFld{phi, g.pitch}, rhs, g, geo);
}
void launch_dc_modrhs(hipStream_t s, double *modrhs, const double *rhs, const double *phi, const Grid &g,
                      const Geo &geo) {
    k_dc_modrhs<<<tiles(g.nx, g.ny), 256, 0, s>>>(modrhs, rhs, CFld{phi, g.pitch}, g, geo);
}
void launch_gs_mc_interior(hipStream_t s, double *phi, const double *rhs, const double *invd, const Grid &g,
                           const Geo &geo, double omega, int ci, int cj) {
    const int na = (g.nx - 2 - (ci ? 1 : 2)) / 2 + 1, nb = (g.ny - 2 - (cj ? 1 : 2)) / 2 + 1;
    if (na <= 0 || nb <= 0) return;
    k_gs_mc_interior<<<tiles(na, nb), 256, 0, s>>>(Fld{phi, g.pitch}, rhs, invd, g, geo, omega, ci, cj);
}
void launch_gs_mc_interior_tiled(hipStream_t s, double *phi, const double *rhs, const double *invd, const Grid &g,
                                 const Geo &geo, double omega) {
    if (g.nx < 3 || g.ny < 3) return;
    const dim3 grid((unsigned)((g.nx - 2 + MC_TX - 1) / MC_TX), (unsigned)((g.ny - 2 + MC_TY - 1) / MC_TY));
    k_gs_mc_tile<<<grid, 256, 0, s>>>(Fld{phi, g.pitch}, rhs, invd, g, geo, omega);
}
void launch_gs_mc_ring(hipStream_t s, double *phi, const double *rhs, const double *invd, const Grid &g,
                       const Geo &geo, const Bc &bc, int homog, double omega) {
    const Fld f{phi, g.pitch};
    if (!wide_perimeter(g)) {
        k_gs_mc_ring<<<1, 1024, 0, s>>>(f, rhs, invd, g, geo, bc, homog, omega);
        return;
    }
    launch_fill_ghosts(s, phi, g, geo, bc, homog, 0);
    const int nring = 2 * (g.nx - 2) + 2 * (g.ny - 2);
    for (int colour = 0; colour < 4; ++colour)
        k_gs_mc_ring_colour<<<(nring + 255) / 256, 256, 0, s>>>(f, rhs, invd, g, geo, bc, homog, omega, colour);
    k_gs_mc_corner_cells<<<1, 64, 0, s>>>(f, rhs, invd, g, geo, bc, homog, omega);
}
void launch_gs_lex(hipStream_t s, double *phi, const double *rhs, const double *invd, const Grid &g,
                   const Geo &geo, const Bc &bc, int homog, double omega, int sweeps) {
    k_gs_lex<<<1, 1024, 0, s>>>(Fld{phi, g.pitch}, rhs, invd, g, geo, bc, homog, omega, sweeps);
}
void launch_dc_rb(hipStream_t s, double *phi, const double *modrhs, const double *invd, const Grid &g,
                  const Geo &geo, double omega, int colour) {
    k_dc_rb<<<tiles((g.nx + 1) / 2, g.ny), 256, 0, s>>>(Fld{phi, g.pitch}, modrhs, invd, g, geo, omega, colour);
}
void launch_dc_lex(hipStream_t s, double *phi, const double *modrhs, const double *invd, const Grid &g,
                   const Geo &geo, const Bc &bc, int homog, double omega, int sweeps) {
    k_dc_lex<<<1, 1024, 0, s>>>(Fld{phi, g.pitch}, modrhs, invd, g, geo, bc, homog, omega, sweeps);
}
void launch_restrict_cc(hipStream_t s, double *crse, const double *fine, const Grid &cg, int fpitch, int rx,
                        int ry, double *e, const double *invd) {
    k_restrict_cc<<<tiles(cg.nx, cg.ny), 256, 0, s>>>(crse, fine, cg, fpitch, rx, ry, e, invd);
}
void launch_restrict_xf(hipStream_t s, double *crse, const double *fine, const Grid &cg, int fpitch, int rx,
                        int ry) {
    k_restrict_xf<<<tiles(cg.nx + 1, cg.ny), 256, 0, s>>>(crse, fine, cg, fpitch, rx, ry);
}
void launch_restrict_yf(hipStream_t s, double *crse, const double *fine, const Grid &cg, int fpitch, int rx,
                        int ry) {
    k_restrict_yf<<<tiles(cg.nx, cg.ny + 1), 256, 0, s>>>(crse, fine, cg, fpitch, rx, ry);
}
void launch_prolong2x2(hipStream_t s, double *fine, int fpitch, const double *crse, const double *crse_x3,
                       const Grid &cg, int order) {
    k_prolong2x2<<<tiles(cg.nx, cg.ny), 256, 0, s>>>(Fld{fine, fpitch}, CFld{crse, cg.pitch},
                                                      CFld{crse_x3 ? crse_x3 : crse, cg.pitch}, cg, order);
}
void launch_extrap3(hipStream_t s, double *dst, const double *src, const Grid &g) {
    k_extrap3<<<1, 256, 0, s>>>(Fld{dst, g.pitch}, CFld{src, g.pitch}, g);
}
void launch_extrapolate(hipStream_t s, double *phi, const Grid &g, int order) {
    k_extrapolate<<<1, 256, 0, s>>>(Fld{phi, g.pitch}, g, order);
}
void launch_einit(hipStream_t s, double *e, const double *r, const double *invd, const Grid &g) {
    k_einit<<<tiles(g.nx, g.ny), 256, 0, s>>>(e, r, invd, g);
}
void launch_axpy_ghosted(hipStream_t s, double *y, const double *x, double a, const Grid &g) {
    k_axpy_ghosted<<<tiles(g.nx + 2, g.ny + 2), 256, 0, s>>>(y, x, a, g);
}
void launch_lincomb(hipStream_t s, double *out, double a, const double *y, double b, const double *x,
                    const Grid &g, int ghosted) {
    const int o = ghosted ? 2 : 0;
    k_lincomb<<<tiles(g.nx + o, g.ny + o), 256, 0, s>>>(out, a, y, b, x, g, ghosted);
}
void launch_lincomb_mul(hipStream_t s, double *out, const double *a, const double *b, const Grid &g, int) {
    k_muldiv<<<tiles(g.nx, g.ny), 256, 0, s>>>(out, a, b, g, 0);
}
void launch_lincomb_div(hipStream_t s, double *out, const double *a, const double *b, const Grid &g, int) {
    k_muldiv<<<tiles(g.nx, g.ny), 256, 0, s>>>(out, a, b, g, 1);
}
void launch_setval(hipStream_t s, double *y, double v, const Grid &g, int ghosted) {
    const int o = ghosted ? 2 : 0;
    k_setval<<<tiles(g.nx + o, g.ny + o), 256, 0, s>>>(y, v, g, ghosted);
}
int reduce_blocks(const Grid &g) {
    const long n = (long)g.nx * g.ny;
    long b = (n + 255) / 256;
    if (b > 1024) b = 1024;
    if (b < 1) b = 1;
    return (int)b;
}
void launch_reduce(hipStream_t s, double *out, double *partial, const double *a, const double *b,
                   const double *J, double vol, const Grid &g, int kind, int p) {
    const int nb = reduce_blocks(g) < g.ny ? reduce_blocks(g) : g.ny;
    k_reduce_partial<<<nb, 256, 0, s>>>(partial, a, b, J, vol, g, kind, p);
    k_reduce_final<<<1, 256, 0, s>>>(out, partial, nb, kind);
}
void launch_bicg_scalars(hipStream_t s, double *sc, int op) { k_bicg_scalars<<<1, 64, 0, s>>>(sc, op); }
void launch_bicg_update(hipStream_t s, double *out, const double *x, const double *y, const double *sc,
                        const Grid &g, int op) {
    k_bicg_update<<<tiles(g.nx, g.ny), 256, 0, s>>>(out, x, y, sc, g, op);
}
void launch_seq_mean(hipStream_t s, double *out, const double *fine, const double *J, const Grid &g, double dxprod) {
    k_seq_mean<<<1, 64, 0, s>>>(out, fine, J, g, dxprod);
}
void launch_sub_mean(hipStream_t s, double *fine, const double *sums, const Grid &g) {
    k_sub_mean<<<tiles(g.nx + 2, g.ny + 2), 256, 0, s>>>(fine, sums, g);
}
void launch_bicgstab2(hipStream_t s, const BottomArgs &a) { k_bicgstab2<<<1, 256, 0, s>>>(a); }
void launch_relax_block(hipStream_t s, const BottomArgs &a, double omega, int sweeps, int homog) {
    k_relax_block<<<1, 256, 0, s>>>(a, omega, sweeps, homog);
}

}  // namespace mgp

namespace mgp {
void launch_prolong_y(hipStream_t s, double *ef, const double *ec, const Grid &gf, const Grid &gc, double slo,
                      double shi, int off) {
    k_prolong_y<<<tiles(gf.nx, gf.ny), 256, 0, s>>>(ef, ec, gf, gc, slo, shi, off);
}
void launch_prolong_y(hipStream_t s, float *ef, const float *ec, const Grid &gf, const Grid &gc, double slo,
                      double shi, int off) {
    k_prolong_y<<<tiles(gf.nx, gf.ny), 256, 0, s>>>(ef, ec, gf, gc, (float)slo, (float)shi, off);
}

void launch_setval(hipStream_t s, float *y, float v, const Grid &g, int ghosted) {
    const int o = ghosted ? 2 : 0;
    k_setval_f<<<tiles(g.nx + o, g.ny + o), 256, 0, s>>>(y, v, g, ghosted);
}

void launch_convert(hipStream_t s, float *out, const double *in, const Grid &g, int ghosted) {
    const int o = ghosted ? 2 : 0;
    k_convert<<<tiles(g.nx + o, g.ny + o), 256, 0, s>>>(out, in, g, ghosted);
}

void launch_convert(hipStream_t s, double *out, const float *in, const Grid &g, int ghosted) {
    const int o = ghosted ? 2 : 0;
    k_convert<<<tiles(g.nx + o, g.ny + o), 256, 0, s>>>(out, in, g, ghosted);
}

void launch_convert_flat(hipStream_t s, float *out, const double *in, size_t n) {
    if (n == 0) return;
    k_convert_flat<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(out, in, n);
}

void launch_begin32(hipStream_t s, float *f32, float *e32, const double *f, const Grid &g) {
    k_begin32<<<tiles(g.nx + 2, g.ny + 2), 256, 0, s>>>(f32, e32, f, g);
}
void launch_axpbypcz(hipStream_t s, double *out, double a, const double *x, double b, const double *y, double c,
                     const double *z, const Grid &g) {
    k_axpbypcz<<<tiles(g.nx, g.ny), 256, 0, s>>>(out, a, x, b, y, c, z, g);
}
}  // namespace mgp
