// mgp_rbtile.hip -- red-black Gauss-Seidel of the deferred-correction
// smoother (MGPoisson2D_DD.f90:227-285: modrhs, then per sweep a ghost fill,
// colour 0 and colour 1) with several sweeps per load of an LDS tile.
//
// A workgroup loads a window of phi (128 cells wide, H rows) into LDS and
// keeps its cells' face metrics, modified rhs and inverse diagonal in
// registers; it then runs S sweeps in LDS and writes the window's inner tile
// to a second array.  Updates of one colour read only the other colour, so
// the result is the same in any order within a colour: a cell's value after
// the S sweeps equals, bit for bit, what the one-colour-per-launch sequence
// gives, as long as every value it depends on lies in the window.  One
// half-sweep moves that dependence one cell outwards, so a window reaching
// R = 2S cells beyond the tile on every side gives the tile exact values.
// At the domain boundary the ghost cells are refreshed in LDS before each
// sweep, as the fill before each sweep does; on the level BCs of the
// deferred-correction cycle (mode LAH) a ghost is the adjacent cell, copied
// (Neumann) or negated (Dirichlet), so that refresh is local to the window.
//
// Windows overlap, so a workgroup must not write where a neighbour still
// reads: the result goes to a second array (ping-pong between phi and a
// scratch level, with a final copy when the count of launches is odd).
//
// The tile shape (rows per window, waves per workgroup) is picked per level
// shape by timing the candidates on first use.  Sweeps per load follow the
// call: 2 per launch, 1 for an odd remainder.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>

#include "mgp_host.h"
#include "mgp_kernels.h"
#include "mgp_stencil.h"

namespace mgp {
namespace {

constexpr int RB_WW = 128;  // window width: each lane owns two adjacent columns

// ghost = sgn[side] * the adjacent domain cell (LAH: Neumann +1, Dirichlet -1)
struct RbGhost {
    double sgn[4];
};

// XCD-aware block order: XCD x (= blockIdx % 8) takes a contiguous run of
// tiles, so neighbouring tiles (which read each other's edges) share an L2.
__device__ __forceinline__ int rb_slot() {
    const int nb = gridDim.x, b = blockIdx.x, x = b & 7;
    const int q8 = nb >> 3, r8 = nb & 7;
    return x * q8 + (x < r8 ? x : r8) + (b >> 3);
}

// S sweeps on the window of tile (tx, ty).  FUSE: src is rhs and the
// modified rhs (rhs - lagged cross terms of the incoming phi) is formed here;
// else src is the modified rhs.
template <int H, int NW, int S, bool FUSE>
__global__ void __launch_bounds__(64 * NW) k_dc_rb_tile(const double *__restrict__ in, double *__restrict__ out,
                                                        const double *__restrict__ src,
                                                        const double *__restrict__ invd, Grid g, Geo geo, RbGhost gh,
                                                        double omega, int ntx) {
    constexpr int WW = RB_WW, R = 2 * S, TX = WW - 2 * R, TY = H - 2 * R, Q = H / NW;
    static_assert(H % NW == 0 && NW % 2 == 0 && TY > 0, "window rows split evenly over an even wave count");
    __shared__ double w[H][WW];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = rb_slot();
    const int tx = t % ntx, ty = t / ntx;
    const int wx0 = tx * TX - R, wy0 = ty * TY - R;  // window origin; both even
    const int nx = g.nx, ny = g.ny, P = g.pitch;
    const double idx = 1.0 / g.dx, idy = 1.0 / g.dy;
    const int lx0 = 2 * lane;

    // this lane's cells: columns lx0 + m, rows wave + NW q
    double jw[Q][2], je[Q][2], js[Q][2], jn[Q][2], md[Q][2], iv[Q][2];
    unsigned upd = 0;                                 // bit 2q+m: the cell is relaxed
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int ly = wave + NW * q, gj = wy0 + ly;
        const int gi0 = wx0 + lx0;
        if (gj >= 0 && gj < ny && gi0 >= 0 && gi0 + 1 < nx) {
            // both columns in the domain: 16-byte loads (the offset of cell i = 0,
            // the pitch and gi0 are even), and Jxx east of column 0 is Jxx west of column 1
            const size_t c = ix(P, gi0, gj);
            const double2 v = *reinterpret_cast<const double2 *>(in + c);
            const double2 x = *reinterpret_cast<const double2 *>(geo.Jxx + c);
            const double2 ys = *reinterpret_cast<const double2 *>(geo.Jyy + c);
            const double2 yn = *reinterpret_cast<const double2 *>(geo.Jyy + c + P);
            const double2 sr = *reinterpret_cast<const double2 *>(src + c);
            const double2 dv = *reinterpret_cast<const double2 *>(invd + c);
            jw[q][0] = x.x;
            jw[q][1] = x.y;
            je[q][0] = x.y;
            je[q][1] = geo.Jxx[c + 2];
            js[q][0] = ys.x;
            js[q][1] = ys.y;
            jn[q][0] = yn.x;
            jn[q][1] = yn.y;
            md[q][0] = sr.x;
            md[q][1] = sr.y;
            iv[q][0] = dv.x;
            iv[q][1] = dv.y;
            w[ly][lx0] = v.x;
            w[ly][lx0 + 1] = v.y;
#pragma unroll
            for (int m = 0; m < 2; ++m)
                if (lx0 + m >= 1 && lx0 + m <= WW - 2 && ly >= 1 && ly <= H - 2) upd |= 1u << (2 * q + m);
            continue;
        }
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const int lx = lx0 + m, gi = wx0 + lx;
            const bool dom = gi >= 0 && gi < nx && gj >= 0 && gj < ny;
            double v = 0.0;
            jw[q][m] = je[q][m] = js[q][m] = jn[q][m] = md[q][m] = iv[q][m] = 0.0;
            if (dom) {
                const size_t c = ix(P, gi, gj);
                v = in[c];
                jw[q][m] = geo.Jxx[c];
                je[q][m] = geo.Jxx[c + 1];
                js[q][m] = geo.Jyy[c];
                jn[q][m] = geo.Jyy[c + P];
                md[q][m] = src[c];
                iv[q][m] = invd[c];
                if (lx >= 1 && lx <= WW - 2 && ly >= 1 && ly <= H - 2) upd |= 1u << (2 * q + m);
            }
            w[ly][lx] = v;
        }
    }
    __syncthreads();

    if (FUSE) {
        // MGPoisson2D_DD.f90:227-250 with the neighbours clamped into the domain
        const double xyscale = 0.5 * idx * idy;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int ly = wave + NW * q, gj = wy0 + ly;
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                if (!(upd >> (2 * q + m) & 1u)) continue;
                const int gi = wx0 + lx0 + m;
                const int le = (gi + 1 < nx ? gi + 1 : nx - 1) - wx0, lw = (gi - 1 >= 0 ? gi - 1 : 0) - wx0;
                const int ln = (gj + 1 < ny ? gj + 1 : ny - 1) - wy0, ls = (gj - 1 >= 0 ? gj - 1 : 0) - wy0;
                // the cross metrics are read here, not held across the load
                const size_t c = ix(P, gi, gj);
                md[q][m] = md[q][m] - dc_lagged_cross_vals(w[ln][le], w[ln][lw], w[ls][lw], w[ls][le], geo.Jxy[c + 1],
                                                           geo.Jxy[c], geo.Jyx[c + P], geo.Jyx[c], xyscale);
            }
        }
    }

    auto relax = [&](int q, int m) {
        if (!(upd >> (2 * q + m) & 1u)) return;
        const int ly = wave + NW * q, lx = lx0 + m;
        const double p = w[ly][lx];
        double ax, ay;
        dc_five_vals(p, w[ly][lx + 1], w[ly][lx - 1], w[ly + 1][lx], w[ly - 1][lx], je[q][m], jw[q][m], jn[q][m],
                     js[q][m], idx, idy, ax, ay);
        const double lphi = ax + ay;
        w[ly][lx] = gs_update(p, md[q][m], lphi, iv[q][m], omega);
    };

#pragma unroll 1
    for (int sw = 0; sw < S; ++sw) {
        // ghost fill (arrayutils.f90 fill_ghosts, LAH): x sides over the
        // domain rows, y sides over the domain columns; corners are never read
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int ly = wave + NW * q, gj = wy0 + ly;
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                const int lx = lx0 + m, gi = wx0 + lx;
                if (gj >= 0 && gj < ny) {
                    if (gi == -1) w[ly][lx] = gh.sgn[XLO] * w[ly][lx + 1];
                    else if (gi == nx) w[ly][lx] = gh.sgn[XHI] * w[ly][lx - 1];
                } else if (gi >= 0 && gi < nx) {
                    if (gj == -1) w[ly][lx] = gh.sgn[YLO] * w[ly + 1][lx];
                    else if (gj == ny) w[ly][lx] = gh.sgn[YHI] * w[ly - 1][lx];
                }
            }
        }
        __syncthreads();
        // colour c relaxes (i + j + c) even; with an even window origin and an
        // even wave count that is column m = (wave + c) mod 2 of every pair
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            if (((wave + c) & 1) == 0) {
#pragma unroll
                for (int q = 0; q < Q; ++q) relax(q, 0);
            } else {
#pragma unroll
                for (int q = 0; q < Q; ++q) relax(q, 1);
            }
            __syncthreads();
        }
    }

    // the tile's cells, and the ghosts next to them (values of the last
    // fill; corners take the fill's BOGUS)
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int ly = wave + NW * q, gj = wy0 + ly;
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const int lx = lx0 + m, gi = wx0 + lx;
            if (gi < -1 || gi > nx || gj < -1 || gj > ny) continue;
            const int ci = gi < 0 ? 0 : (gi >= nx ? nx - 1 : gi), cj = gj < 0 ? 0 : (gj >= ny ? ny - 1 : gj);
            if (ci / TX != tx || cj / TY != ty) continue;
            const bool corner = (gi < 0 || gi >= nx) && (gj < 0 || gj >= ny);
            out[ix(P, gi, gj)] = corner ? BOGUS : w[ly][lx];
        }
    }
}

// dst = src over cells (-1 .. nx, -1 .. ny) of a level (n = nx + 2 per row,
// nbx blocks of 256 per row, one 1-D grid over all rows)
__global__ void __launch_bounds__(256) k_copy_level(double *__restrict__ dst, const double *__restrict__ src,
                                                     int pitch, int n, int nbx) {
    const int bx = blockIdx.x % nbx, j = (int)(blockIdx.x / nbx) - 1;
    const int i = bx * 256 + threadIdx.x - 1;
    if (i + 1 >= n) return;
    const size_t c = ix(pitch, i, j);
    dst[c] = src[c];
}

struct RbShape {
    int H, NW;  // window rows, waves
};
// candidate shapes, in the numbering of the tuning switch (k >= 1 is entry k-1)
constexpr RbShape RB_SHAPES[] = {{16, 4}, {32, 4}, {32, 8}, {64, 8}};
constexpr int RB_NSHAPES = sizeof(RB_SHAPES) / sizeof(RB_SHAPES[0]);

int rb_tiles(const Grid &g, int shape, int S, int &ntx) {
    const int TX = RB_WW - 4 * S, TY = RB_SHAPES[shape].H - 4 * S;
    ntx = (g.nx + TX - 1) / TX;
    return ntx * ((g.ny + TY - 1) / TY);
}

template <int H, int NW, int S, bool F>
void launch_shape(hipStream_t s, int shape, const double *in, double *out, const double *src, const double *invd,
                  const Grid &g, const Geo &geo, const RbGhost &gh, double omega) {
    int ntx;
    const int nb = rb_tiles(g, shape, S, ntx);
    k_dc_rb_tile<H, NW, S, F><<<nb, 64 * NW, 0, s>>>(in, out, src, invd, g, geo, gh, omega, ntx);
}

template <int S, bool F>
void launch_s(hipStream_t s, int shape, const double *in, double *out, const double *src, const double *invd,
              const Grid &g, const Geo &geo, const RbGhost &gh, double omega) {
#define MGP_RB_SHAPE(k)                                                                                              \
    case k:                                                                                                          \
        launch_shape<RB_SHAPES[k].H, RB_SHAPES[k].NW, S, F>(s, k, in, out, src, invd, g, geo, gh, omega);         \
        break;
    switch (shape) {
        MGP_RB_SHAPE(0)
        MGP_RB_SHAPE(1)
        MGP_RB_SHAPE(2)
        MGP_RB_SHAPE(3)
        default: break;
    }
#undef MGP_RB_SHAPE
}

void launch_tile(hipStream_t s, int shape, int S, bool fuse, const double *in, double *out, const double *src,
                 const double *invd, const Grid &g, const Geo &geo, const RbGhost &gh, double omega) {
    if (S == 2) {
        if (fuse) launch_s<2, true>(s, shape, in, out, src, invd, g, geo, gh, omega);
        else launch_s<2, false>(s, shape, in, out, src, invd, g, geo, gh, omega);
    } else {
        if (fuse) launch_s<1, true>(s, shape, in, out, src, invd, g, geo, gh, omega);
        else launch_s<1, false>(s, shape, in, out, src, invd, g, geo, gh, omega);
    }
}

std::mutex rb_mu;
int rb_choice = -1;  // -1: timed pick; 0: one colour per launch; k >= 1: shape k - 1
std::map<std::tuple<int, int, int, int>, int> rb_picked;

// The shape with the fewest window cells over the level (small levels).
int rb_fewest_cells(const Grid &g, int S) {
    int best = 0;
    long long bestc = -1;
    for (int k = 0; k < RB_NSHAPES; ++k) {
        int ntx;
        const long long c = (long long)rb_tiles(g, k, S, ntx) * RB_WW * RB_SHAPES[k].H;
        if (bestc < 0 || c < bestc) best = k, bestc = c;
    }
    return best;
}

// The shape for this level, sweeps per load and form: timed once per
// (nx, ny, S, fuse) on the call's own operands (the timing launches write
// the same values the real launch writes next).  The lock guards only the
// setting and the cache: the timing runs without it, so other threads'
// sweeps do not wait for it.  Two threads that time the same key at once
// both insert; the first insertion stays (every shape gives the same bits).
int rb_pick(hipStream_t s, int S, bool fuse, const double *in, double *out, const double *src, const double *invd,
            const Grid &g, const Geo &geo, const RbGhost &gh, double omega) {
    const auto key = std::make_tuple(g.nx, g.ny, S, (int)fuse);
    {
        std::lock_guard<std::mutex> lk(rb_mu);
        if (rb_choice > 0) return rb_choice - 1;
        auto it = rb_picked.find(key);
        if (it != rb_picked.end()) return it->second;
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    hip_check(hipStreamIsCapturing(s, &cs), "hipStreamIsCapturing");
    if ((long long)g.nx * g.ny < 256LL * 256 || cs != hipStreamCaptureStatusNone) {
        const int k = rb_fewest_cells(g, S);
        if (cs == hipStreamCaptureStatusNone) {
            std::lock_guard<std::mutex> lk(rb_mu);
            rb_picked.emplace(key, k);
        }
        return k;
    }
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    int best = 0;
    float bestms = -1.0f;
    for (int k = 0; k < RB_NSHAPES; ++k) {
        launch_tile(s, k, S, fuse, in, out, src, invd, g, geo, gh, omega);
        hip_check(hipEventRecord(e0, s), "hipEventRecord");
        for (int r = 0; r < 3; ++r) launch_tile(s, k, S, fuse, in, out, src, invd, g, geo, gh, omega);
        hip_check(hipEventRecord(e1, s), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        if (bestms < 0.0f || ms < bestms) best = k, bestms = ms;
    }
    hip_check(hipEventDestroy(e0), "hipEventDestroy");
    hip_check(hipEventDestroy(e1), "hipEventDestroy");
    std::lock_guard<std::mutex> lk(rb_mu);
    return rb_picked.emplace(key, best).first->second;
}

bool rb_tiles_apply(const Bc &bc) {
    for (int k = 0; k < 4; ++k)
        if (bc.mode[k] != MODE_LAH || (bc.type[k] != BC_NEUM && bc.type[k] != BC_DIRI)) return false;
    return true;
}

}  // namespace

int dc_rb_tile_choice(int choice) {
    std::lock_guard<std::mutex> lk(rb_mu);
    const int prev = rb_choice;
    if (choice >= -1 && choice <= RB_NSHAPES) rb_choice = choice;
    return prev;
}

int dc_rb_tile_shapes() { return RB_NSHAPES; }

int dc_rb_tile_once(hipStream_t s, const double *phi, double *out, const double *rhs, const double *invd,
                    const Grid &g, const Geo &geo, const Bc &bc, double omega, int S) {
    if (!rb_tiles_apply(bc) || S < 1 || S > 2) {
        g_last_error = "DC red-black tiles: LAH Neumann/Dirichlet sides and 1 or 2 sweeps";
        throw Failure(MGP_ERR_UNSUPPORTED);
    }
    RbGhost gh;
    for (int k = 0; k < 4; ++k) gh.sgn[k] = bc.type[k] == BC_NEUM ? 1.0 : -1.0;
    const int shape = rb_pick(s, S, true, phi, out, rhs, invd, g, geo, gh, omega);
    launch_tile(s, shape, S, true, phi, out, rhs, invd, g, geo, gh, omega);
    return shape;
}

bool dc_rb_relax(hipStream_t s, double *phi, double *tmp, double *mod, const double *rhs, const double *invd,
                 const Grid &g, const Geo &geo, const Bc &bc, int homog, double omega, int sweeps, bool may_end_in_tmp) {
    if (sweeps <= 0) return false;
    int choice;
    {
        std::lock_guard<std::mutex> lk(rb_mu);
        choice = rb_choice;
    }
    if (choice == 0 || !rb_tiles_apply(bc)) {
        launch_dc_modrhs(s, mod, rhs, phi, g, geo);
        for (int k = 0; k < sweeps; ++k) {
            launch_fill_ghosts(s, phi, g, geo, bc, homog, 1);
            launch_dc_rb(s, phi, mod, invd, g, geo, omega, 0);
            launch_dc_rb(s, phi, mod, invd, g, geo, omega, 1);
        }
        return false;
    }
    RbGhost gh;
    for (int k = 0; k < 4; ++k) gh.sgn[k] = bc.type[k] == BC_NEUM ? 1.0 : -1.0;
    double *cur = phi, *nxt = tmp;
    if (sweeps <= 2) {
        // one launch, modified rhs formed in the kernel
        const int shape = rb_pick(s, sweeps, true, cur, nxt, rhs, invd, g, geo, gh, omega);
        launch_tile(s, shape, sweeps, true, cur, nxt, rhs, invd, g, geo, gh, omega);
        cur = nxt;
    } else {
        launch_dc_modrhs(s, mod, rhs, phi, g, geo);
        for (int left = sweeps; left > 0;) {
            const int S = left >= 2 ? 2 : 1;
            const int shape = rb_pick(s, S, false, cur, nxt, mod, invd, g, geo, gh, omega);
            launch_tile(s, shape, S, false, cur, nxt, mod, invd, g, geo, gh, omega);
            std::swap(cur, nxt);
            left -= S;
        }
    }
    if (cur == phi) return false;
    if (may_end_in_tmp) return true;
    const int n = g.nx + 2, nbx = (n + 255) / 256;
    k_copy_level<<<nbx * (g.ny + 2), 256, 0, s>>>(phi, cur, g.pitch, n, nbx);
    return false;
}

}  // namespace mgp
