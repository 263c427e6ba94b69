// mgp_flux.hip -- flux tools of the projection workflow.
//
//   Poisson2D::compute_pd           poisson2d.f90:404-795
//   Poisson2D::compute_grad         poisson2d.f90:809-850
//   Poisson2D::compute_div          poisson2d.f90:860-879
//   ArrayUtils::fill_boundary_fluxes arrayutils.f90:1571-1622
//
// Face fields live in the padded level layout of mgp_common.h: x-face (i, j)
// with i = 0..nx, j = 0..ny-1 and y-face (i, j) with i = 0..nx-1, j = 0..ny at
// ix(pitch, i, j).  One thread owns the x-face and the y-face that share the
// lower-left node (i, j), so one launch of (nx+1) x (ny+1) threads writes both
// flux components with phi's 3x4 neighbourhood read once through L1.
//
// The transverse derivative at a face (pd with dir != nodedir) is the cross
// term of compute_laplacian: the average of the two centred differences of the
// cells either side of the face, with one-sided 3-point differences in the
// first and last row (column) of faces and, on the two boundary faces of each
// row, the 3(a - b) + c extrapolated neighbours of the reference's boundary
// closures (simple_bdry_stencil = .false.).  Every expression keeps the
// reference's operand order; the library is built with -ffp-contract=off, so
// results are bit-identical to it.
#include <hip/hip_runtime.h>

#include "mgp_common.h"
#include "mgp_kernels.h"

namespace mgp {
namespace {

struct Phi {
    const double *p;
    int pitch;
    __device__ double operator()(int i, int j) const { return p[ix(pitch, i, j)]; }
};

// d/deta at x-face (i, j), i = 0..nx (poisson2d.f90:619-793).
__device__ double pd_eta_at_xface(const Phi &P, int nx, int ny, double scale, int i, int j)
{
    const int ilo = 0, ihi = nx - 1, jlo = 0, jhi = ny - 1;
    if (i > ilo && i <= ihi) {
        if (j == jhi)
            return scale * (3.0 * P(i, jhi) - 4.0 * P(i, jhi - 1) + P(i, jhi - 2) + 3.0 * P(i - 1, jhi) -
                            4.0 * P(i - 1, jhi - 1) + P(i - 1, jhi - 2));
        if (j == jlo)
            return -(scale * (3.0 * P(i, jlo) - 4.0 * P(i, jlo + 1) + P(i, jlo + 2) + 3.0 * P(i - 1, jlo) -
                              4.0 * P(i - 1, jlo + 1) + P(i - 1, jlo + 2)));
        return scale * (P(i, j + 1) - P(i, j - 1) + P(i - 1, j + 1) - P(i - 1, j - 1));
    }
    double p, pn, pw, ee, ew, en, es, ene, enw, ese, esw;
    if (i == ilo) {  // west face of cell ilo: the west column is extrapolated
        p = P(i, j);
        ee = P(i + 1, j);
        ew = 3.0 * (p - ee) + P(i + 2, j);
        if (j == jlo) {
            en = P(i, j + 1);
            es = 3.0 * (p - en) + P(i, j + 2);
            ene = P(i + 1, j + 1);
            enw = 3.0 * (en - ene) + P(i + 2, j + 1);
            ese = 3.0 * (ee - ene) + P(i + 1, j + 2);
            esw = 0.5 * ((3.0 * (es - ese) + P(i + 2, j - 1)) + (3.0 * (ew - enw) + P(i - 1, j + 2)));
        } else if (j == jhi) {
            es = P(i, j - 1);
            en = 3.0 * (p - es) + P(i, j - 2);
            ese = P(i + 1, j - 1);
            ene = 3.0 * (ee - ese) + P(i + 1, j - 2);
            esw = 3.0 * (es - ese) + P(i + 2, j - 1);
            enw = 0.5 * ((3.0 * (ew - esw) + P(i - 1, j - 2)) + (3.0 * (en - ene) + P(i + 2, j + 1)));
        } else {
            en = P(i, j + 1);
            es = P(i, j - 1);
            ene = P(i + 1, j + 1);
            ese = P(i + 1, j - 1);
            enw = 3.0 * (en - ene) + P(i + 2, j + 1);
            esw = 3.0 * (es - ese) + P(i + 2, j - 1);
        }
        return scale * (en - es + enw - esw);
    }
    // east face of cell ihi (face index nx): the east column is extrapolated
    const int c = ihi;
    p = P(c, j);
    ew = P(c - 1, j);
    ee = 3.0 * (p - ew) + P(c - 2, j);
    if (j == jlo) {
        pn = P(c, j + 1);
        pw = P(c - 1, j);
        en = pn;
        es = 3.0 * (p - en) + P(c, j + 2);
        enw = P(c - 1, j + 1);
        ene = 3.0 * (pn - enw) + P(c - 2, j + 1);
        esw = 3.0 * (pw - enw) + P(c - 1, j + 2);
        ese = 0.5 * ((3.0 * (es - esw) + P(c - 2, j - 1)) + (3.0 * (ee - ene) + P(c + 1, j + 2)));
    } else if (j == jhi) {
        es = P(c, j - 1);
        en = 3.0 * (p - es) + P(c, j - 2);
        esw = P(c - 1, j - 1);
        enw = 3.0 * (ew - esw) + P(c - 1, j - 2);
        ese = 3.0 * (es - esw) + P(c - 2, j - 1);
        ene = 0.5 * ((3.0 * (en - enw) + P(c - 2, j + 1)) + (3.0 * (ee - ese) + P(c + 1, j - 2)));
    } else {
        en = P(c, j + 1);
        es = P(c, j - 1);
        enw = P(c - 1, j + 1);
        esw = P(c - 1, j - 1);
        ene = 3.0 * (en - enw) + P(c - 2, j + 1);
        ese = 3.0 * (es - esw) + P(c - 2, j - 1);
    }
    return scale * (ene - ese + en - es);
}

// d/dxi at y-face (i, j), j = 0..ny (poisson2d.f90:450-605): the transpose.
__device__ double pd_xi_at_yface(const Phi &P, int nx, int ny, double scale, int i, int j)
{
    const int ilo = 0, ihi = nx - 1, jlo = 0, jhi = ny - 1;
    if (j > jlo && j <= jhi) {
        if (i == ihi)
            return scale * (3.0 * P(ihi, j) - 4.0 * P(ihi - 1, j) + P(ihi - 2, j) + 3.0 * P(ihi, j - 1) -
                            4.0 * P(ihi - 1, j - 1) + P(ihi - 2, j - 1));
        if (i == ilo)
            return -(scale * (3.0 * P(ilo, j) - 4.0 * P(ilo + 1, j) + P(ilo + 2, j) + 3.0 * P(ilo, j - 1) -
                              4.0 * P(ilo + 1, j - 1) + P(ilo + 2, j - 1)));
        return scale * (P(i + 1, j) - P(i - 1, j) + P(i + 1, j - 1) - P(i - 1, j - 1));
    }
    double p, pn, pw, ee, ew, en, es, ene, enw, ese, esw;
    if (j == jlo) {  // south face of row jlo: the south row is extrapolated
        p = P(i, j);
        en = P(i, j + 1);
        es = 3.0 * (p - en) + P(i, j + 2);
        if (i == ilo) {
            ee = P(i + 1, j);
            ew = 3.0 * (p - ee) + P(i + 2, j);
            ene = P(i + 1, j + 1);
            enw = 3.0 * (en - ene) + P(i + 2, j + 1);
            ese = 3.0 * (ee - ene) + P(i + 1, j + 2);
            esw = 0.5 * ((3.0 * (es - ese) + P(i + 2, j - 1)) + (3.0 * (ew - enw) + P(i - 1, j + 2)));
        } else if (i == ihi) {
            pn = P(i, j + 1);
            pw = P(i - 1, j);
            ew = pw;
            ee = 3.0 * (p - ew) + P(i - 2, j);
            enw = P(i - 1, j + 1);
            ene = 3.0 * (pn - enw) + P(i - 2, j + 1);
            esw = 3.0 * (pw - enw) + P(i - 1, j + 2);
            ese = 0.5 * ((3.0 * (es - esw) + P(i - 2, j - 1)) + (3.0 * (ee - ene) + P(i + 1, j + 2)));
        } else {
            ee = P(i + 1, j);
            ew = P(i - 1, j);
            ene = P(i + 1, j + 1);
            enw = P(i - 1, j + 1);
            ese = 3.0 * (ee - ene) + P(i + 1, j + 2);
            esw = 3.0 * (ew - enw) + P(i - 1, j + 2);
        }
        return scale * (ee - ew + ese - esw);
    }
    // north face of row jhi (face index ny): the north row is extrapolated
    const int r = jhi;
    p = P(i, r);
    es = P(i, r - 1);
    en = 3.0 * (p - es) + P(i, r - 2);
    if (i == ilo) {
        ee = P(i + 1, r);
        ew = 3.0 * (p - ee) + P(i + 2, r);
        ese = P(i + 1, r - 1);
        ene = 3.0 * (ee - ese) + P(i + 1, r - 2);
        esw = 3.0 * (es - ese) + P(i + 2, r - 1);
        enw = 0.5 * ((3.0 * (ew - esw) + P(i - 1, r - 2)) + (3.0 * (en - ene) + P(i + 2, r + 1)));
    } else if (i == ihi) {
        ew = P(i - 1, r);
        ee = 3.0 * (p - ew) + P(i - 2, r);
        esw = P(i - 1, r - 1);
        enw = 3.0 * (ew - esw) + P(i - 1, r - 2);
        ese = 3.0 * (es - esw) + P(i - 2, r - 1);
        ene = 0.5 * ((3.0 * (en - enw) + P(i - 2, r + 1)) + (3.0 * (ee - ese) + P(i + 1, r - 2)));
    } else {
        ee = P(i + 1, r);
        ew = P(i - 1, r);
        ese = P(i + 1, r - 1);
        esw = P(i - 1, r - 1);
        ene = 3.0 * (ee - ese) + P(i + 1, r - 2);
        enw = 3.0 * (ew - esw) + P(i - 1, r - 2);
    }
    return scale * (ene - enw + ee - ew);
}

// Neumann face value of fill_boundary_fluxes: zero when homogeneous or LAH.
__device__ double neumann_flux(const Bc &bc, int side, int k, int homog)
{
    if (homog) return 0.0;
    if (bc.mode[side] == MODE_UNIFORM) return bc.val[side];
    if (bc.mode[side] == MODE_NONUNIFORM) return bc.data[side][k];
    return 0.0;
}

constexpr int FLUX_TX = 64, FLUX_TY = 4;

// xflux = Jxx dphi/dxi + Jxy dphi/deta on x-faces, yflux = Jyx dphi/dxi +
// Jyy dphi/deta on y-faces, then Neumann faces take their BC flux.  phi's
// ghosts must already hold fill_ghosts(do_neum = .false.).  SUB: the fluxes
// are subtracted from xflux/yflux instead (F - J grad phi, the projection's
// correction, test.f90:2990-2993).
template <bool SUB>
__global__ void __launch_bounds__(FLUX_TX * FLUX_TY)
k_grad(double *__restrict__ xflux, double *__restrict__ yflux, const double *__restrict__ phi, Grid g, Geo geo,
       Bc bc, int homog)
{
    const int i = blockIdx.x * FLUX_TX + threadIdx.x;
    const int j = blockIdx.y * FLUX_TY + threadIdx.y;
    if (i > g.nx || j > g.ny) return;
    const Phi P{phi, g.pitch};
    const size_t k = ix(g.pitch, i, j);
    if (j < g.ny) {
        double f;
        const bool west = i == 0 && bc.type[XLO] == BC_NEUM, east = i == g.nx && bc.type[XHI] == BC_NEUM;
        if (west || east) {
            f = neumann_flux(bc, west ? XLO : XHI, j, homog);
        } else {
            const double across = (1.0 / g.dx) * (P(i, j) - P(i - 1, j));
            const double along = pd_eta_at_xface(P, g.nx, g.ny, 0.25 / g.dy, i, j);
            f = geo.Jxx[k] * across + geo.Jxy[k] * along;
        }
        xflux[k] = SUB ? xflux[k] - f : f;
    }
    if (i < g.nx) {
        double f;
        const bool south = j == 0 && bc.type[YLO] == BC_NEUM, north = j == g.ny && bc.type[YHI] == BC_NEUM;
        if (south || north) {
            f = neumann_flux(bc, south ? YLO : YHI, i, homog);
        } else {
            const double along = pd_xi_at_yface(P, g.nx, g.ny, 0.25 / g.dx, i, j);
            const double across = (1.0 / g.dy) * (P(i, j) - P(i, j - 1));
            f = geo.Jyx[k] * along + geo.Jyy[k] * across;
        }
        yflux[k] = SUB ? yflux[k] - f : f;
    }
}

// One compute_pd field: nodedir 1 (x-faces) or 2 (y-faces), derivative dir.
__global__ void __launch_bounds__(FLUX_TX * FLUX_TY)
k_pd(double *__restrict__ pd, const double *__restrict__ phi, Grid g, int dir, int nodedir)
{
    const int i = blockIdx.x * FLUX_TX + threadIdx.x;
    const int j = blockIdx.y * FLUX_TY + threadIdx.y;
    const Phi P{phi, g.pitch};
    double v;
    if (nodedir == 1) {
        if (i > g.nx || j >= g.ny) return;
        v = dir == 1 ? (1.0 / g.dx) * (P(i, j) - P(i - 1, j)) : pd_eta_at_xface(P, g.nx, g.ny, 0.25 / g.dy, i, j);
    } else {
        if (i >= g.nx || j > g.ny) return;
        v = dir == 2 ? (1.0 / g.dy) * (P(i, j) - P(i, j - 1)) : pd_xi_at_yface(P, g.nx, g.ny, 0.25 / g.dx, i, j);
    }
    pd[ix(g.pitch, i, j)] = v;
}

// J div F over valid cells.
__global__ void __launch_bounds__(FLUX_TX * FLUX_TY)
k_div(double *__restrict__ div, const double *__restrict__ xflux, const double *__restrict__ yflux, Grid g)
{
    const int i = blockIdx.x * FLUX_TX + threadIdx.x;
    const int j = blockIdx.y * FLUX_TY + threadIdx.y;
    if (i >= g.nx || j >= g.ny) return;
    const size_t k = ix(g.pitch, i, j);
    const double invdx = 1.0 / g.dx, invdy = 1.0 / g.dy;
    div[k] = invdx * (xflux[k + 1] - xflux[k]) + invdy * (yflux[k + g.pitch] - yflux[k]);
}

// fill_boundary_fluxes alone: thread k of side s sets face k of that side.
__global__ void k_boundary_fluxes(double *__restrict__ xflux, double *__restrict__ yflux, Grid g, Bc bc, int homog)
{
    const int side = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (bc.type[side] != BC_NEUM || k >= (side < 2 ? g.ny : g.nx)) return;
    const double v = neumann_flux(bc, side, k, homog);
    if (side == XLO) xflux[ix(g.pitch, 0, k)] = v;
    else if (side == XHI) xflux[ix(g.pitch, g.nx, k)] = v;
    else if (side == YLO) yflux[ix(g.pitch, k, 0)] = v;
    else yflux[ix(g.pitch, k, g.ny)] = v;
}

// integrate2d_bdry (arrayutils.f90:721): the net outward flux, summed in the
// reference's order (x faces row by row, then y faces column by column) by
// one thread, so the sum is the reference's bit for bit.
__global__ void k_integrate_bdry(double *__restrict__ out, const double *__restrict__ xflux,
                                 const double *__restrict__ yflux, Grid g)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double sum = 0.0;
    for (int j = 0; j < g.ny; ++j) sum = sum + g.dx * (xflux[ix(g.pitch, g.nx, j)] - xflux[ix(g.pitch, 0, j)]);
    for (int i = 0; i < g.nx; ++i) sum = sum + g.dy * (yflux[ix(g.pitch, i, g.ny)] - yflux[ix(g.pitch, i, 0)]);
    *out = sum;
}

// a**p for integer p >= 1 by squaring from the lowest bit (the order of the
// Fortran runtime's integer power).
__device__ __forceinline__ double ipow(double a, int p)
{
    double r = 1.0;
    while (true) {
        if (p & 1) r *= a;
        p >>= 1;
        if (!p) break;
        a *= a;
    }
    return r;
}

// gpnorm (arrayutils.f90:802) before the final root: over columns cx0, cx1
// (rows 0 .. ny-1) and rows ry0, ry1 (columns 0 .. nx-1) of phi, in the
// reference's order, one thread.  p == 0: max |phi|; else sum of w phi^p with
// w = dy on the columns and dx on the rows.
__global__ void k_gpnorm_sum(double *__restrict__ out, const double *__restrict__ phi, Grid g, int p, int cx0,
                             int cx1, int ry0, int ry1)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double pn = 0.0;
    const int cols[2] = {cx0, cx1}, rows[2] = {ry0, ry1};
    if (p == 0) {
        for (int s = 0; s < 2; ++s)
            for (int j = 0; j < g.ny; ++j) pn = fmax(pn, fabs(phi[ix(g.pitch, cols[s], j)]));
        for (int s = 0; s < 2; ++s)
            for (int i = 0; i < g.nx; ++i) pn = fmax(pn, fabs(phi[ix(g.pitch, i, rows[s])]));
    } else {
        for (int s = 0; s < 2; ++s)
            for (int j = 0; j < g.ny; ++j) pn = pn + g.dy * ipow(phi[ix(g.pitch, cols[s], j)], p);
        for (int s = 0; s < 2; ++s)
            for (int i = 0; i < g.nx; ++i) pn = pn + g.dx * ipow(phi[ix(g.pitch, i, rows[s])], p);
    }
    *out = pn;
}

dim3 face_grid(int w, int h) { return dim3((w + FLUX_TX - 1) / FLUX_TX, (h + FLUX_TY - 1) / FLUX_TY); }

}  // namespace

void launch_grad(hipStream_t s, double *xflux, double *yflux, const double *phi, const Grid &g, const Geo &geo,
                 const Bc &bc, int homog)
{
    hipLaunchKernelGGL(k_grad<false>, face_grid(g.nx + 1, g.ny + 1), dim3(FLUX_TX, FLUX_TY), 0, s, xflux, yflux,
                       phi, g, geo, bc, homog);
}

void launch_grad_subtract(hipStream_t s, double *xflux, double *yflux, const double *phi, const Grid &g,
                          const Geo &geo, const Bc &bc, int homog)
{
    hipLaunchKernelGGL(k_grad<true>, face_grid(g.nx + 1, g.ny + 1), dim3(FLUX_TX, FLUX_TY), 0, s, xflux, yflux,
                       phi, g, geo, bc, homog);
}

void launch_pd(hipStream_t s, double *pd, const double *phi, const Grid &g, int dir, int nodedir)
{
    hipLaunchKernelGGL(k_pd, face_grid(g.nx + 1, g.ny + 1), dim3(FLUX_TX, FLUX_TY), 0, s, pd, phi, g, dir, nodedir);
}

void launch_div(hipStream_t s, double *div, const double *xflux, const double *yflux, const Grid &g)
{
    hipLaunchKernelGGL(k_div, face_grid(g.nx, g.ny), dim3(FLUX_TX, FLUX_TY), 0, s, div, xflux, yflux, g);
}

void launch_boundary_fluxes(hipStream_t s, double *xflux, double *yflux, const Grid &g, const Bc &bc, int homog)
{
    const int n = g.nx > g.ny ? g.nx : g.ny;
    hipLaunchKernelGGL(k_boundary_fluxes, dim3((n + 255) / 256, 4), dim3(256), 0, s, xflux, yflux, g, bc, homog);
}

void launch_integrate_bdry(hipStream_t s, double *out, const double *xflux, const double *yflux, const Grid &g)
{
    hipLaunchKernelGGL(k_integrate_bdry, dim3(1), dim3(64), 0, s, out, xflux, yflux, g);
}

void launch_gpnorm_sum(hipStream_t s, double *out, const double *phi, const Grid &g, int p, int cx0, int cx1, int ry0,
                       int ry1)
{
    hipLaunchKernelGGL(k_gpnorm_sum, dim3(1), dim3(64), 0, s, out, phi, g, p, cx0, cx1, ry0, ry1);
}

}  // namespace mgp
