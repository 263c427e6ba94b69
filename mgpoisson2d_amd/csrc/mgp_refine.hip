// mgp_refine.hip -- kernels of the resident solver's refinement loop that are
// not part of a V-cycle:
//
//  * the double-double residual r = f - L (phi_hi + phi_lo) (mgp_dd.h) and the
//    fold (phi_hi, phi_lo) += x of a correction into the double-double
//    solution, so that a solve can reach a relative residual below the
//    rounding of the FP64 operator;
//  * the J-weighted zero-mean projection of an all-Neumann rhs (compensated
//    sums, then rhs -= J sum(rhs) / sum(J), which makes sum(rhs) vanish: the
//    compatibility condition of the singular problem);
//  * transposes between the caller's layout and the transposed level layout of
//    a y-line hierarchy (the preconditioner of grids whose rows are too long
//    for one workgroup's line solve relaxes columns: it runs on the transposed
//    problem, whose rows are the columns).
#include <hip/hip_runtime.h>

#include "mgp_dd.h"
#include "mgp_kernels.h"
#include "mgp_reduce.h"

namespace mgp {

namespace {

inline dim3 tiles64x4(int nx, int ny) { return dim3((unsigned)((nx + 63) / 64), (unsigned)((ny + 3) / 4)); }

__global__ void __launch_bounds__(256) k_residual_dd(double *__restrict__ out, const double *__restrict__ hi,
                                                     const double *__restrict__ lo, const double *__restrict__ rhs,
                                                     Grid g, Geo geo, Bc bc) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= g.nx || j >= g.ny) return;
    DDGet get;
    get.hi = hi;
    get.lo = lo;
    get.pitch = g.pitch;
    get.nx = g.nx;
    get.ny = g.ny;
    for (int k = 0; k < 4; ++k) get.t[k] = bc.type[k];
    const size_t c = ix(g.pitch, i, j);
    const DD r = cell_lap_dd(get, g, geo, bc, i, j);
    // rhs - L phi, rounded once
    const DD d = (-r) + rhs[c];
    out[c] = d.hi + d.lo;
}

__global__ void __launch_bounds__(256) k_fold_dd(double *__restrict__ hi, double *__restrict__ lo,
                                                 double *__restrict__ x, Grid g) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= g.nx || j >= g.ny) return;
    const size_t c = ix(g.pitch, i, j);
    const DD s = DD{hi[c], lo[c]} + x[c];
    hi[c] = s.hi;
    lo[c] = s.lo;
    x[c] = 0.0;
}

// rhs -= J * sums[0] / sums[1] over the valid cells.
__global__ void __launch_bounds__(256) k_sub_jmean(double *__restrict__ rhs, const double *__restrict__ J,
                                                   const double *__restrict__ sums, Grid g) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= g.nx || j >= g.ny) return;
    const size_t c = ix(g.pitch, i, j);
    rhs[c] = rhs[c] - J[c] * (sums[0] / sums[1]);
}

// dst(j, i) = src(i, j) for i < ni, j < nj (index ranges from 0), through a
// 32 x 32 LDS tile: both the reads and the writes are coalesced.
template <typename TI, typename TO>
__global__ void __launch_bounds__(256) k_transpose(TO *__restrict__ dst, int dpitch, const TI *__restrict__ src,
                                                   int spitch, int ni, int nj) {
    __shared__ TI tile[32][33];
    const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int i = i0 + tx, j = j0 + ty + k;
        if (i < ni && j < nj) tile[ty + k][tx] = src[ix(spitch, i, j)];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int j = j0 + tx, i = i0 + ty + k;
        if (i < ni && j < nj) dst[ix(dpitch, j, i)] = (TO)tile[tx][ty + k];
    }
}

}  // namespace

void launch_residual_dd(hipStream_t s, double *out, const double *hi, const double *lo, const double *rhs,
                        const Grid &g, const Geo &geo, const Bc &bc) {
    k_residual_dd<<<tiles64x4(g.nx, g.ny), 256, 0, s>>>(out, hi, lo, rhs, g, geo, bc);
}

void launch_fold_dd(hipStream_t s, double *hi, double *lo, double *x, const Grid &g) {
    k_fold_dd<<<tiles64x4(g.nx, g.ny), 256, 0, s>>>(hi, lo, x, g);
}

void launch_project_jmean(hipStream_t s, double *rhs, const double *J, double *sums, double *partial,
                          const Grid &g) {
    launch_reduce(s, sums, partial, rhs, nullptr, nullptr, 1.0, g, 3);
    launch_reduce(s, sums + 1, partial, J, nullptr, nullptr, 1.0, g, 3);
    k_sub_jmean<<<tiles64x4(g.nx, g.ny), 256, 0, s>>>(rhs, J, sums, g);
}

template <typename TI, typename TO>
static void transpose_impl(hipStream_t s, TO *dst, int dpitch, const TI *src, int spitch, int ni, int nj) {
    const dim3 grid((unsigned)((ni + 31) / 32), (unsigned)((nj + 31) / 32));
    k_transpose<TI, TO><<<grid, 256, 0, s>>>(dst, dpitch, src, spitch, ni, nj);
}
void launch_transpose(hipStream_t s, double *dst, int dpitch, const double *src, int spitch, int ni, int nj) {
    transpose_impl(s, dst, dpitch, src, spitch, ni, nj);
}
void launch_transpose(hipStream_t s, float *dst, int dpitch, const double *src, int spitch, int ni, int nj) {
    transpose_impl(s, dst, dpitch, src, spitch, ni, nj);
}
void launch_transpose(hipStream_t s, double *dst, int dpitch, const float *src, int spitch, int ni, int nj) {
    transpose_impl(s, dst, dpitch, src, spitch, ni, nj);
}

}  // namespace mgp
