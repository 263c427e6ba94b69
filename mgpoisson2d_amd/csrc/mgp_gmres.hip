// mgp_gmres.hip -- vector kernels of the flexible GMRES outer iteration of
// the resident solver (no counterpart in the reference): block
// Gram-Schmidt against the stored Krylov basis.
//
// The basis V_0 .. V_m sits in one allocation, vector i at V + i * stride, in
// the padded level layout; only the valid cells are read or written.  One
// orthogonalisation pass of a new vector w against V_0 .. V_{nv-1} is
//   h = V^T w  (k_mdot_partial + k_mdot_final: w and each V_i read once per
//               group of 8 vectors, fixed-order sums, so the result is
//               deterministic),
//   w -= V h   (k_maxpy: each V_i read once, w read and written once),
// and the coefficients stay on the device between the two kernels.  Each
// kernel is bound by HBM traffic: about (nv + nv / 8) vectors for the dots and
// nv + 2 for the update.
#include <hip/hip_runtime.h>

#include "mgp_kernels.h"

namespace mgp {

namespace {

constexpr int MD_NV = 8;        // vectors per workgroup of the dot kernel
constexpr int MD_THREADS = 256;

inline dim3 tiles64x4(int nx, int ny) { return dim3((unsigned)((nx + 63) / 64), (unsigned)((ny + 3) / 4)); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// partial[i * nb + b] = sum over block b's rows of V_i . w, i = v0 .. v0 + 7 (< nv).
__global__ void __launch_bounds__(MD_THREADS) k_mdot_partial(double *__restrict__ partial,
                                                             const double *__restrict__ V, size_t stride,
                                                             const double *__restrict__ w, int nv, Grid g) {
    const int v0 = blockIdx.y * MD_NV, nb = gridDim.x;
    double acc[MD_NV];
#pragma unroll
    for (int q = 0; q < MD_NV; ++q) acc[q] = 0.0;
    const int nq = nv - v0 < MD_NV ? nv - v0 : MD_NV;
    for (int j = blockIdx.x; j < g.ny; j += nb) {
        const size_t row = ix(g.pitch, 0, j);
        for (int i = threadIdx.x; i < g.nx; i += MD_THREADS) {
            const size_t c = row + i;
            const double wv = w[c];
#pragma unroll
            for (int q = 0; q < MD_NV; ++q)
                if (q < nq) acc[q] = fma(V[(size_t)(v0 + q) * stride + c], wv, acc[q]);
        }
    }
    __shared__ double red[MD_NV][MD_THREADS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < MD_NV; ++q) {
        const double v = wave_sum(acc[q]);
        if (lane == 0) red[q][wave] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < nq) {
        double s = 0.0;
        for (int k = 0; k < MD_THREADS / 64; ++k) s += red[threadIdx.x][k];
        partial[(size_t)(v0 + threadIdx.x) * nb + blockIdx.x] = s;
    }
}

// out[i] = sum_b partial[i * nb + b], one workgroup per vector.
__global__ void __launch_bounds__(256) k_mdot_final(double *__restrict__ out, const double *__restrict__ partial,
                                                    int nb) {
    const int i = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) s += partial[(size_t)i * nb + b];
    __shared__ double red[4];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[i] = ((red[0] + red[1]) + red[2]) + red[3];
}

// y += sign * sum_{i < nv} h[i] V_i over the valid cells.
__global__ void __launch_bounds__(256) k_maxpy(double *__restrict__ y, const double *__restrict__ V, size_t stride,
                                               const double *__restrict__ h, int nv, double sign, Grid g) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= g.nx || j >= g.ny) return;
    const size_t c = ix(g.pitch, i, j);
    double s = 0.0;
    for (int k = 0; k < nv; ++k) s = fma(h[k], V[(size_t)k * stride + c], s);
    y[c] = fma(sign, s, y[c]);
}

// v = src / sqrt(*nrm2) over the valid cells (src may be v).
__global__ void __launch_bounds__(256) k_scale_rsqrt(double *__restrict__ v, const double *src,
                                                     const double *__restrict__ nrm2, Grid g) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= g.nx || j >= g.ny) return;
    const size_t c = ix(g.pitch, i, j);
    v[c] = src[c] / sqrt(*nrm2);
}

}  // namespace

int mdot_blocks(const Grid &g) { return g.ny < 1024 ? g.ny : 1024; }

void launch_mdot(hipStream_t s, double *out, double *partial, const double *V, size_t stride, const double *w, int nv,
                 const Grid &g) {
    if (nv <= 0) return;
    const int nb = mdot_blocks(g);
    const dim3 grid((unsigned)nb, (unsigned)((nv + MD_NV - 1) / MD_NV));
    k_mdot_partial<<<grid, MD_THREADS, 0, s>>>(partial, V, stride, w, nv, g);
    k_mdot_final<<<nv, 256, 0, s>>>(out, partial, nb);
}

void launch_maxpy(hipStream_t s, double *y, const double *V, size_t stride, const double *h, int nv, double sign,
                  const Grid &g) {
    if (nv <= 0) return;
    k_maxpy<<<tiles64x4(g.nx, g.ny), 256, 0, s>>>(y, V, stride, h, nv, sign, g);
}

void launch_scale_rsqrt(hipStream_t s, double *v, const double *src, const double *nrm2, const Grid &g) {
    k_scale_rsqrt<<<tiles64x4(g.nx, g.ny), 256, 0, s>>>(v, src, nrm2, g);
}

}  // namespace mgp
