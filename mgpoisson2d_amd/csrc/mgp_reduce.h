// mgp_reduce.h -- device pieces of the deterministic FP64 reductions: a
// compensated (Neumaier) accumulator and a fixed-order tree combine of the
// per-thread accumulators of one workgroup.
#pragma once

#include <hip/hip_runtime.h>

namespace mgp {

struct Acc {
    double s, c;
    __device__ __forceinline__ void add(double x) {
        const double t = s + x;
        c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ double val() const { return s + c; }
};

// Fixed-order tree combine of per-thread (s, c) pairs in LDS; thread 0 ends
// with the block's compensated sum (or max for kind 2).
__device__ __forceinline__ double block_combine(double *ss, double *sc, Acc acc, double mx, int kind) {
    const int t = threadIdx.x;
    ss[t] = kind == 2 ? mx : acc.s;
    sc[t] = kind == 2 ? 0.0 : acc.c;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if (t < off) {
            if (kind == 2) {
                ss[t] = fmax(ss[t], ss[t + off]);
            } else {
                Acc u{ss[t], sc[t]};
                u.add(ss[t + off]);
                u.c += sc[t + off];
                ss[t] = u.s;
                sc[t] = u.c;
            }
        }
        __syncthreads();
    }
    return kind == 2 ? ss[0] : ss[0] + sc[0];
}

}  // namespace mgp
