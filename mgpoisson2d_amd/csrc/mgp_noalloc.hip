// mgp_noalloc.hip -- mgp_pnorm_noalloc: the norm of mgp_pnorm on a field the
// caller already holds in device memory, with no device allocation.
//
// The reduction's per-block partial sums live in a file-scope device array,
// `scratch`, of NOALLOC_MAX_BLOCKS doubles (8 KiB per GPU, held until the
// process ends).  Host code never passes the array's name to a kernel: on
// the first call on each device, with that device current,
// hipGetSymbolAddress gives the array's device address there, which is kept
// per device and passed to the kernels.  The kernels are mgp_pnorm's own
// (k_reduce_partial over the same row blocks, then k_reduce_final), so the
// sum is the one mgp_pnorm forms.  Every call on a GPU shares that GPU's
// array; each call writes every partial it reads, so nothing needs clearing,
// but calls that overlap on one GPU (different streams, or threads) race.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "mgp2d.h"
#include "mgp_host.h"
#include "mgp_kernels.h"

namespace {

constexpr int NOALLOC_MAX_BLOCKS = 1024;
constexpr int NOALLOC_MAX_DEVICES = 64;

__device__ __align__(256) double scratch[NOALLOC_MAX_BLOCKS];

std::mutex scratch_mu;
double *scratch_ptr[NOALLOC_MAX_DEVICES];  // device address of `scratch` per device (null: not looked up)

__global__ void k_norm_root(double *out, int p) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && p == 2) *out = sqrt(*out);
}

double *scratch_on_current_device() {
    int dev = -1;
    mgp::hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        mgp::g_last_error = "pnorm_noalloc: device index out of range";
        throw mgp::Failure(MGP_ERR_UNSUPPORTED);
    }
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        mgp::hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        scratch_ptr[dev] = static_cast<double *>(p);
    }
    return scratch_ptr[dev];
}

}  // namespace

extern "C" int mgp_pnorm_noalloc(int nx, int ny, double dx, double dy, const double *a, int pitch, int p,
                                 double *out, void *stream) {
    try {
        if (nx <= 0 || ny <= 0 || !(dx > 0) || !(dy > 0) || !a || !out || p < 0 || p > 2 ||
            pitch < nx + mgp::OFFX + 2) {
            mgp::g_last_error = "pnorm_noalloc: need nx, ny > 0, positive spacing, p in 0..2, "
                                "pitch >= nx + 34 and device pointers";
            return MGP_ERR_ARG;
        }
        const mgp::Grid g{nx, ny, pitch, dx, dy};
        if (mgp::reduce_blocks(g) > NOALLOC_MAX_BLOCKS) {
            mgp::g_last_error = "pnorm_noalloc: more row blocks than the scratch array holds";
            return MGP_ERR_ARG;
        }
        double *partial = scratch_on_current_device();
        hipStream_t s = static_cast<hipStream_t>(stream);
        // mgp_pnorm's reductions: max |a| (p = 0), sum w |a| (p = 1), sum w a^2 then the root (p = 2)
        if (p == 0) mgp::launch_reduce(s, out, partial, a, nullptr, nullptr, 1.0, g, 2);
        else if (p == 1) mgp::launch_reduce(s, out, partial, a, nullptr, nullptr, dx * dy, g, 4, 1);
        else mgp::launch_reduce(s, out, partial, a, nullptr, nullptr, dx * dy, g, 1);
        if (p == 2) hipLaunchKernelGGL(k_norm_root, dim3(1), dim3(64), 0, s, out, p);
        mgp::hip_check(hipGetLastError(), "pnorm_noalloc launch");
        return MGP_OK;
    } catch (const mgp::Failure &f) {
        return f.status;
    } catch (...) {
        mgp::g_last_error = "unexpected host exception";
        return MGP_ERR_HIP;
    }
}
