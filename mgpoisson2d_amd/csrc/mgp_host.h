// mgp_host.h -- host-side objects of the library (device buffers, level
// hierarchies, cycle drivers).  Internal; the public interface is mgp2d.h.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "mgp2d.h"
#include "mgp_common.h"

namespace mgp {

extern thread_local std::string g_last_error;

struct Failure {
    int status;
    explicit Failure(int s) : status(s) {}
};

void hip_check(hipError_t e, const char *what);

// Zero-filled device allocation of `bytes` (throws MGP_ERR_NOMEM); dev_free
// releases it.
void *dev_alloc_zero(size_t bytes);
void dev_free(void *p);

// Owning device allocation of `count` elements of T, zero-filled on allocation.
template <typename T>
struct DevArr {
    T *p = nullptr;
    size_t n = 0;
    DevArr() = default;
    DevArr(const DevArr &) = delete;
    DevArr &operator=(const DevArr &) = delete;
    DevArr(DevArr &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevArr &operator=(DevArr &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevArr() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        p = static_cast<T *>(dev_alloc_zero(count * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) dev_free(p);
        p = nullptr;
        n = 0;
    }
};
using DevBuf = DevArr<double>;
using DevBufF = DevArr<float>;

Grid make_grid(int nx, int ny, double dx, double dy);
void upload_block(hipStream_t s, double *dev, const Grid &g, const double *host, int i0, int j0, int w, int h);
void download_block(hipStream_t s, double *host, const double *dev, const Grid &g, int i0, int j0, int w, int h);

struct DevGeo {
    Grid g;
    DevBuf J, Jxx, Jxy, Jyx, Jyy;
    void upload(hipStream_t s, const mgp_geo &h);
    // metrics of a mapping generated on the device (mgp_geometry.hip)
    void generate(hipStream_t s, const mgp_mapping &m, int nx, int ny, double dx, double dy);
    Geo view() const;
};

struct DevBc {
    Bc bc;
    DevBuf data[4];
    void upload(hipStream_t s, const mgp_bc &h, int nx, int ny);
};

Bc mg_bc(const Bc &user, int mode);

int reference_schedule(int nx, int ny, double dx, double dy, int maxdepth, std::vector<int> &refx,
                       std::vector<int> &refy);

// Scratch for reductions (device partials + pinned host scalar).
struct Workspace {
    DevBuf partial;
    DevBuf ones_buf;
    int ones_pitch = -1, ones_nx = -1, ones_ny = -1;
    double *host_scalar = nullptr;
    Workspace() = default;
    Workspace(const Workspace &) = delete;
    ~Workspace();
    double reduce(hipStream_t s, const double *a, const double *b, const double *J, double vol, const Grid &g,
                  int kind, int p = 1);
    const double *ones(hipStream_t s, const Grid &g);
};

double pnorm2(Workspace &W, hipStream_t s, const double *a, const Grid &g);
void remove_average(Workspace &W, hipStream_t s, double *fine, const Grid &fg, const double *fineJ);

struct HLevel {
    Grid g;
    Geo geo;
    Bc bc;  // homogeneous MG BCs of the level
    DevBuf J, Jxx, Jxy, Jyx, Jyy, e, r, w, invd, mod;
};

void prolong_level(Workspace &W, hipStream_t s, HLevel &f, HLevel &c, int order, DevBuf *tmp);

// The reference's level hierarchy (isotropic schedule) and cycle drivers.
struct RefHierarchy {
    int dc = 0, ordering = 0, maxdepth = 0;
    int fmg = 0;  // outer loop of MGPoisson2D::fmg (MGPoisson2D.f90:860) instead of vcycle
    std::vector<int> refx, refy;
    std::vector<HLevel> levels;
    DevBuf bt_pt, bt_st, bt_r, bt_rt, bt_p, bt_t, bt_v, bt_e, bt_mod, bt_scal;

    void build(hipStream_t s, const DevGeo &top, const Bc &user_bc, int dc, int ordering, int maxdepth_user);
    void relax(hipStream_t s, int d, double *phi, const double *rhs, int sweeps, double omega);
    // Measurement of the finest-level DC red-black smoother (2 sweeps):
    // kernel 4 the relax with LDS tiles, 5 with one launch per colour, 6 the
    // tile kernel alone.  Mean ms per call over reps; *bytes the algorithmic
    // bytes of one call (7 fields read, 1 written).
    double time_relax(hipStream_t s, int kernel, int reps, double *bytes);
    void bottom(hipStream_t s, int d);
    void vcycle_noinit(Workspace &W, hipStream_t s, int d, int numcycles, int sd, int su, int sb);
    void fmg_noinit(Workspace &W, hipStream_t s, int d, int numcycles, int sd, int su, int sb);
    int solve(Workspace &W, hipStream_t s, const DevGeo &top, const Bc &user_bc, int homog, double *phi,
              double *rhs, double tol, int maxiters, int numcycles, int sd, int su, int sb, int zerophi,
              double *relres, int *iters_out);
};

}  // namespace mgp
