// mgp_host.hip -- host side of the library: device level hierarchies, the
// reference cycle drivers (MGPoisson2D::vcycle, DeferredMGPoisson2D::vcycle)
// on resident HBM data, and the extern "C" entry points of include/mgp2d.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mgp2d.h"
#include "mgp_host.h"
#include "mgp_geometry.h"
#include "mgp_kernels.h"

namespace mgp {

thread_local std::string g_last_error;

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        g_last_error = std::string(what) + ": " + hipGetErrorString(e);
        throw Failure(MGP_ERR_HIP);
    }
}

// --------------------------------------------------------------------------
// Device buffers
// --------------------------------------------------------------------------
void *dev_alloc_zero(size_t bytes) {
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        g_last_error = "hipMalloc failed";
        throw Failure(MGP_ERR_NOMEM);
    }
    // zero-fill, then wait: the library's streams are non-blocking, so work
    // they enqueue next is not ordered after the null stream otherwise
    if (hipMemset(p, 0, bytes) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        (void)hipFree(p);
        g_last_error = "hipMemset of a new allocation failed";
        throw Failure(MGP_ERR_HIP);
    }
    return p;
}
void dev_free(void *p) { (void)hipFree(p); }

Grid make_grid(int nx, int ny, double dx, double dy) {
    Grid g;
    g.nx = nx;
    g.ny = ny;
    g.pitch = level_pitch(nx);
    g.dx = dx;
    g.dy = dy;
    return g;
}

// Host <-> padded device copies.  (i0, j0) is the first element of the host
// block, w x h its extent.
//
// The copies go through a pinned staging buffer.  The runtime moves a 2D
// region of pageable host memory one row at a time (setting the rhs of the
// 1024 problems of the batched mix took 2.8 s, 2.8 ms for a problem of 400
// rows); from pinned memory the region is one DMA.  One buffer per device,
// grown as needed, under a mutex; an upload returns once its DMA is queued,
// and the next use of the buffer first waits for that DMA (an event on its
// stream).  A download waits for its DMA, then copies out.  A region larger
// than the buffer's 64 MiB moves in blocks of rows, one after the other.  The
// direct copy remains for a stream of another device than the current one,
// for a copy on one stream while the buffer's last DMA is still queued on
// another (so that streams do not wait on each other's work through the
// buffer), and where no pinned memory can be had.
namespace {
struct Staging {
    std::mutex mu;
    void *p = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    hipStream_t last = nullptr;  // the stream of the pending DMA
    bool pending = false;
};
Staging *staging_for(hipStream_t s) {
    int cur = 0, sd = 0;
    if (hipGetDevice(&cur) != hipSuccess || hipStreamGetDevice(s, &sd) != hipSuccess || sd != cur) {
        (void)hipGetLastError();
        return nullptr;
    }
    static std::mutex m;
    static std::map<int, std::unique_ptr<Staging>> all;
    std::lock_guard<std::mutex> lk(m);
    std::unique_ptr<Staging> &x = all[cur];
    if (!x) x = std::make_unique<Staging>();
    return x.get();
}
constexpr size_t STAGE_MAX = (size_t)64 << 20;

// The buffer, at least `bytes` long, once its previous copy has completed;
// nullptr if that copy is still queued on another stream, or if the pinned
// allocation fails.
void *stage_buffer(Staging &st, size_t bytes, hipStream_t s) {
    if (st.pending) {
        if (st.last != s) {
            const hipError_t q = hipEventQuery(st.ev);
            if (q == hipErrorNotReady) {
                (void)hipGetLastError();
                return nullptr;
            }
            hip_check(q, "staging query");
        }
        hip_check(hipEventSynchronize(st.ev), "staging wait");
        st.pending = false;
    }
    if (!st.ev) hip_check(hipEventCreateWithFlags(&st.ev, hipEventDisableTiming), "hipEventCreate");
    if (bytes > st.cap) {
        if (st.p) (void)hipHostFree(st.p);
        st.p = nullptr;
        st.cap = 0;
        if (hipHostMalloc(&st.p, bytes, hipHostMallocPortable) != hipSuccess || !st.p) {
            (void)hipGetLastError();
            st.p = nullptr;
            return nullptr;
        }
        st.cap = bytes;
    }
    return st.p;
}
}  // namespace

void upload_block(hipStream_t s, double *dev, const Grid &g, const double *host, int i0, int j0, int w, int h) {
    const size_t row = (size_t)w * sizeof(double), bytes = row * (size_t)h;
    double *d = dev + ix(g.pitch, i0, j0);
    const size_t dpitch = (size_t)g.pitch * sizeof(double);
    Staging *st = bytes && row <= STAGE_MAX ? staging_for(s) : nullptr;
    std::unique_lock<std::mutex> lk;
    const size_t rows = row ? STAGE_MAX / row : 0;  // rows per block
    void *p = nullptr;
    if (st) {
        lk = std::unique_lock<std::mutex>(st->mu);
        p = stage_buffer(*st, std::min(bytes, rows * row), s);
    }
    if (!p) {
        hip_check(hipMemcpy2DAsync(d, dpitch, host, row, row, (size_t)h, hipMemcpyHostToDevice, s), "upload");
        return;
    }
    for (size_t j = 0; j < (size_t)h; j += rows) {
        const size_t nr = std::min(rows, (size_t)h - j);
        if (j) p = stage_buffer(*st, nr * row, s);  // the previous block's DMA done (same stream)
        if (!p) {
            hip_check(hipMemcpy2DAsync(reinterpret_cast<char *>(d) + j * dpitch, dpitch,
                                       reinterpret_cast<const char *>(host) + j * row, row, row, (size_t)h - j,
                                       hipMemcpyHostToDevice, s),
                      "upload");
            return;
        }
        memcpy(p, reinterpret_cast<const char *>(host) + j * row, nr * row);
        hip_check(hipMemcpy2DAsync(reinterpret_cast<char *>(d) + j * dpitch, dpitch, p, row, row, nr,
                                   hipMemcpyHostToDevice, s),
                  "upload");
        hip_check(hipEventRecord(st->ev, s), "staging event");
        st->last = s;
        st->pending = true;
    }
}
void download_block(hipStream_t s, double *host, const double *dev, const Grid &g, int i0, int j0, int w, int h) {
    const size_t row = (size_t)w * sizeof(double), bytes = row * (size_t)h;
    const double *d = dev + ix(g.pitch, i0, j0);
    const size_t dpitch = (size_t)g.pitch * sizeof(double);
    Staging *st = bytes && row <= STAGE_MAX ? staging_for(s) : nullptr;
    std::unique_lock<std::mutex> lk;
    const size_t rows = row ? STAGE_MAX / row : 0;
    void *p = nullptr;
    if (st) {
        lk = std::unique_lock<std::mutex>(st->mu);
        p = stage_buffer(*st, std::min(bytes, rows * row), s);
    }
    if (!p) {
        hip_check(hipMemcpy2DAsync(host, row, d, dpitch, row, (size_t)h, hipMemcpyDeviceToHost, s), "download");
        return;
    }
    for (size_t j = 0; j < (size_t)h; j += rows) {
        const size_t nr = std::min(rows, (size_t)h - j);
        hip_check(hipMemcpy2DAsync(p, row, reinterpret_cast<const char *>(d) + j * dpitch, dpitch, row, nr,
                                   hipMemcpyDeviceToHost, s),
                  "download");
        hip_check(hipEventRecord(st->ev, s), "staging event");
        hip_check(hipEventSynchronize(st->ev), "download wait");
        memcpy(reinterpret_cast<char *>(host) + j * row, p, nr * row);
    }
}

// --------------------------------------------------------------------------
// Geometry / BC on the device
// --------------------------------------------------------------------------
void DevGeo::upload(hipStream_t s, const mgp_geo &h) {
    g = make_grid(h.nx, h.ny, h.dx, h.dy);
    const size_t n = level_elems(h.nx, h.ny);
    J.alloc(n);
    Jxx.alloc(n);
    Jxy.alloc(n);
    Jyx.alloc(n);
    Jyy.alloc(n);
    upload_block(s, J.p, g, h.J, 0, 0, h.nx, h.ny);
    upload_block(s, Jxx.p, g, h.Jxx, 0, 0, h.nx + 1, h.ny);
    upload_block(s, Jxy.p, g, h.Jxy, 0, 0, h.nx + 1, h.ny);
    upload_block(s, Jyx.p, g, h.Jyx, 0, 0, h.nx, h.ny + 1);
    upload_block(s, Jyy.p, g, h.Jyy, 0, 0, h.nx, h.ny + 1);
}
void DevGeo::generate(hipStream_t s, const mgp_mapping &m, int nx, int ny, double dx, double dy) {
    g = make_grid(nx, ny, dx, dy);
    const size_t n = level_elems(nx, ny);
    J.alloc(n);
    Jxx.alloc(n);
    Jxy.alloc(n);
    Jyx.alloc(n);
    Jyy.alloc(n);
    launch_metric(s, m, g, J.p, Jxx.p, Jxy.p, Jyx.p, Jyy.p);
}
Geo DevGeo::view() const { return Geo{J.p, Jxx.p, Jxy.p, Jyx.p, Jyy.p}; }

void DevBc::upload(hipStream_t s, const mgp_bc &h, int nx, int ny) {
    for (int k = 0; k < 4; ++k) {
        bc.type[k] = h.type[k];
        bc.mode[k] = h.mode[k];
        bc.val[k] = 0.0;
        bc.data[k] = nullptr;
        if (h.type[k] == BC_CF) {
            g_last_error = "BCTYPE_CF is not supported (the reference leaves its coefficients unset)";
            throw Failure(MGP_ERR_UNSUPPORTED);
        }
        if (h.mode[k] < MODE_UNIFORM || h.mode[k] > MODE_LAH) {
            g_last_error = "bad BC mode";
            throw Failure(MGP_ERR_ARG);
        }
        if (h.mode[k] == MODE_NONUNIFORM) {
            const int n = k < 2 ? ny : nx;
            if (!h.data[k]) {
                g_last_error = "NONUNIFORM BC without data";
                throw Failure(MGP_ERR_ARG);
            }
            data[k].alloc(n);
            hip_check(hipMemcpyAsync(data[k].p, h.data[k], n * sizeof(double), hipMemcpyHostToDevice, s), "bc");
            bc.data[k] = data[k].p;
        } else if (h.data[k]) {
            bc.val[k] = h.data[k][0];
        }
    }
}

Bc mg_bc(const Bc &user, int mode) {
    Bc b;
    for (int k = 0; k < 4; ++k) {
        b.type[k] = user.type[k];
        b.mode[k] = mode;
        b.val[k] = 0.0;
        b.data[k] = nullptr;
    }
    return b;
}

// --------------------------------------------------------------------------
// Schedule (MGPoisson2D.f90:38-129)
// --------------------------------------------------------------------------
int reference_schedule(int nx, int ny, double dx, double dy, int maxdepth, std::vector<int> &refx,
                       std::vector<int> &refy) {
    if (maxdepth < 0) maxdepth = 1000;
    int mgnx = nx, mgny = ny, d;
    for (d = 0; d < maxdepth; ++d) {
        if (mgnx % 2 == 0) mgnx /= 2;
        else if (mgny % 2 == 0) mgny /= 2;
        else break;
    }
    maxdepth = std::min(maxdepth, d + 1);
    refx.assign(maxdepth, 0);
    refy.assign(maxdepth, 0);
    mgnx = nx;
    mgny = ny;
    double mgdx = dx, mgdy = dy;
    for (d = 0; d < maxdepth; ++d) {
        if (mgnx % 2 != 0 || mgny % 2 != 0) break;
        const double max_dx = std::max(mgdx, mgdy);
        refx[d] = mgdx <= max_dx / 2 ? 2 : 1;
        refy[d] = mgdy <= max_dx / 2 ? 2 : 1;
        if (refx[d] * refy[d] == 1) refx[d] = refy[d] = 2;
        mgnx /= refx[d];
        mgny /= refy[d];
        if (mgnx < 4 || mgny < 4) break;
        mgdx = mgdx * (double)refx[d];
        mgdy = mgdy * (double)refy[d];
    }
    refx.resize(d);
    refy.resize(d);
    return d;
}

// --------------------------------------------------------------------------
// Scalars
// --------------------------------------------------------------------------
double Workspace::reduce(hipStream_t s, const double *a, const double *b, const double *J, double vol,
                         const Grid &g, int kind, int p) {
    const size_t need = (size_t)reduce_blocks(g) + 8;
    if (partial.n < need) partial.alloc(need);
    if (!host_scalar) hip_check(hipHostMalloc((void **)&host_scalar, 8 * sizeof(double)), "hipHostMalloc");
    // result in partial.p[0], per-block partial sums from partial.p[8] on
    launch_reduce(s, partial.p, partial.p + 8, a, b, J, vol, g, kind, p);
    hip_check(hipMemcpyAsync(host_scalar, partial.p, sizeof(double), hipMemcpyDeviceToHost, s), "scalar");
    hip_check(hipStreamSynchronize(s), "sync");
    return host_scalar[0];
}
Workspace::~Workspace() {
    if (host_scalar) (void)hipHostFree(host_scalar);
}

double pnorm2(Workspace &W, hipStream_t s, const double *a, const Grid &g) {
    return std::sqrt(W.reduce(s, a, nullptr, nullptr, g.dx * g.dy, g, 1));
}

// --------------------------------------------------------------------------
// Reference hierarchy and cycle drivers
// --------------------------------------------------------------------------
void RefHierarchy::build(hipStream_t s, const DevGeo &top, const Bc &user_bc, int dc_, int ordering_,
                         int maxdepth_user) {
    dc = dc_;
    ordering = ordering_;
    maxdepth = reference_schedule(top.g.nx, top.g.ny, top.g.dx, top.g.dy, maxdepth_user, refx, refy);
    levels.clear();
    levels.resize(maxdepth + 1);
    int nx = top.g.nx, ny = top.g.ny;
    double dx = top.g.dx, dy = top.g.dy;
    for (int d = 0; d <= maxdepth; ++d) {
        HLevel &l = levels[d];
        l.g = make_grid(nx, ny, dx, dy);
        const size_t n = level_elems(nx, ny);
        l.J.alloc(n);
        l.Jxx.alloc(n);
        l.Jxy.alloc(n);
        l.Jyx.alloc(n);
        l.Jyy.alloc(n);
        l.e.alloc(n);
        l.r.alloc(n);
        l.w.alloc(n);
        l.invd.alloc(n);
        l.mod.alloc(n);
        l.bc = mg_bc(user_bc, dc ? MODE_LAH : MODE_UNIFORM);
        if (d == 0) {
            const size_t bytes = n * sizeof(double);
            hip_check(hipMemcpyAsync(l.J.p, top.J.p, bytes, hipMemcpyDeviceToDevice, s), "copy");
            hip_check(hipMemcpyAsync(l.Jxx.p, top.Jxx.p, bytes, hipMemcpyDeviceToDevice, s), "copy");
            hip_check(hipMemcpyAsync(l.Jxy.p, top.Jxy.p, bytes, hipMemcpyDeviceToDevice, s), "copy");
            hip_check(hipMemcpyAsync(l.Jyx.p, top.Jyx.p, bytes, hipMemcpyDeviceToDevice, s), "copy");
            hip_check(hipMemcpyAsync(l.Jyy.p, top.Jyy.p, bytes, hipMemcpyDeviceToDevice, s), "copy");
        } else {
            const HLevel &f = levels[d - 1];
            const int rx = refx[d - 1], ry = refy[d - 1];
            launch_restrict_cc(s, l.J.p, f.J.p, l.g, f.g.pitch, rx, ry, nullptr, nullptr);
            launch_restrict_xf(s, l.Jxx.p, f.Jxx.p, l.g, f.g.pitch, rx, ry);
            launch_restrict_xf(s, l.Jxy.p, f.Jxy.p, l.g, f.g.pitch, rx, ry);
            launch_restrict_yf(s, l.Jyx.p, f.Jyx.p, l.g, f.g.pitch, rx, ry);
            launch_restrict_yf(s, l.Jyy.p, f.Jyy.p, l.g, f.g.pitch, rx, ry);
        }
        l.geo = Geo{l.J.p, l.Jxx.p, l.Jxy.p, l.Jyx.p, l.Jyy.p};
        if (dc) launch_dc_invdiags(s, l.invd.p, l.g, l.geo);
        else launch_invdiags(s, l.invd.p, l.g, l.geo, l.bc, 1);
        if (d < maxdepth) {
            nx /= refx[d];
            ny /= refy[d];
            dx *= refx[d];
            dy *= refy[d];
        }
    }
    // bottom-solver work arrays
    const HLevel &b = levels[maxdepth];
    const size_t nb = level_elems(b.g.nx, b.g.ny);
    for (DevBuf *w : {&bt_pt, &bt_st, &bt_r, &bt_rt, &bt_p, &bt_t, &bt_v, &bt_e, &bt_mod}) w->alloc(nb);
    bt_scal.alloc(8);
}

static bool removes_average(const Bc &b) {
    const bool rax = (b.type[XLO] == BC_PERIODIC && b.type[XHI] == BC_PERIODIC) ||
                     (b.type[XLO] == BC_NEUM && b.type[XHI] == BC_NEUM);
    const bool ray = (b.type[YLO] == BC_PERIODIC && b.type[YHI] == BC_PERIODIC) ||
                     (b.type[YLO] == BC_NEUM && b.type[YHI] == BC_NEUM);
    return rax && ray;
}

// J-weighted mean removal of prolong (MGPoisson2D.f90:371-383), applied to
// the whole array including ghosts.  Levels up to 65536 cells sum in the
// reference's sequential order (rows outer, columns inner) so small cases
// match it bit for bit; larger levels use the parallel compensated sum.
void remove_average(Workspace &W, hipStream_t s, double *fine, const Grid &fg, const double *fineJ) {
    const double dxprod = fg.dx * fg.dy;
    const size_t need = (size_t)reduce_blocks(fg) + 8;
    if (W.partial.n < need) W.partial.alloc(need);
    double *sums = W.partial.p;  // [0] = sum, [1] = vol, partials from [8]
    if ((long)fg.nx * fg.ny <= 65536) {
        launch_seq_mean(s, sums, fine, fineJ, fg, dxprod);
    } else {
        launch_reduce(s, sums, sums + 8, fine, nullptr, fineJ, dxprod, fg, 3);
        launch_reduce(s, sums + 1, sums + 8, fineJ, nullptr, nullptr, dxprod, fg, 3);
    }
    launch_sub_mean(s, fine, sums, fg);
}

const double *Workspace::ones(hipStream_t s, const Grid &g) {
    const size_t need = level_elems(g.nx, g.ny);
    if (ones_buf.n < need) {
        ones_buf.alloc(need);
        ones_pitch = -1;
    }
    if (ones_pitch != g.pitch || ones_nx != g.nx || ones_ny != g.ny) {
        launch_setval(s, ones_buf.p, 1.0, g, 1);
        ones_pitch = g.pitch;
        ones_nx = g.nx;
        ones_ny = g.ny;
    }
    return ones_buf.p;
}

void prolong_level(Workspace &W, hipStream_t s, HLevel &f, HLevel &c, int order, DevBuf *tmp) {
    // the coarse ghosts are filled (homog, do_neum) only for the linear pass
    if (order >= 1) launch_fill_ghosts(s, c.e.p, c.g, c.geo, c.bc, 1, 1);
    const double *x3 = nullptr;
    if (order >= 2) {
        launch_extrap3(s, tmp->p, c.e.p, c.g);
        x3 = tmp->p;
    }
    launch_prolong2x2(s, f.e.p, f.g.pitch, c.e.p, x3, c.g, order);
    if (order >= 2)
        hip_check(hipMemcpyAsync(c.e.p, tmp->p, level_elems(c.g.nx, c.g.ny) * sizeof(double),
                                 hipMemcpyDeviceToDevice, s),
                  "copy");
    if (removes_average(c.bc)) remove_average(W, s, f.e.p, f.g, f.J.p);
}

void RefHierarchy::relax(hipStream_t s, int d, double *phi, const double *rhs, int sweeps, double omega) {
    if (sweeps <= 0) return;
    HLevel &l = levels[d];
    if (!dc) {
        if (ordering == MGP_ORDER_LEX) {
            launch_gs_lex(s, phi, rhs, l.invd.p, l.g, l.geo, l.bc, 1, omega, sweeps);
        } else {
            for (int k = 0; k < sweeps; ++k) {
                launch_gs_mc_interior_tiled(s, phi, rhs, l.invd.p, l.g, l.geo, omega);
                launch_gs_mc_ring(s, phi, rhs, l.invd.p, l.g, l.geo, l.bc, 1, omega);
            }
        }
    } else {
        if (ordering == MGP_ORDER_RB) {
            // w is free until the residual after the smoother.  phi is the
            // level's e; when the result ends in w the two swap roles (every
            // later use reads l.e.p afresh), which saves the copy back.
            if (dc_rb_relax(s, phi, l.w.p, l.mod.p, rhs, l.invd.p, l.g, l.geo, l.bc, 1, omega, sweeps,
                            phi == l.e.p))
                std::swap(l.e, l.w);
        } else {
            launch_dc_modrhs(s, l.mod.p, rhs, phi, l.g, l.geo);
            launch_dc_lex(s, phi, l.mod.p, l.invd.p, l.g, l.geo, l.bc, 1, omega, sweeps);
        }
    }
}

double RefHierarchy::time_relax(hipStream_t s, int kernel, int reps, double *bytes) {
    const bool rb = dc && ordering == MGP_ORDER_RB && kernel >= 4 && kernel <= 6;
    const bool mc = !dc && ordering == MGP_ORDER_MC && (kernel == 7 || kernel == 8);
    if (!(rb || mc) || reps <= 0) {
        g_last_error = "time_kernel: kernels 4-6 time the red-black smoother of a DC_VCYCLE solver (ordering RB), "
                       "7-8 the multi-colour sweep of a VCYCLE solver (ordering MC)";
        throw Failure(MGP_ERR_ARG);
    }
    HLevel &l = levels[0];
    const double omega = rb ? 4.0 / 5.0 : 1.0;
    const int prev = dc_rb_tile_choice(kernel == 5 ? 0 : -1);
    auto once = [&] {
        if (kernel == 6) {
            dc_rb_tile_once(s, l.e.p, l.w.p, l.r.p, l.invd.p, l.g, l.geo, l.bc, omega, 2);
        } else if (kernel == 7) {  // one multi-colour sweep: the tiled interior colours, then the ring
            launch_gs_mc_interior_tiled(s, l.e.p, l.r.p, l.invd.p, l.g, l.geo, omega);
            launch_gs_mc_ring(s, l.e.p, l.r.p, l.invd.p, l.g, l.geo, l.bc, 1, omega);
        } else if (kernel == 8) {  // the same sweep with one launch per interior colour
            for (int c = 0; c < 4; ++c) launch_gs_mc_interior(s, l.e.p, l.r.p, l.invd.p, l.g, l.geo, omega, c & 1, c >> 1);
            launch_gs_mc_ring(s, l.e.p, l.r.p, l.invd.p, l.g, l.geo, l.bc, 1, omega);
        } else {
            dc_rb_relax(s, l.e.p, l.w.p, l.mod.p, l.r.p, l.invd.p, l.g, l.geo, l.bc, 1, omega, 2);
        }
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.0f;
    try {
        once();  // first use picks the tile shape
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        hip_check(hipEventRecord(e0, s), "hipEventRecord");
        for (int r = 0; r < reps; ++r) once();
        hip_check(hipEventRecord(e1, s), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    } catch (...) {
        dc_rb_tile_choice(prev);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        throw;
    }
    dc_rb_tile_choice(prev);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // RB: two sweeps of phi, modified rhs, invdiags and four face metrics read
    // once per call and phi written (64 B/cell); MC: one sweep of the same
    // operands (64 B/cell)
    *bytes = 8.0 * 8.0 * (double)l.g.nx * (double)l.g.ny;
    return ms / reps;
}

void RefHierarchy::bottom(hipStream_t s, int d) {
    HLevel &l = levels[d];
    BottomArgs a;
    a.g = l.g;
    a.geo = l.geo;
    a.bc = l.bc;
    a.phi = l.e.p;
    a.rhs = l.r.p;
    a.invd = l.invd.p;
    a.e = bt_e.p;
    a.pt = bt_pt.p;
    a.st = bt_st.p;
    a.r = bt_r.p;
    a.rt = bt_rt.p;
    a.p = bt_p.p;
    a.t = bt_t.p;
    a.v = bt_v.p;
    a.mod = bt_mod.p;
    a.scal = bt_scal.p;
    a.dc = dc;
    a.order = ordering;
    a.max_iters = 80;       // MGPoisson2D.f90:710-712
    a.max_restarts = 5;
    a.homog = 1;
    a.tol = 1.0e-6;
    launch_bicgstab2(s, a);
}

void RefHierarchy::vcycle_noinit(Workspace &W, hipStream_t s, int d, int numcycles, int sd, int su, int sb) {
    HLevel &l = levels[d];
    const double omega = dc ? 4.0 / 5.0 : 1.0;
    if (dc) launch_einit(s, l.e.p, l.r.p, l.invd.p, l.g);  // MGPoisson2D_DD.f90:931,971
    if (d == maxdepth) {
        relax(s, d, l.e.p, l.r.p, sb, omega);
        bottom(s, d);
        return;
    }
    HLevel &c = levels[d + 1];
    relax(s, d, l.e.p, l.r.p, sd, omega);
    if (dc) {
        launch_fill_ghosts(s, l.e.p, l.g, l.geo, l.bc, 1, 1);
        launch_dc_laplacian(s, l.w.p, l.e.p, l.r.p, l.g, l.geo);
        launch_restrict_cc(s, c.r.p, l.w.p, c.g, l.g.pitch, refx[d], refy[d], nullptr, nullptr);
    } else {
        launch_fill_ghosts(s, l.e.p, l.g, l.geo, l.bc, 1, 0);
        launch_laplacian(s, l.w.p, l.e.p, l.r.p, l.g, l.geo, l.bc, 1, 0);
        // e(d+1) = r(d+1) * invdiags(d+1) fused (MGPoisson2D.f90:805-806)
        launch_restrict_cc(s, c.r.p, l.w.p, c.g, l.g.pitch, refx[d], refy[d], c.e.p, c.invd.p);
    }
    for (int k = 0; k < numcycles; ++k) vcycle_noinit(W, s, d + 1, numcycles, sd, su, sb);
    if (refx[d] != 2 || refy[d] != 2) {
        g_last_error = "prolong: only 2x2 coarsening is defined by the reference (MGPoisson2D.f90:385-388)";
        throw Failure(MGP_ERR_UNSUPPORTED);
    }
    prolong_level(W, s, l, c, 1, nullptr);
    relax(s, d, l.e.p, l.r.p, su, omega);
}

// MGPoisson2D::fmg_noinit (MGPoisson2D.f90:1127-1189): e(d) = 0; restrict r(d)
// to r(d+1), solve there recursively and prolong with order 2; then numcycles
// V-cycles at depth d from that initial guess.
void RefHierarchy::fmg_noinit(Workspace &W, hipStream_t s, int d, int numcycles, int sd, int su, int sb) {
    HLevel &l = levels[d];
    launch_setval(s, l.e.p, 0.0, l.g, 1);
    if (d < maxdepth) {
        HLevel &c = levels[d + 1];
        launch_restrict_cc(s, c.r.p, l.r.p, c.g, l.g.pitch, refx[d], refy[d], nullptr, nullptr);
        fmg_noinit(W, s, d + 1, numcycles, sd, su, sb);
        if (refx[d] != 2 || refy[d] != 2) {
            g_last_error = "prolong: only 2x2 coarsening is defined by the reference (MGPoisson2D.f90:385-388)";
            throw Failure(MGP_ERR_UNSUPPORTED);
        }
        // the coarse work array is free here and sized for the order-3 ghost extrapolation
        prolong_level(W, s, l, c, 2, &c.w);
    }
    for (int k = 0; k < numcycles; ++k) vcycle_noinit(W, s, d, numcycles, sd, su, sb);
}

int RefHierarchy::solve(Workspace &W, hipStream_t s, const DevGeo &top, const Bc &user_bc, int homog,
                        double *phi, double *rhs, double tol, int maxiters, int numcycles, int sd, int su,
                        int sb, int zerophi, double *relres, int *iters_out) {
    HLevel &l0 = levels[0];
    const Grid &g = top.g;
    const Geo geo = top.view();
    if (zerophi) launch_setval(s, phi, 0.0, g, 1);
    if (dc) launch_lincomb_mul(s, rhs, rhs, geo.J, g, 0);  // rhs = rhs * J (MGPoisson2D_DD.f90:772-773)
    // residual with the user's BCs (Poisson2D::compute_residual for both drivers)
    launch_fill_ghosts(s, phi, g, geo, user_bc, homog, 0);
    launch_laplacian(s, l0.r.p, phi, rhs, g, geo, user_bc, homog, 0);
    double rscale = pnorm2(W, s, l0.r.p, g);
    std::vector<double> rr(maxiters + 1, 0.0);
    if (rscale == 0.0) {
        rr[0] = rscale;
        rscale = 1.0;
    } else {
        rr[0] = 1.0;
    }
    int outcome = MGP_MAXITERS, it;
    for (it = 1; it <= maxiters; ++it) {
        if (fmg) {
            fmg_noinit(W, s, 0, numcycles, sd, su, sb);  // MGPoisson2D.f90:1083-1088
        } else {
            if (!dc) launch_einit(s, l0.e.p, l0.r.p, l0.invd.p, l0.g);
            vcycle_noinit(W, s, 0, numcycles, sd, su, sb);
        }
        launch_axpy_ghosted(s, phi, l0.e.p, 1.0, g);
        launch_fill_ghosts(s, phi, g, geo, user_bc, homog, 0);
        launch_laplacian(s, l0.r.p, phi, rhs, g, geo, user_bc, homog, 0);
        rr[it] = pnorm2(W, s, l0.r.p, g) / rscale;
        if (rr[it] <= tol) {
            outcome = MGP_CONVERGED;
            break;
        }
        if (rr[it] > rr[it - 1]) {
            launch_axpy_ghosted(s, phi, l0.e.p, -1.0, g);
            outcome = MGP_DIVERGED;
            break;
        }
    }
    const int iters = it > maxiters ? maxiters : it;
    if (dc) launch_lincomb_div(s, rhs, rhs, geo.J, g, 0);  // restore (MGPoisson2D_DD.f90:837-838)
    hip_check(hipStreamSynchronize(s), "sync");
    if (relres)
        for (int k = 0; k <= iters; ++k) relres[k] = rr[k];
    if (iters_out) *iters_out = iters;
    return outcome;
}

}  // namespace mgp
