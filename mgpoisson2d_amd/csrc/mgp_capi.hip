// mgp_capi.hip -- extern "C" entry points of include/mgp2d.h.
//
// Section-2 calls (drop-in replacements of the reference's module
// procedures) copy host arrays into padded HBM buffers, run the kernels on a
// per-thread stream and copy results back.  Section-3 calls operate on a
// resident mgp_solver.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "mgp2d.h"
#include "mgp_batch.h"
#include "mgp_geometry.h"
#include "mgp_host.h"
#include "mgp_kernels.h"
#include "mgp_krylov.h"
#include "mgp_single.h"

using namespace mgp;

namespace {

hipStream_t thread_stream() {
    thread_local hipStream_t s = nullptr;
    if (!s) hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    return s;
}

template <class F>
int guarded(F fn) {
    try {
        fn();
        return MGP_OK;
    } catch (const Failure &f) {
        return f.status;
    } catch (...) {
        g_last_error = "unexpected host exception";
        return MGP_ERR_HIP;
    }
}

void check_geo(const mgp_geo *geo) {
    if (!geo || geo->nx < 3 || geo->ny < 3 || !(geo->dx > 0) || !(geo->dy > 0) || !geo->J || !geo->Jxx ||
        !geo->Jxy || !geo->Jyx || !geo->Jyy) {
        g_last_error = "bad geometry (need nx, ny >= 3, positive spacing, five metric arrays)";
        throw Failure(MGP_ERR_ARG);
    }
}

void check_bc(const mgp_bc *bc) {
    if (!bc) {
        g_last_error = "null bc";
        throw Failure(MGP_ERR_ARG);
    }
    for (int k = 0; k < 4; ++k)
        if (bc->type[k] < BC_NONE || bc->type[k] > BC_EXTRAP2) {
            g_last_error = "bad BC type";
            throw Failure(MGP_ERR_ARG);
        }
}

void no_lah(const mgp_bc *bc) {
    for (int k = 0; k < 4; ++k)
        if (bc->mode[k] == MODE_LAH) {
            g_last_error = "Poisson2D operators reject BCMODE_LAH (poisson2d.f90:35-39)";
            throw Failure(MGP_ERR_ARG);
        }
}

// One-ghost host field <-> device.
void up_phi(hipStream_t s, DevBuf &d, const Grid &g, const double *h) { upload_block(s, d.p, g, h, -1, -1, g.nx + 2, g.ny + 2); }
void down_phi(hipStream_t s, double *h, const DevBuf &d, const Grid &g) { download_block(s, h, d.p, g, -1, -1, g.nx + 2, g.ny + 2); }
void up_cc(hipStream_t s, DevBuf &d, const Grid &g, const double *h) { upload_block(s, d.p, g, h, 0, 0, g.nx, g.ny); }
void down_cc(hipStream_t s, double *h, const DevBuf &d, const Grid &g) { download_block(s, h, d.p, g, 0, 0, g.nx, g.ny); }

struct Call {
    hipStream_t s;
    DevGeo geo;
    DevBc bc;
    Call(const mgp_geo *g, const mgp_bc *b) : s(thread_stream()) {
        check_geo(g);
        geo.upload(s, *g);
        if (b) {
            check_bc(b);
            bc.upload(s, *b, g->nx, g->ny);
        }
    }
    DevBuf field() {
        DevBuf d;
        d.alloc(level_elems(geo.g.nx, geo.g.ny));
        return d;
    }
    void finish() { hip_check(hipStreamSynchronize(s), "sync"); }
};

}  // namespace

struct mgp_batch {
    hipStream_t stream = nullptr;
    mgp::BatchKMG *b = nullptr;
    ~mgp_batch() {
        delete b;
        if (stream) (void)hipStreamDestroy(stream);
    }
};

struct mgp_server {
    mgp::Server *srv = nullptr;
    ~mgp_server() { delete srv; }
};

struct mgp_solver {
    hipStream_t stream = nullptr;
    mgp_solver_opts opts;
    DevGeo geo;
    DevBc ubc;
    DevBuf phi, rhs;
    DevBuf xflux, yflux, div;  // projection fields, allocated by the first set_flux
    Workspace W;
    RefHierarchy ref;
    KrylovMG kmg;
    int last_cycles = 0;
};

extern "C" {

const char *mgp_status_string(int status) {
    switch (status) {
    case MGP_OK: return "ok";
    case MGP_ERR_ARG: return "bad argument";
    case MGP_ERR_UNSUPPORTED: return "unsupported";
    case MGP_ERR_HIP: return "HIP runtime error";
    case MGP_ERR_NOMEM: return "out of device memory";
    case MGP_ERR_BREAKDOWN: return "solver breakdown";
    default: return "unknown status";
    }
}

const char *mgp_last_error(void) { return g_last_error.c_str(); }

int mgp_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int mgp_fill_ghosts(const mgp_geo *geo, const mgp_bc *bc, int homog, int do_neum, double *phi) {
    return guarded([&] {
        Call c(geo, bc);
        DevBuf f = c.field();
        up_phi(c.s, f, c.geo.g, phi);
        launch_fill_ghosts(c.s, f.p, c.geo.g, c.geo.view(), c.bc.bc, homog, do_neum);
        down_phi(c.s, phi, f, c.geo.g);
        c.finish();
    });
}

int mgp_compute_laplacian(const mgp_geo *geo, const mgp_bc *bc, int homog, int jscale, double *phi, double *lap) {
    return guarded([&] {
        no_lah(bc);
        Call c(geo, bc);
        DevBuf f = c.field(), out = c.field();
        up_phi(c.s, f, c.geo.g, phi);
        launch_fill_ghosts(c.s, f.p, c.geo.g, c.geo.view(), c.bc.bc, homog, 0);
        launch_laplacian(c.s, out.p, f.p, nullptr, c.geo.g, c.geo.view(), c.bc.bc, homog, jscale);
        down_phi(c.s, phi, f, c.geo.g);
        down_cc(c.s, lap, out, c.geo.g);
        c.finish();
    });
}

int mgp_compute_residual(const mgp_geo *geo, const mgp_bc *bc, int homog, const double *rhs, double *phi,
                         double *res) {
    return guarded([&] {
        no_lah(bc);
        Call c(geo, bc);
        DevBuf f = c.field(), r = c.field(), out = c.field();
        up_phi(c.s, f, c.geo.g, phi);
        up_cc(c.s, r, c.geo.g, rhs);
        launch_fill_ghosts(c.s, f.p, c.geo.g, c.geo.view(), c.bc.bc, homog, 0);
        launch_laplacian(c.s, out.p, f.p, r.p, c.geo.g, c.geo.view(), c.bc.bc, homog, 0);
        down_phi(c.s, phi, f, c.geo.g);
        down_cc(c.s, res, out, c.geo.g);
        c.finish();
    });
}

int mgp_compute_invdiags(const mgp_geo *geo, const mgp_bc *bc, int homog, double *invd) {
    return guarded([&] {
        no_lah(bc);
        Call c(geo, bc);
        DevBuf out = c.field();
        launch_invdiags(c.s, out.p, c.geo.g, c.geo.view(), c.bc.bc, homog);
        down_cc(c.s, invd, out, c.geo.g);
        c.finish();
    });
}

int mgp_relax_gs(const mgp_geo *geo, const mgp_bc *bc, int homog, const double *invd, double omega, int maxiters,
                 const double *rhs, double *phi, int ordering) {
    return guarded([&] {
        no_lah(bc);
        if (maxiters <= 0) return;
        if (ordering != MGP_ORDER_LEX && ordering != MGP_ORDER_MC) {
            g_last_error = "relax_gs ordering must be LEX or MC";
            throw Failure(MGP_ERR_ARG);
        }
        Call c(geo, bc);
        if (ordering == MGP_ORDER_MC)
            for (int k = 0; k < 4; ++k)
                if (bc->type[k] == BC_PERIODIC) {
                    g_last_error = "MC ordering needs non-periodic sides";
                    throw Failure(MGP_ERR_UNSUPPORTED);
                }
        DevBuf f = c.field(), r = c.field(), iv = c.field();
        up_phi(c.s, f, c.geo.g, phi);
        up_cc(c.s, r, c.geo.g, rhs);
        up_cc(c.s, iv, c.geo.g, invd);
        const Grid &g = c.geo.g;
        if (ordering == MGP_ORDER_LEX) {
            launch_gs_lex(c.s, f.p, r.p, iv.p, g, c.geo.view(), c.bc.bc, homog, omega, maxiters);
        } else {
            for (int k = 0; k < maxiters; ++k) {
                launch_gs_mc_interior_tiled(c.s, f.p, r.p, iv.p, g, c.geo.view(), omega);
                launch_gs_mc_ring(c.s, f.p, r.p, iv.p, g, c.geo.view(), c.bc.bc, homog, omega);
            }
        }
        down_phi(c.s, phi, f, g);
        c.finish();
    });
}

int mgp_xline_relax(const mgp_geo *geo, const mgp_bc *bc, const double *rhs, double *phi, int sweeps) {
    return guarded([&] {
        if (sweeps <= 0) return;
        Call c(geo, bc);
        const Grid &g = c.geo.g;
        for (int k = 0; k < 4; ++k)
            if (bc->type[k] != BC_DIRI && bc->type[k] != BC_NEUM) {
                g_last_error = "xline_relax supports Dirichlet and Neumann sides";
                throw Failure(MGP_ERR_UNSUPPORTED);
            }
        if (g.nx < 4 || g.ny < 4 || !xline_supported(g.nx)) {
            g_last_error = "xline_relax needs ny >= 4 and nx >= 4 with nx <= 4111 or nx divisible into <= 64 "
                           "segments (a power of two) of at most 4111 cells";
            throw Failure(MGP_ERR_ARG);
        }
        const Bc hb = mg_bc(c.bc.bc, MODE_UNIFORM);
        DevBuf e = c.field(), f = c.field(), K;
        K.alloc((size_t)ring_table_elems(g));
        up_phi(c.s, e, g, phi);
        up_cc(c.s, f, g, rhs);
        launch_ring_coef(c.s, K.p, g, c.geo.view(), hb);
        for (int k = 0; k < sweeps; ++k) xline_zebra(c.s, g, c.geo.view(), K.p, e.p, f.p);
        down_phi(c.s, phi, e, g);
        c.finish();
    });
}

int mgp_dc_compute_laplacian(const mgp_geo *geo, const mgp_bc *bc, int homog, double *phi, double *lap) {
    return guarded([&] {
        Call c(geo, bc);
        DevBuf f = c.field(), out = c.field();
        up_phi(c.s, f, c.geo.g, phi);
        launch_fill_ghosts(c.s, f.p, c.geo.g, c.geo.view(), c.bc.bc, homog, 1);
        launch_dc_laplacian(c.s, out.p, f.p, nullptr, c.geo.g, c.geo.view());
        down_phi(c.s, phi, f, c.geo.g);
        down_cc(c.s, lap, out, c.geo.g);
        c.finish();
    });
}

int mgp_dc_compute_invdiags(const mgp_geo *geo, double *invd) {
    return guarded([&] {
        Call c(geo, nullptr);
        DevBuf out = c.field();
        launch_dc_invdiags(c.s, out.p, c.geo.g, c.geo.view());
        down_cc(c.s, invd, out, c.geo.g);
        c.finish();
    });
}

int mgp_dc_relax_gs(const mgp_geo *geo, const mgp_bc *bc, int homog, const double *invd, double omega, int maxiters,
                    const double *rhs, double *phi, int ordering) {
    return guarded([&] {
        if (maxiters <= 0) return;
        if (ordering != MGP_ORDER_LEX && ordering != MGP_ORDER_RB) {
            g_last_error = "DC relax_gs ordering must be LEX or RB";
            throw Failure(MGP_ERR_ARG);
        }
        Call c(geo, bc);
        DevBuf f = c.field(), r = c.field(), iv = c.field(), mod = c.field(), tmp;
        up_phi(c.s, f, c.geo.g, phi);
        up_cc(c.s, r, c.geo.g, rhs);
        up_cc(c.s, iv, c.geo.g, invd);
        const Grid &g = c.geo.g;
        if (ordering == MGP_ORDER_RB) {
            tmp = c.field();
            dc_rb_relax(c.s, f.p, tmp.p, mod.p, r.p, iv.p, g, c.geo.view(), c.bc.bc, homog, omega, maxiters);
        } else {
            launch_dc_modrhs(c.s, mod.p, r.p, f.p, g, c.geo.view());
            launch_dc_lex(c.s, f.p, mod.p, iv.p, g, c.geo.view(), c.bc.bc, homog, omega, maxiters);
        }
        down_phi(c.s, phi, f, g);
        c.finish();
    });
}

int mgp_dc_rb_tiles(int choice, int *previous) {
    return guarded([&] {
        if (choice < -1 || choice > dc_rb_tile_shapes()) {
            g_last_error = "dc_rb_tiles: choice -1 (timed), 0 (one launch per colour) or a tile shape 1..n";
            throw Failure(MGP_ERR_ARG);
        }
        const int prev = dc_rb_tile_choice(choice);
        if (previous) *previous = prev;
    });
}

int mgp_dc_rb_tile_shapes(void) { return dc_rb_tile_shapes(); }

int mgp_restrict(int kind, int fnx, int fny, const double *fine, int cnx, int cny, double *crse) {
    return guarded([&] {
        if (fnx <= 0 || fny <= 0 || cnx <= 0 || cny <= 0 || fnx % cnx || fny % cny || kind < 0 || kind > 2) {
            g_last_error = "restrict: bad shapes";
            throw Failure(MGP_ERR_ARG);
        }
        hipStream_t s = thread_stream();
        const Grid fg = make_grid(fnx, fny, 1.0, 1.0), cg = make_grid(cnx, cny, 1.0, 1.0);
        const int rx = fnx / cnx, ry = fny / cny;
        DevBuf f, c;
        f.alloc(level_elems(fnx, fny));
        c.alloc(level_elems(cnx, cny));
        const int fw = fnx + (kind == 1), fh = fny + (kind == 2), cw = cnx + (kind == 1), ch = cny + (kind == 2);
        upload_block(s, f.p, fg, fine, 0, 0, fw, fh);
        if (kind == 0) launch_restrict_cc(s, c.p, f.p, cg, fg.pitch, rx, ry, nullptr, nullptr);
        else if (kind == 1) launch_restrict_xf(s, c.p, f.p, cg, fg.pitch, rx, ry);
        else launch_restrict_yf(s, c.p, f.p, cg, fg.pitch, rx, ry);
        download_block(s, crse, c.p, cg, 0, 0, cw, ch);
        hip_check(hipStreamSynchronize(s), "sync");
    });
}

int mgp_prolong(const mgp_geo *fine_geo, const mgp_geo *crse_geo, const int *bc_types, const int *bc_modes,
                int order, double *fine, double *crse) {
    return guarded([&] {
        check_geo(fine_geo);
        check_geo(crse_geo);
        if (fine_geo->dx * fine_geo->dy > crse_geo->dx * crse_geo->dy) {
            g_last_error = "prolong: crse and fine geos passed in wrong order";
            throw Failure(MGP_ERR_ARG);
        }
        if (fine_geo->nx != 2 * crse_geo->nx || fine_geo->ny != 2 * crse_geo->ny) {
            g_last_error = "prolong: only 2x2 refinement is defined (MGPoisson2D.f90:385-388)";
            throw Failure(MGP_ERR_UNSUPPORTED);
        }
        if (order < 0 || order > 2) {
            g_last_error = "prolong: order must be 0, 1 or 2";
            throw Failure(MGP_ERR_ARG);
        }
        hipStream_t s = thread_stream();
        HLevel f, c;
        DevGeo fg, cg;
        fg.upload(s, *fine_geo);
        cg.upload(s, *crse_geo);
        f.g = fg.g;
        c.g = cg.g;
        f.geo = fg.view();
        c.geo = cg.view();
        f.J.alloc(1);  // unused placeholders
        mgp_bc hb;
        const double zero = 0.0;
        for (int k = 0; k < 4; ++k) {
            hb.type[k] = bc_types[k];
            hb.mode[k] = bc_modes[k] == MODE_LAH ? MODE_LAH : MODE_UNIFORM;
            hb.data[k] = &zero;
        }
        DevBc dbc;
        dbc.upload(s, hb, crse_geo->nx, crse_geo->ny);
        c.bc = dbc.bc;
        f.e.alloc(level_elems(f.g.nx, f.g.ny));
        c.e.alloc(level_elems(c.g.nx, c.g.ny));
        DevBuf tmp;
        tmp.alloc(level_elems(c.g.nx, c.g.ny));
        up_phi(s, f.e, f.g, fine);
        up_phi(s, c.e, c.g, crse);
        // prolong_level reads the fine J for the mean removal
        Workspace W;
        f.J = std::move(fg.J);
        prolong_level(W, s, f, c, order, &tmp);
        down_phi(s, fine, f.e, f.g);
        down_phi(s, crse, c.e, c.g);
        hip_check(hipStreamSynchronize(s), "sync");
    });
}

int mgp_compute_pd(int nx, int ny, double dx, double dy, const double *phi, int dir, int nodedir, double *pd) {
    return guarded([&] {
        if (nx < 3 || ny < 3 || !(dx > 0) || !(dy > 0) || (dir != 1 && dir != 2) || (nodedir != 1 && nodedir != 2) ||
            !phi || !pd) {
            g_last_error = "compute_pd: need nx, ny >= 3, positive spacing, dir and nodedir in {1, 2}";
            throw Failure(MGP_ERR_ARG);
        }
        hipStream_t s = thread_stream();
        const Grid g = make_grid(nx, ny, dx, dy);
        DevBuf f, out;
        f.alloc(level_elems(nx, ny));
        out.alloc(level_elems(nx, ny));
        up_phi(s, f, g, phi);
        launch_pd(s, out.p, f.p, g, dir, nodedir);
        download_block(s, pd, out.p, g, 0, 0, nx + (nodedir == 1), ny + (nodedir == 2));
        hip_check(hipStreamSynchronize(s), "sync");
    });
}

int mgp_compute_grad(const mgp_geo *geo, const mgp_bc *bc, int homog, double *phi, double *xflux, double *yflux) {
    return guarded([&] {
        if (!phi || !xflux || !yflux) {
            g_last_error = "compute_grad: null array";
            throw Failure(MGP_ERR_ARG);
        }
        Call c(geo, bc);
        const Grid &g = c.geo.g;
        DevBuf f = c.field(), xf = c.field(), yf = c.field();
        up_phi(c.s, f, g, phi);
        launch_fill_ghosts(c.s, f.p, g, c.geo.view(), c.bc.bc, homog, 0);
        launch_grad(c.s, xf.p, yf.p, f.p, g, c.geo.view(), c.bc.bc, homog);
        down_phi(c.s, phi, f, g);
        download_block(c.s, xflux, xf.p, g, 0, 0, g.nx + 1, g.ny);
        download_block(c.s, yflux, yf.p, g, 0, 0, g.nx, g.ny + 1);
        c.finish();
    });
}

int mgp_compute_div(int nx, int ny, double dx, double dy, const double *xflux, const double *yflux, double *div) {
    return guarded([&] {
        if (nx <= 0 || ny <= 0 || !(dx > 0) || !(dy > 0) || !xflux || !yflux || !div) {
            g_last_error = "compute_div: bad shape, spacing or null array";
            throw Failure(MGP_ERR_ARG);
        }
        hipStream_t s = thread_stream();
        const Grid g = make_grid(nx, ny, dx, dy);
        DevBuf xf, yf, out;
        xf.alloc(level_elems(nx, ny));
        yf.alloc(level_elems(nx, ny));
        out.alloc(level_elems(nx, ny));
        upload_block(s, xf.p, g, xflux, 0, 0, nx + 1, ny);
        upload_block(s, yf.p, g, yflux, 0, 0, nx, ny + 1);
        launch_div(s, out.p, xf.p, yf.p, g);
        download_block(s, div, out.p, g, 0, 0, nx, ny);
        hip_check(hipStreamSynchronize(s), "sync");
    });
}

int mgp_fill_boundary_fluxes(int nx, int ny, const mgp_bc *bc, int homog, double *xflux, double *yflux) {
    return guarded([&] {
        if (nx <= 0 || ny <= 0 || !xflux || !yflux) {
            g_last_error = "fill_boundary_fluxes: bad shape or null array";
            throw Failure(MGP_ERR_ARG);
        }
        check_bc(bc);
        hipStream_t s = thread_stream();
        const Grid g = make_grid(nx, ny, 1.0, 1.0);
        DevBc b;
        b.upload(s, *bc, nx, ny);
        DevBuf xf, yf;
        xf.alloc(level_elems(nx, ny));
        yf.alloc(level_elems(nx, ny));
        upload_block(s, xf.p, g, xflux, 0, 0, nx + 1, ny);
        upload_block(s, yf.p, g, yflux, 0, 0, nx, ny + 1);
        launch_boundary_fluxes(s, xf.p, yf.p, g, b.bc, homog);
        download_block(s, xflux, xf.p, g, 0, 0, nx + 1, ny);
        download_block(s, yflux, yf.p, g, 0, 0, nx, ny + 1);
        hip_check(hipStreamSynchronize(s), "sync");
    });
}

int mgp_integrate2d_bdry(int nx, int ny, double dx, double dy, const double *xflux, const double *yflux, double *out) {
    return guarded([&] {
        if (nx <= 0 || ny <= 0 || !xflux || !yflux || !out) {
            g_last_error = "integrate2d_bdry: bad shape or null array";
            throw Failure(MGP_ERR_ARG);
        }
        hipStream_t s = thread_stream();
        const Grid g = make_grid(nx, ny, dx, dy);
        DevBuf xf, yf, sum;
        xf.alloc(level_elems(nx, ny));
        yf.alloc(level_elems(nx, ny));
        sum.alloc(1);
        upload_block(s, xf.p, g, xflux, 0, 0, nx + 1, ny);
        upload_block(s, yf.p, g, yflux, 0, 0, nx, ny + 1);
        launch_integrate_bdry(s, sum.p, xf.p, yf.p, g);
        hip_check(hipMemcpyAsync(out, sum.p, sizeof(double), hipMemcpyDeviceToHost, s), "copy");
        hip_check(hipStreamSynchronize(s), "sync");
    });
}

int mgp_gpnorm(int ilo, int ihi, int jlo, int jhi, double dx, double dy, const double *phi, int p, double *out) {
    return guarded([&] {
        const int nx = ihi - ilo + 1, ny = jhi - jlo + 1;
        if (nx <= 0 || ny <= 0 || !phi || !out || p < 0) {
            g_last_error = "gpnorm: bad box, p < 0 or null array";
            throw Failure(MGP_ERR_ARG);
        }
        // The reference reads the x-boundary ghost columns at i = jlo-1, jhi+1
        // and the y-boundary ghost rows at j = ilo-1, ihi+1 (arrayutils.f90:820-
        // 858); on a box with ilo == jlo and ihi == jhi these are the ghost
        // layer.  Indices relative to the valid cells:
        const int cx0 = (jlo - 1) - ilo, cx1 = (jhi + 1) - ilo;
        const int ry0 = (ilo - 1) - jlo, ry1 = (ihi + 1) - jlo;
        if (cx0 < -1 || cx1 > nx || cx1 < -1 || cx0 > nx || ry0 < -1 || ry1 > ny || ry1 < -1 || ry0 > ny) {
            g_last_error = "gpnorm: the reference's index pattern (columns jlo-1, jhi+1; rows ilo-1, ihi+1) "
                           "falls outside the array on this box";
            throw Failure(MGP_ERR_ARG);
        }
        hipStream_t s = thread_stream();
        const Grid g = make_grid(nx, ny, dx, dy);
        DevBuf f, sum;
        f.alloc(level_elems(nx, ny));
        sum.alloc(1);
        up_phi(s, f, g, phi);
        launch_gpnorm_sum(s, sum.p, f.p, g, p, cx0, cx1, ry0, ry1);
        double pn = 0.0;
        hip_check(hipMemcpyAsync(&pn, sum.p, sizeof(double), hipMemcpyDeviceToHost, s), "copy");
        hip_check(hipStreamSynchronize(s), "sync");
        *out = p == 0 ? pn : std::pow(pn, 1.0 / p);
    });
}

int mgp_compute_residual_dd(const mgp_geo *geo, const mgp_bc *bc, const double *rhs, const double *phi_hi,
                            const double *phi_lo, double *res) {
    return guarded([&] {
        if (!rhs || !phi_hi || !phi_lo || !res) {
            g_last_error = "compute_residual_dd: null array";
            throw Failure(MGP_ERR_ARG);
        }
        for (int k = 0; k < 4; ++k)
            if (!bc || (bc->type[k] != BC_DIRI && bc->type[k] != BC_NEUM) || bc->mode[k] == MODE_LAH) {
                g_last_error = "compute_residual_dd: Dirichlet or Neumann sides (homogeneous, not LAH)";
                throw Failure(MGP_ERR_UNSUPPORTED);
            }
        Call c(geo, bc);
        const Grid &g = c.geo.g;
        DevBuf hi = c.field(), lo = c.field(), f = c.field(), r = c.field();
        up_phi(c.s, hi, g, phi_hi);
        up_phi(c.s, lo, g, phi_lo);
        up_cc(c.s, f, g, rhs);
        launch_residual_dd(c.s, r.p, hi.p, lo.p, f.p, g, c.geo.view(), c.bc.bc);
        down_cc(c.s, res, r, g);
        c.finish();
    });
}

int mgp_lap4(int nx, int ny, double dx, double dy, const double *const *metrics, const double *phi, double *lap) {
    return guarded([&] {
        if (nx < 5 || ny < 5 || !(dx > 0) || !(dy > 0) || !metrics || !phi || !lap) {
            g_last_error = "lap4: need nx, ny >= 5, positive spacing and arrays";
            throw Failure(MGP_ERR_ARG);
        }
        for (int k = 0; k < 8; ++k)
            if (!metrics[k]) {
                g_last_error = "lap4: null metric array";
                throw Failure(MGP_ERR_ARG);
            }
        hipStream_t s = thread_stream();
        const Grid g = make_grid(nx, ny, dx, dy);
        const size_t n = level_elems(nx, ny);
        DevBuf m[8], f, out;
        for (int k = 0; k < 8; ++k) {
            m[k].alloc(n);
            const bool xface = k < 4;
            upload_block(s, m[k].p, g, metrics[k], 0, 0, xface ? nx + 1 : nx, xface ? ny : ny + 1);
        }
        f.alloc(n);
        out.alloc(n);
        upload_block(s, f.p, g, phi, 0, 0, nx, ny);
        launch_lap4(s, out.p, f.p, nullptr, g, Geo4{m[0].p, m[1].p, m[2].p, m[3].p, m[4].p, m[5].p, m[6].p, m[7].p});
        download_block(s, lap, out.p, g, 0, 0, nx, ny);
        hip_check(hipStreamSynchronize(s), "sync");
    });
}

int mgp_metric4_fill(const mgp_mapping *map, int nx, int ny, double dx, double dy, double *const *metrics) {
    return guarded([&] {
        check_mapping(map);
        if (map->kind != MGP_MAP_WARPED || nx < 1 || ny < 1 || !(dx > 0) || !(dy > 0) || !metrics) {
            g_last_error = "metric4_fill: MGP_MAP_WARPED, a positive grid and 8 arrays";
            throw Failure(MGP_ERR_ARG);
        }
        for (int k = 0; k < 8; ++k)
            if (!metrics[k]) {
                g_last_error = "metric4_fill: null metric array";
                throw Failure(MGP_ERR_ARG);
            }
        hipStream_t s = thread_stream();
        const Grid g = make_grid(nx, ny, dx, dy);
        const size_t n = level_elems(nx, ny);
        DevBuf m[8];
        for (int k = 0; k < 8; ++k) m[k].alloc(n);
        launch_metric4(s, *map, g, Geo4W{m[0].p, m[1].p, m[2].p, m[3].p, m[4].p, m[5].p, m[6].p, m[7].p});
        for (int k = 0; k < 8; ++k) {
            const bool xface = k < 4;
            download_block(s, metrics[k], m[k].p, g, 0, 0, xface ? nx + 1 : nx, xface ? ny : ny + 1);
        }
        hip_check(hipStreamSynchronize(s), "sync");
    });
}

int mgp_create_mgschedule(int nx, int ny, double dx, double dy, int maxdepth_in, int *maxdepth_out, int *refx,
                          int *refy) {
    return guarded([&] {
        std::vector<int> rx, ry;
        const int d = reference_schedule(nx, ny, dx, dy, maxdepth_in, rx, ry);
        if (d > 64) {
            g_last_error = "schedule deeper than 64 levels";
            throw Failure(MGP_ERR_ARG);
        }
        *maxdepth_out = d;
        for (int k = 0; k < d; ++k) {
            refx[k] = rx[k];
            refy[k] = ry[k];
        }
    });
}

namespace {

void check_field(int nx, int ny, double dx, double dy, const void *a, const void *out, const char *what) {
    if (nx <= 0 || ny <= 0 || !(dx > 0) || !(dy > 0) || !a || !out) {
        g_last_error = std::string(what) + ": bad shape, spacing or null array";
        throw Failure(MGP_ERR_ARG);
    }
}

}  // namespace

int mgp_pnorm(int nx, int ny, double dx, double dy, const double *a, int p, double *out) {
    return guarded([&] {
        check_field(nx, ny, dx, dy, a, out, "pnorm");
        if (p < 0) {
            g_last_error = "pnorm: p must be >= 0";
            throw Failure(MGP_ERR_ARG);
        }
        hipStream_t s = thread_stream();
        const Grid g = make_grid(nx, ny, dx, dy);
        DevBuf d;
        d.alloc(level_elems(nx, ny));
        upload_block(s, d.p, g, a, 0, 0, nx, ny);
        Workspace W;
        const double vol = dx * dy;
        if (p == 0) *out = W.reduce(s, d.p, nullptr, nullptr, 1.0, g, 2);
        else if (p == 2) *out = pnorm2(W, s, d.p, g);
        else if (p == 1) *out = W.reduce(s, d.p, nullptr, nullptr, vol, g, 4, 1);
        else *out = std::pow(W.reduce(s, d.p, nullptr, nullptr, vol, g, 4, p), 1.0 / p);
    });
}

int mgp_inner_prod(int nx, int ny, double dx, double dy, const double *a, const double *b, double *out) {
    return guarded([&] {
        check_field(nx, ny, dx, dy, a, out, "inner_prod");
        if (!b) {
            g_last_error = "inner_prod: null array";
            throw Failure(MGP_ERR_ARG);
        }
        hipStream_t s = thread_stream();
        const Grid g = make_grid(nx, ny, dx, dy);
        DevBuf da, db;
        da.alloc(level_elems(nx, ny));
        db.alloc(level_elems(nx, ny));
        upload_block(s, da.p, g, a, 0, 0, nx, ny);
        upload_block(s, db.p, g, b, 0, 0, nx, ny);
        Workspace W;
        *out = W.reduce(s, da.p, db.p, nullptr, dx * dy, g, 0);
    });
}

int mgp_integrate2d(int nx, int ny, double dx, double dy, const double *a, const double *J, double *out) {
    return guarded([&] {
        check_field(nx, ny, dx, dy, a, out, "integrate2d");
        hipStream_t s = thread_stream();
        const Grid g = make_grid(nx, ny, dx, dy);
        DevBuf da, dj;
        da.alloc(level_elems(nx, ny));
        upload_block(s, da.p, g, a, 0, 0, nx, ny);
        if (J) {
            dj.alloc(level_elems(nx, ny));
            upload_block(s, dj.p, g, J, 0, 0, nx, ny);
        }
        Workspace W;
        *out = W.reduce(s, da.p, nullptr, J ? dj.p : nullptr, dx * dy, g, 3);
    });
}

int mgp_extrapolate_ghosts(int nx, int ny, int order, double *phi) {
    return guarded([&] {
        if (order < 0 || order > 3 || !phi || nx < order + 1 || ny < order + 1) {
            g_last_error = "extrapolate_ghosts: order must be 0..3 and nx, ny > order";
            throw Failure(order > 3 ? MGP_ERR_UNSUPPORTED : MGP_ERR_ARG);
        }
        hipStream_t s = thread_stream();
        const Grid g = make_grid(nx, ny, 1.0, 1.0);
        DevBuf d;
        d.alloc(level_elems(nx, ny));
        up_phi(s, d, g, phi);
        launch_extrapolate(s, d.p, g, order);
        down_phi(s, phi, d, g);
        hip_check(hipStreamSynchronize(s), "sync");
    });
}

namespace {

void single_ordering(const mgp_bc *bc, int ordering) {
    if (ordering != MGP_ORDER_LEX && ordering != MGP_ORDER_MC) {
        g_last_error = "ordering must be LEX or MC";
        throw Failure(MGP_ERR_ARG);
    }
    if (ordering == MGP_ORDER_MC)
        for (int k = 0; k < 4; ++k)
            if (bc->type[k] == BC_PERIODIC) {
                g_last_error = "MC ordering needs non-periodic sides";
                throw Failure(MGP_ERR_UNSUPPORTED);
            }
}

}  // namespace

int mgp_relax_jacobi(const mgp_geo *geo, const mgp_bc *bc, int homog, const double *invd, double omega, double tol,
                     int maxiters, int zerophi, const double *rhs, double *phi, int ordering, double *relres,
                     int *iters) {
    return guarded([&] {
        no_lah(bc);
        if (maxiters < 0) {
            g_last_error = "relax_jacobi: maxiters < 0";
            throw Failure(MGP_ERR_ARG);
        }
        Call c(geo, bc);
        single_ordering(bc, ordering);
        const Grid &g = c.geo.g;
        DevBuf f = c.field(), r = c.field(), iv = c.field();
        up_phi(c.s, f, g, phi);
        up_cc(c.s, r, g, rhs);
        up_cc(c.s, iv, g, invd);
        SingleLevel S(c.s, g, c.geo.view(), c.bc.bc, ordering);
        std::vector<double> hist;
        const int n = S.jacobi(homog, iv.p, omega, tol, maxiters, zerophi, r.p, f.p, hist);
        down_phi(c.s, phi, f, g);
        c.finish();
        if (relres && tol > 0.0)
            for (int k = 0; k <= n; ++k) relres[k] = hist[k];
        if (iters) *iters = n;
    });
}

int mgp_solve_bicgstab(const mgp_geo *geo, const mgp_bc *bc, int homog, const double *invd, double tol,
                       int maxiters, int maxrestarts, int zerophi, const double *rhs, double *phi, int ordering,
                       double *relres, int *last) {
    return guarded([&] {
        no_lah(bc);
        if (maxiters < 0 || maxrestarts < 0) {
            g_last_error = "solve_bicgstab: negative iteration limits";
            throw Failure(MGP_ERR_ARG);
        }
        Call c(geo, bc);
        single_ordering(bc, ordering);
        const Grid &g = c.geo.g;
        DevBuf f = c.field(), r = c.field(), iv = c.field();
        up_phi(c.s, f, g, phi);
        up_cc(c.s, r, g, rhs);
        up_cc(c.s, iv, g, invd);
        SingleLevel S(c.s, g, c.geo.view(), c.bc.bc, ordering);
        std::vector<double> hist;
        const int n = S.bicgstab1(homog, iv.p, tol, maxiters, maxrestarts, zerophi, r.p, f.p, hist);
        down_phi(c.s, phi, f, g);
        c.finish();
        if (relres)
            for (int k = 0; k <= n; ++k) relres[k] = hist[k];
        if (last) *last = n;
    });
}

int mgp_solve_bicgstab2(int kind, const mgp_geo *geo, const mgp_bc *bc, int homog, const double *invd, double tol,
                        int maxiters, int maxrestarts, int zerophi, const double *rhs, double *phi, int ordering) {
    return guarded([&] {
        if (kind == 0) no_lah(bc);
        Call c(geo, bc);
        const Grid &g = c.geo.g;
        if (kind == 0 && ordering == MGP_ORDER_MC) {
            single_ordering(bc, ordering);
            DevBuf f = c.field(), r = c.field(), iv = c.field();
            up_phi(c.s, f, g, phi);
            up_cc(c.s, r, g, rhs);
            up_cc(c.s, iv, g, invd);
            SingleLevel S(c.s, g, c.geo.view(), c.bc.bc, ordering);
            std::vector<double> norms;
            S.bicgstab2(homog, iv.p, tol, maxiters, maxrestarts, zerophi, r.p, f.p, norms);
            down_phi(c.s, phi, f, g);
            c.finish();
            return;
        }
        std::vector<DevBuf> w(13);
        for (auto &b : w) b.alloc(level_elems(g.nx, g.ny));
        up_phi(c.s, w[0], g, phi);
        up_cc(c.s, w[1], g, rhs);
        up_cc(c.s, w[2], g, invd);
        if (zerophi) launch_setval(c.s, w[0].p, 0.0, g, 1);
        BottomArgs a;
        a.g = g;
        a.geo = c.geo.view();
        a.bc = c.bc.bc;
        a.phi = w[0].p;
        a.rhs = w[1].p;
        a.invd = w[2].p;
        a.e = w[3].p;
        a.pt = w[4].p;
        a.st = w[5].p;
        a.r = w[6].p;
        a.rt = w[7].p;
        a.p = w[8].p;
        a.t = w[9].p;
        a.v = w[10].p;
        a.mod = w[11].p;
        a.scal = w[12].p;
        a.dc = kind;
        a.order = ordering;
        a.max_iters = maxiters;
        a.max_restarts = maxrestarts;
        a.tol = tol;
        a.homog = homog;
        launch_bicgstab2(c.s, a);
        down_phi(c.s, phi, w[0], g);
        c.finish();
    });
}

void mgp_solver_default_opts(mgp_solver_opts *o) {
    o->method = MGP_METHOD_KRYLOV_MG;
    o->ordering = MGP_ORDER_MC;
    o->numcycles = 1;
    o->smooth_down = 2;
    o->smooth_up = 2;
    o->smooth_bottom = 2;
    o->maxdepth = -1;
    o->homog = 1;
    o->use_graph = 1;
    o->lines = MGP_LINES_AUTO;
    o->refine = MGP_REFINE_FP64;
    o->order = 2;
    o->krylov = MGP_KRYLOV_AUTO;
    o->restart = 0;
}

extern "C++" {
namespace {

// Shared body of the two create calls: make_geo(stream) fills s->geo.
template <class MakeGeo>
int create_solver(mgp_solver **out, int nx, int ny, const mgp_bc *bc, const mgp_solver_opts *opts, MakeGeo make_geo,
                  const mgp_mapping *map = nullptr) {
    *out = nullptr;
    mgp_solver *s = new mgp_solver;
    const int st = guarded([&] {
        check_bc(bc);
        if (opts) s->opts = *opts;
        else mgp_solver_default_opts(&s->opts);
        const mgp_solver_opts &o = s->opts;
        const bool plain = o.method == MGP_METHOD_VCYCLE || o.method == MGP_METHOD_FMG;
        if (plain) no_lah(bc);
        if (plain && o.ordering != MGP_ORDER_LEX && o.ordering != MGP_ORDER_MC) {
            g_last_error = "VCYCLE / FMG ordering must be LEX or MC";
            throw Failure(MGP_ERR_ARG);
        }
        if (o.method == MGP_METHOD_DC_VCYCLE && o.ordering != MGP_ORDER_LEX && o.ordering != MGP_ORDER_RB) {
            g_last_error = "DC_VCYCLE ordering must be LEX or RB";
            throw Failure(MGP_ERR_ARG);
        }
        hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
        make_geo(s->stream, s->geo);
        s->ubc.upload(s->stream, *bc, nx, ny);
        const size_t n = level_elems(nx, ny);
        s->phi.alloc(n);
        s->rhs.alloc(n);
        if (plain || o.method == MGP_METHOD_DC_VCYCLE) {
            s->ref.build(s->stream, s->geo, s->ubc.bc, o.method == MGP_METHOD_DC_VCYCLE, o.ordering, o.maxdepth);
            s->ref.fmg = o.method == MGP_METHOD_FMG;
        } else if (o.method == MGP_METHOD_KRYLOV_MG || o.method == MGP_METHOD_KRYLOV_MG32)
            s->kmg.build(s->stream, s->geo, s->ubc.bc, o, map);
        else {
            g_last_error = "unknown method";
            throw Failure(MGP_ERR_ARG);
        }
        hip_check(hipStreamSynchronize(s->stream), "sync");
    });
    if (st != MGP_OK) {
        mgp_solver_destroy(s);
        return st;
    }
    *out = s;
    return MGP_OK;
}

}  // namespace
}  // extern "C++"

int mgp_solver_create(mgp_solver **out, const mgp_geo *geo, const mgp_bc *bc, const mgp_solver_opts *opts) {
    if (!out) return MGP_ERR_ARG;
    const int st = guarded([&] { check_geo(geo); });
    if (st != MGP_OK) {
        *out = nullptr;
        return st;
    }
    return create_solver(out, geo->nx, geo->ny, bc, opts,
                         [&](hipStream_t stream, DevGeo &dg) { dg.upload(stream, *geo); });
}

int mgp_solver_create_mapped(mgp_solver **out, const mgp_mapping *m, int nx, int ny, double dx, double dy,
                             const mgp_bc *bc, const mgp_solver_opts *opts) {
    if (!out) return MGP_ERR_ARG;
    const int st = guarded([&] {
        check_mapping(m);
        if (nx < 3 || ny < 3 || !(dx > 0) || !(dy > 0)) {
            g_last_error = "bad grid (need nx, ny >= 3, positive spacing)";
            throw Failure(MGP_ERR_ARG);
        }
    });
    if (st != MGP_OK) {
        *out = nullptr;
        return st;
    }
    return create_solver(
        out, nx, ny, bc, opts, [&](hipStream_t stream, DevGeo &dg) { dg.generate(stream, *m, nx, ny, dx, dy); }, m);
}

int mgp_metric_fill(const mgp_mapping *m, int nx, int ny, double dx, double dy, double *J, double *Jxx, double *Jxy,
                    double *Jyx, double *Jyy) {
    return guarded([&] {
        check_mapping(m);
        if (nx < 1 || ny < 1 || !(dx > 0) || !(dy > 0) || !J || !Jxx || !Jxy || !Jyx || !Jyy) {
            g_last_error = "metric_fill: bad grid or null array";
            throw Failure(MGP_ERR_ARG);
        }
        hipStream_t s = thread_stream();
        DevGeo dg;
        dg.generate(s, *m, nx, ny, dx, dy);
        const Grid &g = dg.g;
        download_block(s, J, dg.J.p, g, 0, 0, nx, ny);
        download_block(s, Jxx, dg.Jxx.p, g, 0, 0, nx + 1, ny);
        download_block(s, Jxy, dg.Jxy.p, g, 0, 0, nx + 1, ny);
        download_block(s, Jyx, dg.Jyx.p, g, 0, 0, nx, ny + 1);
        download_block(s, Jyy, dg.Jyy.p, g, 0, 0, nx, ny + 1);
        hip_check(hipStreamSynchronize(s), "sync");
    });
}

int mgp_solver_destroy(mgp_solver *s) {
    if (!s) return MGP_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    hipStream_t st = s->stream;
    delete s;
    if (st) (void)hipStreamDestroy(st);
    return MGP_OK;
}

int mgp_solver_set_rhs(mgp_solver *s, const double *rhs) {
    return guarded([&] {
        up_cc(s->stream, s->rhs, s->geo.g, rhs);
        hip_check(hipStreamSynchronize(s->stream), "sync");
    });
}

int mgp_solver_set_phi(mgp_solver *s, const double *phi) {
    return guarded([&] {
        up_phi(s->stream, s->phi, s->geo.g, phi);
        if (double *lo = s->kmg.phi_lo()) launch_setval(s->stream, lo, 0.0, s->geo.g, 1);
        hip_check(hipStreamSynchronize(s->stream), "sync");
    });
}

int mgp_solver_get_phi(mgp_solver *s, double *phi) {
    return guarded([&] {
        down_phi(s->stream, phi, s->phi, s->geo.g);
        hip_check(hipStreamSynchronize(s->stream), "sync");
    });
}

int mgp_solver_get_phi_lo(mgp_solver *s, double *phi_lo) {
    return guarded([&] {
        double *lo = s->kmg.phi_lo();
        if (!lo) {
            g_last_error = "get_phi_lo: the solver does not carry a double-double solution (refine MGP_REFINE_DD)";
            throw Failure(MGP_ERR_ARG);
        }
        download_block(s->stream, phi_lo, lo, s->geo.g, -1, -1, s->geo.g.nx + 2, s->geo.g.ny + 2);
        hip_check(hipStreamSynchronize(s->stream), "sync");
    });
}

int mgp_solver_copy_phi_from(mgp_solver *s, const mgp_solver *src) {
    return guarded([&] {
        if (src->geo.g.nx != s->geo.g.nx || src->geo.g.ny != s->geo.g.ny) {
            g_last_error = "copy_phi_from: shape mismatch";
            throw Failure(MGP_ERR_ARG);
        }
        const size_t bytes = s->phi.n * sizeof(double);
        hip_check(hipMemcpyAsync(s->phi.p, src->phi.p, bytes, hipMemcpyDeviceToDevice, s->stream), "copy");
        if (double *lo = s->kmg.phi_lo()) {
            if (const double *slo = src->kmg.phi_lo())
                hip_check(hipMemcpyAsync(lo, slo, bytes, hipMemcpyDeviceToDevice, s->stream), "copy");
            else
                launch_setval(s->stream, lo, 0.0, s->geo.g, 1);
        }
        hip_check(hipStreamSynchronize(s->stream), "sync");
    });
}

int mgp_solver_residual(mgp_solver *s, double *rnorm, double *rhsnorm) {
    return guarded([&] {
        if (s->opts.method != MGP_METHOD_KRYLOV_MG && s->opts.method != MGP_METHOD_KRYLOV_MG32) {
            g_last_error = "residual: KRYLOV_MG / KRYLOV_MG32 solvers only";
            throw Failure(MGP_ERR_ARG);
        }
        double rh = 0.0;
        const double rn = s->kmg.residual_norm(s->W, s->stream, s->geo, s->ubc.bc, s->opts.homog, s->phi.p, s->rhs.p,
                                               &rh);
        hip_check(hipStreamSynchronize(s->stream), "sync");
        if (rnorm) *rnorm = rn;
        if (rhsnorm) *rhsnorm = rh;
    });
}

int mgp_solver_project_mean(mgp_solver *s, double *removed) {
    return guarded([&] {
        if (s->opts.method != MGP_METHOD_KRYLOV_MG && s->opts.method != MGP_METHOD_KRYLOV_MG32) {
            g_last_error = "project_mean: KRYLOV_MG / KRYLOV_MG32 solvers only";
            throw Failure(MGP_ERR_ARG);
        }
        const double m = s->kmg.project_mean(s->stream, s->geo, s->rhs.p);
        if (removed) *removed = m;
    });
}

int mgp_solver_copy_rhs_from(mgp_solver *s, const mgp_solver *src) {
    return guarded([&] {
        if (src->geo.g.nx != s->geo.g.nx || src->geo.g.ny != s->geo.g.ny) {
            g_last_error = "copy_rhs_from: shape mismatch";
            throw Failure(MGP_ERR_ARG);
        }
        hip_check(hipMemcpyAsync(s->rhs.p, src->rhs.p, s->rhs.n * sizeof(double), hipMemcpyDeviceToDevice,
                                 s->stream),
                  "copy");
        hip_check(hipStreamSynchronize(s->stream), "sync");
    });
}

namespace {

int run_solve(mgp_solver *s, double tol, int maxiters, int zerophi, double *relres, int *iters) {
    const mgp_solver_opts &o = s->opts;
    int it = 0, oc = MGP_MAXITERS;
    if (o.method == MGP_METHOD_KRYLOV_MG || o.method == MGP_METHOD_KRYLOV_MG32) {
        oc = s->kmg.solve(s->W, s->stream, s->geo, s->ubc.bc, o.homog, s->phi.p, s->rhs.p, tol, maxiters, zerophi,
                          relres, &it);
        s->last_cycles = s->kmg.last_cycles;
    } else {
        oc = s->ref.solve(s->W, s->stream, s->geo, s->ubc.bc, o.homog, s->phi.p, s->rhs.p, tol, maxiters,
                          o.numcycles, o.smooth_down, o.smooth_up, o.smooth_bottom, zerophi, relres, &it);
        s->last_cycles = it;
    }
    if (iters) *iters = it;
    return oc;
}

void need_flux(const mgp_solver *s) {
    if (!s->xflux.p) {
        g_last_error = "no resident fluxes: call mgp_solver_set_flux first";
        throw Failure(MGP_ERR_ARG);
    }
}

}  // namespace

int mgp_solver_solve(mgp_solver *s, double tol, int maxiters, int zerophi, double *relres, int *iters,
                     int *outcome) {
    return guarded([&] {
        const int oc = run_solve(s, tol, maxiters, zerophi, relres, iters);
        if (outcome) *outcome = oc;
    });
}

int mgp_solver_set_flux(mgp_solver *s, const double *xflux, const double *yflux) {
    return guarded([&] {
        const Grid &g = s->geo.g;
        if (!s->xflux.p) {
            s->xflux.alloc(level_elems(g.nx, g.ny));
            s->yflux.alloc(level_elems(g.nx, g.ny));
            s->div.alloc(level_elems(g.nx, g.ny));
        }
        upload_block(s->stream, s->xflux.p, g, xflux, 0, 0, g.nx + 1, g.ny);
        upload_block(s->stream, s->yflux.p, g, yflux, 0, 0, g.nx, g.ny + 1);
        hip_check(hipStreamSynchronize(s->stream), "sync");
    });
}

int mgp_solver_get_flux(mgp_solver *s, double *xflux, double *yflux) {
    return guarded([&] {
        need_flux(s);
        const Grid &g = s->geo.g;
        download_block(s->stream, xflux, s->xflux.p, g, 0, 0, g.nx + 1, g.ny);
        download_block(s->stream, yflux, s->yflux.p, g, 0, 0, g.nx, g.ny + 1);
        hip_check(hipStreamSynchronize(s->stream), "sync");
    });
}

int mgp_solver_project(mgp_solver *s, double tol, int maxiters, double *divnorm, int *iters, int *outcome) {
    return guarded([&] {
        need_flux(s);
        const Grid &g = s->geo.g;
        const int homog = s->opts.homog;
        launch_div(s->stream, s->rhs.p, s->xflux.p, s->yflux.p, g);
        const double d0 = s->W.reduce(s->stream, s->rhs.p, nullptr, nullptr, 1.0, g, 2);
        const int oc = run_solve(s, tol, maxiters, 1, nullptr, iters);
        launch_fill_ghosts(s->stream, s->phi.p, g, s->geo.view(), s->ubc.bc, homog, 0);
        launch_grad_subtract(s->stream, s->xflux.p, s->yflux.p, s->phi.p, g, s->geo.view(), s->ubc.bc, homog);
        launch_div(s->stream, s->div.p, s->xflux.p, s->yflux.p, g);
        const double d1 = s->W.reduce(s->stream, s->div.p, nullptr, nullptr, 1.0, g, 2);
        if (divnorm) {
            divnorm[0] = d0;
            divnorm[1] = d1;
        }
        if (outcome) *outcome = oc;
    });
}

int mgp_solver_field(mgp_solver *s, int which, double **ptr, int *pitch) {
    if (!s || which < 0 || which > 3 || (which >= 2 && !s->xflux.p)) return MGP_ERR_ARG;
    const DevBuf *f[4] = {&s->phi, &s->rhs, &s->xflux, &s->yflux};
    *ptr = f[which]->p;
    *pitch = s->geo.g.pitch;
    return MGP_OK;
}

void *mgp_solver_stream(mgp_solver *s) { return s ? (void *)s->stream : nullptr; }

int mgp_solver_last_cycles(const mgp_solver *s) { return s ? s->last_cycles : 0; }

int mgp_solver_time_kernel(mgp_solver *s, int kernel, int reps, int cold, double *avg_ms, double *bytes) {
    return guarded([&] {
        const bool kmg = s->opts.method == MGP_METHOD_KRYLOV_MG || s->opts.method == MGP_METHOD_KRYLOV_MG32;
        if ((kernel >= 4 && kernel <= 6 && s->opts.method == MGP_METHOD_DC_VCYCLE) ||
            ((kernel == 7 || kernel == 8) && s->opts.method == MGP_METHOD_VCYCLE)) {
            *avg_ms = s->ref.time_relax(s->stream, kernel, reps, bytes);
            return;
        }
        if (!kmg || reps <= 0 || kernel < 0 || kernel > 3) {
            g_last_error = "time_kernel: KRYLOV_MG / KRYLOV_MG32 solvers with kernel 0 to 3, or DC_VCYCLE with "
                           "kernel 4 to 6; reps > 0";
            throw Failure(MGP_ERR_ARG);
        }
        *avg_ms = s->kmg.time_kernel(s->stream, kernel, reps, cold, bytes);
    });
}

extern "C++" {
namespace {

// The problem list of a batch or server (validated), and its options.
void batch_problems(int n, const mgp_bc *bcs, const mgp_solver_opts *opts,
                const std::function<void(int, BatchKMG::Problem &)> &fill, std::vector<BatchKMG::Problem> &probs,
                mgp_solver_opts &o) {
    if (n <= 0 || !bcs) {
        g_last_error = "batch: n > 0 problems and their BCs";
        throw Failure(MGP_ERR_ARG);
    }
    if (opts) o = *opts;
    else mgp_solver_default_opts(&o);
    if ((o.method != MGP_METHOD_KRYLOV_MG && o.method != MGP_METHOD_KRYLOV_MG32) || !o.homog) {
        g_last_error = "batch: method KRYLOV_MG or KRYLOV_MG32 with homogeneous BCs (homog = 1)";
        throw Failure(MGP_ERR_UNSUPPORTED);
    }
    probs.assign(n, BatchKMG::Problem());
    for (int k = 0; k < n; ++k) {
        check_bc(&bcs[k]);
        BatchKMG::Problem &q = probs[k];
        q.geo = nullptr;
        q.map = nullptr;
        for (int side = 0; side < 4; ++side) {
            q.bc.type[side] = bcs[k].type[side];
            q.bc.mode[side] = MODE_UNIFORM;
            q.bc.val[side] = 0.0;
            q.bc.data[side] = nullptr;
        }
        fill(k, q);
    }
}

int create_batch(mgp_batch **out, int n, const mgp_bc *bcs, const mgp_solver_opts *opts,
                 const std::function<void(int, BatchKMG::Problem &)> &fill) {
    if (!out) return MGP_ERR_ARG;
    *out = nullptr;
    mgp_batch *b = new mgp_batch;
    const int st = guarded([&] {
        std::vector<BatchKMG::Problem> probs;
        mgp_solver_opts o;
        batch_problems(n, bcs, opts, fill, probs, o);
        hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
        b->b = new BatchKMG(b->stream, probs, o);
    });
    if (st != MGP_OK) {
        delete b;
        return st;
    }
    *out = b;
    return MGP_OK;
}

void batch_index(const mgp_batch *b, int k) {
    if (!b || !b->b || k < 0 || k >= b->b->count()) {
        g_last_error = "batch: problem index out of range";
        throw Failure(MGP_ERR_ARG);
    }
}

}  // namespace
}  // extern "C++"

int mgp_batch_create(mgp_batch **out, int n, const mgp_geo *geos, const mgp_bc *bcs, const mgp_solver_opts *opts) {
    return create_batch(out, n, bcs, opts, [&](int k, BatchKMG::Problem &q) {
        if (!geos) {
            g_last_error = "batch: null geometry array";
            throw Failure(MGP_ERR_ARG);
        }
        check_geo(&geos[k]);
        q.nx = geos[k].nx;
        q.ny = geos[k].ny;
        q.dx = geos[k].dx;
        q.dy = geos[k].dy;
        q.geo = &geos[k];
    });
}

int mgp_batch_create_mapped(mgp_batch **out, int n, const mgp_mapping *maps, const int *nx, const int *ny,
                            const double *dx, const double *dy, const mgp_bc *bcs, const mgp_solver_opts *opts) {
    return create_batch(out, n, bcs, opts, [&](int k, BatchKMG::Problem &q) {
        if (!maps || !nx || !ny || !dx || !dy) {
            g_last_error = "batch: null mapping or grid arrays";
            throw Failure(MGP_ERR_ARG);
        }
        check_mapping(&maps[k]);
        q.nx = nx[k];
        q.ny = ny[k];
        q.dx = dx[k];
        q.dy = dy[k];
        q.map = &maps[k];
    });
}

int mgp_batch_destroy(mgp_batch *b) {
    delete b;
    return MGP_OK;
}

int mgp_batch_count(const mgp_batch *b) { return b && b->b ? b->b->count() : 0; }

int mgp_batch_set_rhs(mgp_batch *b, int k, const double *rhs) {
    return guarded([&] {
        batch_index(b, k);
        if (!rhs) {
            g_last_error = "batch: null rhs";
            throw Failure(MGP_ERR_ARG);
        }
        const Grid &g = b->b->grid(k);
        upload_block(b->stream, b->b->vec(k, BV_RHS), g, rhs, 0, 0, g.nx, g.ny);
        hip_check(hipStreamSynchronize(b->stream), "sync");
    });
}

int mgp_batch_get_phi(mgp_batch *b, int k, double *phi) {
    return guarded([&] {
        batch_index(b, k);
        if (!phi) {
            g_last_error = "batch: null phi";
            throw Failure(MGP_ERR_ARG);
        }
        const Grid &g = b->b->grid(k);
        download_block(b->stream, phi, b->b->vec(k, BV_PHI), g, 0, 0, g.nx, g.ny);
        hip_check(hipStreamSynchronize(b->stream), "sync");
    });
}

int mgp_batch_solve(mgp_batch *b, double tol, int maxiters, int zerophi, int *iters, int *outcomes) {
    return guarded([&] {
        if (!b || !b->b || maxiters < 0 || !(tol >= 0.0)) {
            g_last_error = "batch solve: bad arguments";
            throw Failure(MGP_ERR_ARG);
        }
        b->b->solve(tol, maxiters, zerophi, iters, outcomes);
    });
}

int mgp_batch_time_kernel(mgp_batch *b, int kernel, int reps, double *avg_ms, double *bytes) {
    return guarded([&] {
        if (!b || !b->b || kernel != 0 || reps < 1 || !avg_ms || !bytes) {
            g_last_error = "batch_time_kernel: kernel must be 0, reps >= 1";
            throw Failure(MGP_ERR_ARG);
        }
        *avg_ms = b->b->time_fine_xline(reps, bytes);
    });
}

int mgp_batch_last_cycles(const mgp_batch *b) { return b && b->b ? b->b->last_cycles : 0; }

namespace {
void export_tuning(const std::vector<mgp::BatchKMG::TuneEntry> &t, int *table, int max_entries, int *n_entries) {
    if (n_entries) *n_entries = (int)t.size();
    if (!table) return;
    for (size_t k = 0; k < t.size() && (int)k < max_entries; ++k) {
        table[4 * k] = t[k].nx;
        table[4 * k + 1] = t[k].bc;
        table[4 * k + 2] = t[k].fp32;
        table[4 * k + 3] = t[k].threads;
    }
}
}  // namespace

int mgp_batch_tuning(const mgp_batch *b, int *table, int max_entries, int *n_entries) {
    return guarded([&] {
        if (!b || !b->b || max_entries < 0) {
            g_last_error = "batch_tuning: no batch";
            throw Failure(MGP_ERR_ARG);
        }
        export_tuning(b->b->tuning(), table, max_entries, n_entries);
    });
}

int mgp_batch_set_tuning(mgp_batch *b, const int *table, int n_entries, int *changed) {
    return guarded([&] {
        if (!b || !b->b || n_entries < 0 || (n_entries > 0 && !table)) {
            g_last_error = "batch_set_tuning: no batch or table";
            throw Failure(MGP_ERR_ARG);
        }
        std::vector<mgp::BatchKMG::TuneEntry> t;
        for (int k = 0; k < n_entries; ++k)
            t.push_back(mgp::BatchKMG::TuneEntry{table[4 * k], table[4 * k + 1], table[4 * k + 2], table[4 * k + 3]});
        const int c = b->b->set_tuning(t);
        if (changed) *changed = c;
    });
}

int mgp_batch_tune(mgp_batch *b, int reps, int *table, int max_entries, int *n_entries) {
    return guarded([&] {
        if (!b || !b->b || reps < 1 || max_entries < 0) {
            g_last_error = "batch_tune: reps >= 1";
            throw Failure(MGP_ERR_ARG);
        }
        export_tuning(b->b->tune(reps), table, max_entries, n_entries);
    });
}

int mgp_server_create(mgp_server **out, int n, const mgp_mapping *maps, const int *nx, const int *ny,
                      const double *dx, const double *dy, const mgp_bc *bcs, const mgp_solver_opts *opts,
                      double hbm_budget) {
    if (!out) return MGP_ERR_ARG;
    *out = nullptr;
    mgp_server *v = new mgp_server;
    const int st = guarded([&] {
        std::vector<BatchKMG::Problem> probs;
        mgp_solver_opts o;
        batch_problems(n, bcs, opts, [&](int k, BatchKMG::Problem &q) {
            if (!maps || !nx || !ny || !dx || !dy) {
                g_last_error = "server: null mapping or grid arrays";
                throw Failure(MGP_ERR_ARG);
            }
            check_mapping(&maps[k]);
            q.nx = nx[k];
            q.ny = ny[k];
            q.dx = dx[k];
            q.dy = dy[k];
            q.map = &maps[k];
        }, probs, o);
        v->srv = new Server(probs, o, hbm_budget);
    });
    if (st != MGP_OK) {
        delete v;
        return st;
    }
    *out = v;
    return MGP_OK;
}

int mgp_server_destroy(mgp_server *v) {
    delete v;
    return MGP_OK;
}

double mgp_server_footprint(const mgp_server *v) { return v && v->srv ? (double)v->srv->footprint() : 0.0; }

int mgp_server_add_slot(mgp_server *v, const mgp_mapping *map, int nx, int ny, double dx, double dy, const mgp_bc *bc,
                        int *slot) {
    return guarded([&] {
        if (!v || !v->srv || !slot) {
            g_last_error = "server add_slot: null server or slot";
            throw Failure(MGP_ERR_ARG);
        }
        check_mapping(map);
        std::vector<BatchKMG::Problem> probs;
        mgp_solver_opts o;
        batch_problems(1, bc, nullptr, [&](int, BatchKMG::Problem &q) {
            q.nx = nx;
            q.ny = ny;
            q.dx = dx;
            q.dy = dy;
            q.map = map;
        }, probs, o);
        *slot = v->srv->add_slot(probs[0], map);
    });
}

int mgp_server_evict(mgp_server *v, int slot) {
    return guarded([&] {
        if (!v || !v->srv) {
            g_last_error = "server evict: null server";
            throw Failure(MGP_ERR_ARG);
        }
        v->srv->evict(slot);
    });
}

int mgp_server_submit(mgp_server *v, int slot, const double *rhs, double tol, int maxiters, int warm,
                      long long *ticket) {
    return guarded([&] {
        if (!v || !v->srv || !ticket) {
            g_last_error = "server submit: null server or ticket";
            throw Failure(MGP_ERR_ARG);
        }
        *ticket = v->srv->submit(slot, rhs, tol, maxiters, warm);
    });
}

int mgp_server_step(mgp_server *v, int *running, int *queued) {
    return guarded([&] {
        if (!v || !v->srv) {
            g_last_error = "server step: null server";
            throw Failure(MGP_ERR_ARG);
        }
        const int r = v->srv->step();
        if (running) *running = r;
        if (queued) *queued = v->srv->queued();
    });
}

int mgp_server_result(mgp_server *v, long long ticket, double *phi, int *iters, int *outcome, int *done) {
    return guarded([&] {
        if (!v || !v->srv || !done) {
            g_last_error = "server result: null server or done flag";
            throw Failure(MGP_ERR_ARG);
        }
        *done = v->srv->result(ticket, phi, iters, outcome);
    });
}

int mgp_device_synchronize(void) {
    return guarded([&] { hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize"); });
}

int mgp_vcycle(int kind, const mgp_geo *geo, const mgp_bc *bc, int homog, double tol, int maxiters, int maxdepth,
               int numcycles, int smooth_down, int smooth_up, int smooth_bottom, int zerophi, double *phi,
               double *rhs, int ordering, double *relres, int *iters, int *outcome) {
    mgp_solver_opts o;
    mgp_solver_default_opts(&o);
    if (kind < 0 || kind > 2) {
        g_last_error = "vcycle: kind must be 0 (MGPoisson2D::vcycle), 1 (DeferredMGPoisson2D::vcycle) or 2 (fmg)";
        return MGP_ERR_ARG;
    }
    o.method = kind == 0 ? MGP_METHOD_VCYCLE : (kind == 1 ? MGP_METHOD_DC_VCYCLE : MGP_METHOD_FMG);
    o.ordering = ordering;
    o.numcycles = numcycles;
    o.smooth_down = smooth_down;
    o.smooth_up = smooth_up;
    o.smooth_bottom = smooth_bottom;
    o.maxdepth = maxdepth;
    o.homog = homog;
    mgp_solver *s = nullptr;
    int st = mgp_solver_create(&s, geo, bc, &o);
    if (st != MGP_OK) return st;
    st = guarded([&] {
        up_phi(s->stream, s->phi, s->geo.g, phi);
        up_cc(s->stream, s->rhs, s->geo.g, rhs);
    });
    if (st == MGP_OK) st = mgp_solver_solve(s, tol, maxiters, zerophi, relres, iters, outcome);
    if (st == MGP_OK)
        st = guarded([&] {
            down_phi(s->stream, phi, s->phi, s->geo.g);
            down_cc(s->stream, rhs, s->rhs, s->geo.g);
            hip_check(hipStreamSynchronize(s->stream), "sync");
        });
    mgp_solver_destroy(s);
    return st;
}

}  // extern "C"
