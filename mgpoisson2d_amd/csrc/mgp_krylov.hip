// mgp_krylov.hip -- the resident fast solver (see mgp_krylov.h).
//
// Outer iteration: right-preconditioned BiCGStab in FP64 on A = the
// reference's 9-point operator with the caller's BCs (the same kernels the
// drop-in path uses, bit-identical to poisson2d.f90 compute_laplacian).
// Convergence is declared on the true residual ||rhs - A phi|| relative to
// the initial one, i.e. the relres of MGPoisson2D.f90:614-615.
//
// Preconditioner: one V-cycle on a hierarchy coarsened in y only (1x2, the
// reference's restrict supports any ratio, MGPoisson2D.f90:145-158), each
// level re-discretised from restricted metrics with the reference stencil,
// x-line zebra relaxation (even rows, odd rows, then the two boundary rows,
// whose extrapolated cross terms reach two rows in), 1x2 averaging
// restriction and linear-in-y prolongation (new: the reference has no
// semicoarsened prolongation, MGPoisson2D.f90:385-388).
//
// Mixed precision (MGP_METHOD_KRYLOV_MG32): the same V-cycle on FP32 copies
// of the level metrics and stencil tables, with FP32 corrections and
// residuals.  The Krylov vectors, the operator A, the true residual and the
// solution update stay FP64; the preconditioner's input is rounded to FP32 on
// entry and its output widened on exit.  A preconditioner only has to
// approximate A^-1, so its rounding changes the iteration count, not the
// accuracy of the converged solution, which is set by the FP64 residual
// (with restarts from the FP64 true residual: iterative refinement).
//
// Orientation (MGP_LINES_Y): grids whose rows are too long for one
// workgroup's line solve (nx > 4111) relax columns instead: the hierarchy is
// built on the transposed problem (xi and eta exchanged, the face metrics
// Jxx/Jyy and Jxy/Jyx exchanged and transposed, the sides permuted), whose
// rows are the original columns, so its y-semicoarsening is x-semicoarsening
// of the original grid.  A preconditioner application transposes its input
// into the hierarchy and its output back (two extra passes over one field).
//
// Double-double refinement (MGP_REFINE_DD): the solution is phi + phi_lo,
// each BiCGStab run accumulates its correction in a separate FP64 field x,
// and at every true-residual check x is folded into phi + phi_lo and the
// residual rhs - A (phi + phi_lo) is formed in double-double (mgp_dd.h).
// Each refinement run then starts from an accurate residual, so the relative
// residual can fall below the rounding of the FP64 operator (about 2e-10 of
// the rhs at 2048^2 on the stretched map).
//
// Fourth order (o.order == 4): A is the fourth-order finite-volume operator
// (mgp_fourth.hip) on face-averaged metrics of the mapping; the V-cycle stays
// the second-order one on the point metrics.  As in the reference's deferred
// correction (MGPoisson2D_DD.f90:589-862: outer residual with the full
// operator, inner cycle with a low-order one), the converged solution is the
// high-order operator's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "mgp_kernels.h"
#include "mgp_krylov.h"

namespace mgp {

namespace {

constexpr int MIN_COARSE_NY = 4;


}  // namespace

struct KLevel {
    Grid g;
    int off = 0;  // the fine row of the next coarser level's row 0 (semicoarsen_step)
    Geo geo;
    GeoF geo32;
    Bc bc;
    DevBuf J, Jxx, Jxy, Jyx, Jyy, e, r, w, K;
    DevBufF Jxx32, Jxy32, Jyx32, Jyy32, e32, r32, K32;  // mixed precision only
};

struct KrylovMG::Impl {
    std::vector<KLevel> L;
    bool mixed = false;
    bool tr = false;     // y-lines: the hierarchy is the transposed problem's
    bool dd = false;     // double-double refinement
    int order = 2;       // outer operator: 2 reference, 4 fourth-order
    Grid top_g;          // the caller's finest grid
    DevBuf r, rhat, p, v, s, t, phat, shat, scal, partial;
    DevBuf lo, x;        // dd: low part of the solution, the run's correction
    DevBuf bk, bk_lo;    // the solution where the current run started (undo of a failed run)
    DevBuf m4[8];        // fourth order: face-averaged metrics (A4, Bx4, dA, dBx, C4, By4, dC, dBy)
    DevBuf sums;         // project_mean: the two compensated sums
    double *host = nullptr;  // pinned copy of the device scalars
    int sd = 2, su = 2, sb = 2;
    int ncyc = 1;        // V-cycles per preconditioner application
    // flexible GMRES (the outer iteration of the fourth-order operator): basis
    // V_0..V_m and preconditioned vectors Z_0..Z_{m-1} (level layout, stride
    // r.n), Gram-Schmidt coefficients, dot partials, the solution coefficients
    bool fgmres = false;
    bool fallback = false;  // krylov AUTO with BiCGStab: FGMRES takes over after a breakdown
    int m = 30;
    DevBuf gv, gz, gh, gpart, gy;
    double *ghost = nullptr;  // pinned: 3 (m + 2) coefficients, then m solution coefficients
    bool use_graph = true;
    hipGraphExec_t graph = nullptr;
    ~Impl() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (host) (void)hipHostFree(host);
        if (ghost) (void)hipHostFree(ghost);
    }
};

KrylovMG::KrylovMG() : impl(nullptr) {}
KrylovMG::~KrylovMG() { delete impl; }

static double side_sign(const Bc &bc, int side) { return bc.type[side] == BC_DIRI ? -1.0 : 1.0; }

// The BCs of the transposed problem: its x sides are the original y sides.
static Bc transposed_bc(const Bc &b) {
    Bc t = b;
    const int perm[4] = {YLO, YHI, XLO, XHI};  // side k of the transposed problem is side perm[k]
    for (int k = 0; k < 4; ++k) {
        t.type[k] = b.type[perm[k]];
        t.mode[k] = b.mode[perm[k]];
        t.val[k] = b.val[perm[k]];
        t.data[k] = b.data[perm[k]];
    }
    return t;
}

// The FGMRES basis (m + 1 vectors), preconditioned vectors (m) and
// coefficient buffers, allocated once (at build for krylov FGMRES, at the
// first breakdown for AUTO).
void KrylovMG::ensure_fgmres(const Grid &g) {
    Impl &K = *impl;
    if (K.ghost) return;
    const int m = K.m;
    const size_t n0 = level_elems(g.nx, g.ny);
    K.gv.alloc((size_t)(m + 1) * n0);
    K.gz.alloc((size_t)m * n0);
    K.gh.alloc((size_t)3 * (m + 2));
    K.gpart.alloc((size_t)mdot_blocks(g) * (m + 1));
    K.gy.alloc((size_t)m);
    hip_check(hipHostMalloc((void **)&K.ghost, (size_t)(4 * m + 6) * sizeof(double)), "hipHostMalloc");
}

void KrylovMG::build(hipStream_t s, const DevGeo &top, const Bc &user_bc, const mgp_solver_opts &o,
                     const mgp_mapping *map) {
    if (top.g.ny < 4 || top.g.nx < 4) {
        g_last_error = "KRYLOV_MG needs nx, ny >= 4";
        throw Failure(MGP_ERR_ARG);
    }
    for (int k = 0; k < 4; ++k)
        if (user_bc.type[k] != BC_DIRI && user_bc.type[k] != BC_NEUM) {
            g_last_error = "KRYLOV_MG supports Dirichlet and Neumann sides";
            throw Failure(MGP_ERR_UNSUPPORTED);
        }
    if (o.lines < MGP_LINES_AUTO || o.lines > MGP_LINES_Y || (o.refine != MGP_REFINE_FP64 && o.refine != MGP_REFINE_DD) ||
        (o.order != 2 && o.order != 4) || o.krylov < MGP_KRYLOV_AUTO || o.krylov > MGP_KRYLOV_FGMRES ||
        o.restart < 0 || o.restart > 200) {
        g_last_error = "KRYLOV_MG: lines must be AUTO, X or Y, refine FP64 or DD, order 2 or 4, krylov AUTO, "
                       "BICGSTAB or FGMRES, restart 0..200";
        throw Failure(MGP_ERR_ARG);
    }
    // AUTO: x-lines where the cycle's element type solves the rows whole,
    // else y-lines where it solves the columns whole, else rows in segments
    const size_t celem = o.method == MGP_METHOD_KRYLOV_MG32 ? sizeof(float) : sizeof(double);
    const bool tr = o.lines == MGP_LINES_Y ||
                    (o.lines == MGP_LINES_AUTO && !xline_whole(top.g.nx, celem) && xline_whole(top.g.ny, celem));
    if (xline_segments(tr ? top.g.ny : top.g.nx, celem) == 0) {
        g_last_error = "KRYLOV_MG: no line solve for this size (lines of at most 4111 cells, 8207 in the FP32 "
                       "cycle, or cut into at most 64 equal segments of that length; with y-lines, columns)";
        throw Failure(MGP_ERR_UNSUPPORTED);
    }
    if (o.refine == MGP_REFINE_DD && !o.homog) {
        g_last_error = "KRYLOV_MG: the double-double residual needs homogeneous BCs (homog 1)";
        throw Failure(MGP_ERR_UNSUPPORTED);
    }
    if (o.order == 4) {
        for (int k = 0; k < 4; ++k)
            if (user_bc.type[k] != BC_NEUM) {
                g_last_error = "KRYLOV_MG: the fourth-order operator needs Neumann sides";
                throw Failure(MGP_ERR_UNSUPPORTED);
            }
        if (!map || map->kind != MGP_MAP_WARPED || top.g.nx < 5 || top.g.ny < 5) {
            g_last_error = "KRYLOV_MG: the fourth-order operator needs a solver made from an MGP_MAP_WARPED "
                           "mapping (mgp_solver_create_mapped), nx, ny >= 5";
            throw Failure(MGP_ERR_UNSUPPORTED);
        }
        if (o.refine == MGP_REFINE_DD) {
            g_last_error = "KRYLOV_MG: the double-double residual is built for the second-order operator only";
            throw Failure(MGP_ERR_UNSUPPORTED);
        }
    }
    const bool mixed = o.method == MGP_METHOD_KRYLOV_MG32;
    delete impl;
    impl = new Impl;
    impl->mixed = mixed;
    impl->tr = tr;
    impl->dd = o.refine == MGP_REFINE_DD;
    impl->order = o.order;
    impl->top_g = top.g;
    impl->sd = o.smooth_down > 0 ? o.smooth_down : 1;
    impl->ncyc = o.numcycles > 1 ? o.numcycles : 1;
    impl->fgmres = o.krylov == MGP_KRYLOV_FGMRES || (o.krylov == MGP_KRYLOV_AUTO && o.order == 4);
    impl->m = o.restart > 0 ? o.restart : 30;
    impl->fallback = o.krylov == MGP_KRYLOV_AUTO && !impl->fgmres;
    impl->su = o.smooth_up > 0 ? o.smooth_up : 1;
    impl->sb = o.smooth_bottom > 0 ? o.smooth_bottom : 1;
    std::vector<KLevel> &L = impl->L;
    // the hierarchy's finest grid and BCs: the caller's, or the transposed ones
    const Grid g0 = tr ? make_grid(top.g.ny, top.g.nx, top.g.dy, top.g.dx) : top.g;
    const Bc bc0 = tr ? transposed_bc(user_bc) : user_bc;
    int ny = g0.ny;
    double dy = g0.dy;
    const int maxlev = o.maxdepth < 0 ? 64 : o.maxdepth + 1;
    while (true) {
        L.emplace_back();
        KLevel &l = L.back();
        l.g = make_grid(g0.nx, ny, g0.dx, dy);
        const size_t n = level_elems(l.g.nx, l.g.ny);
        for (DevBuf *b : {&l.J, &l.Jxx, &l.Jxy, &l.Jyx, &l.Jyy}) b->alloc(n);
        // FP64 level fields: the FP64 cycle, and on the finest level the
        // kernel timings (time_kernel)
        if (!mixed || L.size() == 1)
            for (DevBuf *b : {&l.e, &l.r, &l.w}) b->alloc(n);
        l.K.alloc((size_t)ring_table_elems(l.g));
        l.bc = mg_bc(bc0, MODE_UNIFORM);
        const size_t bytes = n * sizeof(double);
        if (L.size() == 1 && tr) {
            // transposed metrics: x faces of the transposed problem are the
            // original y faces (Jxx' = Jyy^T, Jxy' = Jyx^T), y faces the x faces
            const int P = top.g.pitch, nx0 = top.g.nx, ny0 = top.g.ny;
            launch_transpose(s, l.J.p, l.g.pitch, top.J.p, P, nx0, ny0);
            launch_transpose(s, l.Jxx.p, l.g.pitch, top.Jyy.p, P, nx0, ny0 + 1);
            launch_transpose(s, l.Jxy.p, l.g.pitch, top.Jyx.p, P, nx0, ny0 + 1);
            launch_transpose(s, l.Jyx.p, l.g.pitch, top.Jxy.p, P, nx0 + 1, ny0);
            launch_transpose(s, l.Jyy.p, l.g.pitch, top.Jxx.p, P, nx0 + 1, ny0);
        } else if (L.size() == 1) {
            hip_check(hipMemcpyAsync(l.J.p, top.J.p, bytes, hipMemcpyDeviceToDevice, s), "copy");
            hip_check(hipMemcpyAsync(l.Jxx.p, top.Jxx.p, bytes, hipMemcpyDeviceToDevice, s), "copy");
            hip_check(hipMemcpyAsync(l.Jxy.p, top.Jxy.p, bytes, hipMemcpyDeviceToDevice, s), "copy");
            hip_check(hipMemcpyAsync(l.Jyx.p, top.Jyx.p, bytes, hipMemcpyDeviceToDevice, s), "copy");
            hip_check(hipMemcpyAsync(l.Jyy.p, top.Jyy.p, bytes, hipMemcpyDeviceToDevice, s), "copy");
        } else {
            // from the fine level's rows off.. (off = 1: its bottom row is left out)
            const KLevel &f = L[L.size() - 2];
            const size_t sh = (size_t)f.off * f.g.pitch;
            launch_restrict_cc(s, l.J.p, f.J.p + sh, l.g, f.g.pitch, 1, 2, nullptr, nullptr);
            launch_restrict_xf(s, l.Jxx.p, f.Jxx.p + sh, l.g, f.g.pitch, 1, 2);
            launch_restrict_xf(s, l.Jxy.p, f.Jxy.p + sh, l.g, f.g.pitch, 1, 2);
            launch_restrict_yf(s, l.Jyx.p, f.Jyx.p + sh, l.g, f.g.pitch, 1, 2);
            launch_restrict_yf(s, l.Jyy.p, f.Jyy.p + sh, l.g, f.g.pitch, 1, 2);
        }
        l.geo = Geo{l.J.p, l.Jxx.p, l.Jxy.p, l.Jyx.p, l.Jyy.p};
        launch_ring_coef(s, l.K.p, l.g, l.geo, l.bc);
        if (mixed) {
            // FP32 copies of the whole padded arrays (the metrics and the
            // stencil table are formed in FP64 and rounded once)
            DevBuf *src[4] = {&l.Jxx, &l.Jxy, &l.Jyx, &l.Jyy};
            DevBufF *dst[4] = {&l.Jxx32, &l.Jxy32, &l.Jyx32, &l.Jyy32};
            for (int q = 0; q < 4; ++q) {
                dst[q]->alloc(n);
                launch_convert_flat(s, dst[q]->p, src[q]->p, n);
            }
            l.K32.alloc(l.K.n);
            launch_convert_flat(s, l.K32.p, l.K.p, l.K.n);
            l.e32.alloc(n);
            l.r32.alloc(n);
            l.geo32 = GeoF{nullptr, l.Jxx32.p, l.Jxy32.p, l.Jyx32.p, l.Jyy32.p};
        }
        if ((int)L.size() >= maxlev) break;
        const int nyc = semicoarsen_step(ny, MIN_COARSE_NY, bc0.type[YLO] == BC_NEUM, bc0.type[YHI] == BC_NEUM, l.off);
        if (nyc == 0) break;
        ny = nyc;
        dy *= 2.0;
    }
    const size_t n0 = level_elems(top.g.nx, top.g.ny);
    for (DevBuf *b : {&impl->r, &impl->rhat, &impl->p, &impl->v, &impl->s, &impl->t, &impl->phat, &impl->shat})
        b->alloc(n0);
    impl->bk.alloc(n0);
    if (impl->dd) {
        impl->lo.alloc(n0);
        impl->x.alloc(n0);
        impl->bk_lo.alloc(n0);
    }
    if (impl->order == 4) {
        for (DevBuf &b : impl->m4) b.alloc(n0);
        DevBuf *m = impl->m4;
        launch_metric4(s, *map, top.g, Geo4W{m[0].p, m[1].p, m[2].p, m[3].p, m[4].p, m[5].p, m[6].p, m[7].p});
    }
    if (impl->fgmres) ensure_fgmres(top.g);
    impl->sums.alloc(2);
    impl->scal.alloc(SC_COUNT);
    impl->partial.alloc((size_t)reduce_blocks(top.g) + 8);
    hip_check(hipHostMalloc((void **)&impl->host, SC_COUNT * sizeof(double)), "hipHostMalloc");
    impl->use_graph = o.use_graph != 0;
    built = true;
}

template <typename T, typename G>
static void zebra(hipStream_t s, const Grid &g, const G &geo, const T *K, T *e, const T *f) {
    const int ny = g.ny;
    // Even interior rows 2, 4, .. <= ny-2, then odd interior rows 1, 3, ..
    // <= ny-2, then rows 0 and ny-1 (whose extrapolated stencils reach two
    // rows in).  With two or more odd rows, the first odd-row workgroup relaxes
    // row 0 after row 1 and the last one row ny-1 after its own: the same
    // order without a third launch.  The line kernel never reads ghosts (ring
    // cells use the level's stencil table, which folds the Dirichlet ghost
    // rule in).
    const int n_even = ny >= 4 ? (ny - 4) / 2 + 1 : 0;
    const int n_odd = ny >= 3 ? (ny - 3) / 2 + 1 : 0;
    launch_xline(s, e, f, K, g, geo, 2, 2, n_even, -1, -1);
    if (n_odd >= 2) {
        launch_xline(s, e, f, K, g, geo, 1, 2, n_odd, 0, ny - 1);
    } else {
        launch_xline(s, e, f, K, g, geo, 1, 2, n_odd, -1, -1);
        launch_xline(s, e, f, K, g, geo, 0, ny - 1, 2, -1, -1);
    }
}

void xline_zebra(hipStream_t s, const Grid &g, const Geo &geo, const double *K, double *e, const double *f) {
    zebra(s, g, geo, K, e, f);
}

static void zebra_sweep(hipStream_t s, KLevel &l, double *e, const double *f) { zebra(s, l.g, l.geo, l.K.p, e, f); }
static void zebra_sweep(hipStream_t s, KLevel &l, float *e, const float *f) {
    zebra(s, l.g, l.geo32, l.K32.p, e, f);
}

// The coarse level's correction and residual fields of the cycle's type.
static float *coarse_e(KLevel &c, float *) { return c.e32.p; }
static float *coarse_r(KLevel &c, float *) { return c.r32.p; }
static double *coarse_e(KLevel &c, double *) { return c.e.p; }
static double *coarse_r(KLevel &c, double *) { return c.r.p; }
static const float *table(KLevel &l, float *) { return l.K32.p; }
static const double *table(KLevel &l, double *) { return l.K.p; }
static const GeoF &metrics(KLevel &l, float *) { return l.geo32; }
static const Geo &metrics(KLevel &l, double *) { return l.geo; }

// V-cycle from level d down.  Coarse levels start from e = 0 (set by the
// restriction); the finest level's e may hold a previous cycle's result, so
// consecutive calls iterate the cycle.
template <typename T>
static void vcycle_t(hipStream_t s, KrylovMG::Impl &K, size_t d, T *e, const T *f) {
    KLevel &l = K.L[d];
    if (d + 1 == K.L.size()) {
        for (int k = 0; k < coarsest_sweeps(K.sb, l.g.ny, d == 0, false); ++k) zebra_sweep(s, l, e, f);
        return;
    }
    KLevel &c = K.L[d + 1];
    T *ce = coarse_e(c, e), *cr = coarse_r(c, e);
    for (int k = 0; k < K.sd; ++k) zebra_sweep(s, l, e, f);
    launch_residual_restrict_y(s, cr, ce, e, f, table(l, e), l.g, c.g, metrics(l, e), l.off);
    vcycle_t(s, K, d + 1, ce, cr);
    launch_prolong_y(s, e, ce, l.g, c.g, side_sign(l.bc, YLO), side_sign(l.bc, YHI), l.off);
    for (int k = 0; k < K.su; ++k) zebra_sweep(s, l, e, f);
}

// Preconditioner application e = M f (one V-cycle from e = 0), FP64 in and
// out; in mixed precision the cycle runs on FP32 copies.  With y-lines the
// input is transposed into the hierarchy's finest level and the result
// transposed back (the valid cells of e; its ghosts are not read).
static void precondition(hipStream_t s, KrylovMG::Impl &K, double *e, const double *f) {
    KLevel &l = K.L[0];
    const Grid &tg = K.top_g;
    if (K.tr) {
        if (!K.mixed) {
            launch_transpose(s, l.r.p, l.g.pitch, f, tg.pitch, tg.nx, tg.ny);
            launch_setval(s, l.e.p, 0.0, l.g, 1);
            for (int c = 0; c < K.ncyc; ++c) vcycle_t(s, K, 0, l.e.p, (const double *)l.r.p);
            launch_transpose(s, e, tg.pitch, l.e.p, l.g.pitch, l.g.nx, l.g.ny);
        } else {
            launch_transpose(s, l.r32.p, l.g.pitch, f, tg.pitch, tg.nx, tg.ny);
            launch_setval(s, l.e32.p, 0.0f, l.g, 1);
            for (int c = 0; c < K.ncyc; ++c) vcycle_t(s, K, 0, l.e32.p, (const float *)l.r32.p);
            launch_transpose(s, e, tg.pitch, l.e32.p, l.g.pitch, l.g.nx, l.g.ny);
        }
        return;
    }
    if (!K.mixed) {
        launch_setval(s, e, 0.0, l.g, 1);
        for (int c = 0; c < K.ncyc; ++c) vcycle_t(s, K, 0, e, f);
        return;
    }
    launch_begin32(s, l.r32.p, l.e32.p, f, l.g);
    for (int c = 0; c < K.ncyc; ++c) vcycle_t(s, K, 0, l.e32.p, (const float *)l.r32.p);
    launch_convert(s, e, l.e32.p, l.g, 0);
}

static Geo4 metrics4(const KrylovMG::Impl &K) {
    const DevBuf *m = K.m4;
    return Geo4{m[0].p, m[1].p, m[2].p, m[3].p, m[4].p, m[5].p, m[6].p, m[7].p};
}

// out = A_h v (homogeneous user BCs): the reference operator with its
// homogeneous Dirichlet ghosts synthesised from v, or the fourth-order one.
static void apply_h(hipStream_t s, const KrylovMG::Impl &K, const DevGeo &top, const Bc &ubc, double *v,
                    double *out) {
    if (K.order == 4) launch_lap4(s, out, v, nullptr, top.g, metrics4(K));
    else launch_residual_synth(s, out, v, nullptr, top.g, top.view(), ubc);
}

double *KrylovMG::phi_lo() const { return impl && impl->dd ? impl->lo.p : nullptr; }

double KrylovMG::project_mean(hipStream_t s, const DevGeo &top, double *rhs) {
    if (!built) {
        g_last_error = "solver not built";
        throw Failure(MGP_ERR_ARG);
    }
    launch_project_jmean(s, rhs, top.J.p, impl->sums.p, impl->partial.p, top.g);
    double h[2];
    hip_check(hipMemcpyAsync(h, impl->sums.p, sizeof h, hipMemcpyDeviceToHost, s), "copy");
    hip_check(hipStreamSynchronize(s), "sync");
    return h[0] / h[1];
}

double KrylovMG::residual_norm(Workspace &W, hipStream_t s, const DevGeo &top, const Bc &user_bc, int homog,
                               double *phi, double *rhs, double *rhsnorm) {
    if (!built) {
        g_last_error = "solver not built";
        throw Failure(MGP_ERR_ARG);
    }
    Impl &K = *impl;
    const Grid &g = top.g;
    const Geo geo = top.view();
    if (K.dd) {
        launch_fold_dd(s, phi, K.lo.p, K.x.p, g);
        launch_residual_dd(s, K.r.p, phi, K.lo.p, rhs, g, geo, user_bc);
    } else if (K.order == 4) {
        launch_lap4(s, K.r.p, phi, rhs, g, metrics4(K));
    } else {
        launch_fill_ghosts(s, phi, g, geo, user_bc, homog, 0);
        launch_laplacian(s, K.r.p, phi, rhs, g, geo, user_bc, homog, 0);
    }
    if (rhsnorm) *rhsnorm = pnorm2(W, s, rhs, g);
    return pnorm2(W, s, K.r.p, g);
}

// Flexible GMRES(m), right preconditioned (Z_k = M V_k kept, so the FP32
// V-cycle need not be an exact linear map), classical Gram-Schmidt applied
// twice (two block passes: two reductions per step instead of k + 1), Givens
// rotations on the host.  Each cycle starts from the true residual
// (true_residual: FP64, or double-double with the correction folded in) and
// adds Z y to the solution; a cycle that does not lower the true residual is
// undone and the solve stops.  GMRES minimises the residual over the Krylov
// space, where BiCGStab's short recurrences break down on the strongly
// non-normal pairing of the fourth-order operator with the second-order
// preconditioner.
template <typename TR>
int KrylovMG::solve_fgmres(hipStream_t s, const DevGeo &top, const Bc &user_bc, double *phi, double tol,
                           int maxiters, int zerophi, double *relres, int *iters, TR &true_residual) {
    Impl &K = *impl;
    const Grid &g = top.g;
    const int m = K.m;
    const size_t n = K.r.n;
    double *V = K.gv.p, *Z = K.gz.p, *dh = K.gh.p, *hh = K.ghost;
    double *h1 = dh, *h2 = dh + (m + 2), *nw = dh + 2 * (m + 2);
    double *xc = K.dd ? K.x.p : phi;  // where a cycle's correction is added
    const size_t nbytes = n * sizeof(double);
    if (zerophi) launch_setval(s, phi, 0.0, g, 1);
    if (zerophi && K.dd) launch_setval(s, K.lo.p, 0.0, g, 1);
    const double r0 = true_residual();
    std::vector<double> hist{r0 == 0.0 ? 0.0 : 1.0};
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), gg(m + 1), y(m);
    auto Hc = [&](int i, int k) -> double & { return H[(size_t)k * (m + 1) + i]; };
    last_cycles = 0;
    int outcome = r0 == 0.0 ? MGP_CONVERGED : MGP_MAXITERS;
    double base = 1.0;  // relres where the current cycle started
    int it = 0;
    while (outcome == MGP_MAXITERS && it < maxiters) {
        // V_0 = r / ||r||
        launch_mdot(s, nw, K.gpart.p, K.r.p, n, K.r.p, 1, g);
        launch_scale_rsqrt(s, V, K.r.p, nw, g);
        hip_check(hipMemcpyAsync(hh, nw, sizeof(double), hipMemcpyDeviceToHost, s), "copy");
        hip_check(hipStreamSynchronize(s), "sync");
        const double beta = std::sqrt(hh[0]);
        if (!(beta > 0.0) || !std::isfinite(beta)) break;
        std::fill(gg.begin(), gg.end(), 0.0);
        gg[0] = beta;
        int k = 0;
        while (k < m && it < maxiters) {
            double *vk = V + (size_t)k * n, *zk = Z + (size_t)k * n, *w = V + (size_t)(k + 1) * n;
            precondition(s, K, zk, vk);
            apply_h(s, K, top, user_bc, zk, w);
            launch_mdot(s, h1, K.gpart.p, V, n, w, k + 1, g);
            launch_maxpy(s, w, V, n, h1, k + 1, -1.0, g);
            launch_mdot(s, h2, K.gpart.p, V, n, w, k + 1, g);
            launch_maxpy(s, w, V, n, h2, k + 1, -1.0, g);
            launch_mdot(s, nw, K.gpart.p, w, n, w, 1, g);
            launch_scale_rsqrt(s, w, w, nw, g);
            hip_check(hipMemcpyAsync(hh, dh, (size_t)3 * (m + 2) * sizeof(double), hipMemcpyDeviceToHost, s), "copy");
            hip_check(hipStreamSynchronize(s), "sync");
            for (int i = 0; i <= k; ++i) Hc(i, k) = hh[i] + hh[(m + 2) + i];
            const double hn = std::sqrt(hh[2 * (m + 2)]);
            Hc(k + 1, k) = hn;
            for (int i = 0; i < k; ++i) {
                const double a = Hc(i, k), b = Hc(i + 1, k);
                Hc(i, k) = cs[i] * a + sn[i] * b;
                Hc(i + 1, k) = -sn[i] * a + cs[i] * b;
            }
            const double rr = std::hypot(Hc(k, k), Hc(k + 1, k));
            cs[k] = rr > 0.0 ? Hc(k, k) / rr : 1.0;
            sn[k] = rr > 0.0 ? Hc(k + 1, k) / rr : 0.0;
            Hc(k, k) = rr;
            Hc(k + 1, k) = 0.0;
            gg[k + 1] = -sn[k] * gg[k];
            gg[k] = cs[k] * gg[k];
            ++k;
            ++it;
            last_cycles += K.ncyc;
            const double est = base * std::fabs(gg[k]) / beta;
            hist.push_back(est);
            if (!std::isfinite(est) || est <= tol || !(hn > 1e-300)) break;
        }
        // y = R^-1 g, then x += Z y
        for (int i = k - 1; i >= 0; --i) {
            double t = gg[i];
            for (int l = i + 1; l < k; ++l) t -= Hc(i, l) * y[l];
            y[i] = Hc(i, i) != 0.0 ? t / Hc(i, i) : 0.0;
        }
        double *hy = hh + 3 * (m + 2);
        for (int i = 0; i < k; ++i) hy[i] = y[i];
        hip_check(hipMemcpyAsync(K.gy.p, hy, (size_t)k * sizeof(double), hipMemcpyHostToDevice, s), "copy");
        hip_check(hipMemcpyAsync(K.bk.p, phi, nbytes, hipMemcpyDeviceToDevice, s), "backup");
        if (K.dd) hip_check(hipMemcpyAsync(K.bk_lo.p, K.lo.p, nbytes, hipMemcpyDeviceToDevice, s), "backup");
        launch_maxpy(s, xc, Z, n, K.gy.p, k, 1.0, g);
        const double tr = true_residual() / r0;
        hist.back() = tr;
        if (tr <= tol) {
            outcome = MGP_CONVERGED;
        } else if (!std::isfinite(tr) || tr >= base) {
            // no progress over a whole cycle: return the cycle's starting solution
            hip_check(hipMemcpyAsync(phi, K.bk.p, nbytes, hipMemcpyDeviceToDevice, s), "undo");
            if (K.dd) hip_check(hipMemcpyAsync(K.lo.p, K.bk_lo.p, nbytes, hipMemcpyDeviceToDevice, s), "undo");
            outcome = (!std::isfinite(tr) || tr > 10.0 * base) ? MGP_DIVERGED : MGP_MAXITERS;
            hist.back() = base;
            break;
        }
        base = tr;
    }
    hip_check(hipStreamSynchronize(s), "sync");
    if (relres)
        for (size_t k = 0; k < hist.size() && (int)k <= maxiters; ++k) relres[k] = hist[k];
    if (iters) *iters = (int)hist.size() - 1 < maxiters ? (int)hist.size() - 1 : maxiters;
    return outcome;
}

int KrylovMG::solve(Workspace &W, hipStream_t s, const DevGeo &top, const Bc &user_bc, int homog, double *phi,
                    double *rhs, double tol, int maxiters, int zerophi, double *relres, int *iters) {
    if (!built) {
        g_last_error = "solver not built";
        throw Failure(MGP_ERR_ARG);
    }
    Impl &K = *impl;
    const Grid &g = top.g;
    const double vol = g.dx * g.dy;
    double *sc = K.scal.p;
    double *partial = K.partial.p;
    // the solution the BiCGStab runs update: phi itself, or (double-double)
    // the correction x of the current run, folded into phi + lo at each check
    double *xs = K.dd ? K.x.p : phi;
    auto true_residual = [&]() { return residual_norm(W, s, top, user_bc, homog, phi, rhs, nullptr); };
    // One BiCGStab iteration with every scalar on the device (right
    // preconditioning, M = one V-cycle).  Ends with ||r||^2 * vol in sc[SC_RN].
    auto body = [&]() {
        launch_reduce(s, sc + SC_RHO_NEW, partial, K.rhat.p, K.r.p, nullptr, vol, g, 0);
        launch_bicg_scalars(s, sc, 0);
        launch_bicg_update(s, K.p.p, K.r.p, K.v.p, sc, g, 0);
        precondition(s, K, K.phat.p, K.p.p);
        apply_h(s, K, top, user_bc, K.phat.p, K.v.p);
        launch_reduce(s, sc + SC_RV, partial, K.rhat.p, K.v.p, nullptr, vol, g, 0);
        launch_bicg_scalars(s, sc, 1);
        launch_bicg_update(s, K.s.p, K.r.p, K.v.p, sc, g, 1);
        launch_bicg_update(s, xs, K.phat.p, nullptr, sc, g, 2);
        precondition(s, K, K.shat.p, K.s.p);
        apply_h(s, K, top, user_bc, K.shat.p, K.t.p);
        launch_reduce(s, sc + SC_TS, partial, K.t.p, K.s.p, nullptr, vol, g, 0);
        launch_reduce(s, sc + SC_TT, partial, K.t.p, K.t.p, nullptr, vol, g, 0);
        launch_bicg_scalars(s, sc, 2);
        launch_bicg_update(s, xs, K.shat.p, nullptr, sc, g, 3);
        launch_bicg_update(s, K.r.p, K.s.p, K.t.p, sc, g, 4);
        launch_reduce(s, sc + SC_RN, partial, K.r.p, K.r.p, nullptr, vol, g, 0);
    };
    if (K.fgmres) return solve_fgmres(s, top, user_bc, phi, tol, maxiters, zerophi, relres, iters, true_residual);
    if (K.use_graph && !K.graph) {
        hipGraph_t graph = nullptr;
        hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "capture");
        try {
            body();
        } catch (...) {
            (void)hipStreamEndCapture(s, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        hip_check(hipStreamEndCapture(s, &graph), "capture end");
        hip_check(hipGraphInstantiate(&K.graph, graph, nullptr, nullptr, 0), "graph instantiate");
        (void)hipGraphDestroy(graph);
    }

    if (zerophi) launch_setval(s, phi, 0.0, g, 1);
    if (zerophi && K.dd) launch_setval(s, K.lo.p, 0.0, g, 1);
    const double r0 = true_residual();
    std::vector<double> hist;
    hist.push_back(r0 == 0.0 ? 0.0 : 1.0);
    last_cycles = 0;
    int outcome = MGP_MAXITERS;
    if (r0 == 0.0) {
        outcome = MGP_CONVERGED;
    } else {
        // Iterative refinement: each BiCGStab run stops once its recursive
        // residual has fallen by `refine` below the true residual it started
        // from, and the next starts from the FP64 true residual.  The FP32
        // preconditioner is consistent to about 1e-7 of its input, so the
        // recurrences of one run lose their accuracy a few decades down (at
        // 2048^2 a run stalls near 1e-7); 1e-3 per run measured fastest on
        // the 2048^2 stretched-grid problem (74 V-cycles to 1e-9, against 168
        // without restarts).  The FP64 cycle restarts after 1e-6 (62 V-cycles
        // against 64).
        //
        // A run whose recursive residual grows 1e3 above where it started is
        // checked at once.  A run that ends without lowering the true residual
        // is undone (the solution returns to where the run started) and the
        // solve stops as diverged: the preconditioner is then too inaccurate
        // for this operator, and a repeated run would repeat the failure.
        const double refine = K.mixed ? 1e-3 : 1e-6;
        double base = 1.0;  // relres at the last restart
        bool restart = true;
        const size_t nbytes = K.r.n * sizeof(double);
        for (int it = 1; it <= maxiters; ++it) {
            if (restart) {
                if (!K.dd) hip_check(hipMemcpyAsync(K.bk.p, phi, nbytes, hipMemcpyDeviceToDevice, s), "backup");
                hip_check(hipMemcpyAsync(K.rhat.p, K.r.p, K.r.n * sizeof(double), hipMemcpyDeviceToDevice, s), "copy");
                launch_setval(s, K.p.p, 0.0, g, 0);
                launch_setval(s, K.v.p, 0.0, g, 0);
                launch_bicg_scalars(s, sc, 3);
                restart = false;
            }
            if (K.graph) hip_check(hipGraphLaunch(K.graph, s), "graph launch");
            else body();
            last_cycles += 2 * K.ncyc;
            hip_check(hipMemcpyAsync(K.host, sc, SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, s), "scalars");
            hip_check(hipStreamSynchronize(s), "sync");
            const double rn = std::sqrt(K.host[SC_RN]) / r0;
            const double omega = K.host[SC_OMEGA];
            if (rn <= tol || omega == 0.0 || !std::isfinite(rn) || rn <= refine * base || rn > 1e3 * base) {
                if (K.dd) {  // the fold below is undone if the run made no progress
                    hip_check(hipMemcpyAsync(K.bk.p, phi, nbytes, hipMemcpyDeviceToDevice, s), "backup");
                    hip_check(hipMemcpyAsync(K.bk_lo.p, K.lo.p, nbytes, hipMemcpyDeviceToDevice, s), "backup");
                }
                const double tr = true_residual() / r0;
                hist.push_back(tr);
                if (tr <= tol) {
                    outcome = MGP_CONVERGED;
                    break;
                }
                if (!std::isfinite(tr) || tr >= base) {
                    hip_check(hipMemcpyAsync(phi, K.bk.p, nbytes, hipMemcpyDeviceToDevice, s), "undo");
                    if (K.dd) hip_check(hipMemcpyAsync(K.lo.p, K.bk_lo.p, nbytes, hipMemcpyDeviceToDevice, s), "undo");
                    hist.back() = base;  // the solution returned is the run's starting one
                    outcome = MGP_DIVERGED;
                    break;
                }
                restart = true;  // recursive residual drifted: restart from the true one
                base = tr;
                continue;
            }
            hist.push_back(rn);
        }
    }
    if (K.dd) launch_fold_dd(s, phi, K.lo.p, K.x.p, g);  // a run cut off by maxiters
    const int used = (int)hist.size() - 1;
    if (outcome == MGP_DIVERGED && K.fallback && used < maxiters) {
        // BiCGStab broke down (the solution is the last good iterate): flexible
        // GMRES carries on from it with the iterations left, to the same
        // tolerance relative to the initial residual
        const double at = hist.back();
        ensure_fgmres(top.g);
        std::vector<double> h2((size_t)(maxiters - used) + 1);
        int it2 = 0;
        const int cyc = last_cycles;
        outcome = solve_fgmres(s, top, user_bc, phi, tol / at, maxiters - used, 0, h2.data(), &it2, true_residual);
        last_cycles += cyc;
        for (int k = 1; k <= it2; ++k) hist.push_back(at * h2[k]);
    }
    hip_check(hipStreamSynchronize(s), "sync");
    if (relres)
        for (size_t k = 0; k < hist.size() && (int)k <= maxiters; ++k) relres[k] = hist[k];
    if (iters) *iters = (int)hist.size() - 1 < maxiters ? (int)hist.size() - 1 : maxiters;
    return outcome;
}

double KrylovMG::time_kernel(hipStream_t s, int kernel, int reps, int cold, double *bytes) {
    if (!built) {
        g_last_error = "solver not built";
        throw Failure(MGP_ERR_ARG);
    }
    KLevel &l = impl->L[0];
    const Grid &g = l.g;
    const double nx = g.nx;
    hipEvent_t a, b;
    hip_check(hipEventCreate(&a), "event");
    hip_check(hipEventCreate(&b), "event");
    const int n_even = g.ny >= 4 ? (g.ny - 4) / 2 + 1 : 0;
    launch_fill_ghosts(s, l.e.p, g, l.geo, l.bc, 1, 0);
    if (kernel == 0) {
        // rows 2, 4, ..: reads e on rows 1 .. 2*n_even+1 and writes it on the
        // updated rows; f, Jxx, Jxy on the updated rows; Jyx, Jyy on face rows
        // 2 .. 2*n_even+1; the two ring cells of each row read 9 table entries.
        const double R = n_even;
        *bytes = nx * 8.0 * ((2.0 * R + 1.0) + R + R + R + 2.0 * R + 2.0 * R + R) + R * 2.0 * 9.0 * 8.0;
    } else if (kernel == 1) {
        // r = f - L e over all cells: e (with its one-cell halo), f, r, J-free metrics
        *bytes = nx * (double)g.ny * 8.0 * 3.0 + 16.0 * (nx + 1) * g.ny + 16.0 * nx * (g.ny + 1);
    } else if (kernel == 2) {
        // plain copy w = e over the valid cells: the bandwidth reference for the two above
        *bytes = nx * (double)g.ny * 16.0;
    } else if (kernel == 3 && impl->mixed) {
        // the FP32 x-line kernel of the mixed-precision cycle: kernel 0's
        // array elements at 4 bytes each
        const double R = n_even;
        *bytes = nx * 4.0 * ((2.0 * R + 1.0) + R + R + R + 2.0 * R + 2.0 * R + R) + R * 2.0 * 9.0 * 4.0;
    } else {
        g_last_error = "time_kernel: kernel must be 0 (x-line, even rows), 1 (residual), 2 (copy) or, for a "
                       "mixed-precision solver, 3 (FP32 x-line, even rows)";
        throw Failure(MGP_ERR_ARG);
    }
    auto launch = [&]() {
        if (kernel == 0) launch_xline(s, l.e.p, l.r.p, l.K.p, g, l.geo, 2, 2, n_even, -1, -1);
        else if (kernel == 1) launch_laplacian(s, l.w.p, l.e.p, l.r.p, g, l.geo, l.bc, 1, 0);
        else if (kernel == 2) launch_lincomb(s, l.w.p, 1.0, l.e.p, 0.0, nullptr, g, 0);
        else launch_xline(s, l.e32.p, l.r32.p, l.K32.p, g, l.geo32, 2, 2, n_even, -1, -1);
    };
    launch();  // warm
    double total = 0.0;
    if (!cold) {
        // back to back: the operands stay in the L2 and the 256 MB MALL
        hip_check(hipEventRecord(a, s), "event");
        for (int k = 0; k < reps; ++k) launch();
        hip_check(hipEventRecord(b, s), "event");
        hip_check(hipEventSynchronize(b), "event");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, a, b), "event");
        total = ms;
    } else {
        // back to back over four copies of the operands, used in turn: between
        // two uses of a copy the other three launches move more than 500 MB,
        // more than the L2s and the 256 MB MALL hold, so every launch starts
        // on cold caches, and the events around the batch see no per-launch
        // dispatch gap
        constexpr int copies = 4;
        const size_t n = level_elems(g.nx, g.ny), nb = n * sizeof(double);
        struct Copy {
            DevBuf e, r, w, J, Jxx, Jxy, Jyx, Jyy;
            DevBufF e32, r32, Jxx32, Jxy32, Jyx32, Jyy32;
        };
        std::vector<Copy> extra(copies - 1);
        for (Copy &c : extra) {
            if (kernel == 3) {
                DevBufF *dst[6] = {&c.e32, &c.r32, &c.Jxx32, &c.Jxy32, &c.Jyx32, &c.Jyy32};
                const DevBufF *src[6] = {&l.e32, &l.r32, &l.Jxx32, &l.Jxy32, &l.Jyx32, &l.Jyy32};
                for (int q = 0; q < 6; ++q) {
                    dst[q]->alloc(n);
                    hip_check(hipMemcpyAsync(dst[q]->p, src[q]->p, n * sizeof(float), hipMemcpyDeviceToDevice, s),
                              "copy");
                }
                continue;
            }
            DevBuf *dst[8] = {&c.e, &c.r, &c.w, &c.J, &c.Jxx, &c.Jxy, &c.Jyx, &c.Jyy};
            const DevBuf *src[8] = {&l.e, &l.r, &l.w, &l.J, &l.Jxx, &l.Jxy, &l.Jyx, &l.Jyy};
            for (int q = 0; q < 8; ++q) {
                dst[q]->alloc(n);
                hip_check(hipMemcpyAsync(dst[q]->p, src[q]->p, nb, hipMemcpyDeviceToDevice, s), "copy");
            }
        }
        auto launch_copy = [&](int k) {
            if (k == 0) {
                launch();
                return;
            }
            Copy &c = extra[k - 1];
            const Geo geo{c.J.p, c.Jxx.p, c.Jxy.p, c.Jyx.p, c.Jyy.p};
            if (kernel == 3) {
                const GeoF geo32{nullptr, c.Jxx32.p, c.Jxy32.p, c.Jyx32.p, c.Jyy32.p};
                launch_xline(s, c.e32.p, c.r32.p, l.K32.p, g, geo32, 2, 2, n_even, -1, -1);
            } else if (kernel == 0) launch_xline(s, c.e.p, c.r.p, l.K.p, g, geo, 2, 2, n_even, -1, -1);
            else if (kernel == 1) launch_laplacian(s, c.w.p, c.e.p, c.r.p, g, geo, l.bc, 1, 0);
            else launch_lincomb(s, c.w.p, 1.0, c.e.p, 0.0, nullptr, g, 0);
        };
        for (int k = 0; k < copies; ++k) launch_copy(k);  // first touch of every copy
        hip_check(hipEventRecord(a, s), "event");
        for (int k = 0; k < reps; ++k) launch_copy(k % copies);
        hip_check(hipEventRecord(b, s), "event");
        hip_check(hipEventSynchronize(b), "event");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, a, b), "event");
        total = ms;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return total / reps;
}

}  // namespace mgp
