/*
 * mgp2d.h -- C ABI of the MI355X (gfx950) semicoarsening-multigrid library.
 *
 * Drop-in boundary for libMGPoisson (EdSantilli/MGPoisson2D).  Each entry of
 * section 2 replaces one Fortran 90 module procedure of the reference; the
 * Fortran shim in mgpoisson2d_amd/fortran/ (modules Poisson2D, MGPoisson2D,
 * DeferredPoisson2D, DeferredMGPoisson2D) binds them through iso_c_binding,
 * so a caller written against the reference recompiles unchanged.
 *
 * Array conventions (host arrays, the reference's storage):
 *   valid cells i = 0..nx-1, j = 0..ny-1, column-major data(i, j), i fastest
 *   (arrayutils.f90:107-116).  Cell fields without ghosts: nx*ny doubles.
 *   Fields with one ghost layer (phi, the reference's e/p_tilde/s_tilde):
 *   (nx+2)*(ny+2) doubles, element (i, j) at (i+1) + (j+1)*(nx+2).
 *   Metrics (mgp_geo): J nx*ny; Jxx, Jxy (nx+1)*ny; Jyx, Jyy nx*(ny+1).
 *   BC codes are the reference's integers (arrayutils.f90:27-49).
 *   Logical arguments are ints (0 = .false.).
 *
 * All computation runs on the current HIP device; section-2 calls copy their
 * host arguments in and results out (PCIe-inclusive).  Section 3 keeps a
 * problem resident in HBM across solves.  Calls on one mgp_solver are not
 * thread-safe; distinct solvers may be used from distinct threads.
 */
#ifndef MGP2D_H
#define MGP2D_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 1. Types ----------------------------------------------------------- */

typedef enum {
    MGP_OK = 0,
    MGP_ERR_ARG = 1,          /* bad shape / argument (reference: print* + stop) */
    MGP_ERR_UNSUPPORTED = 2,  /* e.g. BCTYPE_CF, non-2x2 prolongation order 2 */
    MGP_ERR_HIP = 3,          /* HIP runtime failure */
    MGP_ERR_NOMEM = 4,        /* device allocation failed (reference: "Out of memory") */
    MGP_ERR_BREAKDOWN = 5     /* Krylov breakdown where the reference stops (print* + stop) */
} mgp_status;

/* Outcome of an iterative solve. */
enum { MGP_CONVERGED = 0, MGP_DIVERGED = 1, MGP_MAXITERS = 2 };

/* Relaxation orderings.
 *   LEX: the reference's lexicographic order (poisson2d.f90:1561-1959,
 *        MGPoisson2D_DD.f90:288-306), run as t = i + 3j (9-point) or t = i + j
 *        (5-point) wavefronts in one workgroup; bit-identical to the reference.
 *   RB:  the DC smoother's red-black branch (MGPoisson2D_DD.f90:262-285).
 *   MC:  4-colour interior, then the boundary ring in 4 colours, then the
 *        corners (9-point operator; the multi-workgroup GPU ordering). */
enum { MGP_ORDER_LEX = 0, MGP_ORDER_RB = 1, MGP_ORDER_MC = 2 };

typedef struct {
    int nx, ny;
    double dx, dy;
    const double *J, *Jxx, *Jxy, *Jyx, *Jyy;
} mgp_geo;

typedef struct {
    int type[4];          /* xlo, xhi, ylo, yhi */
    int mode[4];          /* BCMODE_UNIFORM / NONUNIFORM / LAH */
    const double *data[4];/* 1 value (UNIFORM, LAH) or ny (x sides) / nx (y sides) */
} mgp_bc;

/* A mapping whose metrics the library generates on the GPU (no counterpart in the
 * reference's library; MGP_MAP_TERRAIN is its test geometry).  L, H: domain extents,
 * <= 0 for nx*dx and ny*dy. */
enum { MGP_MAP_TERRAIN = 0, MGP_MAP_WARPED = 1 };
typedef struct {
    int kind;     /* MGP_MAP_TERRAIN: test.f90:320-716, x = xi, y = el + (H - el) eta/H over the bottom
                     el = amp sin^2(pi x / L) (amp 1/4 there), metrics by differencing point positions;
                     MGP_MAP_WARPED: x = s + amp sin(2 pi s) sin(2 pi t), y = (e^(beta t) - 1)/(e^beta - 1),
                     s = xi/L, t = eta/H, exact derivatives */
    double amp, beta;
    double L, H;
} mgp_mapping;

/* ---- 2. Drop-in entry points (host arrays) ------------------------------ */

const char *mgp_status_string(int status);
/* Last HIP error text recorded by the library (per thread). */
const char *mgp_last_error(void);
/* Device the library runs on (the current HIP device). */
int mgp_device_count(void);

/* arrayutils.f90:1277 fill_ghosts(phi, bc, geo, homog, do_neum) */
int mgp_fill_ghosts(const mgp_geo *geo, const mgp_bc *bc, int homog, int do_neum, double *phi);
/* poisson2d.f90:886 compute_laplacian(lap, phi, geo, bc, homog, opt_jscale); phi's ghosts are filled */
int mgp_compute_laplacian(const mgp_geo *geo, const mgp_bc *bc, int homog, int jscale, double *phi,
                          double *lap);
/* poisson2d.f90:1327 compute_residual(res, rhs, phi, geo, bc, homog) */
int mgp_compute_residual(const mgp_geo *geo, const mgp_bc *bc, int homog, const double *rhs, double *phi,
                         double *res);
/* poisson2d.f90:16 compute_invdiags(invdiags, geo, bc, homog) */
int mgp_compute_invdiags(const mgp_geo *geo, const mgp_bc *bc, int homog, double *invd);
/* poisson2d.f90:1460 relax_gs(phi, rhs, geo, bc, homog, invdiags, omega, tol<0, maxiters, .false., 0);
 * ordering MGP_ORDER_LEX (reference order) or MGP_ORDER_MC. */
int mgp_relax_gs(const mgp_geo *geo, const mgp_bc *bc, int homog, const double *invd, double omega,
                 int maxiters, const double *rhs, double *phi, int ordering);
/* x-line zebra relaxation, the smoother of MGP_METHOD_KRYLOV_MG (no counterpart in the
 * reference): each sweep solves, row by row, T_j delta = rhs - L phi and adds delta, where L is
 * the reference operator with homogeneous BCs and T_j its tridiagonal part along row j; even
 * interior rows, then odd ones, then rows 0 and ny-1.  Dirichlet/Neumann sides, nx, ny >= 4;
 * phi's ghost cells are neither read nor written.  Rows over 4111 cells are cut into 2, 4, ..
 * equal segments (the fewest that fit) and each colour relaxes the even segments of its rows,
 * then the odd ones, each segment with the tridiagonal part of its own block (block
 * Gauss-Seidel along the row); MGP_ERR_ARG if nx has no such cut. */
int mgp_xline_relax(const mgp_geo *geo, const mgp_bc *bc, const double *rhs, double *phi, int sweeps);
/* MGPoisson2D_DD.f90:45 DeferredPoisson2D::compute_laplacian */
int mgp_dc_compute_laplacian(const mgp_geo *geo, const mgp_bc *bc, int homog, double *phi, double *lap);
/* MGPoisson2D_DD.f90:16 DeferredPoisson2D::compute_invdiags */
int mgp_dc_compute_invdiags(const mgp_geo *geo, double *invd);
/* MGPoisson2D_DD.f90:153 DeferredPoisson2D::relax_gs; ordering LEX or RB */
int mgp_dc_relax_gs(const mgp_geo *geo, const mgp_bc *bc, int homog, const double *invd, double omega,
                    int maxiters, const double *rhs, double *phi, int ordering);
/* Red-black DC sweeps (mgp_dc_relax_gs with RB, and the DC V-cycle with RB
 * ordering) run several sweeps per LDS tile load when every side is a
 * Neumann or Dirichlet side in BCMODE_LAH (the level BCs of the DC cycle;
 * MGPoisson2D_DD.f90:262-285 is one launch per colour).  No counterpart in the
 * reference: a tuning switch, process-wide.  choice -1: tile shape timed per
 * level shape on first use (default); 0: one launch per colour; k >= 1: tile
 * shape k.  Results are bit-identical for every choice.  *previous (if not
 * null) receives the setting replaced. */
int mgp_dc_rb_tiles(int choice, int *previous);
/* Number of tile shapes the switch above accepts (1 .. n). */
int mgp_dc_rb_tile_shapes(void);
/* MGPoisson2D.f90:135 restrict(fine, crse); kind 0 cell, 1 x-face, 2 y-face */
int mgp_restrict(int kind, int fnx, int fny, const double *fine, int cnx, int cny, double *crse);
/* MGPoisson2D.f90:262 prolong(fine, crse, fine_geo, crse_geo, crse_bc, order), crse_bc
 * homogeneous with the given types/modes; fine and crse carry one ghost layer. */
int mgp_prolong(const mgp_geo *fine_geo, const mgp_geo *crse_geo, const int *bc_types, const int *bc_modes,
                int order, double *fine, double *crse);
/* compute_residual in double-double (no counterpart in the reference): res = rhs - L (phi_hi + phi_lo)
 * with every operation of poisson2d.f90 compute_laplacian's stencil (:886-1319) carried in
 * double-double and the result rounded once; phi_hi, phi_lo (nx+2) x (ny+2) as phi, their homogeneous
 * Dirichlet ghosts synthesised from the valid cells.  Dirichlet / Neumann sides, homogeneous, not LAH.
 * The residual of a solution that FP64 cannot represent accurately enough (see refine in
 * mgp_solver_opts). */
int mgp_compute_residual_dd(const mgp_geo *geo, const mgp_bc *bc, const double *rhs, const double *phi_hi,
                            const double *phi_lo, double *res);
/* The fourth-order finite-volume operator (no counterpart in the reference; the operator of the
 * fourth-order deferred correction): lap = L4 phi on cell averages phi (nx x ny), homogeneous Neumann
 * sides, nx, ny >= 5.  metrics[0..3] on x faces ((nx+1) x ny): <J g^11>, <J g^12>, and (deta/12) times
 * the differences of J g^11 and J g^12 across the face; metrics[4..7] on y faces (nx x (ny+1)):
 * <J g^22>, <J g^21> and (dxi/12) times their differences along the face (mgp_metric4_fill). */
int mgp_lap4(int nx, int ny, double dx, double dy, const double *const *metrics, const double *phi, double *lap);
/* The eight face-averaged metric arrays of mgp_lap4 for an MGP_MAP_WARPED mapping, generated on the
 * GPU (3-point Gauss-Legendre averages over each face). */
int mgp_metric4_fill(const mgp_mapping *map, int nx, int ny, double dx, double dy, double *const *metrics);
/* MGPoisson2D.f90:38,67 estimate_maxdepth + create_mgschedule; refx/refy hold >= 64 entries */
int mgp_create_mgschedule(int nx, int ny, double dx, double dy, int maxdepth_in, int *maxdepth_out,
                          int *refx, int *refy);
/* arrayutils.f90:753 pnorm over the valid cells of a ghost-free field: p = 0 max |a|, p >= 1
 * (sum dx dy |a|^p)^(1/p).  Sums are parallel and compensated (within an ulp or two of the
 * reference's sequential sums); the max norm is exact. */
int mgp_pnorm(int nx, int ny, double dx, double dy, const double *a, int p, double *out);
/* No counterpart in the reference: mgp_pnorm on a field already in device memory, for callers that
 * cannot allocate (code under HIP graph capture, serving loops that forbid allocation).  It allocates
 * nothing: the reduction's per-block partial sums live in a static 8 KiB device array (1024
 * doubles, one per GPU, held until the process ends); every grid size fits, its rows being strided
 * over at most 1024 row blocks.  a: device pointer in the resident layout of mgp_solver_field
 * (element (i, j) at a[(j + 1) * pitch + 32 + i], pitch >= nx + 34); p = 0 (max |a|), 1 or 2, the
 * sums of mgp_pnorm (the same kernels and row blocks, so the same result); *out: device pointer,
 * written on `stream` (a hipStream_t; null: the null stream).  Calls that overlap on one GPU
 * (different streams or threads) race on the shared array. */
int mgp_pnorm_noalloc(int nx, int ny, double dx, double dy, const double *a, int pitch, int p, double *out,
                      void *stream);
/* arrayutils.f90:637 inner_prod(bd1, bd2): sum of dx dy a b over the nx x ny cells. */
int mgp_inner_prod(int nx, int ny, double dx, double dy, const double *a, const double *b, double *out);
/* arrayutils.f90:670 integrate2d(bd, bx, geo, opt_jscale): sum of dx dy J a (J given, jscale true)
 * or dx dy a (J == NULL). */
int mgp_integrate2d(int nx, int ny, double dx, double dy, const double *a, const double *J, double *out);
/* arrayutils.f90:721 integrate2d_bdry(xflux, yflux, bx): sum over j of dx (xflux(nx, j) - xflux(0, j)),
 * then over i of dy (yflux(i, ny) - yflux(i, 0)), in the reference's order (bit-identical); xflux
 * (nx+1) x ny, yflux nx x (ny+1) over the box's faces; dx, dy the box's spacing. */
int mgp_integrate2d_bdry(int nx, int ny, double dx, double dy, const double *xflux, const double *yflux,
                         double *out);
/* arrayutils.f90:802 gpnorm(bd, p) of a one-ghost cell field over the valid box ilo..ihi, jlo..jhi
 * (phi holds data(ilo-1:ihi+1, jlo-1:jhi+1), column-major).  As the reference: p = 0 the max of |.|,
 * else (sum dy phi^p + sum dx phi^p)^(1/p) without absolute values, over the columns i = jlo-1 and
 * jhi+1 and the rows j = ilo-1 and ihi+1 (the ghost layer when ilo == jlo and ihi == jhi), corner
 * ghosts skipped, summed in the reference's order.  MGP_ERR_ARG where that index pattern leaves the
 * array (the reference reads out of bounds there). */
int mgp_gpnorm(int ilo, int ihi, int jlo, int jhi, double dx, double dy, const double *phi, int p, double *out);
/* arrayutils.f90:1497 extrapolate_ghosts(phi, order), order 0..3, on a one-ghost field
 * (nx+2) x (ny+2): x-side ghost columns first (over all rows), then y-side rows. */
int mgp_extrapolate_ghosts(int nx, int ny, int order, double *phi);
/* poisson2d.f90:404 compute_pd(pd, phi, dir): d phi / d(dir) on the faces normal to nodedir
 * (1: x-faces, pd is (nx+1) x ny; 2: y-faces, pd is nx x (ny+1)); phi (nx+2) x (ny+2) is read
 * with its ghosts as given.  Across the face (dir == nodedir) a two-point difference; along it
 * compute_laplacian's cross-term stencil with its boundary closures.  nx, ny >= 3. */
int mgp_compute_pd(int nx, int ny, double dx, double dy, const double *phi, int dir, int nodedir, double *pd);
/* poisson2d.f90:809 compute_grad(xflux, yflux, phi, geo, bc, homog, xwk, ywk): fills phi's
 * ghosts (do_neum false), then J grad phi on the faces -- xflux (nx+1) x ny, yflux nx x (ny+1)
 * -- then fill_boundary_fluxes.  The workspaces xwk/ywk are internal. */
int mgp_compute_grad(const mgp_geo *geo, const mgp_bc *bc, int homog, double *phi, double *xflux,
                     double *yflux);
/* poisson2d.f90:860 compute_div(div, xflux, yflux): J div F on the nx x ny cells */
int mgp_compute_div(int nx, int ny, double dx, double dy, const double *xflux, const double *yflux,
                    double *div);
/* arrayutils.f90:1571 fill_boundary_fluxes(xflux, yflux, bc, homog): Neumann faces take the BC
 * flux (zero when homogeneous or BCMODE_LAH); other faces are left as they are. */
int mgp_fill_boundary_fluxes(int nx, int ny, const mgp_bc *bc, int homog, double *xflux, double *yflux);
/* poisson2d.f90:1358 relax_jacobi(phi, rhs, geo, bc, homog, invdiags, omega, tol, maxiters, zerophi,
 * verbosity).  relres (maxiters+1 doubles, may be NULL) receives the history the reference prints
 * at verbosity >= 3 (written when tol > 0); *iters the sweeps made.  ordering MGP_ORDER_LEX sums
 * the norms in the reference's order (bit-identical stopping decisions); MGP_ORDER_MC uses the
 * parallel reductions (large grids). */
int mgp_relax_jacobi(const mgp_geo *geo, const mgp_bc *bc, int homog, const double *invd, double omega,
                     double tol, int maxiters, int zerophi, const double *rhs, double *phi, int ordering,
                     double *relres, int *iters);
/* poisson2d.f90:1993 solve_bicgstab(phi, rhs, geo, bc, homog, invdiags, tol, max_iters, max_restarts,
 * zerophi, verbosity), the first BiCGStab version (relax_gs-preconditioned, max-norm residuals,
 * restarts on hanging or divergence).  ordering MGP_ORDER_LEX: the reference's preconditioner order
 * and reduction order, bit-identical iterates; MGP_ORDER_MC: multi-colour preconditioner and
 * parallel reductions over the whole grid.  relres (maxiters+maxrestarts+1 doubles, may be NULL)
 * receives the history the reference prints at verbosity >= 3, entries 0..*last (restarts take an
 * entry).  Returns MGP_ERR_BREAKDOWN where the reference stops on beta == 0. */
int mgp_solve_bicgstab(const mgp_geo *geo, const mgp_bc *bc, int homog, const double *invd, double tol,
                       int maxiters, int maxrestarts, int zerophi, const double *rhs, double *phi, int ordering,
                       double *relres, int *last);
/* test.f90:320 define_geometry (MGP_MAP_TERRAIN) or the analytic MGP_MAP_WARPED metrics, computed
 * on the GPU: J (nx x ny), Jxx, Jxy ((nx+1) x ny), Jyx, Jyy (nx x (ny+1)) in the mgp_geo layouts. */
int mgp_metric_fill(const mgp_mapping *m, int nx, int ny, double dx, double dy, double *J, double *Jxx,
                    double *Jxy, double *Jyx, double *Jyy);
/* poisson2d.f90:2253 (kind 0) / MGPoisson2D_DD.f90:339 (kind 1) solve_bicgstab2.  Kind 0 with
 * ordering MGP_ORDER_MC runs grid-wide kernels with parallel reductions (full-size grids); every
 * other combination runs the bottom-solver kernel (one workgroup, the reference's reduction order,
 * bit-identical). */
int mgp_solve_bicgstab2(int kind, const mgp_geo *geo, const mgp_bc *bc, int homog, const double *invd,
                        double tol, int maxiters, int maxrestarts, int zerophi, const double *rhs, double *phi,
                        int ordering);
/* MGPoisson2D.f90:400 vcycle (kind 0) / MGPoisson2D_DD.f90:589 DeferredMGPoisson2D::vcycle (kind 1) /
 * MGPoisson2D.f90:860 fmg (kind 2, same arguments; order-2 prolongation inside, 2x2 schedules only):
 * (phi, rhs, geo, bc, homog, amrrefx, amrrefy, tol, maxiters, maxdepth_user, numcycles,
 *  smooth_down, smooth_up, smooth_bottom, zerophi, verbosity) -- amrref* are unused by the
 * reference and omitted.  relres (maxiters+1 doubles, may be NULL) receives the history the
 * reference prints at verbosity >= 1; *iters the cycles run; *outcome MGP_CONVERGED/DIVERGED/
 * MAXITERS.  The kind-1 rhs is scaled by J and restored, as in the reference. */
int mgp_vcycle(int kind, const mgp_geo *geo, const mgp_bc *bc, int homog, double tol, int maxiters,
               int maxdepth, int numcycles, int smooth_down, int smooth_up, int smooth_bottom, int zerophi,
               double *phi, double *rhs, int ordering, double *relres, int *iters, int *outcome);

/* ---- 3. Resident solver (no counterpart in the reference: keeps the level
 *         hierarchy and phi/rhs in HBM across solves) ---------------------- */

typedef struct mgp_solver mgp_solver;

enum {
    MGP_METHOD_VCYCLE = 0,     /* MGPoisson2D::vcycle on device data */
    MGP_METHOD_DC_VCYCLE = 1,  /* DeferredMGPoisson2D::vcycle on device data */
    MGP_METHOD_KRYLOV_MG = 2,  /* FP64 BiCGStab on the reference operator, preconditioned by one
                                  semicoarsening (y) x-line-relaxation V-cycle per application */
    MGP_METHOD_FMG = 3,        /* MGPoisson2D::fmg on device data */
    MGP_METHOD_KRYLOV_MG32 = 4 /* KRYLOV_MG with the preconditioner V-cycle in FP32 (metrics,
                                  stencil tables, corrections, residuals); the Krylov vectors, the
                                  operator, the true residual and the solution stay FP64 */
};

typedef struct {
    int method;
    int ordering;       /* smoother ordering for methods 0/1 */
    int numcycles;      /* methods 0-3: 1 = V, 2 = W; KRYLOV_MG: V-cycles per preconditioner application
                           (the cycle iterated from e = 0), 1 by default */
    int smooth_down, smooth_up, smooth_bottom;  /* KRYLOV_MG: x-line sweeps before and after the
                           coarse correction, and at least smooth_bottom on the coarsest level (the
                           single-problem solver makes half that level's rows, up to 64; mgp_batch 8
                           below a finer level) */
    int maxdepth;       /* -1: as deep as the schedule allows */
    int homog;          /* BCs of the top-level residual */
    int use_graph;      /* KRYLOV_MG: replay each iteration as a captured hipGraph */
    /* KRYLOV_MG / KRYLOV_MG32 only (set by mgp_solver_default_opts): */
    int lines;          /* MGP_LINES_AUTO: x-lines with y-semicoarsening while one workgroup solves a
                           row whole in the cycle's precision (nx <= 4111 in FP64, 8207 in FP32), else
                           y-lines where it solves a column whole, else x-lines in segments;
                           MGP_LINES_X; MGP_LINES_Y: the preconditioner relaxes columns and semicoarsens
                           in x (it runs on the transposed problem).  Lines longer than a whole-line
                           solve are cut into 2, 4, .. equal segments (block Gauss-Seidel along the
                           line) */
    int refine;         /* MGP_REFINE_FP64: residual and solution in FP64; MGP_REFINE_DD: the solution
                           is carried as phi + phi_lo (two FP64 fields) and the refinement residual
                           rhs - L(phi + phi_lo) is evaluated in double-double, so the relative residual
                           can fall below the FP64 operator's rounding (homog 1 only) */
    int order;          /* 2: the reference 9-point operator; 4: the fourth-order finite-volume operator
                           (all-Neumann sides, mapped solvers only): the outer iteration's residual is
                           the fourth-order one, the preconditioner V-cycle stays second order (the
                           deferred-correction pairing of MGPoisson2D_DD.f90:589-862, one order up) */
    int krylov;         /* outer iteration: MGP_KRYLOV_AUTO (BiCGStab for order 2, flexible GMRES for
                           order 4), MGP_KRYLOV_BICGSTAB, MGP_KRYLOV_FGMRES */
    int restart;        /* FGMRES: basis vectors per cycle (restart length); 0: 30 */
} mgp_solver_opts;

enum { MGP_LINES_AUTO = 0, MGP_LINES_X = 1, MGP_LINES_Y = 2 };
enum { MGP_REFINE_FP64 = 0, MGP_REFINE_DD = 1 };
enum { MGP_KRYLOV_AUTO = 0, MGP_KRYLOV_BICGSTAB = 1, MGP_KRYLOV_FGMRES = 2 };

void mgp_solver_default_opts(mgp_solver_opts *opts);
int mgp_solver_create(mgp_solver **out, const mgp_geo *geo, const mgp_bc *bc, const mgp_solver_opts *opts);
int mgp_solver_destroy(mgp_solver *s);
/* As mgp_solver_create, with the metrics generated on the GPU from a mapping (nothing of the
 * geometry crosses PCIe; removes host metric assembly from set-up). */
int mgp_solver_create_mapped(mgp_solver **s, const mgp_mapping *m, int nx, int ny, double dx, double dy,
                             const mgp_bc *bc, const mgp_solver_opts *opts);
/* Upload / download in the host layouts above (phi: one ghost layer). */
int mgp_solver_set_rhs(mgp_solver *s, const double *rhs);
int mgp_solver_set_phi(mgp_solver *s, const double *phi);
int mgp_solver_get_phi(mgp_solver *s, double *phi);
/* Low part of the double-double solution (solvers with refine MGP_REFINE_DD): the solution is
 * phi + phi_lo; layout of phi. */
int mgp_solver_get_phi_lo(mgp_solver *s, double *phi_lo);
/* The compatibility condition of an all-Neumann problem on the resident rhs: rhs -= J m with
 * m = sum(rhs) / sum(J) (compensated sums in FP64), which makes sum(rhs) vanish -- the J-weighted
 * zero-mean projection of f for rhs = -J f.  *removed (may be NULL) receives m.  KRYLOV_MG /
 * KRYLOV_MG32 solvers. */
int mgp_solver_project_mean(mgp_solver *s, double *removed);
/* Device copies between resident fields of the same shape (no PCIe). */
int mgp_solver_copy_rhs_from(mgp_solver *s, const mgp_solver *src);
/* The resident solution of src copied into s: phi, and phi_lo when both carry a double-double
 * solution (else s's phi_lo is cleared; a double-double phi is its value rounded to FP64).  The
 * warm start of one discretisation from another: the fourth-order solve from the second-order
 * solution (deferred correction).  No counterpart in the reference. */
int mgp_solver_copy_phi_from(mgp_solver *s, const mgp_solver *src);
/* KRYLOV_MG / KRYLOV_MG32: ||rhs - A phi|| and ||rhs|| on the resident fields (either pointer may
 * be NULL), with the operator and precision of the solver's refinement residual (fourth order
 * for order 4; double-double on phi + phi_lo for MGP_REFINE_DD); norms as pnorm p = 2.  Turns a
 * target relative to ||rhs|| into solve's tolerance, which is relative to the initial residual.
 * No counterpart in the reference. */
int mgp_solver_residual(mgp_solver *s, double *rnorm, double *rhsnorm);
/* Solve in place on the resident phi/rhs; relres may be NULL (else maxiters+1 doubles). */
int mgp_solver_solve(mgp_solver *s, double tol, int maxiters, int zerophi, double *relres, int *iters,
                     int *outcome);
/* Resident face fluxes for the projection step, in the host layouts of compute_grad:
 * xflux (nx+1) x ny, yflux nx x (ny+1).  The first set_flux allocates them. */
int mgp_solver_set_flux(mgp_solver *s, const double *xflux, const double *yflux);
int mgp_solver_get_flux(mgp_solver *s, double *xflux, double *yflux);
/* The divergence-cleaning step of the reference's projection workflow (test.f90:2817-3024)
 * on the resident fields: rhs = J div F (compute_div); solve L phi = rhs from phi = 0 with the
 * solver's method and BCs (homog from the options); F -= J grad phi (compute_grad, same BCs);
 * divnorm (2 doubles, may be NULL) receives max|J div F| before and after.  The reference
 * workflow's solver is fmg(tol, maxiters 10, numcycles 1, smoothing 8/8/8) with all-Neumann
 * BCs: an MGP_METHOD_FMG solver created with those options reproduces it. */
int mgp_solver_project(mgp_solver *s, double tol, int maxiters, double *divnorm, int *iters, int *outcome);
/* Device pointer + pitch (doubles) of the resident fields, padded layout:
 * element (i, j) at ptr[(j + 1) * pitch + 32 + i], i, j = -1 .. n.  which: 0 phi, 1 rhs,
 * 2 xflux (x-face (i, j) at the same offset), 3 yflux (after the first set_flux). */
int mgp_solver_field(mgp_solver *s, int which, double **ptr, int *pitch);
/* HIP stream the solver enqueues on (a hipStream_t). */
void *mgp_solver_stream(mgp_solver *s);
/* Number of V-cycle applications made by the last solve. */
int mgp_solver_last_cycles(const mgp_solver *s);
/* Measurement hook (KRYLOV_MG, KRYLOV_MG32): times `reps` launches of one finest-level
 * kernel with HIP events on the solver's stream.  kernel 0: x-line
 * relaxation of the even interior rows; 1: residual; 2: plain copy of one
 * finest-level field (bandwidth reference); 3 (KRYLOV_MG32 only): the FP32
 * x-line relaxation of the even interior rows.  DC_VCYCLE solvers with RB
 * ordering: kernel 4: one red-black smoother call of 2 sweeps on the finest
 * level with LDS tiles (tile kernel and copy back); 5: the same call with one
 * launch per colour; 6: the tile kernel alone (cold is ignored; *bytes is 7
 * fields read and 1 written).  VCYCLE solvers with MC ordering: kernel 7: one
 * multi-colour sweep of the finest level (the four interior colours in one
 * LDS-tiled launch, then the ring phase); 8: the same sweep with one launch per
 * interior colour (cold is ignored; *bytes as for 6).  cold != 0: back-to-back launches
 * that take four copies of the operands in turn, so that each launch finds its
 * operands evicted from the L2s and the MALL (as the kernels run inside a
 * solve); cold == 0: back-to-back launches on cache-resident operands.
 * *avg_ms is the mean duration per launch, *bytes the algorithmic bytes one
 * launch moves. */
int mgp_solver_time_kernel(mgp_solver *s, int kernel, int reps, int cold, double *avg_ms, double *bytes);
/* hipDeviceSynchronize on the current device. */
int mgp_device_synchronize(void);

/* ---- 4. Batched solver (no counterpart in the reference, whose callers solve
 *         one problem per call: many independent problems of uneven sizes,
 *         each solved as MGP_METHOD_KRYLOV_MG solves it, with one GPU launch
 *         per level operation for all of them) ---------------------------- */

typedef struct mgp_batch mgp_batch;

/* n problems with host metrics (mgp_geo layouts) and BCs (Dirichlet or Neumann sides,
 * homogeneous).  opts: method MGP_METHOD_KRYLOV_MG or _KRYLOV_MG32, homog 1; smoothing counts, maxdepth and
 * use_graph as for mgp_solver_create.  Each problem keeps its own hierarchy in HBM. */
int mgp_batch_create(mgp_batch **b, int n, const mgp_geo *geos, const mgp_bc *bcs, const mgp_solver_opts *opts);
/* As mgp_batch_create, with problem k's metrics generated on the GPU from maps[k] on an
 * nx[k] x ny[k] grid of spacing dx[k], dy[k]. */
int mgp_batch_create_mapped(mgp_batch **b, int n, const mgp_mapping *maps, const int *nx, const int *ny,
                            const double *dx, const double *dy, const mgp_bc *bcs, const mgp_solver_opts *opts);
int mgp_batch_destroy(mgp_batch *b);
int mgp_batch_count(const mgp_batch *b);
/* Problem k's rhs (nx x ny, the mgp_solver_set_rhs layout) / phi (valid cells, nx x ny). */
int mgp_batch_set_rhs(mgp_batch *b, int k, const double *rhs);
int mgp_batch_get_phi(mgp_batch *b, int k, double *phi);
/* Solves every problem to ||rhs - L phi|| <= tol ||rhs|| from phi = 0, or (zerophi == 0) from
 * the resident phi; iters[k], outcomes[k] (may be NULL) per problem.  A problem's iterates do
 * not depend on the others in the batch. */
int mgp_batch_solve(mgp_batch *b, double tol, int maxiters, int zerophi, int *iters, int *outcomes);
/* V-cycle applications of the longest-running problem in the last solve. */
int mgp_batch_last_cycles(const mgp_batch *b);
/* Measurement hook (between solves): times `reps` back-to-back launch sets of the finest-level
 * line relaxation of the even interior rows of every problem (kernel 0; one launch per chunk
 * class, side by side, as inside a solve) with HIP events on the batch's stream.  *avg_ms is the
 * mean per set, *bytes the algorithmic bytes of one set (each array element it needs once, at
 * the element size of the cycle: 8 bytes, or 4 with KRYLOV_MG32).  Changes the resident
 * finest-level corrections, nothing else. */
int mgp_batch_time_kernel(mgp_batch *b, int kernel, int reps, double *avg_ms, double *bytes);
/* Autotuner of the batched line solve (no counterpart in the reference).  The shape of a row's
 * line solve -- threads per row (64, 128 or 256) and with it the chunk length -- is chosen per key
 * (nx, BC class, precision); BC class: bit k set when side k is Dirichlet; precision: 1 for
 * KRYLOV_MG32.  Tables hold 4 ints per entry: nx, BC class, precision, threads.
 * mgp_batch_tuning exports the batch's current choice per key present (at most max_entries are
 * written; *n_entries receives the count).  mgp_batch_set_tuning applies the entries that match
 * problems of the batch (*changed, may be NULL: problems whose shape changed) and rebuilds the
 * launch tables; MGP_ERR_ARG for a shape with no kernel.  mgp_batch_tune searches one band of
 * row lengths at a time (up to 128, 256, 512, 1024, 4111 cells): each width the band's rows all
 * have a kernel for is timed with `reps` launch sets of the whole batch's finest-level even-row
 * relaxation (HIP events, the other bands as chosen so far) and kept when 2 % faster; it applies
 * the result and exports the choice per key.  A shape moves a solution by rounding
 * only; a committed table of choices keeps results reproducible.  Between solves only (tune
 * relaxes the resident finest-level corrections). */
int mgp_batch_tuning(const mgp_batch *b, int *table, int max_entries, int *n_entries);
int mgp_batch_set_tuning(mgp_batch *b, const int *table, int n_entries, int *changed);
int mgp_batch_tune(mgp_batch *b, int reps, int *table, int max_entries, int *n_entries);

/* ---- 5. Serving loop (no counterpart in the reference): resident geometries
 *         ("slots") with their hierarchies and last solutions in HBM, requests
 *         queued per slot, continuous batching: every step admits queued
 *         requests to idle slots and advances all running solves by one
 *         iteration in the batched launches. ------------------------------ */

typedef struct mgp_server mgp_server;

/* n slots with metrics generated on the GPU (as mgp_batch_create_mapped).  hbm_budget > 0:
 * MGP_ERR_NOMEM if the resident hierarchies would need more device bytes than that. */
int mgp_server_create(mgp_server **v, int n, const mgp_mapping *maps, const int *nx, const int *ny,
                      const double *dx, const double *dy, const mgp_bc *bcs, const mgp_solver_opts *opts,
                      double hbm_budget);
int mgp_server_destroy(mgp_server *v);
/* Device bytes the resident slots hold. */
double mgp_server_footprint(const mgp_server *v);
/* Admits a geometry while serving (as the slots of mgp_server_create, its metrics generated on the
 * GPU): its hierarchy is built, in a batch of its own, if the resident footprint stays within the
 * server's HBM budget; otherwise MGP_ERR_NOMEM and nothing changes.  *slot receives its index. */
int mgp_server_add_slot(mgp_server *v, const mgp_mapping *map, int nx, int ny, double dx, double dy,
                        const mgp_bc *bc, int *slot);
/* Frees an idle slot's hierarchy and warm start (no request of it queued or running, nor of a slot
 * sharing its batch: else MGP_ERR_ARG, retry later); the slot's other batch members keep their warm
 * starts.  Submits to an evicted slot fail with MGP_ERR_ARG. */
int mgp_server_evict(mgp_server *v, int slot);
/* Queues a solve of slot `slot` with this rhs (nx x ny, copied): ||rhs - L phi|| <= tol ||rhs||
 * within maxiters iterations, from phi = 0 or (warm != 0) from the slot's last solution. */
int mgp_server_submit(mgp_server *v, int slot, const double *rhs, double tol, int maxiters, int warm,
                      long long *ticket);
/* Slots are grouped by size (up to 2^22 cells, and larger), each group a batch on a stream of
 * its own, so the groups' work overlaps and a group advances at its own pace.  A step takes the
 * results of every group whose iteration has finished, admits the oldest queued request of each
 * of its idle slots and launches its next iteration; it returns once at least one group has
 * finished an iteration (or none has work).  *running / *queued (may be NULL) receive the solves
 * on the device and waiting. */
int mgp_server_step(mgp_server *v, int *running, int *queued);
/* *done = 1 and phi (nx x ny), *iters, *outcome when the request has finished (its result is
 * then released), else *done = 0.  A result does not depend on the other requests in flight. */
int mgp_server_result(mgp_server *v, long long ticket, double *phi, int *iters, int *outcome, int *done);

#ifdef __cplusplus
}
#endif
#endif
