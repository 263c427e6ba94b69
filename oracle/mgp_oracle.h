/*
 * mgp_oracle.h -- CPU restatement of the EdSantilli/MGPoisson2D hot path.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load liboracle.so, and only as the checker.  The
 * product library (mgpoisson2d_amd/) never calls into it.
 *
 * Every entry point takes the same arguments, in the same order, as the
 * matching ref_* entry of oracle/ref_bridge.f90, so a test can run both
 * through one helper.  Layout conventions (all column-major, i fastest):
 *   valid cells i = 0..nx-1, j = 0..ny-1; spacing dx, dy
 *   J [nx*ny]; Jxx, Jxy [(nx+1)*ny]; Jyx, Jyy [nx*(ny+1)]
 *   one-ghost cell fields [(nx+2)*(ny+2)], element (i,j) at (i+1)+(j+1)*(nx+2)
 *   ghost-free cell fields [nx*ny]
 *   bct[4], bcm[4]: BC type / mode for xlo, xhi, ylo, yhi (codes below)
 *   bd_*: face data (1 value for UNIFORM / LAH, ny or nx values for NONUNIFORM)
 *
 * Parity pins (see tests/test_oracle_vs_reference.py): every function is
 * compared against the unmodified reference built into oracle/_ref, and the
 * operators also against closed forms (5-point Laplacian on the identity map,
 * polynomial exactness of the ghost extrapolations, linearity).
 */
#ifndef MGP_ORACLE_H
#define MGP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* arrayutils.f90:27-35, 48-50 */
enum { MGO_BC_NONE = -1, MGO_BC_NEUM = 0, MGO_BC_DIRI = 1, MGO_BC_PERIODIC = 2,
       MGO_BC_CF = 3, MGO_BC_EXTRAP0 = 4, MGO_BC_EXTRAP1 = 5, MGO_BC_EXTRAP2 = 6 };
enum { MGO_MODE_UNIFORM = 0, MGO_MODE_NONUNIFORM = 1, MGO_MODE_LAH = 2 };

/* Relaxation orderings.  LEX is the reference's (poisson2d.f90:1561-1959,
 * MGPoisson2D_DD.f90:288-306); RB is the reference DC smoother's disabled
 * red-black branch (MGPoisson2D_DD.f90:262-285); MC is this project's
 * multi-colour ordering for the 9-point operator (interior 4 colours, then
 * the boundary ring in 4 colours, then the 4 corners; ghosts refreshed
 * between interior and ring). */
enum { MGO_ORDER_LEX = 0, MGO_ORDER_RB = 1, MGO_ORDER_MC = 2 };

#define MGO_GEO_ARGS int nx, int ny, double dx, double dy, const double *J, \
    const double *Jxx, const double *Jxy, const double *Jyx, const double *Jyy
#define MGO_BC_ARGS const int *bct, const int *bcm, const double *bd_xlo, \
    const double *bd_xhi, const double *bd_ylo, const double *bd_yhi

void mgo_fill_ghosts(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, int do_neum, double *phi);
void mgo_extrapolate_ghosts(int nx, int ny, int order, double *phi);
void mgo_laplacian(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, int jscale, double *phi, double *lap);
void mgo_residual(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, const double *rhs, double *phi, double *res);
void mgo_invdiags(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, double *invd);
void mgo_relax_gs(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, const double *invd, double omega,
                  int maxiters, const double *rhs, double *phi, int order);
/* One relax_gs point update at (i, j) on phi as it stands (no ghost fill). */
void mgo_gs_point(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, const double *invd, double omega,
                  const double *rhs, double *phi, int i, int j);
void mgo_dc_laplacian(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, double *phi, double *lap);
void mgo_dc_invdiags(MGO_GEO_ARGS, double *invd);
void mgo_dc_relax_gs(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, const double *invd, double omega,
                     int maxiters, const double *rhs, double *phi, int order);
void mgo_restrict(int kind, int fnx, int fny, const double *fine, int cnx, int cny, double *crse);
void mgo_prolong(int fnx, int fny, double fdx, double fdy, const double *fJ, const double *fJxx,
                 const double *fJxy, const double *fJyx, const double *fJyy,
                 int cnx, int cny, double cdx, double cdy, const double *cJ, const double *cJxx,
                 const double *cJxy, const double *cJyx, const double *cJyy,
                 const int *bct, const int *bcm, int order, double *fine, double *crse);
void mgo_schedule(int nx, int ny, double dx, double dy, int maxdepth_in, int *maxdepth_out,
                  int *refx, int *refy);
double mgo_pnorm(int nx, int ny, double dx, double dy, const double *a, int p);
/* inner_prod (arrayutils.f90:637-661) and integrate2d (:670-710, J == NULL: no J scaling)
 * over ghost-free nx x ny fields, in the reference's summation order. */
double mgo_inner_prod(int nx, int ny, double dx, double dy, const double *a, const double *b);
double mgo_integrate2d(int nx, int ny, double dx, double dy, const double *a, const double *J);
/* integrate2d_bdry (arrayutils.f90:721-748) and gpnorm (:802-859, the columns
 * cx0, cx1 and rows ry0, ry1 that the reference reads, -1 .. n), in the
 * reference's summation order. */
double mgo_integrate2d_bdry(int nx, int ny, double dx, double dy, const double *xflux, const double *yflux);
double mgo_gpnorm(int nx, int ny, double dx, double dy, const double *phi, int p, int cx0, int cx1, int ry0,
                  int ry1);
void mgo_bicgstab2(int kind, MGO_GEO_ARGS, MGO_BC_ARGS, int homog, const double *invd, double tol,
                   int maxiters, int maxrestarts, int zerophi, const double *rhs, double *phi,
                   int order);
/* kind 0: MGPoisson2D::vcycle, kind 1: DeferredMGPoisson2D::vcycle,
 * kind 2: MGPoisson2D::fmg (MGPoisson2D.f90:860).
 * relres[0..maxiters] receives the relative-residual history, *iters the
 * number of cycles run (the reference prints this history at verbosity >= 1,
 * MGPoisson2D.f90:590-594,616-618).  Returns 0 converged, 1 diverged (last
 * correction undone), 2 iteration limit reached. */
int mgo_vcycle(int kind, MGO_GEO_ARGS, MGO_BC_ARGS, int homog, double tol, int maxiters,
               int maxdepth, int numcycles, int sdown, int sup, int sbottom, int zerophi,
               double *phi, double *rhs, int order, double *relres, int *iters);

/* Flux tools (poisson2d.f90:404-879).  compute_pd: pd nodal in direction
 * nodedir (1: x-face field (nx+1) x ny, 2: y-face field nx x (ny+1)), derivative
 * along dir.  compute_grad fills phi's ghosts, then J grad phi with Neumann
 * boundary fluxes.  compute_div: J div F into an nx x ny field. */
void mgo_compute_pd(int nx, int ny, double dx, double dy, const double *phi, int dir, int nodedir, double *pd);
void mgo_compute_grad(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, double *phi, double *xflux, double *yflux);
/* relax_jacobi (poisson2d.f90:1358-1451): returns the sweeps made; relres (maxiters+1)
 * gets the relative 2-norm history when tol > 0. */
int mgo_relax_jacobi(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, const double *invd, double omega, double tol,
                     int maxiters, int zerophi, const double *rhs, double *phi, double *relres);
/* solve_bicgstab, version 1 (poisson2d.f90:1993-2248): returns the last bookkeeping
 * index i (relres, max_iters+max_restarts+1 doubles, holds 0..i), or -1 where the
 * reference stops on beta == 0.  order: relax_gs ordering of the preconditioner. */
int mgo_bicgstab1(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, const double *invd, double tol, int max_iters,
                  int max_restarts, int zerophi, const double *rhs, double *phi, int order, double *relres);
/* Metric assembly of the reference's test geometry (test.f90:320-716): terrain-following
 * y over el(x) = amp sin^2(pi x / L), J and J g^ij from differences of point positions. */
void mgo_metric_terrain(int nx, int ny, double dx, double dy, double amp, double *J, double *Jxx, double *Jxy,
                        double *Jyx, double *Jyy);
/* fill_boundary_fluxes (arrayutils.f90:1571-1622): Neumann faces take the BC flux. */
void mgo_fill_boundary_fluxes(int nx, int ny, MGO_BC_ARGS, int homog, double *xflux, double *yflux);
void mgo_compute_div(int nx, int ny, double dx, double dy, const double *xflux, const double *yflux,
                     double *div);

/* Fourth-order finite-volume operator (mgp_oracle4.c; no reference
 * counterpart).  Face arrays: x faces (nx+1)*ny, y faces nx*(ny+1), i fastest;
 * phi and lap are nx*ny cell averages (no ghosts); nx, ny >= 5. */
void mgo_metric4(int nx, int ny, double dx, double dy, double amp, double beta, double L, double H, double *A4,
                 double *Bx4, double *dA, double *dBx, double *C4, double *By4, double *dC, double *dBy);
void mgo_lap4(int nx, int ny, double dx, double dy, const double *A4, const double *Bx4, const double *dA,
              const double *dBx, const double *C4, const double *By4, const double *dC, const double *dBy,
              const double *phi, double *lap);

#ifdef __cplusplus
}
#endif
#endif
