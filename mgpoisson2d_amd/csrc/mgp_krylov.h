// mgp_krylov.h -- the resident fast solver: FP64 BiCGStab on the reference
// operator, preconditioned by one semicoarsening (y) V-cycle with x-line
// zebra relaxation.  No counterpart in the reference (its own cycle drivers
// stall or diverge on strongly stretched mapped grids); the operator it
// solves is the reference's, bit for bit.
#pragma once

#include "mgp_host.h"

namespace mgp {

struct KrylovMG {
    struct Impl;
    Impl *impl;
    int last_cycles = 0;
    bool built = false;
    KrylovMG();
    ~KrylovMG();
    KrylovMG(const KrylovMG &) = delete;
    KrylovMG &operator=(const KrylovMG &) = delete;
    // map: the mapping the metrics came from (null for uploaded metrics); the
    // fourth-order operator (o.order == 4) needs it for its face-averaged metrics.
    void build(hipStream_t s, const DevGeo &top, const Bc &user_bc, const mgp_solver_opts &o,
               const mgp_mapping *map = nullptr);
    // Low part of the double-double solution (MGP_REFINE_DD; null otherwise).
    double *phi_lo() const;
    // rhs -= J sum(rhs) / sum(J) over the valid cells (the compatibility
    // condition of an all-Neumann problem); returns sum(rhs) / sum(J) removed.
    double project_mean(hipStream_t s, const DevGeo &top, double *rhs);
    int solve(Workspace &W, hipStream_t s, const DevGeo &top, const Bc &user_bc, int homog, double *phi,
              double *rhs, double tol, int maxiters, int zerophi, double *relres, int *iters);
    // ||rhs - A phi|| (the operator and precision of the refinement residual:
    // fourth order, double-double with phi + lo) and, if rhsnorm, ||rhs||; the
    // residual is left in the solver's residual field.
    double residual_norm(Workspace &W, hipStream_t s, const DevGeo &top, const Bc &user_bc, int homog, double *phi,
                         double *rhs, double *rhsnorm);
    void ensure_fgmres(const Grid &g);
    // The flexible GMRES outer iteration of solve (opts.krylov); true_residual()
    // leaves rhs - A phi in the solver's residual field and returns its norm.
    template <typename TR>
    int solve_fgmres(hipStream_t s, const DevGeo &top, const Bc &user_bc, double *phi, double tol, int maxiters,
                     int zerophi, double *relres, int *iters, TR &true_residual);
    // Times `reps` launches of one fine-level kernel with HIP events on stream s.
    // kernel 0: x-line relaxation of the even interior rows; 1: residual
    // (9-point operator); 2: plain copy of one field.  Returns the mean ms per launch; *bytes receives the
    // algorithmic bytes one launch moves (each array element it needs, once).
    // cold != 0: each launch timed alone after the caches have been flushed.
    double time_kernel(hipStream_t s, int kernel, int reps, int cold, double *bytes);
};

// One x-line zebra sweep of the homogeneous operator on one level: even
// interior rows, odd interior rows, then rows 0 and ny-1.  K is the level's
// ring stencil table (launch_ring_coef).  e's ghost cells are not read.
void xline_zebra(hipStream_t s, const Grid &g, const Geo &geo, const double *K, double *e, const double *f);

}  // namespace mgp
