// mgp_single.h -- stand-alone single-level solvers (mgp_single.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "mgp_common.h"
#include "mgp_host.h"

namespace mgp {

// One grid, device fields in the level layout.  ordering MGP_ORDER_LEX: the
// reference's relax_gs order and sequential-order reductions (bit-identical
// iterates); MGP_ORDER_MC: multi-colour sweeps and parallel reductions.
class SingleLevel {
  public:
    SingleLevel(hipStream_t s, const Grid &g, const Geo &geo, const Bc &bc, int ordering);

    // relax_jacobi: returns the sweeps made; relres[0..sweeps] when tol > 0.
    int jacobi(int homog, const double *invd, double omega, double tol, int maxiters, int zerophi,
               const double *rhs, double *phi, std::vector<double> &relres);
    // solve_bicgstab (version 1): returns the last bookkeeping index i
    // (relres[0..i]); throws MGP_ERR_BREAKDOWN where the reference stops.
    int bicgstab1(int homog, const double *invd, double tol, int max_iters, int max_restarts, int zerophi,
                  const double *rhs, double *phi, std::vector<double> &relres);
    // solve_bicgstab2 (plain operator): returns the iterations; norms gets the
    // residual 2-norm after each one (entry 0: initial).
    int bicgstab2(int homog, const double *invd, double tol, int max_iters, int max_restarts, int zerophi,
                  const double *rhs, double *phi, std::vector<double> &norms);

  private:
    void field(DevBuf &b);
    void zero(double *x);
    double reduce(const double *a, const double *b, int kind);
    double dot(const double *a, const double *b);
    double norm(const double *a, int p);
    void residual(const double *rhs, double *phi, double *out, int homog);
    void laplacian(double *phi, double *out, int homog);
    void relax(double *phi, const double *rhs, const double *invd, double omega, int sweeps, int homog);
    void vec(int op, double *x, double a, const double *y, double b, const double *z, int ghosted);

    hipStream_t s;
    Grid g;
    Geo geo;
    Bc bc;
    int order;
    bool exact;
    Workspace W;
    DevBuf seq;
};

}  // namespace mgp
