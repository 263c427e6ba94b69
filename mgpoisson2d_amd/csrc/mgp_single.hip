// mgp_single.hip -- stand-alone single-level solvers on one grid.
//
//   Poisson2D::relax_jacobi     poisson2d.f90:1358-1451
//   Poisson2D::solve_bicgstab   poisson2d.f90:1993-2248  (version 1)
//   Poisson2D::solve_bicgstab2  poisson2d.f90:2253-2477  (grid-wide, for full-size grids)
//
// Every field operation is one grid-wide kernel; the scalar control flow
// (convergence, restart and divergence tests) runs on the host, which reads
// each reduction back.  Two reduction modes:
//   exact -- one thread sums in the reference's order (inner_prod: per-row dot
//            products, each scaled by the cell volume and accumulated; pnorm:
//            cell by cell).  With the lexicographic relax_gs preconditioner the
//            iterates are bit-identical to the reference.
//   fast  -- the library's deterministic parallel compensated reductions; with
//            the multi-colour preconditioner this is the full-size path.
// Work fields the reference defines like phi carry the ghost ring and are
// updated there too (its whole-array syntax); all start at zero.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "mgp2d.h"
#include "mgp_host.h"
#include "mgp_kernels.h"
#include "mgp_single.h"

namespace mgp {
namespace {

constexpr int TX = 64, TY = 4;

dim3 cell_grid(const Grid &g, int o) { return dim3((g.nx + 2 * o + TX - 1) / TX, (g.ny + 2 * o + TY - 1) / TY); }

// phi = omega * (phi + r * invd) + (1 - omega) * phi on the valid cells
// (poisson2d.f90:1420-1425).
__global__ void __launch_bounds__(TX * TY) k_jacobi(double *__restrict__ phi, const double *__restrict__ r,
                                                    const double *__restrict__ invd, Grid g, double omega) {
    const int i = blockIdx.x * TX + threadIdx.x, j = blockIdx.y * TY + threadIdx.y;
    if (i >= g.nx || j >= g.ny) return;
    const size_t c = ix(g.pitch, i, j);
    const double newphi = phi[c] + r[c] * invd[c];
    phi[c] = omega * newphi + (1.0 - omega) * phi[c];
}

enum VecOp : int {
    V_AXPY = 0,     // x = x + a*y
    V_AXPY_BZ = 1,  // x = (x + a*y) + b*z
    V_P1 = 2,       // x = (a*x - b*y) + z   solve_bicgstab's p update (:2100-2103), a*x rounded first
    V_P2 = 3        // x = a*(x - b*y) + z   solve_bicgstab2's p update (:2374)
};

// Field update over the valid cells (o = 0) or cells and ghost ring (o = 1).
// x - a*y is computed as x + (-a)*y, which rounds identically.
__global__ void __launch_bounds__(TX * TY) k_vec(int op, double *__restrict__ x, double a, const double *__restrict__ y,
                                                 double b, const double *__restrict__ z, Grid g, int o) {
    const int i = blockIdx.x * TX + threadIdx.x - o, j = blockIdx.y * TY + threadIdx.y - o;
    if (i >= g.nx + o || j >= g.ny + o) return;
    const size_t c = ix(g.pitch, i, j);
    switch (op) {
    case V_AXPY: x[c] = x[c] + a * y[c]; break;
    case V_AXPY_BZ: x[c] = x[c] + a * y[c] + b * z[c]; break;
    case V_P1: {
        const double p1 = a * x[c];
        x[c] = p1 - b * y[c] + z[c];
        break;
    }
    default: x[c] = a * (x[c] - b * y[c]) + z[c]; break;
    }
}

// Reductions in the reference's order, one thread: kind 0 inner_prod(a, b)
// (arrayutils.f90:637-661), 1 sum of vol*a^2 (pnorm p = 2 before the root,
// :770-775), 2 max |a| (pnorm p = 0, :763-768).
__global__ void k_seq_reduce(double *out, const double *__restrict__ a, const double *__restrict__ b, Grid g,
                             int kind) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double vol = g.dx * g.dy;
    double acc = 0.0;
    for (int j = 0; j < g.ny; ++j) {
        const double *ra = a + ix(g.pitch, 0, j);
        if (kind == 0) {
            const double *rb = b + ix(g.pitch, 0, j);
            double dot = 0.0;
            for (int i = 0; i < g.nx; ++i) dot = dot + ra[i] * rb[i];
            acc = acc + vol * dot;
        } else if (kind == 1) {
            for (int i = 0; i < g.nx; ++i) acc = acc + vol * (ra[i] * ra[i]);
        } else {
            for (int i = 0; i < g.nx; ++i) acc = fmax(acc, fabs(ra[i]));
        }
    }
    *out = acc;
}

}  // namespace

SingleLevel::SingleLevel(hipStream_t s_, const Grid &g_, const Geo &geo_, const Bc &bc_, int ordering)
    : s(s_), g(g_), geo(geo_), bc(bc_), order(ordering), exact(ordering == MGP_ORDER_LEX) {
    seq.alloc(8);
}

void SingleLevel::field(DevBuf &b) {
    b.alloc(level_elems(g.nx, g.ny));
    zero(b.p);
}

void SingleLevel::zero(double *x) { launch_setval(s, x, 0.0, g, 1); }

double SingleLevel::reduce(const double *a, const double *b, int kind) {
    if (!exact) return W.reduce(s, a, b, nullptr, g.dx * g.dy, g, kind);
    k_seq_reduce<<<1, 64, 0, s>>>(seq.p, a, b, g, kind);
    double v = 0.0;
    hip_check(hipMemcpyAsync(&v, seq.p, sizeof(double), hipMemcpyDeviceToHost, s), "scalar");
    hip_check(hipStreamSynchronize(s), "sync");
    return v;
}

double SingleLevel::dot(const double *a, const double *b) { return reduce(a, b, 0); }

double SingleLevel::norm(const double *a, int p) { return p == 0 ? reduce(a, nullptr, 2) : std::sqrt(reduce(a, nullptr, 1)); }

void SingleLevel::residual(const double *rhs, double *phi, double *out, int homog) {
    launch_fill_ghosts(s, phi, g, geo, bc, homog, 0);
    launch_laplacian(s, out, phi, rhs, g, geo, bc, homog, 0);
}

void SingleLevel::laplacian(double *phi, double *out, int homog) {
    launch_fill_ghosts(s, phi, g, geo, bc, homog, 0);
    launch_laplacian(s, out, phi, nullptr, g, geo, bc, homog, 0);
}

void SingleLevel::relax(double *phi, const double *rhs, const double *invd, double omega, int sweeps, int homog) {
    if (sweeps <= 0) return;
    if (order == MGP_ORDER_LEX) {
        launch_gs_lex(s, phi, rhs, invd, g, geo, bc, homog, omega, sweeps);
        return;
    }
    for (int k = 0; k < sweeps; ++k) {
        launch_gs_mc_interior_tiled(s, phi, rhs, invd, g, geo, omega);
        launch_gs_mc_ring(s, phi, rhs, invd, g, geo, bc, homog, omega);
    }
}

void SingleLevel::vec(int op, double *x, double a, const double *y, double b, const double *z, int ghosted) {
    const int o = ghosted ? 1 : 0;
    k_vec<<<cell_grid(g, o), dim3(TX, TY), 0, s>>>(op, x, a, y, b, z, g, o);
}

// --------------------------------------------------------------------------
// relax_jacobi
// --------------------------------------------------------------------------
int SingleLevel::jacobi(int homog, const double *invd, double omega, double tol, int maxiters, int zerophi,
                        const double *rhs, double *phi, std::vector<double> &relres) {
    relres.assign((size_t)maxiters + 1, 0.0);
    if (maxiters == 0) return 0;
    DevBuf r;
    field(r);
    if (zerophi) zero(phi);
    residual(rhs, phi, r.p, homog);
    double rscale = 1.0;
    if (tol > 0.0) {
        rscale = norm(r.p, 2);
        if (rscale == 0.0) {
            relres[0] = rscale;
            rscale = 1.0;
        } else {
            relres[0] = 1.0;
        }
    }
    int it;
    for (it = 1; it <= maxiters; ++it) {
        k_jacobi<<<cell_grid(g, 0), dim3(TX, TY), 0, s>>>(phi, r.p, invd, g, omega);
        residual(rhs, phi, r.p, homog);
        if (tol > 0.0) {
            relres[it] = norm(r.p, 2) / rscale;
            if (relres[it] <= tol) break;
        }
    }
    hip_check(hipStreamSynchronize(s), "sync");
    return it > maxiters ? maxiters : it;
}

// --------------------------------------------------------------------------
// solve_bicgstab (version 1): two relax_gs sweeps (omega 1 - 1/3) precondition
// each half step; max-norm residuals; restarts on hanging or divergence.
// --------------------------------------------------------------------------
int SingleLevel::bicgstab1(int homog, const double *invd, double tol, int max_iters, int max_restarts, int zerophi,
                           const double *rhs, double *phi, std::vector<double> &relres) {
    const double hang = 1.0e-8, relax_omega = 1.0 - 1.0 / 3.0;
    relres.assign((size_t)max_iters + max_restarts + 1, 0.0);
    if (max_iters == 0) return 0;
    DevBuf r, r0, nu, p, t, y, z;
    for (DevBuf *b : {&r, &r0, &nu, &p, &t, &y, &z}) field(*b);
    const size_t bytes = r.n * sizeof(double);
    if (zerophi) zero(phi);
    std::vector<double> rho(relres.size(), 0.0), om(relres.size(), 0.0);
    int i = 0, is_restart = 0, num_restarts = 0;
    residual(rhs, phi, r.p, homog);
    hip_check(hipMemcpyAsync(r0.p, r.p, bytes, hipMemcpyDeviceToDevice, s), "copy");
    const double rscale = norm(r0.p, 0);
    relres[0] = 1.0;
    if (rscale == 0.0) return 0;
    double alpha = 1.0;
    rho[0] = 1.0;
    om[0] = 1.0;
    for (int iter = 1; iter <= max_iters; ++iter) {
        ++i;
        rho[i] = dot(r0.p, r.p);
        const double beta = (rho[i] / rho[i - 1]) * (alpha / om[i - 1]);
        if (beta == 0.0) {
            g_last_error = "solve_bicgstab: beta is zero. Probably due to a zero rhs.";
            throw Failure(MGP_ERR_BREAKDOWN);
        }
        vec(V_P1, p.p, beta, nu.p, beta * om[i - 1], r.p, 1);
        launch_lincomb_mul(s, y.p, p.p, invd, g, 0);
        relax(y.p, p.p, invd, relax_omega, 2, homog);
        laplacian(y.p, nu.p, homog);
        alpha = dot(r0.p, nu.p);
        alpha = rho[i] / alpha;
        vec(V_AXPY, r.p, -alpha, nu.p, 0.0, nullptr, 1);
        launch_lincomb_mul(s, z.p, r.p, invd, g, 0);
        relax(z.p, r.p, invd, relax_omega, 2, homog);
        laplacian(z.p, t.p, homog);
        om[i] = dot(t.p, r.p) / dot(t.p, t.p);
        vec(V_AXPY_BZ, phi, alpha, y.p, om[i], z.p, 1);
        vec(V_AXPY, r.p, -om[i], t.p, 0.0, nullptr, 1);
        const double lastres = is_restart ? 1.0e200 : relres[i - 1];
        relres[i] = norm(r.p, 0) / rscale;
        if (relres[i] <= tol) break;
        const bool hanging = std::fabs(relres[i] - lastres) < tol * hang;
        const bool diverging = !hanging && relres[i] > lastres;
        if (hanging || diverging) {
            if (num_restarts < max_restarts) {
                // a fresh residual takes the next bookkeeping slot
                ++num_restarts;
                ++i;
                residual(rhs, phi, r.p, homog);
                hip_check(hipMemcpyAsync(r0.p, r.p, bytes, hipMemcpyDeviceToDevice, s), "copy");
                relres[i] = norm(r.p, 0) / rscale;
                alpha = 1.0;
                rho[i] = 1.0;
                om[i] = 1.0;
                zero(nu.p);
                zero(p.p);
                is_restart = 1;
                continue;
            }
            if (diverging) {  // undo the last correction
                vec(V_AXPY, r.p, om[i], t.p, 0.0, nullptr, 1);
                vec(V_AXPY_BZ, phi, -alpha, y.p, -om[i], z.p, 1);
            }
            break;
        }
        is_restart = 0;
    }
    hip_check(hipStreamSynchronize(s), "sync");
    return i;
}

// --------------------------------------------------------------------------
// solve_bicgstab2 on the plain operator (homogeneous preconditioning sweeps
// and operator inside, as in the reference); returns the iterations.
// --------------------------------------------------------------------------
int SingleLevel::bicgstab2(int homog, const double *invd, double tol, int max_iters, int max_restarts, int zerophi,
                           const double *rhs, double *phi, std::vector<double> &norms) {
    const double reps = 1.0e-12, hang = 1.0e-7, small = 1.0e-30, relax_omega = 1.0 - 1.0 / 3.0;
    norms.clear();
    if (max_iters == 0) return 0;
    DevBuf e, pt, st, r, rt, p, t, v;
    for (DevBuf *b : {&e, &pt, &st, &r, &rt, &p, &t, &v}) field(*b);
    const size_t bytes = r.n * sizeof(double);
    if (zerophi) zero(phi);
    int recount = 0, it = 0, restarts = 0, init = 1;
    residual(rhs, phi, r.p, homog);
    hip_check(hipMemcpyAsync(rt.p, r.p, bytes, hipMemcpyDeviceToDevice, s), "copy");
    double norm0 = norm(r.p, 2);
    const double initial_norm = norm0, initial_rnorm = norm0;
    double norm1 = norm0;
    double rho1 = 0, rho2 = 0, alpha0 = 0, alpha1 = 0, omega0 = 0, omega1 = 0;
    norms.push_back(norm0);
    while (it < max_iters && norm0 > tol * norm1 && norm1 > 0.0) {
        ++it;
        norm1 = norm0;
        alpha1 = alpha0;
        omega1 = omega0;
        rho2 = rho1;
        rho1 = dot(rt.p, r.p);
        if (rho1 == 0.0) break;  // "rho = 0": apply e and return
        if (init) {
            hip_check(hipMemcpyAsync(p.p, r.p, bytes, hipMemcpyDeviceToDevice, s), "copy");
            init = 0;
        } else {
            const double beta1 = (rho1 / rho2) * (alpha1 / omega1);
            vec(V_P2, p.p, beta1, v.p, omega1, r.p, 0);
        }
        launch_lincomb_mul(s, pt.p, p.p, invd, g, 0);
        relax(pt.p, p.p, invd, relax_omega, 2, 1);
        laplacian(pt.p, v.p, 1);
        const double m = dot(rt.p, v.p);
        alpha0 = rho1 / m;
        if (std::fabs(m) > small * std::fabs(rho1)) {
            vec(V_AXPY, r.p, -alpha0, v.p, 0.0, nullptr, 0);
            norm0 = norm(r.p, 2);
            vec(V_AXPY, e.p, alpha0, pt.p, 0.0, nullptr, 1);
        } else {
            launch_setval(s, r.p, 0.0, g, 0);
            norm0 = 0.0;
        }
        if (norm0 > tol * initial_norm && norm0 > reps * initial_rnorm) {
            launch_lincomb_mul(s, st.p, r.p, invd, g, 0);
            relax(st.p, r.p, invd, relax_omega, 2, 1);
            laplacian(st.p, t.p, 1);
            omega0 = dot(t.p, r.p) / dot(t.p, t.p);
            vec(V_AXPY, e.p, omega0, st.p, 0.0, nullptr, 1);
            vec(V_AXPY, r.p, -omega0, t.p, 0.0, nullptr, 0);
            norm0 = norm(r.p, 2);
        }
        norms.push_back(norm0);
        if (norm0 <= tol * initial_norm || norm0 <= reps * initial_rnorm) break;
        if (omega0 == 0.0 || norm0 > (1.0 - hang) * norm1) {
            if (recount == 0) {
                recount = 1;
            } else {
                recount = 0;
                vec(V_AXPY, phi, 1.0, e.p, 0.0, nullptr, 1);
                // the reference exits here and adds e once more after the loop
                if (restarts == max_restarts) break;
                residual(rhs, phi, r.p, homog);
                norm0 = norm(r.p, 2);
                rho1 = rho2 = 0.0;
                alpha0 = 0.0;
                omega0 = 0.0;
                hip_check(hipMemcpyAsync(rt.p, r.p, bytes, hipMemcpyDeviceToDevice, s), "copy");
                zero(e.p);
                ++restarts;
                init = 1;
            }
        }
    }
    vec(V_AXPY, phi, 1.0, e.p, 0.0, nullptr, 1);
    hip_check(hipStreamSynchronize(s), "sync");
    return it;
}

}  // namespace mgp
