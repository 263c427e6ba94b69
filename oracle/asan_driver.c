/*
 * asan_driver.c -- runs every oracle entry point on small grids, every side
 * type and both homogeneous modes, so that a build of the oracle with
 * -fsanitize=address,undefined (oracle/Makefile `asan`) checks it for
 * out-of-bounds accesses, use of freed memory, leaks and undefined behaviour.
 * TEST INFRASTRUCTURE (tests/test_oracle_sanitizers.py runs it); it prints
 * one checksum line so a run that the sanitizers let through is visible.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mgp_oracle.h"

static unsigned long long rng_state = 88172645463325252ull;
static double urand(void)
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) * (1.0 / 9007199254740992.0) - 0.5;
}

static double *vec(size_t n, int random)
{
    double *a = malloc(n * sizeof(double));
    if (!a) abort();
    for (size_t k = 0; k < n; ++k) a[k] = random ? urand() : 0.0;
    return a;
}

static double checksum = 0.0;
static void acc(const double *a, size_t n)
{
    for (size_t k = 0; k < n; ++k)
        if (isfinite(a[k])) checksum += fabs(a[k]) > 1e6 ? 1.0 : a[k];
}

/* metrics: a mild warp so the cross terms are non-zero */
static void metrics(int nx, int ny, double *J, double *Jxx, double *Jxy, double *Jyx, double *Jyy)
{
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) J[i + (size_t)j * nx] = 1.0 + 0.1 * urand();
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            Jxx[i + (size_t)j * (nx + 1)] = 1.0 + 0.1 * urand();
            Jxy[i + (size_t)j * (nx + 1)] = 0.1 * urand();
        }
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i) {
            Jyx[i + (size_t)j * nx] = 0.1 * urand();
            Jyy[i + (size_t)j * nx] = 1.0 + 0.1 * urand();
        }
}

static void one_case(int nx, int ny, const int bct[4], int mode, int homog)
{
    const double dx = 1.0 / nx, dy = 1.0 / ny;
    const size_t ncc = (size_t)nx * ny, ngh = (size_t)(nx + 2) * (ny + 2);
    const size_t nxf = (size_t)(nx + 1) * ny, nyf = (size_t)nx * (ny + 1);
    double *J = vec(ncc, 0), *Jxx = vec(nxf, 0), *Jxy = vec(nxf, 0), *Jyx = vec(nyf, 0), *Jyy = vec(nyf, 0);
    metrics(nx, ny, J, Jxx, Jxy, Jyx, Jyy);
    const int bcm[4] = {mode, mode, mode, mode};
    double *bd[4];
    for (int s = 0; s < 4; ++s) bd[s] = vec(mode == MGO_MODE_NONUNIFORM ? (size_t)(s < 2 ? ny : nx) : 1, 1);
#define GEO nx, ny, dx, dy, J, Jxx, Jxy, Jyx, Jyy
#define BC bct, bcm, bd[0], bd[1], bd[2], bd[3]
    double *phi = vec(ngh, 1), *rhs = vec(ncc, 1), *out = vec(ncc, 0), *invd = vec(ncc, 0);
    double *xf = vec(nxf, 1), *yf = vec(nyf, 1);
    mgo_fill_ghosts(GEO, BC, homog, 1, phi);
    acc(phi, ngh);
    for (int order = 0; order <= 3; ++order) mgo_extrapolate_ghosts(nx, ny, order, phi);
    if (mode != MGO_MODE_LAH) {  /* the Poisson2D operators reject LAH (poisson2d.f90:35-39) */
        mgo_laplacian(GEO, BC, homog, 0, phi, out);
        acc(out, ncc);
        mgo_residual(GEO, BC, homog, rhs, phi, out);
        acc(out, ncc);
        mgo_invdiags(GEO, BC, homog, invd);
        acc(invd, ncc);
        mgo_relax_gs(GEO, BC, homog, invd, 1.0, 2, rhs, phi, MGO_ORDER_LEX);
        if (bct[0] != MGO_BC_PERIODIC && bct[2] != MGO_BC_PERIODIC)
            mgo_relax_gs(GEO, BC, homog, invd, 1.0, 2, rhs, phi, MGO_ORDER_MC);
        acc(phi, ngh);
        double relres[8];
        mgo_relax_jacobi(GEO, BC, homog, invd, 2.0 / 3.0, 1e-3, 6, 0, rhs, phi, relres);
        double hist[16];
        mgo_bicgstab1(GEO, BC, homog, invd, 1e-6, 8, 2, 0, rhs, phi, MGO_ORDER_LEX, hist);
        mgo_bicgstab2(0, GEO, BC, homog, invd, 1e-6, 8, 2, 1, rhs, phi, MGO_ORDER_LEX);
        acc(phi, ngh);
        mgo_compute_grad(GEO, BC, homog, phi, xf, yf);
        acc(xf, nxf);
        mgo_fill_boundary_fluxes(nx, ny, BC, homog, xf, yf);
        mgo_compute_div(nx, ny, dx, dy, xf, yf, out);
        acc(out, ncc);
    }
    mgo_dc_laplacian(GEO, BC, homog, phi, out);
    mgo_dc_invdiags(GEO, invd);
    mgo_dc_relax_gs(GEO, BC, homog, invd, 0.8, 2, rhs, phi, MGO_ORDER_LEX);
    mgo_dc_relax_gs(GEO, BC, homog, invd, 0.8, 2, rhs, phi, MGO_ORDER_RB);
    acc(phi, ngh);
    for (int dir = 1; dir <= 2; ++dir)
        for (int nd = 1; nd <= 2; ++nd) {
            double *pd = vec(nd == 1 ? nxf : nyf, 0);
            mgo_compute_pd(nx, ny, dx, dy, phi, dir, nd, pd);
            acc(pd, nd == 1 ? nxf : nyf);
            free(pd);
        }
    checksum += mgo_pnorm(nx, ny, dx, dy, rhs, 2) + mgo_inner_prod(nx, ny, dx, dy, rhs, rhs) +
                mgo_integrate2d(nx, ny, dx, dy, rhs, J) + mgo_integrate2d_bdry(nx, ny, dx, dy, xf, yf);
    if (nx == ny) checksum += mgo_gpnorm(nx, ny, dx, dy, phi, 2, -1, ny, -1, nx);
    free(J), free(Jxx), free(Jxy), free(Jyx), free(Jyy);
    for (int s = 0; s < 4; ++s) free(bd[s]);
    free(phi), free(rhs), free(out), free(invd), free(xf), free(yf);
#undef GEO
#undef BC
}

static void cycles(int nx, int ny, int kind, const int bct[4])
{
    const double dx = 1.0 / nx, dy = 1.0 / ny;
    const size_t ncc = (size_t)nx * ny, ngh = (size_t)(nx + 2) * (ny + 2);
    const size_t nxf = (size_t)(nx + 1) * ny, nyf = (size_t)nx * (ny + 1);
    double *J = vec(ncc, 0), *Jxx = vec(nxf, 0), *Jxy = vec(nxf, 0), *Jyx = vec(nyf, 0), *Jyy = vec(nyf, 0);
    metrics(nx, ny, J, Jxx, Jxy, Jyx, Jyy);
    const int bcm[4] = {0, 0, 0, 0};
    double z = 0.0;
    double *phi = vec(ngh, 0), *rhs = vec(ncc, 1), relres[5];
    int iters = 0;
    mgo_vcycle(kind, nx, ny, dx, dy, J, Jxx, Jxy, Jyx, Jyy, bct, bcm, &z, &z, &z, &z, 1, 1e-10, 4, -1, 1, 2, 2, 2, 1,
               phi, rhs, kind == 1 ? MGO_ORDER_RB : MGO_ORDER_LEX, relres, &iters);
    acc(phi, ngh);
    int md = 0, rx[64], ry[64];
    mgo_schedule(nx, ny, dx, dy, -1, &md, rx, ry);
    double *crse = vec((size_t)(nx / 2) * (ny / 2), 0);
    mgo_restrict(0, nx, ny, rhs, nx / 2, ny / 2, crse);
    acc(crse, (size_t)(nx / 2) * (ny / 2));
    free(crse);
    free(J), free(Jxx), free(Jxy), free(Jyx), free(Jyy), free(phi), free(rhs);
}

static void fourth(int nx, int ny)
{
    const double dx = 1.0 / nx, dy = 1.0 / ny;
    const size_t nxf = (size_t)(nx + 1) * ny, nyf = (size_t)nx * (ny + 1), ncc = (size_t)nx * ny;
    double *m[8];
    for (int k = 0; k < 8; ++k) m[k] = vec(k < 4 ? nxf : nyf, 0);
    mgo_metric4(nx, ny, dx, dy, 0.12, 8.0, 0.0, 0.0, m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7]);
    double *phi = vec(ncc, 1), *lap = vec(ncc, 0);
    mgo_lap4(nx, ny, dx, dy, m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], phi, lap);
    acc(lap, ncc);
    for (int k = 0; k < 8; ++k) free(m[k]);
    free(phi), free(lap);
}

int main(void)
{
    const int sides[5][4] = {{1, 1, 1, 1}, {1, 1, 0, 0}, {0, 0, 0, 0}, {0, 1, 1, 0}, {2, 2, 1, 1}};
    const int modes[3] = {MGO_MODE_UNIFORM, MGO_MODE_NONUNIFORM, MGO_MODE_LAH};
    for (int c = 0; c < 5; ++c)
        for (int m = 0; m < 3; ++m)
            for (int h = 0; h <= 1; ++h) {
                one_case(12, 12, sides[c], modes[m], h);
                one_case(13, 9, sides[c], modes[m], h);
            }
    for (int kind = 0; kind <= 2; ++kind) {
        cycles(32, 32, kind, sides[0]);
        cycles(32, 32, kind, sides[1]);
    }
    cycles(32, 32, 1, sides[2]);
    fourth(5, 5);
    fourth(17, 11);
    printf("oracle sanitizer run: checksum %.6e\n", checksum);
    return 0;
}
