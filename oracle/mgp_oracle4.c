/*
 * mgp_oracle4.c -- CPU restatement of the fourth-order finite-volume operator
 * that the fourth-order deferred correction solves (TEST INFRASTRUCTURE, part
 * of liboracle.so; only tests/ and bench.py's CPU leg load it).
 *
 * The reference has no fourth-order operator: its deferred correction pairs
 * the second-order 9-point operator outside with a low-order one inside
 * (MGPoisson2D_DD.f90:589-862, residual at :777,808).  This file states the
 * operator the fourth-order variant evaluates outside, so that the GPU kernels
 * have a checker; it is pinned by closed forms and convergence rates, not by
 * the reference (tests/test_oracle4.py).
 *
 * Method: the mapped-grid finite-volume discretisation of Colella, Dorr,
 * Hittinger and Martin, "High-order, finite-volume methods in mapped
 * coordinates", J. Comput. Phys. 230 (2011) 2952-2976.  Unknowns are cell
 * averages over computational cells; L4 phi is the divergence of face-averaged
 * fluxes,
 *   <F^xi>  = <A> <phi_xi> + <B> <phi_eta> + (deta^2/12) (dA/deta d(phi_xi)/deta + dB/deta d(phi_eta)/deta)
 *   <F^eta> = <C> <phi_eta> + <B> <phi_xi> + (dxi^2/12) (dC/dxi d(phi_eta)/dxi + dB/dxi d(phi_xi)/dxi)
 * with A = J g^11, B = J g^12 = J g^21, C = J g^22 (the product rule for face
 * averages, their eq. (13)).  Face averages of phi and of its normal
 * derivative come from cell averages by the fourth-order interpolants: centred
 * (-1, 7, 7, -1)/12 and (1, -15, 15, -1)/(12 h); on the first interior face and
 * the boundary face, the unique quartic through five cell averages.  Tangential
 * derivatives of face averages are differences of node values interpolated the
 * same way along the face.  The correction terms use second-order differences
 * (one-sided at the ends).  Boundary faces carry the homogeneous Neumann flux,
 * zero.  Requires nx, ny >= 5.
 *
 * Metric set (mgo_metric4, the warped mapping of the bench geometry):
 *   x faces (i = 0..nx, j = 0..ny-1), (nx+1)*ny, i fastest:
 *     A4 = <A>, Bx4 = <B> (3-point Gauss-Legendre over the face),
 *     dA = (deta/12) (A(xi_i, eta_j+1) - A(xi_i, eta_j)), dBx likewise
 *   y faces (i = 0..nx-1, j = 0..ny), nx*(ny+1): C4, By4, dC, dBy
 * Every sum is evaluated left to right as written here; the GPU kernels use
 * the same order, so on the same metric arrays the results agree bit for bit.
 */
#include <math.h>
#include <stddef.h>

#include "mgp_oracle.h"

typedef struct {
    int nx, ny;
    double dx, dy;
} G4;

#define CA(f, i, j) (f)[(size_t)(i) + (size_t)(j) * (size_t)g->nx]
#define XA(f, i, j) (f)[(size_t)(i) + (size_t)(j) * (size_t)(g->nx + 1)]
#define YA(f, i, j) (f)[(size_t)(i) + (size_t)(j) * (size_t)g->nx]

/* A line of n cell values c(k) = base[k * stride]. */
typedef struct {
    const double *base;
    ptrdiff_t stride;
    int n;
} Line;

static double lc(const Line *l, int k) { return l->base[(ptrdiff_t)k * l->stride]; }
/* mirrored index: the k-th cell from the high end */
static double lm(const Line *l, int k) { return l->base[(ptrdiff_t)(l->n - 1 - k) * l->stride]; }

/* Point value at node f (0..n) of the line from its cell averages. */
static double face_val(const Line *l, int f)
{
    const int n = l->n;
    if (f == 0)
        return ((((137.0 * lc(l, 0) - 163.0 * lc(l, 1)) + 137.0 * lc(l, 2)) - 63.0 * lc(l, 3)) + 12.0 * lc(l, 4)) *
               (1.0 / 60.0);
    if (f == n)
        return ((((137.0 * lm(l, 0) - 163.0 * lm(l, 1)) + 137.0 * lm(l, 2)) - 63.0 * lm(l, 3)) + 12.0 * lm(l, 4)) *
               (1.0 / 60.0);
    if (f == 1)
        return ((((12.0 * lc(l, 0) + 77.0 * lc(l, 1)) - 43.0 * lc(l, 2)) + 17.0 * lc(l, 3)) - 3.0 * lc(l, 4)) *
               (1.0 / 60.0);
    if (f == n - 1)
        return ((((12.0 * lm(l, 0) + 77.0 * lm(l, 1)) - 43.0 * lm(l, 2)) + 17.0 * lm(l, 3)) - 3.0 * lm(l, 4)) *
               (1.0 / 60.0);
    return (7.0 * (lc(l, f - 1) + lc(l, f)) - (lc(l, f - 2) + lc(l, f + 1))) * (1.0 / 12.0);
}

/* Derivative at interior node f (1..n-1); ih12 = 1/(12 h), ih60 = 1/(60 h). */
static double face_der(const Line *l, int f, double ih12, double ih60)
{
    const int n = l->n;
    if (f == 1)
        return ((((-50.0 * lc(l, 0) + 25.0 * lc(l, 1)) + 45.0 * lc(l, 2)) - 25.0 * lc(l, 3)) + 5.0 * lc(l, 4)) * ih60;
    if (f == n - 1)
        return -(((((-50.0 * lm(l, 0) + 25.0 * lm(l, 1)) + 45.0 * lm(l, 2)) - 25.0 * lm(l, 3)) + 5.0 * lm(l, 4)) *
                 ih60);
    return (15.0 * (lc(l, f) - lc(l, f - 1)) - (lc(l, f + 1) - lc(l, f - 2))) * ih12;
}

/* A face quantity along the tangential direction, by a callback. */
typedef double (*TanFn)(const void *ctx, int t);

/* Tangential node value at node m (0..n) of the face averages q(0..n-1). */
static double tan_node(TanFn q, const void *ctx, int n, int m)
{
    double c[5];
    if (m == 0 || m == 1) {
        for (int k = 0; k < 5; ++k) c[k] = q(ctx, k);
        Line l = {c, 1, 5};
        return face_val(&l, m);
    }
    if (m == n || m == n - 1) {
        for (int k = 0; k < 5; ++k) c[k] = q(ctx, n - 5 + k);
        Line l = {c, 1, 5};
        return face_val(&l, m - (n - 5));
    }
    const double a = q(ctx, m - 2), b = q(ctx, m - 1), d = q(ctx, m), e = q(ctx, m + 1);
    return (7.0 * (b + d) - (a + e)) * (1.0 / 12.0);
}

typedef struct {
    const G4 *g;
    const double *phi;
    int face;  /* node index of the face across its normal direction */
} FaceCtx;

/* x faces: the face value along row t of face i */
static double xface_val(const void *vc, int t)
{
    const FaceCtx *c = (const FaceCtx *)vc;
    const G4 *g = c->g;
    Line l = {&CA(c->phi, 0, t), 1, g->nx};
    return face_val(&l, c->face);
}
/* y faces: the face value along column t of face j */
static double yface_val(const void *vc, int t)
{
    const FaceCtx *c = (const FaceCtx *)vc;
    const G4 *g = c->g;
    Line l = {&CA(c->phi, t, 0), g->nx, g->ny};
    return face_val(&l, c->face);
}

static double flux_x(const G4 *g, const double *phi, const double *A4, const double *Bx4, const double *dA,
                     const double *dBx, int i, int j)
{
    const int nx = g->nx, ny = g->ny;
    if (i == 0 || i == nx) return 0.0;
    const double idx = 1.0 / g->dx, idy = 1.0 / g->dy;
    const double ih12 = 1.0 / (12.0 * g->dx), ih60 = 1.0 / (60.0 * g->dx);
    const double hidy = 0.5 * idy, idy2 = idy * idy;
    Line row = {&CA(phi, 0, j), 1, nx};
    const double Dx = face_der(&row, i, ih12, ih60);
    FaceCtx fc = {g, phi, i};
    const double T = (tan_node(xface_val, &fc, ny, j + 1) - tan_node(xface_val, &fc, ny, j)) * idy;
    double D2[5], S2[5];
    int j0 = j - 2;
    if (j0 < 0) j0 = 0;
    if (j0 > ny - 5) j0 = ny - 5;
    for (int k = 0; k < 5; ++k) {
        D2[k] = (CA(phi, i, j0 + k) - CA(phi, i - 1, j0 + k)) * idx;
        S2[k] = 0.5 * (CA(phi, i - 1, j0 + k) + CA(phi, i, j0 + k));
    }
    /* the five-row window [j0, j0+4] holds every row the differences read;
     * positions relative to it */
    double Cx, Cy;
    {
        const int k = j - j0;
        if (j == 0) {
            Cx = ((4.0 * D2[k + 1] - 3.0 * D2[k]) - D2[k + 2]) * hidy;
            Cy = (((2.0 * S2[k] - 5.0 * S2[k + 1]) + 4.0 * S2[k + 2]) - S2[k + 3]) * idy2;
        } else if (j == ny - 1) {
            Cx = ((3.0 * D2[k] - 4.0 * D2[k - 1]) + D2[k - 2]) * hidy;
            Cy = (((2.0 * S2[k] - 5.0 * S2[k - 1]) + 4.0 * S2[k - 2]) - S2[k - 3]) * idy2;
        } else {
            Cx = (D2[k + 1] - D2[k - 1]) * hidy;
            Cy = ((S2[k + 1] - 2.0 * S2[k]) + S2[k - 1]) * idy2;
        }
    }
    return ((XA(A4, i, j) * Dx + XA(Bx4, i, j) * T) + XA(dA, i, j) * Cx) + XA(dBx, i, j) * Cy;
}

static double flux_y(const G4 *g, const double *phi, const double *C4, const double *By4, const double *dC,
                     const double *dBy, int i, int j)
{
    const int nx = g->nx, ny = g->ny;
    if (j == 0 || j == ny) return 0.0;
    const double idx = 1.0 / g->dx, idy = 1.0 / g->dy;
    const double ih12 = 1.0 / (12.0 * g->dy), ih60 = 1.0 / (60.0 * g->dy);
    const double hidx = 0.5 * idx, idx2 = idx * idx;
    Line col = {&CA(phi, i, 0), nx, ny};
    const double Dy = face_der(&col, j, ih12, ih60);
    FaceCtx fc = {g, phi, j};
    const double T = (tan_node(yface_val, &fc, nx, i + 1) - tan_node(yface_val, &fc, nx, i)) * idx;
    double E2[5], R2[5];
    int i0 = i - 2;
    if (i0 < 0) i0 = 0;
    if (i0 > nx - 5) i0 = nx - 5;
    for (int k = 0; k < 5; ++k) {
        E2[k] = (CA(phi, i0 + k, j) - CA(phi, i0 + k, j - 1)) * idy;
        R2[k] = 0.5 * (CA(phi, i0 + k, j - 1) + CA(phi, i0 + k, j));
    }
    double Ey, Ry;
    {
        const int k = i - i0;
        if (i == 0) {
            Ey = ((4.0 * E2[k + 1] - 3.0 * E2[k]) - E2[k + 2]) * hidx;
            Ry = (((2.0 * R2[k] - 5.0 * R2[k + 1]) + 4.0 * R2[k + 2]) - R2[k + 3]) * idx2;
        } else if (i == nx - 1) {
            Ey = ((3.0 * E2[k] - 4.0 * E2[k - 1]) + E2[k - 2]) * hidx;
            Ry = (((2.0 * R2[k] - 5.0 * R2[k - 1]) + 4.0 * R2[k - 2]) - R2[k - 3]) * idx2;
        } else {
            Ey = (E2[k + 1] - E2[k - 1]) * hidx;
            Ry = ((R2[k + 1] - 2.0 * R2[k]) + R2[k - 1]) * idx2;
        }
    }
    return ((YA(C4, i, j) * Dy + YA(By4, i, j) * T) + YA(dC, i, j) * Ey) + YA(dBy, i, j) * Ry;
}

void mgo_lap4(int nx, int ny, double dx, double dy, const double *A4, const double *Bx4, const double *dA,
              const double *dBx, const double *C4, const double *By4, const double *dC, const double *dBy,
              const double *phi, double *lap)
{
    const G4 G = {nx, ny, dx, dy};
    const G4 *g = &G;
    const double idx = 1.0 / dx, idy = 1.0 / dy;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const double fe = flux_x(g, phi, A4, Bx4, dA, dBx, i + 1, j);
            const double fw = flux_x(g, phi, A4, Bx4, dA, dBx, i, j);
            const double fn = flux_y(g, phi, C4, By4, dC, dBy, i, j + 1);
            const double fs = flux_y(g, phi, C4, By4, dC, dBy, i, j);
            CA(lap, i, j) = (fe - fw) * idx + (fn - fs) * idy;
        }
}

/* J g^ij of the warped mapping x = s + amp sin(2 pi s) sin(2 pi t),
 * y = (exp(beta t) - 1)/(exp(beta) - 1), s = xi/L, t = eta/H (the analytic
 * metrics of the bench geometry, with the chain rule for L and H). */
static void warped_agc(double amp, double beta, double L, double H, double em1, double xi, double eta, double *A,
                       double *B, double *C)
{
    const double two_pi = 6.283185307179586;
    const double s = xi / L, t = eta / H;
    const double x_xi = 1.0 + amp * two_pi * cos(two_pi * s) * sin(two_pi * t);
    const double x_eta = amp * two_pi * sin(two_pi * s) * cos(two_pi * t);
    const double y_eta = beta == 0.0 ? 1.0 : beta * exp(beta * t) / em1;
    const double J = x_xi * y_eta;
    *A = (x_eta * x_eta + y_eta * y_eta) / J * (L / H);
    *B = -(x_xi * x_eta) / J;
    *C = (x_xi * x_xi) / J * (H / L);
}

/* 3-point Gauss-Legendre on [0, 1]: nodes 1/2 -+ sqrt(15)/10, 1/2; weights 5/18, 8/18, 5/18 */
#define GL_T0 0.1127016653792583
#define GL_T2 0.8872983346207417

void mgo_metric4(int nx, int ny, double dx, double dy, double amp, double beta, double L, double H, double *A4,
                 double *Bx4, double *dA, double *dBx, double *C4, double *By4, double *dC, double *dBy)
{
    const G4 G = {nx, ny, dx, dy};
    const G4 *g = &G;
    if (L <= 0.0) L = nx * dx;
    if (H <= 0.0) H = ny * dy;
    const double em1 = expm1(beta);
    const double tg[3] = {GL_T0, 0.5, GL_T2};
    const double w12x = dy * (1.0 / 12.0), w12y = dx * (1.0 / 12.0);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            const double xi = i * dx;
            double a[3], b[3], c, a0, b0, a1, b1;
            for (int k = 0; k < 3; ++k) warped_agc(amp, beta, L, H, em1, xi, (j + tg[k]) * dy, &a[k], &b[k], &c);
            warped_agc(amp, beta, L, H, em1, xi, j * dy, &a0, &b0, &c);
            warped_agc(amp, beta, L, H, em1, xi, (j + 1) * dy, &a1, &b1, &c);
            XA(A4, i, j) = ((5.0 * a[0] + 8.0 * a[1]) + 5.0 * a[2]) * (1.0 / 18.0);
            XA(Bx4, i, j) = ((5.0 * b[0] + 8.0 * b[1]) + 5.0 * b[2]) * (1.0 / 18.0);
            XA(dA, i, j) = (a1 - a0) * w12x;
            XA(dBx, i, j) = (b1 - b0) * w12x;
        }
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const double eta = j * dy;
            double a, b[3], c[3], b0, c0, b1, c1;
            for (int k = 0; k < 3; ++k) warped_agc(amp, beta, L, H, em1, (i + tg[k]) * dx, eta, &a, &b[k], &c[k]);
            warped_agc(amp, beta, L, H, em1, i * dx, eta, &a, &b0, &c0);
            warped_agc(amp, beta, L, H, em1, (i + 1) * dx, eta, &a, &b1, &c1);
            YA(C4, i, j) = ((5.0 * c[0] + 8.0 * c[1]) + 5.0 * c[2]) * (1.0 / 18.0);
            YA(By4, i, j) = ((5.0 * b[0] + 8.0 * b[1]) + 5.0 * b[2]) * (1.0 / 18.0);
            YA(dC, i, j) = (c1 - c0) * w12y;
            YA(dBy, i, j) = (b1 - b0) * w12y;
        }
}
