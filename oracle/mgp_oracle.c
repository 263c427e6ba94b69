/*
 * mgp_oracle.c -- CPU restatement of the EdSantilli/MGPoisson2D hot path
 * (FP64, serial).  TEST INFRASTRUCTURE: the checker for the HIP library, never
 * part of it.  See mgp_oracle.h for the layout conventions.
 *
 * Each function cites the reference routine whose arithmetic it restates.
 * Expressions keep the reference's operand order so that, built with
 * -ffp-contract=off, results match the reference bit for bit; the boundary
 * extrapolations of the nine boundary stencil variants are expressed through
 * one rule (cell_lap) instead of nine copies.
 */
#include "mgp_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define BOGUS 1.2345e300 /* precision.f90:53 */

typedef struct {
    int nx, ny;
    double dx, dy;
    const double *J, *Jxx, *Jxy, *Jyx, *Jyy;
} Geo;

typedef struct {
    int type[4]; /* xlo, xhi, ylo, yhi */
    int mode[4];
    const double *d[4];
} Bc;

enum { XLO = 0, XHI = 1, YLO = 2, YHI = 3 };

static void die(const char *msg)
{
    fprintf(stderr, "mgp_oracle: %s\n", msg);
    abort();
}

/* one-ghost cell field */
#define PH(f, i, j) (f)[(size_t)((i) + 1) + (size_t)((j) + 1) * (size_t)(g->nx + 2)]
/* ghost-free cell field */
#define C0(f, i, j) (f)[(size_t)(i) + (size_t)(j) * (size_t)g->nx]
/* x-face and y-face metric fields */
#define XF(f, i, j) (f)[(size_t)(i) + (size_t)(j) * (size_t)(g->nx + 1)]
#define YF(f, i, j) (f)[(size_t)(i) + (size_t)(j) * (size_t)g->nx]

static Geo mkgeo(int nx, int ny, double dx, double dy, const double *J, const double *Jxx,
                 const double *Jxy, const double *Jyx, const double *Jyy)
{
    Geo g = {nx, ny, dx, dy, J, Jxx, Jxy, Jyx, Jyy};
    if (nx < 3 || ny < 3) die("grids need at least 3 cells per direction");
    return g;
}

static Bc mkbc(const int *bct, const int *bcm, const double *a, const double *b,
               const double *c, const double *d)
{
    Bc bc;
    for (int s = 0; s < 4; ++s) {
        bc.type[s] = bct[s];
        bc.mode[s] = bcm[s];
    }
    bc.d[0] = a;
    bc.d[1] = b;
    bc.d[2] = c;
    bc.d[3] = d;
    return bc;
}

/* ---------------------------------------------------------------------------
 * Ghost filling: arrayutils.f90 fill_ghosts (:1277-1492) and its side
 * helpers fill_ghosts_neum_dir_side (:871-1064), fill_ghosts_diri_dir_side
 * (:1069-1272, Dirichlet order 3 as fixed at :1284-1287).
 * ------------------------------------------------------------------------- */

static double bcval(const Bc *bc, int side, int k)
{
    if (bc->mode[side] == MGO_MODE_NONUNIFORM) return bc->d[side][k];
    return bc->d[side] ? bc->d[side][0] : 0.0;
}

/* Dirichlet, order 3 (arrayutils.f90:1160-1177 for x, :1237-1254 for y) */
static void fill_diri_side(const Geo *g, const Bc *bc, int homog, int side, double *phi)
{
    const double sixteenfifth = 16.0 / 5.0, fifth = 1.0 / 5.0;
    const int nx = g->nx, ny = g->ny;
    const int lah = bc->mode[side] == MGO_MODE_LAH;
    if (side == XLO || side == XHI) {
        const int e = side == XLO ? -1 : 1, i = side == XLO ? 0 : nx - 1;
        for (int j = 0; j < ny; ++j) {
            double *gh = &PH(phi, i + e, j);
            if (lah) { *gh = -PH(phi, i, j); continue; }
            const double p0 = PH(phi, i, j), p1 = PH(phi, i - e, j), p2 = PH(phi, i - 2 * e, j);
            if (homog) *gh = -3.0 * p0 + p1 - fifth * p2;
            else *gh = sixteenfifth * bcval(bc, side, j) - 3.0 * p0 + p1 - fifth * p2;
        }
    } else {
        const int e = side == YLO ? -1 : 1, j = side == YLO ? 0 : ny - 1;
        for (int i = 0; i < nx; ++i) {
            double *gh = &PH(phi, i, j + e);
            if (lah) { *gh = -PH(phi, i, j); continue; }
            const double p0 = PH(phi, i, j), p1 = PH(phi, i, j - e), p2 = PH(phi, i, j - 2 * e);
            /* the NONUNIFORM y branch indexes bcd(jlo:jhi) (arrayutils.f90:1251) */
            if (homog) *gh = -3.0 * p0 + p1 - fifth * p2;
            else *gh = sixteenfifth * bcval(bc, side, bc->mode[side] == MGO_MODE_NONUNIFORM ? i : 0)
                        - 3.0 * p0 + p1 - fifth * p2;
        }
    }
}

/* Neumann (arrayutils.f90:871-1064), including its one-sided end stencils,
 * the lo-side sign flip of the normal spacing and the hi side reading the
 * metric at the last interior face index. */
static void fill_neum_side(const Geo *g, const Bc *bc, int homog, int side, double *phi)
{
    const int nx = g->nx, ny = g->ny;
    const double half = 0.5, threehalves = 1.5;
    double dx = g->dx, dy = g->dy;
    const int mode = bc->mode[side];
    if (side == XLO || side == XHI) {
        int e, i;
        if (side == XLO) { e = -1; i = 0; dx = -dx; }
        else { e = 1; i = nx - 1; }
        if (mode == MGO_MODE_LAH) {
            for (int j = 0; j < ny; ++j) PH(phi, i + e, j) = PH(phi, i, j);
            return;
        }
        const double *Jxy = g->Jxy, *Jxx = g->Jxx;
#define NEUM_X_END_LO(jj, gv)                                                                     \
    do {                                                                                          \
        const int j_ = (jj);                                                                      \
        double dpn = -(3.0 * PH(phi, i, j_) - 4.0 * PH(phi, i, j_ + 1) + PH(phi, i, j_ + 2)) * half / dy; \
        double dpf = -(3.0 * PH(phi, i - e, j_) - 4.0 * PH(phi, i - e, j_ + 1) + PH(phi, i - e, j_ + 2)) * half / dy; \
        double cross = half * (3.0 * dpn - dpf) * XF(Jxy, i, j_);                                 \
        PH(phi, i + e, j_) = PH(phi, i, j_) - ((gv) - cross) * dx / XF(Jxx, i, j_);               \
    } while (0)
#define NEUM_X_END_HI(jj, gv)                                                                     \
    do {                                                                                          \
        const int j_ = (jj);                                                                      \
        double dpn = (3.0 * PH(phi, i, j_) - 4.0 * PH(phi, i, j_ - 1) + PH(phi, i, j_ - 2)) * half / dy; \
        double dpf = (3.0 * PH(phi, i - e, j_) - 4.0 * PH(phi, i - e, j_ - 1) + PH(phi, i - e, j_ - 2)) * half / dy; \
        double cross = half * (3.0 * dpn - dpf) * XF(Jxy, i, j_);                                 \
        PH(phi, i + e, j_) = PH(phi, i, j_) - ((gv) - cross) * dx / XF(Jxx, i, j_);               \
    } while (0)
#define NEUM_X_MID(jj, gv)                                                                        \
    do {                                                                                          \
        const int j_ = (jj);                                                                      \
        double dpn = (PH(phi, i, j_ + 1) - PH(phi, i, j_ - 1)) * half / dy;                       \
        double dpf = (PH(phi, i - e, j_ + 1) - PH(phi, i - e, j_ - 1)) * half / dy;               \
        double cross = (threehalves * dpn - half * dpf) * XF(Jxy, i, j_);                         \
        PH(phi, i + e, j_) = PH(phi, i, j_) - ((gv) - cross) * dx / XF(Jxx, i, j_);               \
    } while (0)
        if (homog) {
            /* homogeneous "(-cross)": (-0.0 - cross) equals -cross for every cross */
            NEUM_X_END_LO(0, -0.0);
            for (int j = 1; j <= ny - 2; ++j) NEUM_X_MID(j, -0.0);
            NEUM_X_END_HI(ny - 1, -0.0);
        } else {
            const int nu = mode == MGO_MODE_NONUNIFORM;
            NEUM_X_END_LO(0, bcval(bc, side, 0));
            for (int j = 0; j <= ny - 1; ++j) NEUM_X_MID(j, bcval(bc, side, nu ? j : 0));
            NEUM_X_END_HI(ny - 1, bcval(bc, side, nu ? ny - 1 : 0));
        }
#undef NEUM_X_END_LO
#undef NEUM_X_END_HI
#undef NEUM_X_MID
    } else {
        int e, j;
        if (side == YLO) { e = -1; j = 0; dy = -dy; }
        else { e = 1; j = ny - 1; }
        if (mode == MGO_MODE_LAH) {
            for (int i = 0; i < nx; ++i) PH(phi, i, j + e) = PH(phi, i, j);
            return;
        }
        const double *Jyx = g->Jyx, *Jyy = g->Jyy;
#define NEUM_Y_END_LO(ii, gv)                                                                     \
    do {                                                                                          \
        const int i_ = (ii);                                                                      \
        double dpn = -(3.0 * PH(phi, i_, j) - 4.0 * PH(phi, i_ + 1, j) + PH(phi, i_ + 2, j)) * half / dx; \
        double dpf = -(3.0 * PH(phi, i_, j - e) - 4.0 * PH(phi, i_ + 1, j - e) + PH(phi, i_ + 2, j - e)) * half / dx; \
        double cross = half * (3.0 * dpn - dpf) * YF(Jyx, i_, j);                                 \
        PH(phi, i_, j + e) = PH(phi, i_, j) - ((gv) - cross) * dy / YF(Jyy, i_, j);               \
    } while (0)
#define NEUM_Y_END_HI(ii, gv)                                                                     \
    do {                                                                                          \
        const int i_ = (ii);                                                                      \
        double dpn = (3.0 * PH(phi, i_, j) - 4.0 * PH(phi, i_ - 1, j) + PH(phi, i_ - 2, j)) * half / dx; \
        double dpf = (3.0 * PH(phi, i_, j - e) - 4.0 * PH(phi, i_ - 1, j - e) + PH(phi, i_ - 2, j - e)) * half / dx; \
        double cross = half * (3.0 * dpn - dpf) * YF(Jyx, i_, j);                                 \
        PH(phi, i_, j + e) = PH(phi, i_, j) - ((gv) - cross) * dy / YF(Jyy, i_, j);               \
    } while (0)
#define NEUM_Y_MID(ii, gv)                                                                        \
    do {                                                                                          \
        const int i_ = (ii);                                                                      \
        double dpn = (PH(phi, i_ + 1, j) - PH(phi, i_ - 1, j)) * half / dx;                       \
        double dpf = (PH(phi, i_ + 1, j - e) - PH(phi, i_ - 1, j - e)) * half / dx;               \
        double cross = (threehalves * dpn - half * dpf) * YF(Jyx, i_, j);                         \
        PH(phi, i_, j + e) = PH(phi, i_, j) - ((gv) - cross) * dy / YF(Jyy, i_, j);               \
    } while (0)
        /* The y-side loops run over ilo..ihi in every branch (:1008, :1028,
         * :1048), so the ilo end keeps the centred stencil and reads the
         * x ghosts filled just before. */
        if (homog) {
            NEUM_Y_END_LO(0, -0.0);
            for (int i = 0; i <= nx - 1; ++i) NEUM_Y_MID(i, -0.0);
            NEUM_Y_END_HI(nx - 1, -0.0);
        } else {
            const int nu = mode == MGO_MODE_NONUNIFORM;
            NEUM_Y_END_LO(0, bcval(bc, side, 0));
            for (int i = 0; i <= nx - 1; ++i) NEUM_Y_MID(i, bcval(bc, side, nu ? i : 0));
            NEUM_Y_END_HI(nx - 1, bcval(bc, side, nu ? nx - 1 : 0));
        }
#undef NEUM_Y_END_LO
#undef NEUM_Y_END_HI
#undef NEUM_Y_MID
    }
}

static void fill_ghosts(const Geo *g, const Bc *bc, int homog, int do_neum, double *phi)
{
    const int nx = g->nx, ny = g->ny;
    for (int side = 0; side < 4; ++side) {
        const int t = bc->type[side];
        const int xs = side < 2;
        const int lo = side == XLO || side == YLO;
        switch (t) {
        case MGO_BC_NEUM:
            if (do_neum) fill_neum_side(g, bc, homog, side, phi);
            break;
        case MGO_BC_DIRI:
            fill_diri_side(g, bc, homog, side, phi);
            break;
        case MGO_BC_PERIODIC:
            if (xs) {
                for (int j = 0; j < ny; ++j)
                    PH(phi, lo ? -1 : nx, j) = PH(phi, lo ? nx - 1 : 0, j);
            } else {
                for (int i = 0; i < nx; ++i)
                    PH(phi, i, lo ? -1 : ny) = PH(phi, i, lo ? ny - 1 : 0);
            }
            break;
        case MGO_BC_EXTRAP0:
        case MGO_BC_EXTRAP1:
        case MGO_BC_EXTRAP2:
            /* whole ghost line including the corner cells (:1348-1356) */
            if (xs) {
                const int gi = lo ? -1 : nx, a = lo ? 0 : nx - 1, s = lo ? 1 : -1;
                for (int j = -1; j <= ny; ++j) {
                    const double p0 = PH(phi, a, j), p1 = PH(phi, a + s, j);
                    if (t == MGO_BC_EXTRAP0) PH(phi, gi, j) = p0;
                    else if (t == MGO_BC_EXTRAP1) PH(phi, gi, j) = 2.0 * p0 - p1;
                    else PH(phi, gi, j) = 3.0 * (p0 - p1) + PH(phi, a + 2 * s, j);
                }
            } else {
                const int gj = lo ? -1 : ny, a = lo ? 0 : ny - 1, s = lo ? 1 : -1;
                for (int i = -1; i <= nx; ++i) {
                    const double p0 = PH(phi, i, a), p1 = PH(phi, i, a + s);
                    if (t == MGO_BC_EXTRAP0) PH(phi, i, gj) = p0;
                    else if (t == MGO_BC_EXTRAP1) PH(phi, i, gj) = 2.0 * p0 - p1;
                    else PH(phi, i, gj) = 3.0 * (p0 - p1) + PH(phi, i, a + 2 * s);
                }
            }
            break;
        case MGO_BC_NONE:
            break;
        default:
            /* BCTYPE_CF reads uninitialised coefficients in the reference
             * (arrayutils.f90:1296-1302); it is rejected here. */
            die("fill_ghosts: unsupported BC type");
        }
    }
    /* corners (:1459-1490) */
    const int px = bc->type[XLO] == MGO_BC_PERIODIC, py = bc->type[YLO] == MGO_BC_PERIODIC;
    if (!px && !py) {
        PH(phi, -1, -1) = BOGUS;
        PH(phi, -1, ny) = BOGUS;
        PH(phi, nx, -1) = BOGUS;
        PH(phi, nx, ny) = BOGUS;
    } else if (px && py) {
        PH(phi, -1, -1) = PH(phi, nx - 1, ny - 1);
        PH(phi, -1, ny) = PH(phi, nx - 1, 0);
        PH(phi, nx, -1) = PH(phi, 0, ny - 1);
        PH(phi, nx, ny) = PH(phi, 0, 0);
    } else if (px) {
        PH(phi, -1, -1) = PH(phi, nx - 1, -1);
        PH(phi, -1, ny) = PH(phi, nx - 1, ny);
        PH(phi, nx, -1) = PH(phi, 0, -1);
        PH(phi, nx, ny) = PH(phi, 0, ny);
    } else {
        PH(phi, -1, -1) = PH(phi, -1, ny - 1);
        PH(phi, nx, -1) = PH(phi, nx, ny - 1);
        PH(phi, -1, ny) = PH(phi, -1, 0);
        PH(phi, nx, ny) = PH(phi, nx, 0);
    }
}

/* arrayutils.f90:1497-1566 */
static void extrapolate_ghosts(const Geo *g, int order, double *phi)
{
    const int nx = g->nx, ny = g->ny;
    for (int pass = 0; pass < 2; ++pass) {
        for (int side = 0; side < 2; ++side) {
            const int lo = side == 0;
            const int n = pass == 0 ? ny : nx;
            for (int k = -1; k <= n; ++k) {
#define AT(m) (pass == 0 ? &PH(phi, lo ? (m) : nx - 1 - (m), k) : &PH(phi, k, lo ? (m) : ny - 1 - (m)))
                double *gh = pass == 0 ? &PH(phi, lo ? -1 : nx, k) : &PH(phi, k, lo ? -1 : ny);
                const double p0 = *AT(0);
                switch (order) {
                case 0: *gh = p0; break;
                case 1: *gh = 2.0 * p0 - *AT(1); break;
                case 2: *gh = 3.0 * (p0 - *AT(1)) + *AT(2); break;
                case 3: *gh = 4.0 * (p0 + *AT(2)) - 6.0 * *AT(1) - *AT(3); break;
                default: die("extrapolate_ghosts: bad order");
                }
#undef AT
            }
        }
    }
}

/* ---------------------------------------------------------------------------
 * The full 9-point mapped operator at one cell: poisson2d.f90
 * compute_laplacian (:886-1319); the identical stencil is inlined in relax_gs
 * (:1561-1959) and, applied to a unit impulse with zero ghosts, in
 * compute_invdiags (:16-388).
 *
 * Cross-term neighbours that fall outside a non-periodic side are replaced by
 * quadratic extrapolation through the two interior neighbours on the same
 * line and the (possibly ghost) cell beyond them: "3(a - b) + c".  A corner
 * neighbour outside two sides is the mean of its x- and y-extrapolations.
 * Normal-flux neighbours are the stored values (ghosts included); a Neumann
 * side replaces its whole face flux with the BC data.
 * ------------------------------------------------------------------------- */
typedef struct {
    int ring_lo[2], ring_hi[2]; /* [0]=x, [1]=y: non-periodic side has a ring */
    double idx, idy;
} Ring;

static Ring mkring(const Geo *g, const Bc *bc)
{
    Ring r;
    r.ring_lo[0] = bc->type[XLO] != MGO_BC_PERIODIC;
    r.ring_hi[0] = bc->type[XHI] != MGO_BC_PERIODIC;
    r.ring_lo[1] = bc->type[YLO] != MGO_BC_PERIODIC;
    r.ring_hi[1] = bc->type[YHI] != MGO_BC_PERIODIC;
    r.idx = 1.0 / g->dx;
    r.idy = 1.0 / g->dy;
    return r;
}

static double flux_data(const Bc *bc, int side, int homog, int k)
{
    if (homog) return 0.0;
    if (bc->mode[side] == MGO_MODE_UNIFORM) return bc->d[side][0];
    return bc->d[side][k];
}

/* phi: one-ghost field (or the impulse patch accessor for invdiags). */
static double cell_lap_at(const Geo *g, const Bc *bc, const Ring *rg, int homog, int i, int j,
                          double (*get)(const void *, int, int), const void *ctx)
{
    const int nx = g->nx, ny = g->ny;
    int ox = 0, oy = 0;
    if (i == 0 && rg->ring_lo[0]) ox = -1;
    else if (i == nx - 1 && rg->ring_hi[0]) ox = 1;
    if (j == 0 && rg->ring_lo[1]) oy = -1;
    else if (j == ny - 1 && rg->ring_hi[1]) oy = 1;

    double e[3][3]; /* e[dj+1][di+1] */
    for (int dj = -1; dj <= 1; ++dj)
        for (int di = -1; di <= 1; ++di)
            e[dj + 1][di + 1] = (di == ox && ox) || (dj == oy && oy) ? 0.0 : get(ctx, i + di, j + dj);
    if (ox)
        for (int dj = -1; dj <= 1; ++dj) {
            if (oy && dj == oy) continue;
            e[dj + 1][ox + 1] = 3.0 * (e[dj + 1][1] - e[dj + 1][1 - ox]) + get(ctx, i - 2 * ox, j + dj);
        }
    if (oy)
        for (int di = -1; di <= 1; ++di) {
            if (ox && di == ox) continue;
            e[oy + 1][di + 1] = 3.0 * (e[1][di + 1] - e[1 - oy][di + 1]) + get(ctx, i + di, j - 2 * oy);
        }
    if (ox && oy) {
        const double X = 3.0 * (e[oy + 1][1] - e[oy + 1][1 - ox]) + get(ctx, i - 2 * ox, j + oy);
        const double Y = 3.0 * (e[1][ox + 1] - e[1 - oy][ox + 1]) + get(ctx, i + ox, j - 2 * oy);
        e[oy + 1][ox + 1] = 0.5 * (X + Y);
    }
    const double p = get(ctx, i, j), pe = get(ctx, i + 1, j), pw = get(ctx, i - 1, j);
    const double pn = get(ctx, i, j + 1), ps = get(ctx, i, j - 1);
    const double ee = e[1][2], ew = e[1][0], en = e[2][1], es = e[0][1];
    const double ene = e[2][2], enw = e[2][0], ese = e[0][2], esw = e[0][0];
    const double idx = rg->idx, idy = rg->idy, fourth = 0.25;

    double xfe = XF(g->Jxx, i + 1, j) * (pe - p) * idx + XF(g->Jxy, i + 1, j) * fourth * (ene - ese + en - es) * idy;
    double xfw = XF(g->Jxx, i, j) * (p - pw) * idx + XF(g->Jxy, i, j) * fourth * (en - es + enw - esw) * idy;
    double yfn = YF(g->Jyy, i, j + 1) * (pn - p) * idy + YF(g->Jyx, i, j + 1) * fourth * (ene - enw + ee - ew) * idx;
    double yfs = YF(g->Jyy, i, j) * (p - ps) * idy + YF(g->Jyx, i, j) * fourth * (ee - ew + ese - esw) * idx;

    if (ox == -1 && bc->type[XLO] == MGO_BC_NEUM) xfw = flux_data(bc, XLO, homog, j);
    if (ox == 1 && bc->type[XHI] == MGO_BC_NEUM) xfe = flux_data(bc, XHI, homog, j);
    if (oy == -1 && bc->type[YLO] == MGO_BC_NEUM) yfs = flux_data(bc, YLO, homog, i);
    /* the reference indexes the yhi data with j (poisson2d.f90:1216 etc.) */
    if (oy == 1 && bc->type[YHI] == MGO_BC_NEUM) yfn = flux_data(bc, YHI, homog, j);

    return (xfe - xfw) * idx + (yfn - yfs) * idy;
}

typedef struct {
    const Geo *g;
    const double *phi;
} PhiCtx;

static double get_phi(const void *c, int i, int j)
{
    const PhiCtx *pc = (const PhiCtx *)c;
    const Geo *g = pc->g;
    return PH(pc->phi, i, j);
}

typedef struct {
    int ci, cj;
} ImpulseCtx;

static double get_impulse(const void *c, int i, int j)
{
    const ImpulseCtx *ic = (const ImpulseCtx *)c;
    return (i == ic->ci && j == ic->cj) ? 1.0 : 0.0;
}

static void check_lah(const Bc *bc)
{
    for (int s = 0; s < 4; ++s)
        if (bc->mode[s] == MGO_MODE_LAH) die("Poisson2D operators reject BCMODE_LAH");
}

static void laplacian(const Geo *g, const Bc *bc, int homog, int jscale, double *phi, double *lap)
{
    check_lah(bc);
    fill_ghosts(g, bc, homog, 0, phi); /* neum_ghosts = .false. (poisson2d.f90:894,935) */
    Ring rg = mkring(g, bc);
    PhiCtx pc = {g, phi};
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i) {
            double v = cell_lap_at(g, bc, &rg, homog, i, j, get_phi, &pc);
            if (jscale) v = v / C0(g->J, i, j);
            C0(lap, i, j) = v;
        }
}

static void residual(const Geo *g, const Bc *bc, int homog, const double *rhs, double *phi,
                     double *res)
{
    laplacian(g, bc, homog, 0, phi, res);
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i) C0(res, i, j) = C0(rhs, i, j) - C0(res, i, j);
}

static void invdiags(const Geo *g, const Bc *bc, int homog, double *invd)
{
    check_lah(bc);
    Ring rg = mkring(g, bc);
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i) {
            ImpulseCtx ic = {i, j};
            C0(invd, i, j) = 1.0 / cell_lap_at(g, bc, &rg, homog, i, j, get_impulse, &ic);
        }
}

static void gs_point(const Geo *g, const Bc *bc, const Ring *rg, int homog, const double *invd,
                     double omega, const double *rhs, double *phi, int i, int j)
{
    PhiCtx pc = {g, phi};
    const double lphi = cell_lap_at(g, bc, rg, homog, i, j, get_phi, &pc);
    const double newphi = PH(phi, i, j) + (C0(rhs, i, j) - lphi) * C0(invd, i, j);
    PH(phi, i, j) = omega * newphi + (1.0 - omega) * PH(phi, i, j);
}

static int colour4(int i, int j) { return (i & 1) | ((j & 1) << 1); }

/* poisson2d.f90 relax_gs (:1460-1988) with tol < 0 (no residual monitoring). */
static void relax_gs(const Geo *g, const Bc *bc, int homog, const double *invd, double omega,
                     int maxiters, const double *rhs, double *phi, int order)
{
    if (maxiters == 0) return;
    check_lah(bc);
    const int nx = g->nx, ny = g->ny;
    Ring rg = mkring(g, bc);
    for (int it = 0; it < maxiters; ++it) {
        if (order == MGO_ORDER_LEX) {
            fill_ghosts(g, bc, homog, 0, phi);
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) gs_point(g, bc, &rg, homog, invd, omega, rhs, phi, i, j);
        } else if (order == MGO_ORDER_MC) {
            for (int s = 0; s < 4; ++s)
                if (bc->type[s] == MGO_BC_PERIODIC) die("MC ordering needs non-periodic sides");
            for (int c = 0; c < 4; ++c)
                for (int j = 1; j <= ny - 2; ++j)
                    for (int i = 1; i <= nx - 2; ++i)
                        if (colour4(i, j) == c) gs_point(g, bc, &rg, homog, invd, omega, rhs, phi, i, j);
            fill_ghosts(g, bc, homog, 0, phi);
            for (int c = 0; c < 4; ++c)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i) {
                        const int ring = i == 0 || i == nx - 1 || j == 0 || j == ny - 1;
                        const int corner = (i == 0 || i == nx - 1) && (j == 0 || j == ny - 1);
                        if (ring && !corner && colour4(i, j) == c)
                            gs_point(g, bc, &rg, homog, invd, omega, rhs, phi, i, j);
                    }
            gs_point(g, bc, &rg, homog, invd, omega, rhs, phi, 0, 0);
            gs_point(g, bc, &rg, homog, invd, omega, rhs, phi, nx - 1, 0);
            gs_point(g, bc, &rg, homog, invd, omega, rhs, phi, 0, ny - 1);
            gs_point(g, bc, &rg, homog, invd, omega, rhs, phi, nx - 1, ny - 1);
        } else {
            die("relax_gs: ordering not available for the 9-point operator");
        }
    }
}

/* ---------------------------------------------------------------------------
 * DeferredPoisson2D (MGPoisson2D_DD.f90:3-568): low-order inner operator.
 * ------------------------------------------------------------------------- */

/* MGPoisson2D_DD.f90:16-39 */
static void dc_invdiags(const Geo *g, double *invd)
{
    const double idx2 = 1.0 / (g->dx * g->dx), idy2 = 1.0 / (g->dy * g->dy);
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i) {
            const double d = (XF(g->Jxx, i + 1, j) + XF(g->Jxx, i, j)) * idx2
                           + (YF(g->Jyy, i, j + 1) + YF(g->Jyy, i, j)) * idy2;
            C0(invd, i, j) = -1.0 / d;
        }
}

/* The 5-point part of MGPoisson2D_DD.f90:99-117 and :288-306, returned as
 * its x and y halves because the operator accumulates them one at a time. */
static void dc_five_point(const Geo *g, const double *phi, int i, int j, double idx, double idy,
                          double *ax, double *ay)
{
    const double p = PH(phi, i, j), pe = PH(phi, i + 1, j), pw = PH(phi, i - 1, j);
    const double pn = PH(phi, i, j + 1), ps = PH(phi, i, j - 1);
    const double xfe = XF(g->Jxx, i + 1, j) * (pe - p) * idx;
    const double xfw = XF(g->Jxx, i, j) * (p - pw) * idx;
    const double yfn = YF(g->Jyy, i, j + 1) * (pn - p) * idy;
    const double yfs = YF(g->Jyy, i, j) * (p - ps) * idy;
    *ax = (xfe - xfw) * idx;
    *ay = (yfn - yfs) * idy;
}

/* MGPoisson2D_DD.f90:45-119 */
static void dc_laplacian(const Geo *g, const Bc *bc, int homog, double *phi, double *lap)
{
    const double idx = 1.0 / g->dx, idy = 1.0 / g->dy, xyscale = 0.25 * idx * idy;
    extrapolate_ghosts(g, 0, phi);
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i) {
            const double pe = PH(phi, i + 1, j), pw = PH(phi, i - 1, j);
            const double pn = PH(phi, i, j + 1), ps = PH(phi, i, j - 1);
            const double pne = PH(phi, i + 1, j + 1), pnw = PH(phi, i - 1, j + 1);
            const double psw = PH(phi, i - 1, j - 1), pse = PH(phi, i + 1, j - 1);
            const double gxye = XF(g->Jxy, i + 1, j) * (pne - pse + pn - ps);
            const double gxyw = XF(g->Jxy, i, j) * (pn - ps + pnw - psw);
            const double gyxn = YF(g->Jyx, i, j + 1) * (pne - pnw + pe - pw);
            const double gyxs = YF(g->Jyx, i, j) * (pe - pw + pse - psw);
            C0(lap, i, j) = (gxye - gxyw + gyxn - gyxs) * xyscale;
        }
    fill_ghosts(g, bc, homog, 1, phi);
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i) {
            double ax, ay;
            dc_five_point(g, phi, i, j, idx, idy, &ax, &ay);
            C0(lap, i, j) = C0(lap, i, j) + ax + ay;
        }
}

static void dc_residual(const Geo *g, const Bc *bc, int homog, const double *rhs, double *phi,
                        double *res)
{
    dc_laplacian(g, bc, homog, phi, res);
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i) C0(res, i, j) = C0(rhs, i, j) - C0(res, i, j);
}

static void dc_point(const Geo *g, const double *invd, double omega, const double *modrhs,
                     double *phi, int i, int j, double idx, double idy)
{
    double ax, ay;
    dc_five_point(g, phi, i, j, idx, idy, &ax, &ay);
    const double lphi = ax + ay;
    const double newphi = PH(phi, i, j) + (C0(modrhs, i, j) - lphi) * C0(invd, i, j);
    PH(phi, i, j) = omega * newphi + (1.0 - omega) * PH(phi, i, j);
}

/* MGPoisson2D_DD.f90 relax_gs (:153-334), tol < 0.  The lagged cross terms
 * (scaled by half*idx*idy) are folded into modrhs once per call from the phi
 * given on entry; each sweep refills all ghosts (neum_ghosts = .true.) before
 * the lexicographic or red-black point updates. */
static void dc_relax_gs(const Geo *g, const Bc *bc, int homog, const double *invd, double omega,
                        int maxiters, const double *rhs, double *phi, int order)
{
    if (maxiters == 0) return;
    const int nx = g->nx, ny = g->ny;
    const double idx = 1.0 / g->dx, idy = 1.0 / g->dy, xyscale = 0.5 * idx * idy;
    double *modrhs = (double *)malloc(sizeof(double) * (size_t)nx * (size_t)ny);
    if (!modrhs) die("out of memory");
    memcpy(modrhs, rhs, sizeof(double) * (size_t)nx * (size_t)ny);
    extrapolate_ghosts(g, 0, phi);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const double pne = PH(phi, i + 1, j + 1), pnw = PH(phi, i - 1, j + 1);
            const double psw = PH(phi, i - 1, j - 1), pse = PH(phi, i + 1, j - 1);
            const double gxye = XF(g->Jxy, i + 1, j) * (pne - pse);
            const double gxyw = XF(g->Jxy, i, j) * (pnw - psw);
            const double gyxn = YF(g->Jyx, i, j + 1) * (pne - pnw);
            const double gyxs = YF(g->Jyx, i, j) * (pse - psw);
            C0(modrhs, i, j) = C0(modrhs, i, j) - (gxye - gxyw + gyxn - gyxs) * xyscale;
        }
    for (int it = 0; it < maxiters; ++it) {
        fill_ghosts(g, bc, homog, 1, phi);
        if (order == MGO_ORDER_RB) {
            /* colour 0 first; imin = ilo + mod(color + j, 2) with ilo = 0 */
            for (int colour = 0; colour < 2; ++colour)
                for (int j = 0; j < ny; ++j)
                    for (int i = (colour + j) % 2; i < nx; i += 2)
                        dc_point(g, invd, omega, modrhs, phi, i, j, idx, idy);
        } else if (order == MGO_ORDER_LEX) {
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) dc_point(g, invd, omega, modrhs, phi, i, j, idx, idy);
        } else {
            die("dc_relax_gs: ordering must be LEX or RB");
        }
    }
    free(modrhs);
}

/* ---------------------------------------------------------------------------
 * Transfers, schedule, norms (MGPoisson2D.f90:38-394, arrayutils.f90:637-796)
 * ------------------------------------------------------------------------- */

/* MGPoisson2D.f90 restrict (:135-257), full_weighting = .false. */
static void restrict_field(int kind, int fnx, int fny, const double *fine, int cnx, int cny,
                           double *crse)
{
    if (kind == 0) {
        const int rx = fnx / cnx, ry = fny / cny;
        const double scale = 1.0 / (rx * ry);
        memset(crse, 0, sizeof(double) * (size_t)cnx * (size_t)cny);
        for (int fj = 0; fj < fny; ++fj)
            for (int fi = 0; fi < fnx; ++fi)
                crse[(fi / rx) + (size_t)(fj / ry) * cnx] =
                    crse[(fi / rx) + (size_t)(fj / ry) * cnx] + scale * fine[fi + (size_t)fj * fnx];
    } else if (kind == 1) { /* nodal in x: (n+1) x n */
        const int rx = fnx / cnx, ry = fny / cny;
        const double scale = 1.0 / ry;
        memset(crse, 0, sizeof(double) * (size_t)(cnx + 1) * (size_t)cny);
        for (int fj = 0; fj < fny; ++fj)
            for (int ci = 0; ci <= cnx; ++ci)
                crse[ci + (size_t)(fj / ry) * (cnx + 1)] =
                    crse[ci + (size_t)(fj / ry) * (cnx + 1)] + scale * fine[ci * rx + (size_t)fj * (fnx + 1)];
    } else { /* nodal in y: n x (n+1) */
        const int rx = fnx / cnx, ry = fny / cny;
        const double scale = 1.0 / rx;
        memset(crse, 0, sizeof(double) * (size_t)cnx * (size_t)(cny + 1));
        for (int cj = 0; cj <= cny; ++cj)
            for (int fi = 0; fi < fnx; ++fi)
                crse[(fi / rx) + (size_t)cj * cnx] =
                    crse[(fi / rx) + (size_t)cj * cnx] + scale * fine[fi + (size_t)(cj * ry) * fnx];
    }
}

/* MGPoisson2D.f90 prolong (:262-394), cell-centred 2x2 only as in the
 * reference (any other ratio stops at :385-388). */
static void prolong(const Geo *fg, const Geo *cg, const Bc *cbc, int order, double *fine,
                    double *crse)
{
    const Geo *g;
    const int cnx = cg->nx, cny = cg->ny, fnx = fg->nx, fny = fg->ny;
    const int rax = (cbc->type[XLO] == MGO_BC_PERIODIC && cbc->type[XHI] == MGO_BC_PERIODIC)
                 || (cbc->type[XLO] == MGO_BC_NEUM && cbc->type[XHI] == MGO_BC_NEUM);
    const int ray = (cbc->type[YLO] == MGO_BC_PERIODIC && cbc->type[YHI] == MGO_BC_PERIODIC)
                 || (cbc->type[YLO] == MGO_BC_NEUM && cbc->type[YHI] == MGO_BC_NEUM);
    const double dxprod = fg->dx * fg->dy;
    if (dxprod > cg->dx * cg->dy) die("prolong: crse and fine geos passed in wrong order");
    if (fnx != 2 * cnx || fny != 2 * cny) die("prolong: only 2x2 refinement (reference stops otherwise)");
#define FP(i, j) fine[(size_t)((i) + 1) + (size_t)((j) + 1) * (size_t)(fnx + 2)]
#define CP(i, j) crse[(size_t)((i) + 1) + (size_t)((j) + 1) * (size_t)(cnx + 2)]
    for (int cj = 0; cj < cny; ++cj)
        for (int ci = 0; ci < cnx; ++ci) {
            const int fi = 2 * ci, fj = 2 * cj;
            const double c = CP(ci, cj);
            FP(fi, fj) = FP(fi, fj) + c;
            FP(fi + 1, fj) = FP(fi + 1, fj) + c;
            FP(fi, fj + 1) = FP(fi, fj + 1) + c;
            FP(fi + 1, fj + 1) = FP(fi + 1, fj + 1) + c;
        }
    if (order >= 1) {
        fill_ghosts(cg, cbc, 1, 1, crse);
        for (int cj = 0; cj < cny; ++cj)
            for (int ci = 0; ci < cnx; ++ci) {
                const int fi = 2 * ci, fj = 2 * cj;
                const double mx = 0.5 * (CP(ci + 1, cj) - CP(ci - 1, cj));
                const double my = 0.5 * (CP(ci, cj + 1) - CP(ci, cj - 1));
                FP(fi, fj) = FP(fi, fj) + 0.25 * (-mx - my);
                FP(fi + 1, fj) = FP(fi + 1, fj) + 0.25 * (mx - my);
                FP(fi, fj + 1) = FP(fi, fj + 1) + 0.25 * (-mx + my);
                FP(fi + 1, fj + 1) = FP(fi + 1, fj + 1) + 0.25 * (mx + my);
            }
    }
    if (order >= 2) {
        for (int cj = 0; cj < cny; ++cj)
            for (int ci = 0; ci < cnx; ++ci) {
                const int fi = 2 * ci, fj = 2 * cj;
                const double mx = 0.25 * (CP(ci + 1, cj) - 2.0 * CP(ci, cj) + CP(ci - 1, cj));
                const double my = 0.25 * (CP(ci, cj + 1) - 2.0 * CP(ci, cj) + CP(ci, cj - 1));
                const double q = 0.125 * (mx + my);
                FP(fi, fj) = FP(fi, fj) + q;
                FP(fi + 1, fj) = FP(fi + 1, fj) + q;
                FP(fi, fj + 1) = FP(fi, fj + 1) + q;
                FP(fi + 1, fj + 1) = FP(fi + 1, fj + 1) + q;
            }
        g = cg;
        extrapolate_ghosts(g, 3, crse);
        for (int cj = 0; cj < cny; ++cj)
            for (int ci = 0; ci < cnx; ++ci) {
                const int fi = 2 * ci, fj = 2 * cj;
                const double mc = (1.0 / 16.0) * (CP(ci + 1, cj + 1) - CP(ci + 1, cj - 1)
                                                 - CP(ci - 1, cj + 1) + CP(ci - 1, cj - 1));
                FP(fi, fj) = FP(fi, fj) + 0.25 * mc;
                FP(fi + 1, fj) = FP(fi + 1, fj) - 0.25 * mc;
                FP(fi, fj + 1) = FP(fi, fj + 1) - 0.25 * mc;
                FP(fi + 1, fj + 1) = FP(fi + 1, fj + 1) + 0.25 * mc;
            }
    }
    if (rax && ray) {
        double sum = 0.0, vol = 0.0;
        for (int fj = 0; fj < fny; ++fj)
            for (int fi = 0; fi < fnx; ++fi) {
                const double dvol = dxprod * fg->J[fi + (size_t)fj * fnx];
                sum = sum + dvol * FP(fi, fj);
                vol = vol + dvol;
            }
        const double mean = sum / vol;
        const size_t n = (size_t)(fnx + 2) * (size_t)(fny + 2);
        for (size_t k = 0; k < n; ++k) fine[k] = fine[k] - mean; /* whole array, ghosts too */
    }
#undef FP
#undef CP
}

/* MGPoisson2D.f90 estimate_maxdepth (:38-62) + create_mgschedule (:67-129) */
static int schedule(int nx, int ny, double dx, double dy, int maxdepth, int *refx, int *refy)
{
    int mgnx = nx, mgny = ny, d;
    for (d = 0; d < maxdepth; ++d) {
        if (mgnx % 2 == 0) mgnx /= 2;
        else if (mgny % 2 == 0) mgny /= 2;
        else break;
    }
    maxdepth = maxdepth < d + 1 ? maxdepth : d + 1;

    for (d = 0; d < maxdepth; ++d) refx[d] = refy[d] = 0;
    mgnx = nx;
    mgny = ny;
    double mgdx = dx, mgdy = dy;
    for (d = 0; d < maxdepth; ++d) {
        if (mgnx % 2 != 0 || mgny % 2 != 0) break;
        const double max_dx = mgdx > mgdy ? mgdx : mgdy;
        refx[d] = mgdx <= max_dx / 2 ? 2 : 1;
        refy[d] = mgdy <= max_dx / 2 ? 2 : 1;
        if (refx[d] * refy[d] == 1) refx[d] = refy[d] = 2;
        mgnx /= refx[d];
        mgny /= refy[d];
        if (mgnx < 4 || mgny < 4) break;
        mgdx = mgdx * (double)refx[d];
        mgdy = mgdy * (double)refy[d];
    }
    return d;
}

/* Fortran's real**integer (abs(x)**p in pnorm): binary powering, squaring the
 * base and multiplying the set bits into the result from the lowest up. */
static double powi(double x, int n)
{
    double r = 1.0;
    while (1) {
        if (n & 1) r *= x;
        n /= 2;
        if (n == 0) break;
        x *= x;
    }
    return r;
}

/* arrayutils.f90 pnorm (:753-796) over a ghost-free field */
static double pnorm(int nx, int ny, double dx, double dy, const double *a, int p)
{
    const double vs = dx * dy;
    double pn = 0.0;
    if (p == 0) {
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double v = fabs(a[i + (size_t)j * nx]);
                pn = pn > v ? pn : v;
            }
    } else if (p == 1) {
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) pn = pn + vs * fabs(a[i + (size_t)j * nx]);
    } else if (p == 2) {
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double v = a[i + (size_t)j * nx];
                pn = pn + vs * (v * v);
            }
        pn = sqrt(pn);
    } else {
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) pn = pn + vs * powi(fabs(a[i + (size_t)j * nx]), p);
        pn = pow(pn, 1.0 / p);
    }
    return pn;
}

/* arrayutils.f90 inner_prod (:637-661): per-row dot products, scaled by the
 * cell volume and accumulated row by row. */
static double inner_prod(int nx, int ny, double dx, double dy, const double *a, const double *b)
{
    const double vol = dx * dy;
    double ip = 0.0;
    for (int j = 0; j < ny; ++j) {
        double dot = 0.0;
        for (int i = 0; i < nx; ++i) dot = dot + a[i + (size_t)j * nx] * b[i + (size_t)j * nx];
        ip = ip + vol * dot;
    }
    return ip;
}

/* ---------------------------------------------------------------------------
 * Bottom solver: poisson2d.f90 solve_bicgstab2 (:2253-2477) when dc == 0,
 * MGPoisson2D_DD.f90 solve_bicgstab2 (:339-566) when dc == 1 (max norm,
 * omega 4/5 preconditioning sweeps, low-order operator).
 * ------------------------------------------------------------------------- */
static void op_residual(int dc, const Geo *g, const Bc *bc, int homog, const double *rhs,
                        double *phi, double *res)
{
    if (dc) dc_residual(g, bc, homog, rhs, phi, res);
    else residual(g, bc, homog, rhs, phi, res);
}

static void op_laplacian(int dc, const Geo *g, const Bc *bc, int homog, double *phi, double *lap)
{
    if (dc) dc_laplacian(g, bc, homog, phi, lap);
    else laplacian(g, bc, homog, 0, phi, lap);
}

static void op_relax(int dc, const Geo *g, const Bc *bc, int homog, const double *invd,
                     double omega, int iters, const double *rhs, double *phi, int order)
{
    if (dc) dc_relax_gs(g, bc, homog, invd, omega, iters, rhs, phi, order);
    else relax_gs(g, bc, homog, invd, omega, iters, rhs, phi, order);
}

static void bicgstab2(int dc, const Geo *g, const Bc *bc, int homog, const double *invd, double tol,
                      int max_iters, int max_restarts, int zerophi, const double *rhs, double *phi,
                      int order)
{
    const double reps = 1.0e-12, hang = 1.0e-7, small = 1.0e-30;
    const int norm_type = dc ? 0 : 2;
    const double relax_omega = dc ? 4.0 / 5.0 : 1.0 - 1.0 / 3.0;
    const int relax_iters = 2;
    const int nx = g->nx, ny = g->ny;
    const size_t n1 = (size_t)(nx + 2) * (size_t)(ny + 2), n0 = (size_t)nx * (size_t)ny;
    if (max_iters == 0) return;
    if (zerophi) memset(phi, 0, sizeof(double) * n1);

    double *e = calloc(n1, sizeof(double)), *pt = calloc(n1, sizeof(double));
    double *st = calloc(n1, sizeof(double));
    double *r = calloc(n0, sizeof(double)), *rt = calloc(n0, sizeof(double));
    double *p = calloc(n0, sizeof(double)), *t = calloc(n0, sizeof(double));
    double *v = calloc(n0, sizeof(double));
    if (!e || !pt || !st || !r || !rt || !p || !t || !v) die("out of memory");

    int recount = 0, it = 0, restarts = 0, init = 1;
    op_residual(dc, g, bc, homog, rhs, phi, r);
    memcpy(rt, r, sizeof(double) * n0);
    double norm0 = pnorm(nx, ny, g->dx, g->dy, r, norm_type);
    const double initial_norm = norm0, initial_rnorm = norm0;
    double norm1 = norm0;
    double rho1 = 0, rho2 = 0, alpha0 = 0, alpha1 = 0, beta1 = 0, omega0 = 0, omega1 = 0;
    double beta0 = 0;
    (void)beta0;

    while (it < max_iters && norm0 > tol * norm1 && norm1 > 0.0) {
        ++it;
        norm1 = norm0;
        alpha1 = alpha0;
        beta1 = beta0;
        omega1 = omega0;
        rho2 = rho1;
        rho1 = inner_prod(nx, ny, g->dx, g->dy, rt, r);
        if (rho1 == 0.0) break; /* "rho = 0": apply e and return */
        if (init) {
            memcpy(p, r, sizeof(double) * n0);
            init = 0;
        } else {
            beta1 = (rho1 / rho2) * (alpha1 / omega1);
            for (size_t k = 0; k < n0; ++k) p[k] = beta1 * (p[k] - omega1 * v[k]) + r[k];
        }
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) PH(pt, i, j) = C0(p, i, j) * C0(invd, i, j);
        op_relax(dc, g, bc, 1, invd, relax_omega, relax_iters, p, pt, order);
        op_laplacian(dc, g, bc, 1, pt, v);
        const double m = inner_prod(nx, ny, g->dx, g->dy, rt, v);
        alpha0 = rho1 / m;
        if (fabs(m) > small * fabs(rho1)) {
            for (size_t k = 0; k < n0; ++k) r[k] = r[k] - alpha0 * v[k];
            norm0 = pnorm(nx, ny, g->dx, g->dy, r, norm_type);
            for (size_t k = 0; k < n1; ++k) e[k] = e[k] + alpha0 * pt[k];
        } else {
            memset(r, 0, sizeof(double) * n0);
            norm0 = 0.0;
        }
        if (norm0 > tol * initial_norm && norm0 > reps * initial_rnorm) {
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) PH(st, i, j) = C0(r, i, j) * C0(invd, i, j);
            op_relax(dc, g, bc, 1, invd, relax_omega, relax_iters, r, st, order);
            op_laplacian(dc, g, bc, 1, st, t);
            omega0 = inner_prod(nx, ny, g->dx, g->dy, t, r) / inner_prod(nx, ny, g->dx, g->dy, t, t);
            for (size_t k = 0; k < n1; ++k) e[k] = e[k] + omega0 * st[k];
            for (size_t k = 0; k < n0; ++k) r[k] = r[k] - omega0 * t[k];
            norm0 = pnorm(nx, ny, g->dx, g->dy, r, norm_type);
        }
        if (norm0 <= tol * initial_norm || norm0 <= reps * initial_rnorm) break;
        if (omega0 == 0.0 || norm0 > (1.0 - hang) * norm1) {
            if (recount == 0) {
                recount = 1;
            } else {
                recount = 0;
                for (size_t k = 0; k < n1; ++k) phi[k] = phi[k] + e[k];
                /* the reference exits here and then adds e a second time
                 * after the loop (poisson2d.f90:2434-2441, :2465); kept */
                if (restarts == max_restarts) break;
                op_residual(dc, g, bc, homog, rhs, phi, r);
                norm0 = pnorm(nx, ny, g->dx, g->dy, r, norm_type);
                rho1 = rho2 = 0.0;
                alpha0 = 0.0;
                beta0 = 0.0;
                omega0 = 0.0;
                memcpy(rt, r, sizeof(double) * n0);
                memset(e, 0, sizeof(double) * n1);
                ++restarts;
                init = 1;
            }
        }
    }
    for (size_t k = 0; k < n1; ++k) phi[k] = phi[k] + e[k];
    free(e); free(pt); free(st); free(r); free(rt); free(p); free(t); free(v);
}

/* ---------------------------------------------------------------------------
 * Single-level solvers: Poisson2D::relax_jacobi (poisson2d.f90:1358-1451) and
 * Poisson2D::solve_bicgstab (:1993-2248).  Work fields the reference defines
 * like phi carry a ghost ring; the ones it never writes there stay zero (the
 * reference build allocates zeroed memory).
 * ------------------------------------------------------------------------- */

/* inner_prod (arrayutils.f90:637-661) over the valid cells of ghosted fields */
static double ip_ph(const Geo *g, const double *a, const double *b)
{
    const double vol = g->dx * g->dy;
    double ip = 0.0;
    for (int j = 0; j < g->ny; ++j) {
        double dot = 0.0;
        for (int i = 0; i < g->nx; ++i) dot = dot + PH(a, i, j) * PH(b, i, j);
        ip = ip + vol * dot;
    }
    return ip;
}

/* pnorm(.., 0) (arrayutils.f90:763-768) over the valid cells of a ghosted field */
static double maxnorm_ph(const Geo *g, const double *a)
{
    double pn = 0.0;
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i) {
            const double v = fabs(PH(a, i, j));
            pn = pn > v ? pn : v;
        }
    return pn;
}

static void valid_to_c0(const Geo *g, const double *ph, double *c0)
{
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i) C0(c0, i, j) = PH(ph, i, j);
}

static void c0_to_valid(const Geo *g, const double *c0, double *ph)
{
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i) PH(ph, i, j) = C0(c0, i, j);
}

/* relax_jacobi: returns the sweeps made; relres[0..sweeps] is written when
 * tol > 0 (2-norm of the residual relative to the initial one). */
static int relax_jacobi(const Geo *g, const Bc *bc, int homog, const double *invd, double omega, double tol,
                        int maxiters, int zerophi, const double *rhs, double *phi, double *relres)
{
    const int nx = g->nx, ny = g->ny;
    const size_t n1 = (size_t)(nx + 2) * (size_t)(ny + 2), n0 = (size_t)nx * (size_t)ny;
    if (maxiters == 0) return 0;
    double *r = calloc(n0, sizeof(double));
    if (!r) die("out of memory");
    if (zerophi) memset(phi, 0, sizeof(double) * n1);
    residual(g, bc, homog, rhs, phi, r);
    double rscale = 1.0;
    if (tol > 0.0) {
        rscale = pnorm(nx, ny, g->dx, g->dy, r, 2);
        if (rscale == 0.0) {
            relres[0] = rscale;
            rscale = 1.0;
        } else {
            relres[0] = 1.0;
        }
    }
    int it;
    for (it = 1; it <= maxiters; ++it) {
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double newphi = PH(phi, i, j) + C0(r, i, j) * C0(invd, i, j);
                PH(phi, i, j) = omega * newphi + (1.0 - omega) * PH(phi, i, j);
            }
        residual(g, bc, homog, rhs, phi, r);
        if (tol > 0.0) {
            relres[it] = pnorm(nx, ny, g->dx, g->dy, r, 2) / rscale;
            if (relres[it] <= tol) break;
        }
    }
    free(r);
    return it > maxiters ? maxiters : it;
}

/* solve_bicgstab (version 1): returns the last bookkeeping index i (relres
 * holds entries 0..i; restarts take an index of their own), or -1 where the
 * reference stops the program (beta == 0). */
static int bicgstab1(const Geo *g, const Bc *bc, int homog, const double *invd, double tol, int max_iters,
                     int max_restarts, int zerophi, const double *rhs, double *phi, int order, double *relres)
{
    const double hang = 1.0e-8;
    const double relax_omega = 1.0 - 1.0 / 3.0;
    const int nx = g->nx, ny = g->ny;
    const size_t n1 = (size_t)(nx + 2) * (size_t)(ny + 2), n0 = (size_t)nx * (size_t)ny;
    if (max_iters == 0) return 0;
    const int nb = max_iters + max_restarts + 1;
    double *r = calloc(n1, sizeof(double)), *r0 = calloc(n1, sizeof(double)), *nu = calloc(n1, sizeof(double));
    double *p = calloc(n1, sizeof(double)), *t = calloc(n1, sizeof(double)), *y = calloc(n1, sizeof(double));
    double *z = calloc(n1, sizeof(double)), *c0a = calloc(n0, sizeof(double)), *c0b = calloc(n0, sizeof(double));
    double *rho = calloc((size_t)nb, sizeof(double)), *om = calloc((size_t)nb, sizeof(double));
    if (!r || !r0 || !nu || !p || !t || !y || !z || !c0a || !c0b || !rho || !om) die("out of memory");
    if (zerophi) memset(phi, 0, sizeof(double) * n1);

    int i = 0, is_restart = 0, num_restarts = 0;
    residual(g, bc, homog, rhs, phi, c0a);
    c0_to_valid(g, c0a, r);
    memcpy(r0, r, sizeof(double) * n1);
    const double rscale = maxnorm_ph(g, r0);
    relres[0] = 1.0;
    if (rscale != 0.0) {
        double alpha = 1.0, lastres;
        rho[0] = 1.0;
        om[0] = 1.0;
        for (int iter = 1; iter <= max_iters; ++iter) {
            ++i;
            rho[i] = ip_ph(g, r0, r);
            const double beta = (rho[i] / rho[i - 1]) * (alpha / om[i - 1]);
            if (beta == 0.0) {
                i = -1;
                break;
            }
            for (size_t k = 0; k < n1; ++k) p[k] = beta * p[k];
            for (size_t k = 0; k < n1; ++k) p[k] = p[k] - beta * om[i - 1] * nu[k] + r[k];
            for (int jj = 0; jj < ny; ++jj)
                for (int ii = 0; ii < nx; ++ii) PH(y, ii, jj) = PH(p, ii, jj) * C0(invd, ii, jj);
            valid_to_c0(g, p, c0a);
            relax_gs(g, bc, homog, invd, relax_omega, 2, c0a, y, order);
            laplacian(g, bc, homog, 0, y, c0b);
            c0_to_valid(g, c0b, nu);
            alpha = ip_ph(g, r0, nu);
            alpha = rho[i] / alpha;
            for (size_t k = 0; k < n1; ++k) r[k] = r[k] - alpha * nu[k];
            for (int jj = 0; jj < ny; ++jj)
                for (int ii = 0; ii < nx; ++ii) PH(z, ii, jj) = PH(r, ii, jj) * C0(invd, ii, jj);
            valid_to_c0(g, r, c0a);
            relax_gs(g, bc, homog, invd, relax_omega, 2, c0a, z, order);
            laplacian(g, bc, homog, 0, z, c0b);
            c0_to_valid(g, c0b, t);
            om[i] = ip_ph(g, t, r) / ip_ph(g, t, t);
            for (size_t k = 0; k < n1; ++k) phi[k] = phi[k] + alpha * y[k] + om[i] * z[k];
            for (size_t k = 0; k < n1; ++k) r[k] = r[k] - om[i] * t[k];
            lastres = is_restart ? 1.0e200 : relres[i - 1];
            relres[i] = maxnorm_ph(g, r) / rscale;
            if (relres[i] <= tol) break;
            const int hanging = fabs(relres[i] - lastres) < tol * hang;
            const int diverging = !hanging && relres[i] > lastres;
            if (hanging || diverging) {
                if (num_restarts < max_restarts) {
                    /* restart: a fresh residual takes the next bookkeeping slot */
                    ++num_restarts;
                    ++i;
                    residual(g, bc, homog, rhs, phi, c0a);
                    c0_to_valid(g, c0a, r);
                    memcpy(r0, r, sizeof(double) * n1);
                    relres[i] = maxnorm_ph(g, r) / rscale;
                    alpha = 1.0;
                    rho[i] = 1.0;
                    om[i] = 1.0;
                    memset(nu, 0, sizeof(double) * n1);
                    memset(p, 0, sizeof(double) * n1);
                    is_restart = 1;
                    continue;
                }
                if (diverging) {
                    /* undo the last correction */
                    for (size_t k = 0; k < n1; ++k) r[k] = r[k] + om[i] * t[k];
                    for (size_t k = 0; k < n1; ++k) phi[k] = phi[k] - alpha * y[k] - om[i] * z[k];
                }
                break;
            }
            is_restart = 0;
        }
    }
    free(r); free(r0); free(nu); free(p); free(t); free(y); free(z); free(c0a); free(c0b); free(rho); free(om);
    return i;
}

/* ---------------------------------------------------------------------------
 * Cycle drivers: MGPoisson2D::vcycle / vcycle_noinit (MGPoisson2D.f90:400-854)
 * and DeferredMGPoisson2D::vcycle / vcycle_noinit (MGPoisson2D_DD.f90:589-1045).
 * Level workspace is zero-initialised, so ghosts that the reference leaves
 * unset read as zero here.
 * ------------------------------------------------------------------------- */
typedef struct {
    Geo g;
    double *J, *Jxx, *Jxy, *Jyx, *Jyy;
    double *e, *r, *w, *invd;
    Bc bc;
} Lvl;

static const double ZERO_DATA[1] = {0.0};

typedef struct {
    int dc, maxdepth, numcycles, sd, su, sb, order;
    const int *refx, *refy;
    Lvl *L;
} Cyc;

/* e = r * invdiag over the valid cells; ghosts keep their values */
static void einit(const Geo *g, double *e, const double *r, const double *invd)
{
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i) PH(e, i, j) = C0(r, i, j) * C0(invd, i, j);
}

static void vcycle_noinit(const Cyc *c, int d)
{
    Lvl *l = &c->L[d];
    const Geo *g = &l->g;
    const double relax_omega = c->dc ? 4.0 / 5.0 : 1.0;
    const int rorder = c->order;
    if (c->dc) einit(g, l->e, l->r, l->invd); /* every visit (MGPoisson2D_DD.f90:931,971) */
    if (d == c->maxdepth) {
        op_relax(c->dc, g, &l->bc, 1, l->invd, relax_omega, c->sb, l->r, l->e, rorder);
        bicgstab2(c->dc, g, &l->bc, 1, l->invd, 1.0e-6, 80, 5, 0, l->r, l->e, rorder);
        return;
    }
    Lvl *lc = &c->L[d + 1];
    op_relax(c->dc, g, &l->bc, 1, l->invd, relax_omega, c->sd, l->r, l->e, rorder);
    op_residual(c->dc, g, &l->bc, 1, l->r, l->e, l->w);
    restrict_field(0, g->nx, g->ny, l->w, lc->g.nx, lc->g.ny, lc->r);
    if (!c->dc) einit(&lc->g, lc->e, lc->r, lc->invd); /* MGPoisson2D.f90:805-806 */
    for (int k = 0; k < c->numcycles; ++k) vcycle_noinit(c, d + 1);
    prolong(g, &lc->g, &lc->bc, 1, l->e, lc->e);
    op_relax(c->dc, g, &l->bc, 1, l->invd, relax_omega, c->su, l->r, l->e, rorder);
}

/* MGPoisson2D::fmg_noinit (MGPoisson2D.f90:1127-1189): e(d) = 0; restrict r(d)
 * to r(d+1), solve there recursively, prolong with order 2; then numcycles
 * V-cycles at depth d from that initial guess. */
static void fmg_noinit(const Cyc *c, int d)
{
    Lvl *l = &c->L[d];
    const Geo *g = &l->g;
    memset(l->e, 0, (size_t)(g->nx + 2) * (g->ny + 2) * sizeof(double));
    if (d < c->maxdepth) {
        Lvl *lc = &c->L[d + 1];
        restrict_field(0, g->nx, g->ny, l->r, lc->g.nx, lc->g.ny, lc->r);
        fmg_noinit(c, d + 1);
        prolong(g, &lc->g, &lc->bc, 2, l->e, lc->e);
    }
    for (int k = 0; k < c->numcycles; ++k) vcycle_noinit(c, d);
}

/* kind 0: MGPoisson2D::vcycle (MGPoisson2D.f90:400-670); 1: DeferredMGPoisson2D::vcycle
 * (MGPoisson2D_DD.f90:589-862); 2: MGPoisson2D::fmg (MGPoisson2D.f90:860-1122), whose outer
 * loop is the plain one with fmg_noinit in place of e-init + vcycle_noinit and no rhs scaling
 * (the J scaling is commented out there, :1071-1072). */
static int vcycle(int kind, const Geo *g0, const Bc *bc, int homog, double tol, int maxiters,
                  int maxdepth_user, int numcycles, int sd, int su, int sb, int zerophi,
                  double *phi, double *rhs, int order, double *relres, int *iters)
{
    const int dc = kind == 1, fmg = kind == 2;
    const Geo *g = g0;
    int md = maxdepth_user < 0 ? 1000 : maxdepth_user;
    int *refx = calloc((size_t)md + 1, sizeof(int)), *refy = calloc((size_t)md + 1, sizeof(int));
    if (!refx || !refy) die("out of memory");
    const int maxdepth = schedule(g->nx, g->ny, g->dx, g->dy, md, refx, refy);
    Lvl *L = calloc((size_t)maxdepth + 1, sizeof(Lvl));
    if (!L) die("out of memory");
    int nx = g->nx, ny = g->ny;
    double dx = g->dx, dy = g->dy;
    for (int d = 0; d <= maxdepth; ++d) {
        Lvl *l = &L[d];
        const size_t ncc = (size_t)nx * ny, nxf = (size_t)(nx + 1) * ny, nyf = (size_t)nx * (ny + 1);
        l->J = malloc(ncc * sizeof(double));
        l->Jxx = malloc(nxf * sizeof(double));
        l->Jxy = malloc(nxf * sizeof(double));
        l->Jyx = malloc(nyf * sizeof(double));
        l->Jyy = malloc(nyf * sizeof(double));
        l->e = calloc((size_t)(nx + 2) * (ny + 2), sizeof(double));
        l->r = calloc(ncc, sizeof(double));
        l->w = calloc(ncc, sizeof(double));
        l->invd = calloc(ncc, sizeof(double));
        if (!l->J || !l->Jxx || !l->Jxy || !l->Jyx || !l->Jyy || !l->e || !l->r || !l->w || !l->invd)
            die("out of memory");
        if (d == 0) {
            memcpy(l->J, g->J, ncc * sizeof(double));
            memcpy(l->Jxx, g->Jxx, nxf * sizeof(double));
            memcpy(l->Jxy, g->Jxy, nxf * sizeof(double));
            memcpy(l->Jyx, g->Jyx, nyf * sizeof(double));
            memcpy(l->Jyy, g->Jyy, nyf * sizeof(double));
        } else {
            const Lvl *f = &L[d - 1];
            restrict_field(0, f->g.nx, f->g.ny, f->J, nx, ny, l->J);
            restrict_field(1, f->g.nx, f->g.ny, f->Jxx, nx, ny, l->Jxx);
            restrict_field(1, f->g.nx, f->g.ny, f->Jxy, nx, ny, l->Jxy);
            restrict_field(2, f->g.nx, f->g.ny, f->Jyx, nx, ny, l->Jyx);
            restrict_field(2, f->g.nx, f->g.ny, f->Jyy, nx, ny, l->Jyy);
        }
        l->g = mkgeo(nx, ny, dx, dy, l->J, l->Jxx, l->Jxy, l->Jyx, l->Jyy);
        for (int s = 0; s < 4; ++s) {
            l->bc.type[s] = bc->type[s];
            l->bc.mode[s] = dc ? MGO_MODE_LAH : MGO_MODE_UNIFORM;
            l->bc.d[s] = ZERO_DATA;
        }
        if (dc) dc_invdiags(&l->g, l->invd);
        else invdiags(&l->g, &l->bc, 1, l->invd);
        if (d < maxdepth) {
            nx /= refx[d];
            ny /= refy[d];
            dx *= refx[d];
            dy *= refy[d];
        }
    }
    /* coarsen_box scales dx by the integer ratio: dx*refx, as above */

    const size_t n1 = (size_t)(g->nx + 2) * (g->ny + 2);
    if (zerophi) memset(phi, 0, n1 * sizeof(double));
    if (dc)
        for (int j = 0; j < g->ny; ++j)
            for (int i = 0; i < g->nx; ++i) C0(rhs, i, j) = C0(rhs, i, j) * C0(g->J, i, j);

    Lvl *l0 = &L[0];
    residual(g, bc, homog, rhs, phi, l0->r);
    double rscale = pnorm(g->nx, g->ny, g->dx, g->dy, l0->r, 2);
    for (int k = 0; k <= maxiters; ++k) relres[k] = 0.0;
    if (rscale == 0.0) {
        relres[0] = rscale;
        rscale = 1.0;
    } else {
        relres[0] = 1.0;
    }
    Cyc cyc = {dc, maxdepth, numcycles, sd, su, sb, order, refx, refy, L};
    int status = 2, it;
    for (it = 1; it <= maxiters; ++it) {
        if (fmg) {
            fmg_noinit(&cyc, 0); /* MGPoisson2D.f90:1083-1088 */
        } else {
            if (!dc) einit(g, l0->e, l0->r, l0->invd); /* MGPoisson2D.f90:600 */
            vcycle_noinit(&cyc, 0);
        }
        for (size_t k = 0; k < n1; ++k) phi[k] = phi[k] + l0->e[k];
        residual(g, bc, homog, rhs, phi, l0->r);
        relres[it] = pnorm(g->nx, g->ny, g->dx, g->dy, l0->r, 2) / rscale;
        if (relres[it] <= tol) { status = 0; break; }
        if (relres[it] > relres[it - 1]) {
            for (size_t k = 0; k < n1; ++k) phi[k] = phi[k] - l0->e[k];
            status = 1;
            break;
        }
    }
    *iters = it > maxiters ? maxiters : it;
    if (dc)
        for (int j = 0; j < g->ny; ++j)
            for (int i = 0; i < g->nx; ++i) C0(rhs, i, j) = C0(rhs, i, j) / C0(g->J, i, j);
    for (int d = 0; d <= maxdepth; ++d) {
        Lvl *l = &L[d];
        free(l->J); free(l->Jxx); free(l->Jxy); free(l->Jyx); free(l->Jyy);
        free(l->e); free(l->r); free(l->w); free(l->invd);
    }
    free(L);
    free(refx);
    free(refy);
    return status;
}

/* ---------------------------------------------------------------------------
 * Flux tools of the projection workflow: Poisson2D::compute_pd
 * (poisson2d.f90:404-795), compute_grad (:809-850), compute_div (:860-879)
 * and ArrayUtils::fill_boundary_fluxes (arrayutils.f90:1571-1622).
 * ------------------------------------------------------------------------- */

/* compute_pd with pd nodal in x (an x-face field, (nx+1) x ny).
 * dir 1: d/dxi across the face, (phi(i) - phi(i-1))/dx, ghosts included (:446-449).
 * dir 2: d/deta at the face, the compute_laplacian cross-term stencil with
 * its boundary closures (:619-793, simple_bdry_stencil = .false.). */
static void pd_xface(const Geo *g, const double *phi, int dir, double *pd)
{
    const int nx = g->nx, ny = g->ny, ilo = 0, ihi = nx - 1, jlo = 0, jhi = ny - 1;
    if (dir == 1) {
        const double scale = 1.0 / g->dx;
        for (int j = jlo; j <= jhi; ++j)
            for (int i = ilo; i <= ihi + 1; ++i) XF(pd, i, j) = scale * (PH(phi, i, j) - PH(phi, i - 1, j));
        return;
    }
    const double scale = 0.25 / g->dy;
    /* away from the x boundaries: interior rows, then the two y-boundary rows */
    for (int j = jlo + 1; j <= jhi - 1; ++j)
        for (int i = ilo + 1; i <= ihi; ++i)
            XF(pd, i, j) = scale * (PH(phi, i, j + 1) - PH(phi, i, j - 1) + PH(phi, i - 1, j + 1) -
                                    PH(phi, i - 1, j - 1));
    for (int i = ilo + 1; i <= ihi; ++i) {
        XF(pd, i, jlo) = -(scale * (3.0 * PH(phi, i, jlo) - 4.0 * PH(phi, i, jlo + 1) + PH(phi, i, jlo + 2) +
                                    3.0 * PH(phi, i - 1, jlo) - 4.0 * PH(phi, i - 1, jlo + 1) +
                                    PH(phi, i - 1, jlo + 2)));
        XF(pd, i, jhi) = scale * (3.0 * PH(phi, i, jhi) - 4.0 * PH(phi, i, jhi - 1) + PH(phi, i, jhi - 2) +
                                  3.0 * PH(phi, i - 1, jhi) - 4.0 * PH(phi, i - 1, jhi - 1) +
                                  PH(phi, i - 1, jhi - 2));
    }
    /* x boundaries: the west face of cell ilo, the east face of cell ihi */
    double p, pn, pw, ee, ew, en, es, ene, enw, ese, esw;
    int i = ilo, j;
    for (j = jlo; j <= jhi; ++j) {
        p = PH(phi, i, j);
        ee = PH(phi, i + 1, j);
        ew = 3.0 * (p - ee) + PH(phi, i + 2, j);
        if (j == jlo) {
            en = PH(phi, i, j + 1);
            es = 3.0 * (p - en) + PH(phi, i, j + 2);
            ene = PH(phi, i + 1, j + 1);
            enw = 3.0 * (en - ene) + PH(phi, i + 2, j + 1);
            ese = 3.0 * (ee - ene) + PH(phi, i + 1, j + 2);
            esw = 0.5 * ((3.0 * (es - ese) + PH(phi, i + 2, j - 1)) + (3.0 * (ew - enw) + PH(phi, i - 1, j + 2)));
        } else if (j == jhi) {
            es = PH(phi, i, j - 1);
            en = 3.0 * (p - es) + PH(phi, i, j - 2);
            ese = PH(phi, i + 1, j - 1);
            ene = 3.0 * (ee - ese) + PH(phi, i + 1, j - 2);
            esw = 3.0 * (es - ese) + PH(phi, i + 2, j - 1);
            enw = 0.5 * ((3.0 * (ew - esw) + PH(phi, i - 1, j - 2)) + (3.0 * (en - ene) + PH(phi, i + 2, j + 1)));
        } else {
            en = PH(phi, i, j + 1);
            es = PH(phi, i, j - 1);
            ene = PH(phi, i + 1, j + 1);
            ese = PH(phi, i + 1, j - 1);
            enw = 3.0 * (en - ene) + PH(phi, i + 2, j + 1);
            esw = 3.0 * (es - ese) + PH(phi, i + 2, j - 1);
        }
        XF(pd, i, j) = scale * (en - es + enw - esw);
    }
    i = ihi;
    for (j = jlo; j <= jhi; ++j) {
        p = PH(phi, i, j);
        ew = PH(phi, i - 1, j);
        ee = 3.0 * (p - ew) + PH(phi, i - 2, j);
        if (j == jlo) {
            pn = PH(phi, i, j + 1);
            pw = PH(phi, i - 1, j);
            en = pn;
            es = 3.0 * (p - en) + PH(phi, i, j + 2);
            enw = PH(phi, i - 1, j + 1);
            ene = 3.0 * (pn - enw) + PH(phi, i - 2, j + 1);
            esw = 3.0 * (pw - enw) + PH(phi, i - 1, j + 2);
            ese = 0.5 * ((3.0 * (es - esw) + PH(phi, i - 2, j - 1)) + (3.0 * (ee - ene) + PH(phi, i + 1, j + 2)));
        } else if (j == jhi) {
            es = PH(phi, i, j - 1);
            en = 3.0 * (p - es) + PH(phi, i, j - 2);
            esw = PH(phi, i - 1, j - 1);
            enw = 3.0 * (ew - esw) + PH(phi, i - 1, j - 2);
            ese = 3.0 * (es - esw) + PH(phi, i - 2, j - 1);
            ene = 0.5 * ((3.0 * (en - enw) + PH(phi, i - 2, j + 1)) + (3.0 * (ee - ese) + PH(phi, i + 1, j - 2)));
        } else {
            en = PH(phi, i, j + 1);
            es = PH(phi, i, j - 1);
            enw = PH(phi, i - 1, j + 1);
            esw = PH(phi, i - 1, j - 1);
            ene = 3.0 * (en - enw) + PH(phi, i - 2, j + 1);
            ese = 3.0 * (es - esw) + PH(phi, i - 2, j - 1);
        }
        XF(pd, i + 1, j) = scale * (ene - ese + en - es);
    }
}

/* compute_pd with pd nodal in y (a y-face field, nx x (ny+1)).
 * dir 2: (phi(j) - phi(j-1))/dy (:610-613).  dir 1: d/dxi at the face
 * (:450-605), the transpose of pd_xface's dir 2. */
static void pd_yface(const Geo *g, const double *phi, int dir, double *pd)
{
    const int nx = g->nx, ny = g->ny, ilo = 0, ihi = nx - 1, jlo = 0, jhi = ny - 1;
    if (dir == 2) {
        const double scale = 1.0 / g->dy;
        for (int j = jlo; j <= jhi + 1; ++j)
            for (int i = ilo; i <= ihi; ++i) YF(pd, i, j) = scale * (PH(phi, i, j) - PH(phi, i, j - 1));
        return;
    }
    const double scale = 0.25 / g->dx;
    for (int j = jlo + 1; j <= jhi; ++j)
        for (int i = ilo + 1; i <= ihi - 1; ++i)
            YF(pd, i, j) = scale * (PH(phi, i + 1, j) - PH(phi, i - 1, j) + PH(phi, i + 1, j - 1) -
                                    PH(phi, i - 1, j - 1));
    for (int j = jlo + 1; j <= jhi; ++j) {
        YF(pd, ilo, j) = -(scale * (3.0 * PH(phi, ilo, j) - 4.0 * PH(phi, ilo + 1, j) + PH(phi, ilo + 2, j) +
                                    3.0 * PH(phi, ilo, j - 1) - 4.0 * PH(phi, ilo + 1, j - 1) +
                                    PH(phi, ilo + 2, j - 1)));
        YF(pd, ihi, j) = scale * (3.0 * PH(phi, ihi, j) - 4.0 * PH(phi, ihi - 1, j) + PH(phi, ihi - 2, j) +
                                  3.0 * PH(phi, ihi, j - 1) - 4.0 * PH(phi, ihi - 1, j - 1) +
                                  PH(phi, ihi - 2, j - 1));
    }
    /* y boundaries: the south face of row jlo, the north face of row jhi */
    double p, pn, pw, ee, ew, en, es, ene, enw, ese, esw;
    int i, j = jlo;
    for (i = ilo; i <= ihi; ++i) {
        p = PH(phi, i, j);
        en = PH(phi, i, j + 1);
        es = 3.0 * (p - en) + PH(phi, i, j + 2);
        if (i == ilo) {
            ee = PH(phi, i + 1, j);
            ew = 3.0 * (p - ee) + PH(phi, i + 2, j);
            ene = PH(phi, i + 1, j + 1);
            enw = 3.0 * (en - ene) + PH(phi, i + 2, j + 1);
            ese = 3.0 * (ee - ene) + PH(phi, i + 1, j + 2);
            esw = 0.5 * ((3.0 * (es - ese) + PH(phi, i + 2, j - 1)) + (3.0 * (ew - enw) + PH(phi, i - 1, j + 2)));
        } else if (i == ihi) {
            pn = PH(phi, i, j + 1);
            pw = PH(phi, i - 1, j);
            ew = pw;
            ee = 3.0 * (p - ew) + PH(phi, i - 2, j);
            enw = PH(phi, i - 1, j + 1);
            ene = 3.0 * (pn - enw) + PH(phi, i - 2, j + 1);
            esw = 3.0 * (pw - enw) + PH(phi, i - 1, j + 2);
            ese = 0.5 * ((3.0 * (es - esw) + PH(phi, i - 2, j - 1)) + (3.0 * (ee - ene) + PH(phi, i + 1, j + 2)));
        } else {
            ee = PH(phi, i + 1, j);
            ew = PH(phi, i - 1, j);
            ene = PH(phi, i + 1, j + 1);
            enw = PH(phi, i - 1, j + 1);
            ese = 3.0 * (ee - ene) + PH(phi, i + 1, j + 2);
            esw = 3.0 * (ew - enw) + PH(phi, i - 1, j + 2);
        }
        YF(pd, i, j) = scale * (ee - ew + ese - esw);
    }
    j = jhi;
    for (i = ilo; i <= ihi; ++i) {
        p = PH(phi, i, j);
        es = PH(phi, i, j - 1);
        en = 3.0 * (p - es) + PH(phi, i, j - 2);
        if (i == ilo) {
            ee = PH(phi, i + 1, j);
            ew = 3.0 * (p - ee) + PH(phi, i + 2, j);
            ese = PH(phi, i + 1, j - 1);
            ene = 3.0 * (ee - ese) + PH(phi, i + 1, j - 2);
            esw = 3.0 * (es - ese) + PH(phi, i + 2, j - 1);
            enw = 0.5 * ((3.0 * (ew - esw) + PH(phi, i - 1, j - 2)) + (3.0 * (en - ene) + PH(phi, i + 2, j + 1)));
        } else if (i == ihi) {
            ew = PH(phi, i - 1, j);
            ee = 3.0 * (p - ew) + PH(phi, i - 2, j);
            esw = PH(phi, i - 1, j - 1);
            enw = 3.0 * (ew - esw) + PH(phi, i - 1, j - 2);
            ese = 3.0 * (es - esw) + PH(phi, i - 2, j - 1);
            ene = 0.5 * ((3.0 * (en - enw) + PH(phi, i - 2, j + 1)) + (3.0 * (ee - ese) + PH(phi, i + 1, j - 2)));
        } else {
            ee = PH(phi, i + 1, j);
            ew = PH(phi, i - 1, j);
            ese = PH(phi, i + 1, j - 1);
            esw = PH(phi, i - 1, j - 1);
            ene = 3.0 * (ee - ese) + PH(phi, i + 1, j - 2);
            enw = 3.0 * (ew - esw) + PH(phi, i - 1, j - 2);
        }
        YF(pd, i, j + 1) = scale * (ene - enw + ee - ew);
    }
}

/* fill_boundary_fluxes (arrayutils.f90:1571-1622): Neumann faces take the BC
 * data (homogeneous: zero; LAH: zero). */
static void boundary_fluxes(const Geo *g, const Bc *bc, int homog, double *xflux, double *yflux)
{
    const int nx = g->nx, ny = g->ny;
    for (int side = 0; side < 4; ++side) {
        if (bc->type[side] != MGO_BC_NEUM) continue;
        const int n = side < 2 ? ny : nx;
        for (int k = 0; k < n; ++k) {
            double v = 0.0;
            if (!homog) {
                if (bc->mode[side] == MGO_MODE_UNIFORM) v = bc->d[side][0];
                else if (bc->mode[side] == MGO_MODE_NONUNIFORM) v = bc->d[side][k];
            }
            if (side == XLO) XF(xflux, 0, k) = v;
            else if (side == XHI) XF(xflux, nx, k) = v;
            else if (side == YLO) YF(yflux, k, 0) = v;
            else YF(yflux, k, ny) = v;
        }
    }
}

/* compute_grad (poisson2d.f90:809-850): fills phi's ghosts (do_neum off), then
 * J grad phi on the faces, then the Neumann boundary fluxes. */
static void compute_grad(const Geo *g, const Bc *bc, int homog, double *phi, double *xflux, double *yflux)
{
    const int nx = g->nx, ny = g->ny;
    double *xwk = malloc((size_t)(nx + 1) * ny * sizeof(double));
    double *ywk = malloc((size_t)nx * (ny + 1) * sizeof(double));
    if (!xwk || !ywk) die("out of memory");
    fill_ghosts(g, bc, homog, 0, phi);
    pd_xface(g, phi, 1, xflux);
    pd_xface(g, phi, 2, xwk);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i)
            XF(xflux, i, j) = XF(g->Jxx, i, j) * XF(xflux, i, j) + XF(g->Jxy, i, j) * XF(xwk, i, j);
    pd_yface(g, phi, 1, ywk);
    pd_yface(g, phi, 2, yflux);
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i)
            YF(yflux, i, j) = YF(g->Jyx, i, j) * YF(ywk, i, j) + YF(g->Jyy, i, j) * YF(yflux, i, j);
    boundary_fluxes(g, bc, homog, xflux, yflux);
    free(xwk);
    free(ywk);
}

/* compute_div (poisson2d.f90:860-879): J div F. */
static void compute_div(const Geo *g, const double *xflux, const double *yflux, double *div)
{
    const double invdx = 1.0 / g->dx, invdy = 1.0 / g->dy;
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i)
            C0(div, i, j) = invdx * (XF(xflux, i + 1, j) - XF(xflux, i, j)) +
                            invdy * (YF(yflux, i, j + 1) - YF(yflux, i, j));
}

/* ---------------------------------------------------------------------------
 * Metric assembly of the reference's test geometry (test.f90:320-716): x = xi,
 * y = el(x) + (H - el(x)) eta/H over the bottom el(x) = amp sin^2(pi x / L)
 * (amp = 1/4 there), J and J g^ij by differencing positions of the staggered
 * points (fill_dxdXi, :566-646):
 *   J        = y_eta at cells,  from y at (cell, node) points
 *   J g^11   = y_eta at x-faces, from y at (node, node)
 *   J g^12   = -y_xi at x-faces, from y at (cell, cell)
 *   J g^21   = -y_xi at y-faces, from y at (node, node)
 *   J g^22   = (1 + y_xi^2) / y_eta at y-faces, y_eta from y at (cell, cell)
 * L and H are nx dx and ny dy, as the reference's box_data L and H are.
 * ------------------------------------------------------------------------- */

/* fill_x / fill_y / fill_elevation (test.f90:370-469, sinusoidal bottom at :530) */
static double terrain_y(double dx, double dy, double Lx, double Hy, double amp, int i, double offx, int j,
                        double offy)
{
    const double pi = 4.0 * atan(1.0);
    const double xx = (i + offx) * dx;
    const double sn = sin(pi * xx / Lx);
    const double el = amp * (sn * sn);
    const double yfrac = (j + offy) * dy / Hy;
    return el + (Hy - el) * yfrac;
}

static void metric_terrain(const Geo *g, double amp, double *J, double *Jxx, double *Jxy, double *Jyx, double *Jyy)
{
    const int nx = g->nx, ny = g->ny;
    const double dx = g->dx, dy = g->dy, Lx = nx * dx, Hy = ny * dy;
    const double sx = 1.0 / dx, sy = 1.0 / dy;
#define TY(i, ox, j, oy) terrain_y(dx, dy, Lx, Hy, amp, (i), (ox), (j), (oy))
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) C0(J, i, j) = (TY(i, 0.5, j + 1, 0.0) - TY(i, 0.5, j, 0.0)) * sy;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            XF(Jxx, i, j) = (TY(i, 0.0, j + 1, 0.0) - TY(i, 0.0, j, 0.0)) * sy;
            XF(Jxy, i, j) = -((TY(i, 0.5, j, 0.5) - TY(i - 1, 0.5, j, 0.5)) * sx);
        }
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const double yxi = (TY(i + 1, 0.0, j, 0.0) - TY(i, 0.0, j, 0.0)) * sx;
            const double yeta = (TY(i, 0.5, j, 0.5) - TY(i, 0.5, j - 1, 0.5)) * sy;
            YF(Jyx, i, j) = -yxi;
            YF(Jyy, i, j) = (1.0 + yxi * yxi) / yeta;
        }
#undef TY
}

/* ---------------------------------------------------------------------------
 * Exported entry points (argument order of oracle/ref_bridge.f90).
 * ------------------------------------------------------------------------- */
#define MKG Geo G = mkgeo(nx, ny, dx, dy, J, Jxx, Jxy, Jyx, Jyy)
#define MKB Bc B = mkbc(bct, bcm, bd_xlo, bd_xhi, bd_ylo, bd_yhi)

void mgo_fill_ghosts(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, int do_neum, double *phi)
{
    MKG; MKB;
    fill_ghosts(&G, &B, homog, do_neum, phi);
}

void mgo_extrapolate_ghosts(int nx, int ny, int order, double *phi)
{
    Geo G = mkgeo(nx, ny, 1.0, 1.0, 0, 0, 0, 0, 0);
    extrapolate_ghosts(&G, order, phi);
}

void mgo_laplacian(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, int jscale, double *phi, double *lap)
{
    MKG; MKB;
    laplacian(&G, &B, homog, jscale, phi, lap);
}

void mgo_residual(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, const double *rhs, double *phi, double *res)
{
    MKG; MKB;
    residual(&G, &B, homog, rhs, phi, res);
}

void mgo_invdiags(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, double *invd)
{
    MKG; MKB;
    invdiags(&G, &B, homog, invd);
}

void mgo_relax_gs(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, const double *invd, double omega,
                  int maxiters, const double *rhs, double *phi, int order)
{
    MKG; MKB;
    relax_gs(&G, &B, homog, invd, omega, maxiters, rhs, phi, order);
}

/* One point update of relax_gs (poisson2d.f90:1770-1771) at cell (i, j) on phi
 * as it stands, ghosts included (no fill): the update every ordering of
 * relax_gs is made of.  Exported so that a test can compose an ordering from
 * it and the fill without going through relax_gs's own loops. */
void mgo_gs_point(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, const double *invd, double omega,
                  const double *rhs, double *phi, int i, int j)
{
    MKG; MKB;
    check_lah(&B);
    Ring rg = mkring(&G, &B);
    gs_point(&G, &B, &rg, homog, invd, omega, rhs, phi, i, j);
}

void mgo_dc_laplacian(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, double *phi, double *lap)
{
    MKG; MKB;
    dc_laplacian(&G, &B, homog, phi, lap);
}

void mgo_dc_invdiags(MGO_GEO_ARGS, double *invd)
{
    MKG;
    dc_invdiags(&G, invd);
}

void mgo_dc_relax_gs(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, const double *invd, double omega,
                     int maxiters, const double *rhs, double *phi, int order)
{
    MKG; MKB;
    dc_relax_gs(&G, &B, homog, invd, omega, maxiters, rhs, phi, order);
}

void mgo_restrict(int kind, int fnx, int fny, const double *fine, int cnx, int cny, double *crse)
{
    restrict_field(kind, fnx, fny, fine, cnx, cny, crse);
}

void mgo_prolong(int fnx, int fny, double fdx, double fdy, const double *fJ, const double *fJxx,
                 const double *fJxy, const double *fJyx, const double *fJyy,
                 int cnx, int cny, double cdx, double cdy, const double *cJ, const double *cJxx,
                 const double *cJxy, const double *cJyx, const double *cJyy,
                 const int *bct, const int *bcm, int order, double *fine, double *crse)
{
    Geo F = mkgeo(fnx, fny, fdx, fdy, fJ, fJxx, fJxy, fJyx, fJyy);
    Geo C = mkgeo(cnx, cny, cdx, cdy, cJ, cJxx, cJxy, cJyx, cJyy);
    Bc B = mkbc(bct, bcm, ZERO_DATA, ZERO_DATA, ZERO_DATA, ZERO_DATA);
    prolong(&F, &C, &B, order, fine, crse);
}

void mgo_schedule(int nx, int ny, double dx, double dy, int maxdepth_in, int *maxdepth_out,
                  int *refx, int *refy)
{
    const int md = maxdepth_in < 0 ? 1000 : maxdepth_in;
    int *rx = calloc((size_t)md + 1, sizeof(int)), *ry = calloc((size_t)md + 1, sizeof(int));
    if (!rx || !ry) die("out of memory");
    const int d = schedule(nx, ny, dx, dy, md, rx, ry);
    *maxdepth_out = d;
    for (int k = 0; k < d; ++k) {
        refx[k] = rx[k];
        refy[k] = ry[k];
    }
    free(rx);
    free(ry);
}

double mgo_pnorm(int nx, int ny, double dx, double dy, const double *a, int p)
{
    return pnorm(nx, ny, dx, dy, a, p);
}

double mgo_inner_prod(int nx, int ny, double dx, double dy, const double *a, const double *b)
{
    return inner_prod(nx, ny, dx, dy, a, b);
}

/* arrayutils.f90 integrate2d (:670-710): sum volScale*J*a (J given) or volScale*a */
double mgo_integrate2d(int nx, int ny, double dx, double dy, const double *a, const double *J)
{
    const double vs = dx * dy;
    double sum = 0.0;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const size_t c = (size_t)i + (size_t)j * (size_t)nx;
            sum = J ? sum + vs * J[c] * a[c] : sum + vs * a[c];
        }
    return sum;
}

/* arrayutils.f90 integrate2d_bdry (:721-748): x faces row by row, then y
 * faces column by column; xflux (nx+1) x ny, yflux nx x (ny+1), i fastest */
double mgo_integrate2d_bdry(int nx, int ny, double dx, double dy, const double *xflux, const double *yflux)
{
    double sum = 0.0;
    for (int j = 0; j < ny; ++j)
        sum = sum + dx * (xflux[nx + (size_t)j * (nx + 1)] - xflux[(size_t)j * (nx + 1)]);
    for (int i = 0; i < nx; ++i) sum = sum + dy * (yflux[i + (size_t)ny * nx] - yflux[i]);
    return sum;
}

/* arrayutils.f90 gpnorm (:802-859) of a one-ghost field (nx+2) x (ny+2) on the
 * box 0..nx-1 x 0..ny-1, where the reference's index pattern (columns jlo-1,
 * jhi+1, rows ilo-1, ihi+1) is the ghost layer when nx == ny; the caller
 * passes the columns cx0, cx1 and rows ry0, ry1 (-1 .. n) it reads. */
double mgo_gpnorm(int nx, int ny, double dx, double dy, const double *phi, int p, int cx0, int cx1, int ry0,
                  int ry1)
{
    const size_t w = (size_t)nx + 2;
#define GP(i, j) phi[(size_t)((i) + 1) + (size_t)((j) + 1) * w]
    const int cols[2] = {cx0, cx1}, rows[2] = {ry0, ry1};
    double pn = 0.0;
    if (p == 0) {
        for (int s = 0; s < 2; ++s)
            for (int j = 0; j < ny; ++j) {
                const double v = fabs(GP(cols[s], j));
                pn = pn > v ? pn : v;
            }
        for (int s = 0; s < 2; ++s)
            for (int i = 0; i < nx; ++i) {
                const double v = fabs(GP(i, rows[s]));
                pn = pn > v ? pn : v;
            }
        return pn;
    }
    for (int s = 0; s < 2; ++s)
        for (int j = 0; j < ny; ++j) pn = pn + dy * powi(GP(cols[s], j), p);
    for (int s = 0; s < 2; ++s)
        for (int i = 0; i < nx; ++i) pn = pn + dx * powi(GP(i, rows[s]), p);
#undef GP
    return pow(pn, 1.0 / p);
}

void mgo_bicgstab2(int kind, MGO_GEO_ARGS, MGO_BC_ARGS, int homog, const double *invd, double tol,
                   int maxiters, int maxrestarts, int zerophi, const double *rhs, double *phi,
                   int order)
{
    MKG; MKB;
    bicgstab2(kind, &G, &B, homog, invd, tol, maxiters, maxrestarts, zerophi, rhs, phi, order);
}

int mgo_vcycle(int kind, MGO_GEO_ARGS, MGO_BC_ARGS, int homog, double tol, int maxiters,
               int maxdepth, int numcycles, int sdown, int sup, int sbottom, int zerophi,
               double *phi, double *rhs, int order, double *relres, int *iters)
{
    MKG; MKB;
    return vcycle(kind, &G, &B, homog, tol, maxiters, maxdepth, numcycles, sdown, sup, sbottom,
                  zerophi, phi, rhs, order, relres, iters);
}

int mgo_relax_jacobi(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, const double *invd, double omega, double tol,
                     int maxiters, int zerophi, const double *rhs, double *phi, double *relres)
{
    MKG; MKB;
    check_lah(&B);
    return relax_jacobi(&G, &B, homog, invd, omega, tol, maxiters, zerophi, rhs, phi, relres);
}

int mgo_bicgstab1(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, const double *invd, double tol, int max_iters,
                  int max_restarts, int zerophi, const double *rhs, double *phi, int order, double *relres)
{
    MKG; MKB;
    check_lah(&B);
    return bicgstab1(&G, &B, homog, invd, tol, max_iters, max_restarts, zerophi, rhs, phi, order, relres);
}

void mgo_metric_terrain(int nx, int ny, double dx, double dy, double amp, double *J, double *Jxx, double *Jxy,
                        double *Jyx, double *Jyy)
{
    Geo G = mkgeo(nx, ny, dx, dy, 0, 0, 0, 0, 0);
    metric_terrain(&G, amp, J, Jxx, Jxy, Jyx, Jyy);
}

void mgo_compute_pd(int nx, int ny, double dx, double dy, const double *phi, int dir, int nodedir, double *pd)
{
    Geo G = mkgeo(nx, ny, dx, dy, 0, 0, 0, 0, 0);
    if (nodedir == 1) pd_xface(&G, phi, dir, pd);
    else pd_yface(&G, phi, dir, pd);
}

void mgo_compute_grad(MGO_GEO_ARGS, MGO_BC_ARGS, int homog, double *phi, double *xflux, double *yflux)
{
    MKG; MKB;
    compute_grad(&G, &B, homog, phi, xflux, yflux);
}

void mgo_fill_boundary_fluxes(int nx, int ny, MGO_BC_ARGS, int homog, double *xflux, double *yflux)
{
    Geo G = mkgeo(nx, ny, 1.0, 1.0, 0, 0, 0, 0, 0);
    MKB;
    boundary_fluxes(&G, &B, homog, xflux, yflux);
}

void mgo_compute_div(int nx, int ny, double dx, double dy, const double *xflux, const double *yflux,
                     double *div)
{
    Geo G = mkgeo(nx, ny, dx, dy, 0, 0, 0, 0, 0);
    compute_div(&G, xflux, yflux, div);
}
