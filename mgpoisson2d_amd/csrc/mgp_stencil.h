// mgp_stencil.h -- device-side restatement of the reference's point stencils
// and ghost fills, written once and shared by every kernel.
//
//  * cell_lap      the full 9-point mapped operator at one cell
//                  (poisson2d.f90 compute_laplacian :886-1319; the same stencil
//                  is inlined in relax_gs :1561-1959 and, on a unit impulse,
//                  in compute_invdiags :16-388)
//  * dc_*          DeferredPoisson2D's low-order operator and lagged cross
//                  terms (MGPoisson2D_DD.f90:45-119, :227-250)
//  * fill_*        arrayutils.f90 fill_ghosts (:1277-1492) and its side
//                  helpers (:871-1272); Dirichlet ghosts are order 3
//                  (:1284-1287)
//
// Operand order follows the reference expression by expression; all HIP code
// is compiled with -ffp-contract=off, so results equal the reference's bit
// for bit.
#pragma once

#include <hip/hip_runtime.h>

#include "mgp_common.h"

namespace mgp {

// Accessor for a padded field.
struct Fld {
    double *p;
    int pitch;
    __device__ __forceinline__ double &operator()(int i, int j) const { return p[ix(pitch, i, j)]; }
};
struct CFld {
    const double *p;
    int pitch;
    __device__ __forceinline__ double operator()(int i, int j) const { return p[ix(pitch, i, j)]; }
};

// Field accessor that synthesises the homogeneous Dirichlet ghosts
// (order 3, or LAH negation) from the interior values at the time of the read,
// and returns the stored value for every other ghost.  Equivalent to
// fill_ghosts(homog, do_neum = .false.) immediately before the read, which is
// what every MG-level stencil of the fast solver is preceded by.
struct SynthGet {
    const double *p;
    int pitch, nx, ny;
    int t[4], m[4];
    __device__ __forceinline__ double raw(int i, int j) const { return p[ix(pitch, i, j)]; }
    // Side indices are compile-time constants below (a runtime index into
    // t[] / m[] would place them in scratch memory).
    __device__ __forceinline__ static double dg(int mode, double p0, double p1, double p2) {
        if (mode == MODE_LAH) return -p0;
        return -3.0 * p0 + p1 - (1.0 / 5.0) * p2;
    }
    __device__ __forceinline__ double operator()(int i, int j) const {
        const bool xin = i >= 0 && i < nx, yin = j >= 0 && j < ny;
        if (xin == yin) return raw(i, j);  // valid cell, or a corner (never read)
        if (i < 0) return t[XLO] != BC_DIRI ? raw(i, j) : dg(m[XLO], raw(0, j), raw(1, j), raw(2, j));
        if (i >= nx)
            return t[XHI] != BC_DIRI ? raw(i, j) : dg(m[XHI], raw(nx - 1, j), raw(nx - 2, j), raw(nx - 3, j));
        if (j < 0) return t[YLO] != BC_DIRI ? raw(i, j) : dg(m[YLO], raw(i, 0), raw(i, 1), raw(i, 2));
        return t[YHI] != BC_DIRI ? raw(i, j) : dg(m[YHI], raw(i, ny - 1), raw(i, ny - 2), raw(i, ny - 3));
    }
};

__device__ __forceinline__ SynthGet synth(const double *p, const Grid &g, const Bc &bc) {
    SynthGet s;
    s.p = p;
    s.pitch = g.pitch;
    s.nx = g.nx;
    s.ny = g.ny;
    for (int k = 0; k < 4; ++k) {
        s.t[k] = bc.type[k];
        s.m[k] = bc.mode[k];
    }
    return s;
}

// Unit impulse at (ci, cj), zero everywhere else, ghosts included: what
// compute_invdiags evaluates the stencil on (poisson2d.f90:29-33).
struct Impulse {
    int ci, cj;
    __device__ __forceinline__ double operator()(int i, int j) const {
        return (i == ci && j == cj) ? 1.0 : 0.0;
    }
};

__device__ __forceinline__ double flux_data(const Bc &bc, int side, int homog, int k) {
    if (homog) return 0.0;
    if (bc.mode[side] == MODE_UNIFORM) return bc.val[side];
    return bc.data[side][k];
}

// Which side rings (non-periodic sides) a cell belongs to: ox, oy in {-1,0,1}.
__device__ __forceinline__ void ring_of(const Grid &g, const Bc &bc, int i, int j, int &ox, int &oy) {
    ox = 0;
    oy = 0;
    if (i == 0 && bc.type[XLO] != BC_PERIODIC) ox = -1;
    else if (i == g.nx - 1 && bc.type[XHI] != BC_PERIODIC) ox = 1;
    if (j == 0 && bc.type[YLO] != BC_PERIODIC) oy = -1;
    else if (j == g.ny - 1 && bc.type[YHI] != BC_PERIODIC) oy = 1;
}

// The 9-point operator at (i, j).  Cross-term neighbours outside a
// non-periodic side are quadratic extrapolations "3(a - b) + c" along the
// line through the two interior neighbours and the cell beyond them; a corner
// neighbour outside two sides is the mean of its x- and y-extrapolations.
// Normal fluxes use the stored (ghost) values; a Neumann side's face flux is
// replaced by its BC data.
// Stencil neighbours of a ring cell on the outer side (OX, OY) (each -1, 0
// or +1, not both 0): values outside a non-periodic side are extrapolated
// 3(a - b) + c along the grid line, the corner one is the mean of its x and
// y extrapolations (poisson2d.f90:170-1430).  The side is a template
// argument so every e[][] index is a compile-time constant (registers, not
// scratch memory).
template <int OX, int OY, class Get>
__device__ __forceinline__ void ring_neighbours(const Get &get, int i, int j, double (&e)[3][3]) {
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
        for (int di = -1; di <= 1; ++di)
            e[dj + 1][di + 1] = ((OX && di == OX) || (OY && dj == OY)) ? 0.0 : get(i + di, j + dj);
    if (OX) {
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj) {
            if (OY && dj == OY) continue;
            e[dj + 1][OX + 1] = 3.0 * (e[dj + 1][1] - e[dj + 1][1 - OX]) + get(i - 2 * OX, j + dj);
        }
    }
    if (OY) {
#pragma unroll
        for (int di = -1; di <= 1; ++di) {
            if (OX && di == OX) continue;
            e[OY + 1][di + 1] = 3.0 * (e[1][di + 1] - e[1 - OY][di + 1]) + get(i + di, j - 2 * OY);
        }
    }
    if (OX && OY) {
        const double X = 3.0 * (e[OY + 1][1] - e[OY + 1][1 - OX]) + get(i - 2 * OX, j + OY);
        const double Y = 3.0 * (e[1][OX + 1] - e[1 - OY][OX + 1]) + get(i + OX, j - 2 * OY);
        e[OY + 1][OX + 1] = 0.5 * (X + Y);
    }
}

template <class Get>
__device__ __forceinline__ double cell_lap(const Get &get, const Grid &g, const Geo &geo,
                                           const Bc &bc, int homog, int i, int j) {
    int ox, oy;
    ring_of(g, bc, i, j, ox, oy);
    const int P = g.pitch;
    const double idx = 1.0 / g.dx, idy = 1.0 / g.dy, fourth = 0.25;
    const double p = get(i, j), pe = get(i + 1, j), pw = get(i - 1, j);
    const double pn = get(i, j + 1), ps = get(i, j - 1);
    double ee, ew, en, es, ene, enw, ese, esw;
    if ((ox | oy) == 0) {
        ee = pe;
        ew = pw;
        en = pn;
        es = ps;
        ene = get(i + 1, j + 1);
        enw = get(i - 1, j + 1);
        ese = get(i + 1, j - 1);
        esw = get(i - 1, j - 1);
    } else {
        double e[3][3];  // e[dj+1][di+1]
        switch ((ox + 1) * 3 + (oy + 1)) {
            case 0: ring_neighbours<-1, -1>(get, i, j, e); break;
            case 1: ring_neighbours<-1, 0>(get, i, j, e); break;
            case 2: ring_neighbours<-1, 1>(get, i, j, e); break;
            case 3: ring_neighbours<0, -1>(get, i, j, e); break;
            case 5: ring_neighbours<0, 1>(get, i, j, e); break;
            case 6: ring_neighbours<1, -1>(get, i, j, e); break;
            case 7: ring_neighbours<1, 0>(get, i, j, e); break;
            default: ring_neighbours<1, 1>(get, i, j, e); break;
        }
        ee = e[1][2];
        ew = e[1][0];
        en = e[2][1];
        es = e[0][1];
        ene = e[2][2];
        enw = e[2][0];
        ese = e[0][2];
        esw = e[0][0];
    }
    const size_t c = ix(P, i, j);
    const double Jxx_e = geo.Jxx[c + 1], Jxx_w = geo.Jxx[c];
    const double Jxy_e = geo.Jxy[c + 1], Jxy_w = geo.Jxy[c];
    const double Jyy_n = geo.Jyy[c + P], Jyy_s = geo.Jyy[c];
    const double Jyx_n = geo.Jyx[c + P], Jyx_s = geo.Jyx[c];
    double xfe = Jxx_e * (pe - p) * idx + Jxy_e * fourth * (ene - ese + en - es) * idy;
    double xfw = Jxx_w * (p - pw) * idx + Jxy_w * fourth * (en - es + enw - esw) * idy;
    double yfn = Jyy_n * (pn - p) * idy + Jyx_n * fourth * (ene - enw + ee - ew) * idx;
    double yfs = Jyy_s * (p - ps) * idy + Jyx_s * fourth * (ee - ew + ese - esw) * idx;
    if (ox == -1 && bc.type[XLO] == BC_NEUM) xfw = flux_data(bc, XLO, homog, j);
    if (ox == 1 && bc.type[XHI] == BC_NEUM) xfe = flux_data(bc, XHI, homog, j);
    if (oy == -1 && bc.type[YLO] == BC_NEUM) yfs = flux_data(bc, YLO, homog, i);
    // the reference indexes the yhi face data with j (poisson2d.f90:1216, :1284)
    if (oy == 1 && bc.type[YHI] == BC_NEUM) yfn = flux_data(bc, YHI, homog, j);
    return (xfe - xfw) * idx + (yfn - yfs) * idy;
}

// Interior fast path: no extrapolation, no BC replacement.
// The interior stencil on the 3x3 neighbourhood values (centre p, east,
// west, north, south, then the corners), the cell's metrics at c taken from
// geo: one expression for every caller, so that values read from global
// memory or from an LDS tile give the same bits.
__device__ __forceinline__ double interior_lap_values(const Geo &geo, size_t c, int P, double idx, double idy,
                                                      double p, double pe, double pw, double pn, double ps,
                                                      double ene, double enw, double ese, double esw) {
    const double fourth = 0.25;
    const double ee = pe, ew = pw, en = pn, es = ps;
    const double xfe = geo.Jxx[c + 1] * (pe - p) * idx + geo.Jxy[c + 1] * fourth * (ene - ese + en - es) * idy;
    const double xfw = geo.Jxx[c] * (p - pw) * idx + geo.Jxy[c] * fourth * (en - es + enw - esw) * idy;
    const double yfn = geo.Jyy[c + P] * (pn - p) * idy + geo.Jyx[c + P] * fourth * (ene - enw + ee - ew) * idx;
    const double yfs = geo.Jyy[c] * (p - ps) * idy + geo.Jyx[c] * fourth * (ee - ew + ese - esw) * idx;
    return (xfe - xfw) * idx + (yfn - yfs) * idy;
}

__device__ __forceinline__ double cell_lap_interior(const CFld &f, const Grid &g, const Geo &geo,
                                                    int i, int j) {
    const int P = g.pitch;
    const size_t c = ix(P, i, j);
    const double *q = f.p + c;
    const double idx = 1.0 / g.dx, idy = 1.0 / g.dy;
    return interior_lap_values(geo, c, P, idx, idy, q[0], q[1], q[-1], q[P], q[-P], q[P + 1], q[P - 1], q[-P + 1],
                               q[-P - 1]);
}

// GS/SOR point update (poisson2d.f90:1770-1771, MGPoisson2D_DD.f90:303-304).
__device__ __forceinline__ double gs_update(double phi, double rhs, double lphi, double invd,
                                            double omega) {
    const double newphi = phi + (rhs - lphi) * invd;
    return omega * newphi + (1.0 - omega) * phi;
}

// ---------------------------------------------------------------------------
// Deferred-correction low-order operator (MGPoisson2D_DD.f90:45-119).
// ---------------------------------------------------------------------------

// Order-0 extrapolated value (extrapolate_ghosts(phi, 0), arrayutils.f90:1517-1527):
// ghosts, corners included, copy the nearest valid cell.
__device__ __forceinline__ double clamped(const CFld &f, const Grid &g, int i, int j) {
    i = i < 0 ? 0 : (i >= g.nx ? g.nx - 1 : i);
    j = j < 0 ? 0 : (j >= g.ny ? g.ny - 1 : j);
    return f(i, j);
}

// Four-point cross terms of the DC operator, scaled by 1/4 idx idy.
__device__ __forceinline__ double dc_cross(const CFld &f, const Grid &g, const Geo &geo, int i, int j) {
    const int P = g.pitch;
    const size_t c = ix(P, i, j);
    const double idx = 1.0 / g.dx, idy = 1.0 / g.dy, xyscale = 0.25 * idx * idy;
    const double pe = clamped(f, g, i + 1, j), pw = clamped(f, g, i - 1, j);
    const double pn = clamped(f, g, i, j + 1), ps = clamped(f, g, i, j - 1);
    const double pne = clamped(f, g, i + 1, j + 1), pnw = clamped(f, g, i - 1, j + 1);
    const double psw = clamped(f, g, i - 1, j - 1), pse = clamped(f, g, i + 1, j - 1);
    const double gxye = geo.Jxy[c + 1] * (pne - pse + pn - ps);
    const double gxyw = geo.Jxy[c] * (pn - ps + pnw - psw);
    const double gyxn = geo.Jyx[c + P] * (pne - pnw + pe - pw);
    const double gyxs = geo.Jyx[c] * (pe - pw + pse - psw);
    return (gxye - gxyw + gyxn - gyxs) * xyscale;
}

// Lagged corner-difference cross terms of the DC smoother (scale half idx idy,
// MGPoisson2D_DD.f90:227-250).
// The same from the four diagonal values and the face metrics Jxy(i+1, j),
// Jxy(i, j), Jyx(i, j+1), Jyx(i, j); xyscale = 0.5 idx idy.
__device__ __forceinline__ double dc_lagged_cross_vals(double pne, double pnw, double psw, double pse, double jxye,
                                                       double jxyw, double jyxn, double jyxs, double xyscale) {
    const double gxye = jxye * (pne - pse);
    const double gxyw = jxyw * (pnw - psw);
    const double gyxn = jyxn * (pne - pnw);
    const double gyxs = jyxs * (pse - psw);
    return (gxye - gxyw + gyxn - gyxs) * xyscale;
}

__device__ __forceinline__ double dc_lagged_cross(const CFld &f, const Grid &g, const Geo &geo, int i, int j) {
    const int P = g.pitch;
    const size_t c = ix(P, i, j);
    const double idx = 1.0 / g.dx, idy = 1.0 / g.dy, xyscale = 0.5 * idx * idy;
    const double pne = clamped(f, g, i + 1, j + 1), pnw = clamped(f, g, i - 1, j + 1);
    const double psw = clamped(f, g, i - 1, j - 1), pse = clamped(f, g, i + 1, j - 1);
    return dc_lagged_cross_vals(pne, pnw, psw, pse, geo.Jxy[c + 1], geo.Jxy[c], geo.Jyx[c + P], geo.Jyx[c], xyscale);
}

// 5-point part, as its x and y halves (accumulated one at a time by the
// operator, summed first by the smoother).
// The same from the five values and the face metrics Jxx(i+1, j), Jxx(i, j),
// Jyy(i, j+1), Jyy(i, j).
__device__ __forceinline__ void dc_five_vals(double p, double pe, double pw, double pn, double ps, double jxe,
                                             double jxw, double jyn, double jys, double idx, double idy, double &ax,
                                             double &ay) {
    const double xfe = jxe * (pe - p) * idx;
    const double xfw = jxw * (p - pw) * idx;
    const double yfn = jyn * (pn - p) * idy;
    const double yfs = jys * (p - ps) * idy;
    ax = (xfe - xfw) * idx;
    ay = (yfn - yfs) * idy;
}

__device__ __forceinline__ void dc_five(const CFld &f, const Grid &g, const Geo &geo, int i, int j,
                                        double &ax, double &ay) {
    const int P = g.pitch;
    const size_t c = ix(P, i, j);
    const double *q = f.p + c;
    const double idx = 1.0 / g.dx, idy = 1.0 / g.dy;
    dc_five_vals(q[0], q[1], q[-1], q[P], q[-P], geo.Jxx[c + 1], geo.Jxx[c], geo.Jyy[c + P], geo.Jyy[c], idx, idy,
                 ax, ay);
}

// ---------------------------------------------------------------------------
// Ghost fills, one boundary cell at a time.  A kernel runs the x sides, then
// (after a barrier) the y sides, then the corners, as fill_ghosts does.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double bcval(const Bc &bc, int side, int k) {
    return bc.mode[side] == MODE_NONUNIFORM ? bc.data[side][k] : bc.val[side];
}

// Dirichlet order 3 (arrayutils.f90:1160-1177, :1237-1254) or LAH (negate).
__device__ __forceinline__ double diri_ghost(const Bc &bc, int side, int homog, double p0, double p1,
                                             double p2, int k) {
    if (bc.mode[side] == MODE_LAH) return -p0;
    const double sixteenfifth = 16.0 / 5.0, fifth = 1.0 / 5.0;
    if (homog) return -3.0 * p0 + p1 - fifth * p2;
    return sixteenfifth * bcval(bc, side, k) - 3.0 * p0 + p1 - fifth * p2;
}

// x-side ghost (side XLO/XHI) at row j: fill_ghosts_neum/diri/periodic.
__device__ void fill_x_ghost(Fld f, const Grid &g, const Geo &geo, const Bc &bc, int homog,
                             int do_neum, int side, int j);
// y-side ghost (side YLO/YHI) at column i.
__device__ void fill_y_ghost(Fld f, const Grid &g, const Geo &geo, const Bc &bc, int homog,
                             int do_neum, int side, int i);
// Corners after the four sides (arrayutils.f90:1459-1490); one thread.
__device__ void fill_corners(Fld f, const Grid &g, const Bc &bc);

// Whole fill by one workgroup (any blockDim), with barriers between phases.
__device__ void block_fill_ghosts(Fld f, const Grid &g, const Geo &geo, const Bc &bc, int homog,
                                  int do_neum);

}  // namespace mgp
