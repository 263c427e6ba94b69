// mgp_dd.h -- double-double arithmetic and the reference operator evaluated
// in it, for the refinement residual f - L (phi_hi + phi_lo).
//
// A double-double value is an unevaluated sum hi + lo with |lo| <= ulp(hi)/2,
// about 106 significant bits.  The error-free transformations are the
// classical ones: TwoSum (Knuth), FastTwoSum (Dekker) and TwoProduct with a
// fused multiply-add (Ogita, Rump and Oishi, "Accurate sum and dot product",
// SIAM J. Sci. Comput. 26 (2005)); the compound operations follow the
// double-double algorithms of Hida, Li and Bailey (QD library report, 2000).
// All HIP code is compiled with -ffp-contract=off, so no other operation is
// fused behind these sequences' backs.
//
// cell_lap_dd evaluates the 9-point operator of cell_lap (mgp_stencil.h) --
// the same terms (regrouped only where exact arithmetic cannot tell), the same
// constants (1/dx, 1/dy, 1/4, 1/5 as the FP64 program forms them), the same
// boundary rules -- with every operation in double-double, for homogeneous
// Dirichlet / Neumann sides.  Its result is the
// real-arithmetic value of the reference's expression on the FP64 metrics to
// about 1e-30 relative, so the residual of a solution carried as hi + lo is no
// longer limited by the rounding of the FP64 operator (at 2048^2 on the
// stretched map that rounding alone is about 2e-10 of the rhs).
#pragma once

#include <hip/hip_runtime.h>

#include "mgp_common.h"

namespace mgp {

struct DD {
    double hi, lo;
};

__device__ __forceinline__ DD dd_two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return DD{s, e};
}
__device__ __forceinline__ DD dd_fast_two_sum(double a, double b) {  // |a| >= |b| or a == 0
    const double s = a + b;
    const double e = b - (s - a);
    return DD{s, e};
}
__device__ __forceinline__ DD dd_two_prod(double a, double b) {
    const double p = a * b;
    const double e = __builtin_fma(a, b, -p);
    return DD{p, e};
}
__device__ __forceinline__ DD dd_from(double a) { return DD{a, 0.0}; }

__device__ __forceinline__ DD operator+(const DD &a, const DD &b) {
    DD s = dd_two_sum(a.hi, b.hi);
    const DD t = dd_two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = dd_fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return dd_fast_two_sum(s.hi, s.lo);
}
__device__ __forceinline__ DD operator-(const DD &a) { return DD{-a.hi, -a.lo}; }
__device__ __forceinline__ DD operator-(const DD &a, const DD &b) { return a + (-b); }
__device__ __forceinline__ DD operator+(const DD &a, double b) {
    DD s = dd_two_sum(a.hi, b);
    s.lo += a.lo;
    return dd_fast_two_sum(s.hi, s.lo);
}
__device__ __forceinline__ DD operator*(const DD &a, double b) {
    DD p = dd_two_prod(a.hi, b);
    p.lo += a.lo * b;
    return dd_fast_two_sum(p.hi, p.lo);
}
__device__ __forceinline__ DD operator*(double b, const DD &a) { return a * b; }
__device__ __forceinline__ DD operator*(const DD &a, const DD &b) {
    DD p = dd_two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return dd_fast_two_sum(p.hi, p.lo);
}

// hi + lo of a field carried as two FP64 arrays, with the homogeneous
// Dirichlet ghosts (order 3, -3 p0 + p1 - (1/5) p2) synthesised in
// double-double; other ghosts read as stored (a Neumann face's flux is
// replaced, so its ghost value never enters).
struct DDGet {
    const double *hi, *lo;
    int pitch, nx, ny;
    int t[4];
    __device__ __forceinline__ DD raw(int i, int j) const {
        const size_t c = ix(pitch, i, j);
        return DD{hi[c], lo[c]};
    }
    __device__ __forceinline__ static DD dg(const DD &p0, const DD &p1, const DD &p2) {
        return (p0 * -3.0 + p1) - p2 * (1.0 / 5.0);
    }
    __device__ __forceinline__ DD operator()(int i, int j) const {
        const bool xin = i >= 0 && i < nx, yin = j >= 0 && j < ny;
        if (xin == yin) return raw(i, j);
        if (i < 0) return t[XLO] != BC_DIRI ? raw(i, j) : dg(raw(0, j), raw(1, j), raw(2, j));
        if (i >= nx) return t[XHI] != BC_DIRI ? raw(i, j) : dg(raw(nx - 1, j), raw(nx - 2, j), raw(nx - 3, j));
        if (j < 0) return t[YLO] != BC_DIRI ? raw(i, j) : dg(raw(i, 0), raw(i, 1), raw(i, 2));
        return t[YHI] != BC_DIRI ? raw(i, j) : dg(raw(i, ny - 1), raw(i, ny - 2), raw(i, ny - 3));
    }
};

// Neighbours of a ring cell on outer side (OX, OY): cross-term values beyond
// a side are 3(a - b) + c along the grid line, the corner one the mean of its
// x and y extrapolations (the rule of ring_neighbours in mgp_stencil.h).
template <int OX, int OY>
__device__ __forceinline__ void ring_neighbours_dd(const DDGet &get, int i, int j, DD (&e)[3][3]) {
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
        for (int di = -1; di <= 1; ++di)
            e[dj + 1][di + 1] = ((OX && di == OX) || (OY && dj == OY)) ? dd_from(0.0) : get(i + di, j + dj);
    if (OX) {
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj) {
            if (OY && dj == OY) continue;
            e[dj + 1][OX + 1] = (e[dj + 1][1] - e[dj + 1][1 - OX]) * 3.0 + get(i - 2 * OX, j + dj);
        }
    }
    if (OY) {
#pragma unroll
        for (int di = -1; di <= 1; ++di) {
            if (OX && di == OX) continue;
            e[OY + 1][di + 1] = (e[1][di + 1] - e[1 - OY][di + 1]) * 3.0 + get(i + di, j - 2 * OY);
        }
    }
    if (OX && OY) {
        const DD X = (e[OY + 1][1] - e[OY + 1][1 - OX]) * 3.0 + get(i - 2 * OX, j + OY);
        const DD Y = (e[1][OX + 1] - e[1 - OY][OX + 1]) * 3.0 + get(i + OX, j - 2 * OY);
        e[OY + 1][OX + 1] = (X + Y) * 0.5;
    }
}

// The operator at (i, j) on homogeneous Dirichlet / Neumann sides (no
// periodic sides).
__device__ __forceinline__ DD cell_lap_dd(const DDGet &get, const Grid &g, const Geo &geo, const Bc &bc, int i,
                                          int j) {
    const int ox = i == 0 ? -1 : (i == g.nx - 1 ? 1 : 0);
    const int oy = j == 0 ? -1 : (j == g.ny - 1 ? 1 : 0);
    const int P = g.pitch;
    const double idx = 1.0 / g.dx, idy = 1.0 / g.dy, fourth = 0.25;
    const DD p = get(i, j), pe = get(i + 1, j), pw = get(i - 1, j);
    const DD pn = get(i, j + 1), ps = get(i, j - 1);
    DD ee, ew, en, es, ene, enw, ese, esw;
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
        DD e[3][3];
        switch ((ox + 1) * 3 + (oy + 1)) {
            case 0: ring_neighbours_dd<-1, -1>(get, i, j, e); break;
            case 1: ring_neighbours_dd<-1, 0>(get, i, j, e); break;
            case 2: ring_neighbours_dd<-1, 1>(get, i, j, e); break;
            case 3: ring_neighbours_dd<0, -1>(get, i, j, e); break;
            case 5: ring_neighbours_dd<0, 1>(get, i, j, e); break;
            case 6: ring_neighbours_dd<1, -1>(get, i, j, e); break;
            case 7: ring_neighbours_dd<1, 0>(get, i, j, e); break;
            default: ring_neighbours_dd<1, 1>(get, i, j, e); break;
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
    DD xfe = (pe - p) * Jxx_e * idx + ((ene - ese) + (en - es)) * Jxy_e * fourth * idy;
    DD xfw = (p - pw) * Jxx_w * idx + ((en - es) + (enw - esw)) * Jxy_w * fourth * idy;
    DD yfn = (pn - p) * Jyy_n * idy + ((ene - enw) + (ee - ew)) * Jyx_n * fourth * idx;
    DD yfs = (p - ps) * Jyy_s * idy + ((ee - ew) + (ese - esw)) * Jyx_s * fourth * idx;
    if (ox == -1 && bc.type[XLO] == BC_NEUM) xfw = dd_from(0.0);
    if (ox == 1 && bc.type[XHI] == BC_NEUM) xfe = dd_from(0.0);
    if (oy == -1 && bc.type[YLO] == BC_NEUM) yfs = dd_from(0.0);
    if (oy == 1 && bc.type[YHI] == BC_NEUM) yfn = dd_from(0.0);
    return (xfe - xfw) * idx + (yfn - yfs) * idy;
}

}  // namespace mgp
