"""The reference's 9-point operator (poisson2d.f90 compute_laplacian
:886-1319, homogeneous Dirichlet / Neumann sides) written once over any
number type, for pinning the double-double residual.

With Python floats (IEEE doubles, the same rounding as the FP64 program) the
operations below run in the reference's order, so lap_float equals the oracle's
and the reference's compute_laplacian bit for bit -- that pins this
restatement.  With fractions.Fraction the same code evaluates the operator
exactly on the FP64 inputs (metrics, 1/dx, 1/dy and the constant 1/5 as the
FP64 program forms them); residual_exact rounds the exact rhs - L phi once,
which is what a double-double evaluation must reproduce to within an ulp.

Ghosts: a Dirichlet side's ghost is the order-3 extrapolation
-3 p0 + p1 - (1/5) p2 (arrayutils.f90:1160-1177, homogeneous); a Neumann
side's face flux is zero and its ghost never enters.  Cross-term neighbours
beyond a side are 3(a - b) + c along the grid line, a corner neighbour the
mean of its x and y extrapolations (poisson2d.f90:170-1430).
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np

XLO, XHI, YLO, YHI = 0, 1, 2, 3
DIRI, NEUM = 1, 0


def _getter(phi, nx, ny, types, num):
    """phi: (ny+2, nx+2) with ghosts; values as `num` (float or Fraction)."""
    fifth = num(1.0 / 5.0)

    def raw(i, j):
        return num(phi[j + 1][i + 1])

    def dg(p0, p1, p2):
        return (num(-3.0) * p0 + p1) - fifth * p2

    def get(i, j):
        xin, yin = 0 <= i < nx, 0 <= j < ny
        if xin == yin:
            return raw(i, j)
        if i < 0:
            return raw(i, j) if types[XLO] != DIRI else dg(raw(0, j), raw(1, j), raw(2, j))
        if i >= nx:
            return raw(i, j) if types[XHI] != DIRI else dg(raw(nx - 1, j), raw(nx - 2, j), raw(nx - 3, j))
        if j < 0:
            return raw(i, j) if types[YLO] != DIRI else dg(raw(i, 0), raw(i, 1), raw(i, 2))
        return raw(i, j) if types[YHI] != DIRI else dg(raw(i, ny - 1), raw(i, ny - 2), raw(i, ny - 3))

    return get


def cell_lap(get, geo, types, i, j, num):
    nx, ny = geo.nx, geo.ny
    ox = -1 if i == 0 else (1 if i == nx - 1 else 0)
    oy = -1 if j == 0 else (1 if j == ny - 1 else 0)
    idx, idy, fourth = num(1.0 / geo.dx), num(1.0 / geo.dy), num(0.25)
    three, half = num(3.0), num(0.5)
    p, pe, pw, pn, ps = get(i, j), get(i + 1, j), get(i - 1, j), get(i, j + 1), get(i, j - 1)
    if ox == 0 and oy == 0:
        ee, ew, en, es = pe, pw, pn, ps
        ene, enw, ese, esw = get(i + 1, j + 1), get(i - 1, j + 1), get(i + 1, j - 1), get(i - 1, j - 1)
    else:
        e = [[None] * 3 for _ in range(3)]  # e[dj+1][di+1]
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                e[dj + 1][di + 1] = num(0.0) if ((ox and di == ox) or (oy and dj == oy)) else get(i + di, j + dj)
        if ox:
            for dj in (-1, 0, 1):
                if oy and dj == oy:
                    continue
                e[dj + 1][ox + 1] = three * (e[dj + 1][1] - e[dj + 1][1 - ox]) + get(i - 2 * ox, j + dj)
        if oy:
            for di in (-1, 0, 1):
                if ox and di == ox:
                    continue
                e[oy + 1][di + 1] = three * (e[1][di + 1] - e[1 - oy][di + 1]) + get(i + di, j - 2 * oy)
        if ox and oy:
            X = three * (e[oy + 1][1] - e[oy + 1][1 - ox]) + get(i - 2 * ox, j + oy)
            Y = three * (e[1][ox + 1] - e[1 - oy][ox + 1]) + get(i + ox, j - 2 * oy)
            e[oy + 1][ox + 1] = half * (X + Y)
        ee, ew, en, es = e[1][2], e[1][0], e[2][1], e[0][1]
        ene, enw, ese, esw = e[2][2], e[2][0], e[0][2], e[0][0]
    Jxx_e, Jxx_w = num(geo.Jxx[j][i + 1]), num(geo.Jxx[j][i])
    Jxy_e, Jxy_w = num(geo.Jxy[j][i + 1]), num(geo.Jxy[j][i])
    Jyy_n, Jyy_s = num(geo.Jyy[j + 1][i]), num(geo.Jyy[j][i])
    Jyx_n, Jyx_s = num(geo.Jyx[j + 1][i]), num(geo.Jyx[j][i])
    xfe = Jxx_e * (pe - p) * idx + Jxy_e * fourth * (ene - ese + en - es) * idy
    xfw = Jxx_w * (p - pw) * idx + Jxy_w * fourth * (en - es + enw - esw) * idy
    yfn = Jyy_n * (pn - p) * idy + Jyx_n * fourth * (ene - enw + ee - ew) * idx
    yfs = Jyy_s * (p - ps) * idy + Jyx_s * fourth * (ee - ew + ese - esw) * idx
    zero = num(0.0)
    if ox == -1 and types[XLO] == NEUM:
        xfw = zero
    if ox == 1 and types[XHI] == NEUM:
        xfe = zero
    if oy == -1 and types[YLO] == NEUM:
        yfs = zero
    if oy == 1 and types[YHI] == NEUM:
        yfn = zero
    return (xfe - xfw) * idx + (yfn - yfs) * idy


def lap_float(geo, types, phi):
    """L phi in FP64 (the reference's operation order); phi (ny+2, nx+2)."""
    get = _getter(phi, geo.nx, geo.ny, types, float)
    return np.array([[cell_lap(get, geo, types, i, j, float) for i in range(geo.nx)] for j in range(geo.ny)])


def residual_exact(geo, types, rhs, phi_hi, phi_lo):
    """rhs - L (phi_hi + phi_lo) evaluated exactly, rounded once to FP64."""
    ny, nx = rhs.shape
    tot = [[Fraction(phi_hi[j][i]) + Fraction(phi_lo[j][i]) for i in range(nx + 2)] for j in range(ny + 2)]
    get = _getter_fraction(tot, nx, ny, types)
    out = np.zeros((ny, nx))
    for j in range(ny):
        for i in range(nx):
            out[j, i] = float(Fraction(rhs[j][i]) - cell_lap(get, geo, types, i, j, Fraction))
    return out


def _getter_fraction(tot, nx, ny, types):
    # the same rule as _getter over exact values
    class _P:
        def __getitem__(self, j):
            return tot[j]
    return _getter(_P(), nx, ny, types, lambda v: v if isinstance(v, Fraction) else Fraction(v))
