"""Helpers for the fourth-order operator's tests: the manufactured all-Neumann
problem u = cos(pi x) cos(2 pi y) on the warped, stretched map, cell averages
by Gauss-Legendre quadrature, the probed operator matrix and a direct solve.

Sign convention of the library (the reference's): L = +J Laplacian, so the
rhs for -Laplacian u = f is -<J f> (cell averages over computational cells).
For u above, f = 5 pi^2 u, and u's conormal derivative vanishes on all four
sides of the unit square the map preserves (x(0, eta) = 0, x(1, eta) = 1,
y(xi, 0) = 0, y(xi, 1) = 1), so the homogeneous Neumann problem holds.
"""
from __future__ import annotations

import math

import numpy as np

import refcall as rc


def mapping_points(xi, eta, amp, beta):
    x = xi + amp * np.sin(2 * math.pi * xi) * np.sin(2 * math.pi * eta)
    y = (np.expm1(beta * eta) / math.expm1(beta) + 0 * xi) if beta != 0.0 else eta + 0 * xi
    return x, y


def jacobian(xi, eta, amp, beta):
    x_xi = 1 + amp * 2 * math.pi * np.cos(2 * math.pi * xi) * np.sin(2 * math.pi * eta)
    y_eta = (beta * np.exp(beta * eta) / math.expm1(beta) + 0 * xi) if beta != 0.0 else 1 + 0 * xi
    return x_xi * y_eta


def cell_average(fun, nx, ny, npts=4):
    """Average of fun(xi, eta) over each computational cell of [0,1]^2 (ny, nx)."""
    t, w = np.polynomial.legendre.leggauss(npts)
    t, w = (t + 1) / 2, w / 2
    hx, hy = 1.0 / nx, 1.0 / ny
    out = np.zeros((ny, nx))
    for a, wa in zip(t, w):
        for b, wb in zip(t, w):
            X, E = np.meshgrid((np.arange(nx) + a) * hx, (np.arange(ny) + b) * hy)
            out += wa * wb * fun(X, E)
    return out


def u_exact(amp, beta):
    def f(xi, eta):
        x, y = mapping_points(xi, eta, amp, beta)
        return np.cos(math.pi * x) * np.cos(2 * math.pi * y)
    return f


def rhs_mms(nx, ny, amp, beta, noise=0.0, seed=0):
    """-<J f> for u = cos(pi x) cos(2 pi y), plus noise * J N(0,1) per cell,
    projected to zero sum (the all-Neumann compatibility condition)."""
    def Jf(xi, eta):
        x, y = mapping_points(xi, eta, amp, beta)
        return jacobian(xi, eta, amp, beta) * 5 * math.pi ** 2 * np.cos(math.pi * x) * np.cos(2 * math.pi * y)
    rhs = -cell_average(Jf, nx, ny)
    if noise:
        Jc = cell_average(lambda a, b: jacobian(a, b, amp, beta), nx, ny, 2)
        rhs -= noise * Jc * np.random.default_rng(seed).standard_normal((ny, nx))
    return rhs - rhs.mean()


def probed_lap4(nx, ny, met, spacing=11):
    """Sparse matrix of the oracle's fourth-order operator (stencil reach <= 5)."""
    import scipy.sparse as sp
    o = rc.oracle()
    rows, cols, vals = [], [], []
    for ci in range(spacing):
        for cj in range(spacing):
            phi = np.zeros((ny, nx))
            phi[cj::spacing, ci::spacing] = 1.0
            L = o.lap4(nx, ny, 1.0 / nx, 1.0 / ny, met, phi)
            J, I = np.nonzero(L)
            si = ci + np.round((I - ci) / spacing).astype(int) * spacing
            sj = cj + np.round((J - cj) / spacing).astype(int) * spacing
            ok = (si >= 0) & (si < nx) & (sj >= 0) & (sj < ny)
            rows.append(J[ok] * nx + I[ok])
            cols.append(sj[ok] * nx + si[ok])
            vals.append(L[J[ok], I[ok]])
    n = nx * ny
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


def solve_neumann(A, rhs):
    """The zero-mean solution of the singular all-Neumann system (bordered)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    n = A.shape[0]
    ones = np.ones((n, 1))
    Ab = sp.bmat([[A, ones], [ones.T, None]]).tocsc()
    x = spla.spsolve(Ab, np.concatenate([rhs.ravel(), [0.0]]))[:n]
    return x.reshape(rhs.shape)


def error_vs_u(phi, nx, ny, amp, beta):
    """max and rms of phi - <u> after removing the mean difference."""
    ua = cell_average(u_exact(amp, beta), nx, ny)
    d = phi - ua
    d = d - d.mean()
    return float(np.max(np.abs(d))), float(np.sqrt(np.mean(d * d)))
