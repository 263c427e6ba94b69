"""Shared problem set-ups for the parity tests (seeded, sized for the CPU
checkers to finish in seconds)."""
from __future__ import annotations

import numpy as np

from mgpoisson2d_amd import geometry as G

# (nx, ny, warp amplitude, stretching beta, BC types xlo, xhi, ylo, yhi)
OPERATOR_CASES = [
    (16, 12, 0.0, 0.0, (1, 1, 1, 1)),     # identity, all Dirichlet
    (16, 16, 0.10, 3.0, (1, 1, 0, 0)),    # config-2 family: Dirichlet xi, Neumann eta
    (20, 12, 0.12, 2.0, (0, 0, 0, 0)),    # all Neumann
    (12, 16, 0.08, 1.0, (0, 1, 1, 0)),    # mixed corners (cross metric at Dirichlet faces)
    (33, 17, 0.05, 4.0, (1, 0, 0, 1)),    # odd sizes
]

CYCLE_CASES = [
    (64, 64, 0.0, 0.0, (1, 1, 1, 1)),     # the 64x64 identity Dirichlet reference case
    (32, 32, 0.10, 3.0, (1, 1, 0, 0)),
    (32, 32, 0.12, 2.0, (0, 0, 0, 0)),
    (48, 48, 0.08, 1.0, (0, 1, 1, 0)),
]


def setup(nx, ny, amp, beta, types, homog=True, seed=0):
    geo = G.mapped_geometry(nx, ny, amp, beta)
    bc = G.BoundaryConditions.uniform(types, value=0.0 if homog else 0.37)
    rng = np.random.default_rng(seed)
    phi = rng.standard_normal((ny + 2, nx + 2))
    rhs = rng.standard_normal((ny, nx))
    if all(t == 0 for t in types):
        # the operator's fluxes telescope, so a J-scaled rhs is compatible when
        # its plain sum vanishes (the J-weighted mean of f = rhs/J is zero)
        rhs = rhs - np.mean(rhs)
    return geo, bc, phi, rhs


def coarse_geo(checker_restrict, geo):
    cnx, cny = geo.nx // 2, geo.ny // 2
    arrs = [checker_restrict(k, geo.nx, geo.ny, a, cnx, cny) for k, a in zip((0, 1, 1, 2, 2), geo.arrays())]
    return G.Geometry(cnx, cny, geo.dx * 2, geo.dy * 2, *arrs)


def probed_matrix(geo, bc):
    """Sparse matrix of the reference operator (homogeneous BCs) by probing the
    oracle with 25 interleaved impulse patterns (the stencil reaches 2 cells)."""
    import scipy.sparse as sp
    import refcall as rc
    o = rc.oracle()
    nx, ny = geo.nx, geo.ny
    rows, cols, vals = [], [], []
    for ci in range(5):
        for cj in range(5):
            phi = np.zeros((ny + 2, nx + 2))
            phi[1 + cj:ny + 1:5, 1 + ci:nx + 1:5] = 1.0
            lap, _ = o.laplacian(geo, bc, 1, phi)
            J, I = np.nonzero(lap)
            d = (I - ci) % 5
            d = np.where(d > 2, d - 5, d)
            e = (J - cj) % 5
            e = np.where(e > 2, e - 5, e)
            si, sj = I - d, J - e
            ok = (si >= 0) & (si < nx) & (sj >= 0) & (sj < ny)
            rows.append(J[ok] * nx + I[ok])
            cols.append(sj[ok] * nx + si[ok])
            vals.append(lap[J[ok], I[ok]])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nx * ny,) * 2)
