"""Full-size GPU checks against the oracle.

* The 2048^2 stretched, warped single-solve problem (Dirichlet xi faces,
  Neumann eta faces, rhs = -J (2 pi^2 sin pi x cos pi y + 1e-2 N(0, 1))):
  KRYLOV_MG and KRYLOV_MG32 solve it to 1e-9 (its FP64 residual floor lies
  near 2e-10, see test_residual_floor.py), and the oracle's operator
  (bit-identical to the reference's compute_residual) confirms the relative
  residual of the returned phi independently of the GPU's own norm.
* The multi-colour sweep at a size that takes the multi-workgroup ring path
  (nx + ny >= 512) relaxes each sampled cell exactly once: with the inverse
  diagonal zero except at that cell and omega = 1, a sweep changes that cell
  alone, by one oracle point update (the single-cell update the golden
  lexicographic vectors pin).
"""
import math

import numpy as np
import pytest

import refcall as rc
from mgpoisson2d_amd import geometry as G

pytestmark = pytest.mark.gpu

mg = pytest.importorskip("mgpoisson2d_amd.capi")


def problem_2048(n=2048, seed=20261202):
    geo = G.mapped_geometry(n, n, 0.1, 5.0)
    x, y = G.cell_centres(geo, 0.1, 5.0)
    f = 2.0 * math.pi ** 2 * np.sin(math.pi * x) * np.cos(math.pi * y)
    rhs = -geo.J * (f + 1e-2 * np.random.default_rng(seed).standard_normal((n, n)))
    bc = G.BoundaryConditions.uniform((1, 1, 0, 0))
    return geo, bc, np.ascontiguousarray(rhs)


@pytest.fixture(scope="module")
def big():
    return problem_2048()


@pytest.mark.parametrize("method", [mg.METHOD_KRYLOV_MG, mg.METHOD_KRYLOV_MG32])
def test_2048_mapped_solve_residual_by_the_oracle(big, method):
    geo, bc, rhs = big
    s = mg.Solver(geo, bc, method=method)
    s.set_rhs(rhs)
    hist, iters, outcome = s.solve(1e-9, 200, zerophi=True)
    assert outcome == mg.CONVERGED, (hist[-4:], iters)
    phi = s.get_phi()
    s.close()
    o = rc.oracle()
    r, _ = o.residual(geo, bc, 1, rhs, phi)
    r0, _ = o.residual(geo, bc, 1, rhs, np.zeros_like(phi))
    rel = np.linalg.norm(r) / np.linalg.norm(r0)
    print(f"method {method}: {iters} iterations, oracle relres {rel:.3e}, GPU relres {hist[-1]:.3e}")
    assert rel <= 1.5e-9
    # the GPU's own residual norm agrees with the oracle's to its rounding
    assert abs(rel - hist[-1]) <= 0.2 * rel + 1e-11


def test_mc_sweep_relaxes_each_cell_once_on_the_wide_path():
    o = rc.oracle()
    nx, ny = 300, 240
    geo = G.mapped_geometry(nx, ny, 0.08, 2.0)
    bc = G.BoundaryConditions.uniform((1, 0, 0, 1), value=0.37)
    rng = np.random.default_rng(9)
    phi = rng.standard_normal((ny + 2, nx + 2))
    rhs = rng.standard_normal((ny, nx))
    invd_full = o.invdiags(geo, bc, 0)
    filled = o.fill_ghosts(geo, bc, 0, 0, phi)
    cells = [(0, 0), (nx - 1, 0), (0, ny - 1), (nx - 1, ny - 1), (1, 0), (nx - 2, ny - 1), (0, 1), (nx - 1, 2),
             (0, ny - 2), (5, 0), (6, ny - 1), (1, 1), (nx - 2, ny - 2), (2, 1), (150, 120), (151, 121)]
    for i, j in cells:
        invd = np.zeros_like(invd_full)
        invd[j, i] = invd_full[j, i]
        got = mg.relax_gs(geo, bc, 0, invd, 1.0, 1, rhs, phi, mg.ORDER_MC)
        once = filled.copy()
        o.gs_point(geo, bc, 0, invd, 1.0, rhs, once, i, j)
        changed = np.argwhere(got[1:-1, 1:-1] != phi[1:-1, 1:-1])
        assert changed.tolist() == [[j, i]], (i, j, changed.tolist())
        assert got[j + 1, i + 1] == once[j + 1, i + 1], (i, j)
