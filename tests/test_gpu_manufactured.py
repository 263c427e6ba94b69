"""Discretization error against a manufactured solution: the CPU-runnable
reference case (unit square, identity mapping, homogeneous Dirichlet BCs,
u = sin(2 pi x) sin(3 pi y), rhs = Laplacian(u) = -13 pi^2 u at cell centres,
zero initial guess, V(2,2) to 1e-10 relative residual).

* The DC V-cycle on the GPU reaches the oracle's (the reference's) discrete
  solution, LEX and RB orderings alike: phi equal to 1e-10 relative, so the
  discretization error matches the reference's.
* Its per-cycle convergence factors agree with the oracle's within 5%.
* The error is second order, a check that does not go through the oracle:
  it falls by 4 (within 3.8 to 4.2) per halving of h, 32^2 -> 64^2 -> 128^2.
* KRYLOV_MG and KRYLOV_MG32 reach the same discrete solution (1e-8 relative),
  hence the same error.
"""
import numpy as np
import pytest

import refcall as rc

pytestmark = pytest.mark.gpu

mg = pytest.importorskip("mgpoisson2d_amd.capi")
from mgpoisson2d_amd import geometry as G  # noqa: E402


def problem(n):
    geo = G.mapped_geometry(n, n, 0.0, 0.0)
    bc = G.BoundaryConditions.uniform((1, 1, 1, 1))
    xc = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(xc, xc)
    u = np.sin(2 * np.pi * X) * np.sin(3 * np.pi * Y)
    return geo, bc, u, -13.0 * np.pi ** 2 * u


def gpu_solve(geo, bc, rhs, method, ordering=mg.ORDER_LEX, maxiters=60):
    s = mg.Solver(geo, bc, method=method, ordering=ordering)
    s.set_rhs(rhs)
    hist, iters, outcome = s.solve(1e-10, maxiters, zerophi=True)
    phi = s.get_phi()[1:-1, 1:-1].copy()
    s.close()
    assert outcome == mg.CONVERGED, (iters, hist)
    return phi, np.asarray(hist)


@pytest.mark.parametrize("ordering", [mg.ORDER_LEX, mg.ORDER_RB])
def test_dc_vcycle_discretization_error_matches_reference(ordering):
    errs = []
    for n in (32, 64, 128):
        geo, bc, u, rhs = problem(n)
        phi, hist = gpu_solve(geo, bc, rhs, mg.METHOD_DC_VCYCLE, ordering)
        ref, _, h2, _, _ = rc.oracle().vcycle(1, geo, bc, 1, 1e-10, 60, np.zeros((n + 2, n + 2)), rhs,
                                              zerophi=True, order=ordering)
        ref = ref[1:-1, 1:-1]
        np.testing.assert_allclose(phi, ref, rtol=0, atol=1e-10 * np.max(np.abs(ref)))
        f, f2 = hist[1:] / hist[:-1], h2[1:] / h2[:-1]
        assert len(f) == len(f2)
        np.testing.assert_allclose(f, f2, rtol=0.05)
        errs.append(np.max(np.abs(phi - u)))
    for a, b in zip(errs, errs[1:]):
        assert 3.8 < a / b < 4.2, errs


@pytest.mark.parametrize("method", [mg.METHOD_KRYLOV_MG, mg.METHOD_KRYLOV_MG32])
def test_krylov_mg_reaches_the_same_discrete_solution(method):
    geo, bc, u, rhs = problem(64)
    ref, _, _, _, _ = rc.oracle().vcycle(1, geo, bc, 1, 1e-12, 80, np.zeros((66, 66)), rhs, zerophi=True)
    ref = ref[1:-1, 1:-1]
    phi, _ = gpu_solve(geo, bc, rhs, method, maxiters=300)
    np.testing.assert_allclose(phi, ref, rtol=0, atol=1e-8 * np.max(np.abs(ref)))
    e, e_ref = np.max(np.abs(phi - u)), np.max(np.abs(ref - u))
    assert abs(e - e_ref) <= 1e-6 * e_ref
