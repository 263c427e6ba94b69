"""Metric assembly on the GPU (mgp_metric_fill, mgp_solver_create_mapped).

* MAP_TERRAIN, the reference's test geometry (test.f90:320-716), against the
  oracle's restatement: the device and host sine may differ in the last bit
  of y = O(1), and the metrics are differences of positions divided by the
  spacing, so the bound is a few ulps of 1 times max(nx, ny): 1e-15 n.
* The discrete geometric conservation law holds on the device result.
* MAP_WARPED against the analytic numpy metrics of geometry.py (1e-14
  relative).
* A resident solver built from the mapping on the device solves exactly as
  one built from the same metrics uploaded from the host.
"""
import numpy as np
import pytest

import refcall as rc

pytestmark = pytest.mark.gpu

mg = pytest.importorskip("mgpoisson2d_amd.capi")


@pytest.mark.parametrize("n", [(16, 12), (129, 64), (512, 384)])
def test_terrain_metrics_match_oracle(n):
    nx, ny = n
    m = mg.Mapping(mg.MAP_TERRAIN, 0.25, 0.0, 0.0, 0.0)
    g = mg.metric_fill(m, nx, ny, 1.0 / nx, 1.0 / ny)
    o = rc.oracle().metric_terrain(nx, ny, 1.0 / nx, 1.0 / ny, 0.25)
    for a, b in zip(g.arrays(), o.arrays()):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-15 * max(nx, ny))
    gcl = (g.Jxx[:, 1:] - g.Jxx[:, :-1]) / g.dx + (g.Jyx[1:, :] - g.Jyx[:-1, :]) / g.dy
    assert np.max(np.abs(gcl)) <= 1e-16 * nx * ny


@pytest.mark.parametrize("amp,beta", [(0.0, 0.0), (0.1, 5.0), (0.05, 2.0)])
def test_warped_metrics_match_analytic(amp, beta):
    from mgpoisson2d_amd import geometry as G
    nx, ny = 300, 200
    m = mg.Mapping(mg.MAP_WARPED, amp, beta, 1.0, 1.0)
    g = mg.metric_fill(m, nx, ny, 1.0 / nx, 1.0 / ny)
    want = G.mapped_geometry(nx, ny, amp, beta)
    for a, b in zip(g.arrays(), want.arrays()):
        np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-14 * np.max(np.abs(b)))


@pytest.mark.parametrize("kind,amp,beta,types", [
    (0, 0.25, 0.0, (1, 1, 0, 0)),       # the reference's terrain geometry
    (1, 0.1, 5.0, (1, 1, 0, 0)),        # the benchmark mapping
])
def test_solver_from_mapping_equals_uploaded_metrics(kind, amp, beta, types):
    from cases import setup
    nx = ny = 256
    m = mg.Mapping(kind, amp, beta, 0.0 if kind == 0 else 1.0, 0.0 if kind == 0 else 1.0)
    geo = mg.metric_fill(m, nx, ny, 1.0 / nx, 1.0 / ny)
    _, bc, _, rhs = setup(nx, ny, 0.0, 0.0, types, seed=31)
    a = mg.Solver(geo, bc, method=mg.METHOD_KRYLOV_MG)
    b = mg.Solver(geo, bc, method=mg.METHOD_KRYLOV_MG, mapping=m)
    for s in (a, b):
        s.set_rhs(rhs)
    ha, ia, oa = a.solve(1e-10, 200, zerophi=True)
    hb, ib, ob = b.solve(1e-10, 200, zerophi=True)
    assert oa == mg.CONVERGED and (ia, oa) == (ib, ob)
    np.testing.assert_array_equal(ha, hb)
    np.testing.assert_array_equal(a.get_phi()[1:-1, 1:-1], b.get_phi()[1:-1, 1:-1])
    # independent check of the converged solution with the oracle's operator
    r, _ = rc.oracle().residual(geo, bc, 1, rhs, a.get_phi())
    r0, _ = rc.oracle().residual(geo, bc, 1, rhs, np.zeros((ny + 2, nx + 2)))
    assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(r0)


def test_reference_fmg_on_terrain_geometry():
    """The reference's own geometry and solver: fmg on the terrain metrics
    (generated on the GPU, handed to both sides) against the oracle's fmg."""
    from cases import setup
    nx = ny = 64
    m = mg.Mapping(mg.MAP_TERRAIN, 0.25, 0.0, 0.0, 0.0)
    geo = mg.metric_fill(m, nx, ny, 1.0 / nx, 1.0 / ny)
    _, bc, _, rhs = setup(nx, ny, 0.0, 0.0, (0, 0, 0, 0), seed=32)
    s = mg.Solver(geo, bc, method=mg.METHOD_FMG, ordering=mg.ORDER_LEX, mapping=m)
    s.set_rhs(rhs)
    hist, iters, outcome = s.solve(1e-30, 6, zerophi=True)
    _, _, h2, it2, oc2 = rc.oracle().vcycle(2, geo, bc, 1, 1e-30, 6, np.zeros((ny + 2, nx + 2)), rhs, zerophi=True)
    assert (iters, outcome) == (it2, oc2)
    np.testing.assert_allclose(hist, h2, rtol=1e-12, atol=0)
