"""GPU (HIP library, through the C ABI) against the CPU checkers.

Kernel-level results must equal the oracle bit for bit (all code is built
with -ffp-contract=off and keeps the reference's operand order); where the
reference build is present (oracle/_ref) it is compared too.  Cycle-level
results use parallel compensated reductions for the norms, so relres
histories agree to 1e-12 relative and solutions to 1e-12 of their scale.
"""
import numpy as np
import pytest

import refcall as rc
from cases import CYCLE_CASES, OPERATOR_CASES, coarse_geo, setup

pytestmark = pytest.mark.gpu

mg = pytest.importorskip("mgpoisson2d_amd.capi")


def checkers():
    out = [rc.oracle()]
    if rc.reference_available():
        out.append(rc.reference())
    return out


def same(a, b):
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("case", OPERATOR_CASES)
@pytest.mark.parametrize("homog", [1, 0])
def test_operator_fill_invdiags(case, homog):
    geo, bc, phi, rhs = setup(*case, homog=homog, seed=1)
    lap, phig = mg.compute_laplacian(geo, bc, homog, phi)
    res, _ = mg.compute_residual(geo, bc, homog, rhs, phi)
    inv = mg.compute_invdiags(geo, bc, homog)
    fn = mg.fill_ghosts(geo, bc, homog, 1, phi)
    for c in checkers():
        l2, p2 = c.laplacian(geo, bc, homog, phi)
        same(lap, l2)
        same(phig, p2)
        same(res, c.residual(geo, bc, homog, rhs, phi)[0])
        same(inv, c.invdiags(geo, bc, homog))
        same(fn, c.fill_ghosts(geo, bc, homog, 1, phi))


@pytest.mark.parametrize("case", OPERATOR_CASES)
@pytest.mark.parametrize("homog", [1, 0])
def test_flux_tools(case, homog):
    """compute_pd (all four face/derivative pairs), compute_grad, compute_div."""
    geo, bc, phi, _ = setup(*case, homog=homog, seed=11)
    nx, ny, dx, dy = geo.nx, geo.ny, geo.dx, geo.dy
    pds = {(nd, d): mg.compute_pd(nx, ny, dx, dy, phi, d, nd) for nd in (1, 2) for d in (1, 2)}
    phig, xf, yf = mg.compute_grad(geo, bc, homog, phi)
    div = mg.compute_div(nx, ny, dx, dy, xf, yf)
    for c in checkers():
        for (nd, d), pd in pds.items():
            same(pd, c.compute_pd(nx, ny, dx, dy, phi, d, nd))
        p2, x2, y2 = c.compute_grad(geo, bc, homog, phi)
        same(phig, p2)
        same(xf, x2)
        same(yf, y2)
        same(div, c.compute_div(nx, ny, dx, dy, xf, yf))


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("homog", [1, 0])
def test_fill_boundary_fluxes(mode, homog):
    from mgpoisson2d_amd import geometry as G
    rng = np.random.default_rng(mode * 2 + homog)
    for nx, ny, types in ((16, 12, (0, 0, 0, 0)), (33, 17, (1, 0, 0, 1)), (5, 300, (0, 1, 1, 0))):
        bc = G.BoundaryConditions(types, (mode,) * 4, tuple(rng.standard_normal(ny if k < 2 else nx) for k in range(4)))
        xf, yf = rng.standard_normal((ny, nx + 1)), rng.standard_normal((ny + 1, nx))
        a = mg.fill_boundary_fluxes(nx, ny, bc, homog, xf, yf)
        for c in checkers():
            b = c.fill_boundary_fluxes(nx, ny, bc, homog, xf, yf)
            same(a[0], b[0])
            same(a[1], b[1])


@pytest.mark.parametrize("n", [(3, 3), (3, 9), (300, 4), (1030, 700)])
def test_flux_tools_sizes(n):
    """Smallest legal grids (every face is a boundary-closure face) and grids
    wider than one workgroup tile, against the oracle."""
    geo, bc, phi, _ = setup(n[0], n[1], 0.07, 2.5, (1, 0, 0, 1), homog=False, seed=12)
    o = rc.oracle()
    for a, b in zip(mg.compute_grad(geo, bc, 0, phi), o.compute_grad(geo, bc, 0, phi)):
        same(a, b)
    for nd in (1, 2):
        for d in (1, 2):
            same(mg.compute_pd(geo.nx, geo.ny, geo.dx, geo.dy, phi, d, nd),
                 o.compute_pd(geo.nx, geo.ny, geo.dx, geo.dy, phi, d, nd))


def test_div_of_grad_is_laplacian_2048():
    """Full size (2048^2 stretched mapped grid, Dirichlet xi / Neumann eta):
    J div(J grad phi) from the flux kernels equals the Laplacian kernel's
    result to rounding (both use the same face stencils; only the operation
    order differs).  Bound: 1e-14 of the largest flux difference quotient."""
    import os, sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    geo, bc, _ = bench.config2_problem(2048, 2048)
    rng = np.random.default_rng(0)
    phi = np.zeros((2050, 2050))
    phi[1:-1, 1:-1] = rng.standard_normal((2048, 2048))
    for homog in (1, 0):
        lap, _ = mg.compute_laplacian(geo, bc, homog, phi)
        _, xf, yf = mg.compute_grad(geo, bc, homog, phi)
        div = mg.compute_div(geo.nx, geo.ny, geo.dx, geo.dy, xf, yf)
        scale = np.max(np.abs(xf)) / geo.dx + np.max(np.abs(yf)) / geo.dy
        assert np.max(np.abs(div - lap)) <= 1e-14 * scale


@pytest.mark.parametrize("n", [(16, 12), (33, 17), (1000, 700)])
def test_reductions_and_extrapolation(n):
    """pnorm p = 0..4, inner_prod, integrate2d against the oracle (sequential
    sums; the GPU's are parallel and compensated: 1e-13 relative, the max norm
    exact); extrapolate_ghosts orders 0..3 bit for bit."""
    nx, ny = n
    dx, dy = 1.0 / nx, 1.0 / ny
    rng = np.random.default_rng(nx)
    a, b, J = rng.standard_normal((ny, nx)), rng.standard_normal((ny, nx)), 1 + rng.random((ny, nx))
    o = rc.oracle()
    assert mg.pnorm(nx, ny, dx, dy, a, 0) == o.pnorm(nx, ny, dx, dy, a, 0)
    for p in (1, 2, 3, 4):
        np.testing.assert_allclose(mg.pnorm(nx, ny, dx, dy, a, p), o.pnorm(nx, ny, dx, dy, a, p), rtol=1e-13)
    np.testing.assert_allclose(mg.inner_prod(nx, ny, dx, dy, a, b), o.inner_prod(nx, ny, dx, dy, a, b),
                               rtol=1e-12, atol=1e-14 * np.sqrt(nx * ny) * dx * dy)
    np.testing.assert_allclose(mg.integrate2d(nx, ny, dx, dy, a, J), o.integrate2d(nx, ny, dx, dy, a, J),
                               rtol=1e-12, atol=1e-14 * np.sqrt(nx * ny) * dx * dy)
    np.testing.assert_allclose(mg.integrate2d(nx, ny, dx, dy, a), o.integrate2d(nx, ny, dx, dy, a),
                               rtol=1e-12, atol=1e-14 * np.sqrt(nx * ny) * dx * dy)
    phi = rng.standard_normal((ny + 2, nx + 2))
    for order in range(4):
        same(mg.extrapolate_ghosts(nx, ny, order, phi), o.extrapolate_ghosts(nx, ny, order, phi))


@pytest.mark.parametrize("case", OPERATOR_CASES)
def test_relax_gs_lex_and_mc(case):
    geo, bc, phi, rhs = setup(*case, seed=2)
    inv = mg.compute_invdiags(geo, bc, 1)
    lex = mg.relax_gs(geo, bc, 1, inv, 1.0, 3, rhs, phi, mg.ORDER_LEX)
    for c in checkers():
        same(lex, c.relax_gs(geo, bc, 1, inv, 1.0, 3, rhs, phi))
    mc = mg.relax_gs(geo, bc, 1, inv, 2.0 / 3.0, 3, rhs, phi, mg.ORDER_MC)
    same(mc, rc.oracle().relax_gs(geo, bc, 1, inv, 2.0 / 3.0, 3, rhs, phi, rc.MC))


# perimeters long enough (nx + ny >= 512) for the multi-workgroup ghost fill and ring sweep
WIDE_CASES = [(300, 240, 0.10, 3.0, (1, 1, 0, 0)), (288, 256, 0.08, 1.0, (0, 1, 1, 0))]


@pytest.mark.parametrize("case", WIDE_CASES)
def test_wide_fill_ghosts_and_mc_ring(case):
    geo, bc, phi, rhs = setup(*case, seed=21)
    o = rc.oracle()
    for homog in (1, 0):
        for do_neum in (1, 0):
            same(mg.fill_ghosts(geo, bc, homog, do_neum, phi), o.fill_ghosts(geo, bc, homog, do_neum, phi))
    inv = mg.compute_invdiags(geo, bc, 1)
    mc = mg.relax_gs(geo, bc, 1, inv, 2.0 / 3.0, 2, rhs, phi, mg.ORDER_MC)
    same(mc, o.relax_gs(geo, bc, 1, inv, 2.0 / 3.0, 2, rhs, phi, rc.MC))
    lap, phig = mg.compute_laplacian(geo, bc, 1, phi)
    l2, p2 = o.laplacian(geo, bc, 1, phi)
    same(lap, l2)
    same(phig, p2)


@pytest.mark.parametrize("case", OPERATOR_CASES)
def test_dc_operator_and_smoothers(case):
    from mgpoisson2d_amd import geometry as G
    geo, bc, phi, rhs = setup(*case, seed=3)
    lah = G.BoundaryConditions.uniform(case[4], mode=G.BCMODE_LAH)
    lap, p1 = mg.dc_compute_laplacian(geo, lah, 1, phi)
    inv = mg.dc_compute_invdiags(geo)
    lex = mg.dc_relax_gs(geo, lah, 1, inv, 0.8, 3, rhs, phi, mg.ORDER_LEX)
    rb = mg.dc_relax_gs(geo, lah, 1, inv, 0.8, 3, rhs, phi, mg.ORDER_RB)
    o = rc.oracle()
    l2, p2 = o.dc_laplacian(geo, lah, 1, phi)
    same(lap, l2)
    same(p1, p2)
    same(inv, o.dc_invdiags(geo))
    same(lex, o.dc_relax_gs(geo, lah, 1, inv, 0.8, 3, rhs, phi, rc.LEX))
    same(rb, o.dc_relax_gs(geo, lah, 1, inv, 0.8, 3, rhs, phi, rc.RB))
    if rc.reference_available(rb=True):
        same(lex, rc.reference().dc_relax_gs(geo, lah, 1, inv, 0.8, 3, rhs, phi))
        same(rb, rc.reference(rb=True).dc_relax_gs(geo, lah, 1, inv, 0.8, 3, rhs, phi))


@pytest.mark.parametrize("case", OPERATOR_CASES[:4])
@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("order", [0, 1, 2])
def test_restrict_prolong(case, mode, order):
    geo, bc, phi, rhs = setup(*case, seed=4)
    o = rc.oracle()
    cgeo = coarse_geo(mg.restrict, geo)
    cgeo_o = coarse_geo(o.restrict, geo)
    for a, b in zip(cgeo.arrays(), cgeo_o.arrays()):
        same(a, b)
    rng = np.random.default_rng(5)
    fine = rng.standard_normal((geo.ny + 2, geo.nx + 2))
    crse = rng.standard_normal((cgeo.ny + 2, cgeo.nx + 2))
    f1, c1 = mg.prolong(geo, cgeo, case[4], (mode,) * 4, order, fine, crse)
    f2, c2 = o.prolong(geo, cgeo, case[4], (mode,) * 4, order, fine, crse)
    if all(t == 0 for t in case[4]):
        # mean removal: parallel compensated sum vs the reference's sequential sum
        np.testing.assert_allclose(f1, f2, rtol=0, atol=1e-13 * np.max(np.abs(f2[1:-1, 1:-1])))
    else:
        same(f1, f2)
    same(c1, c2)


def test_schedule():
    o = rc.oracle()
    for nx, ny in [(64, 64), (256, 64), (8192, 2048), (96, 40), (2048, 2048), (12, 7)]:
        assert mg.create_mgschedule(nx, ny, 1.0 / nx, 1.0 / ny) == o.schedule(nx, ny, 1.0 / nx, 1.0 / ny)


@pytest.mark.parametrize("case", CYCLE_CASES)
@pytest.mark.parametrize("kind", [0, 1])
def test_bicgstab2_bottom(case, kind):
    from mgpoisson2d_amd import geometry as G
    nx = ny = 8
    geo, bc, phi, rhs = setup(nx, ny, *case[2:], seed=6)
    if kind == 1:
        bc = G.BoundaryConditions.uniform(case[4], mode=G.BCMODE_LAH)
        inv = mg.dc_compute_invdiags(geo)
    else:
        inv = mg.compute_invdiags(geo, bc, 1)
    p1 = mg.solve_bicgstab2(kind, geo, bc, 1, inv, 1e-6, 80, 5, 1, rhs, phi)
    p2 = rc.oracle().bicgstab2(kind, geo, bc, 1, inv, 1e-6, 80, 5, 1, rhs, phi)
    same(p1, p2)


@pytest.mark.parametrize("homog", [1, 0])
@pytest.mark.parametrize("case", OPERATOR_CASES + CYCLE_CASES[:2])
def test_relax_jacobi(case, homog):
    """Reference reduction order (LEX): phi and the relres history bit for bit;
    parallel reductions (MC): phi bit for bit, history to 1e-13."""
    geo, bc, phi, rhs = setup(*case, homog=homog, seed=22)
    o = rc.oracle()
    inv = o.invdiags(geo, bc, homog)
    for omega, tol, its, zero in ((0.8, -1.0, 5, 0), (0.6, 1e-3, 40, 1), (1.0, 0.5, 10, 0)):
        want, hist = o.relax_jacobi(geo, bc, homog, inv, omega, tol, its, zero, rhs, phi)
        got, ghist, n = mg.relax_jacobi(geo, bc, homog, inv, omega, tol, its, zero, rhs, phi, mg.ORDER_LEX)
        same(got, want)
        if tol > 0:
            same(ghist, hist)
        got, ghist, n = mg.relax_jacobi(geo, bc, homog, inv, omega, tol, its, zero, rhs, phi, mg.ORDER_MC)
        same(got, want)
        if tol > 0:
            np.testing.assert_allclose(ghist, hist, rtol=1e-13, atol=0)
        if rc.reference_available():
            same(want, rc.reference().relax_jacobi(geo, bc, homog, inv, omega, tol, its, zero, rhs, phi)[0])


@pytest.mark.parametrize("homog", [1, 0])
@pytest.mark.parametrize("case", OPERATOR_CASES + CYCLE_CASES[:2])
def test_solve_bicgstab_v1(case, homog):
    """LEX: the reference's preconditioner and reduction order, bit for bit
    (restarts and the divergence undo included).  MC: multi-colour sweeps and
    parallel reductions against the oracle's MC run, 1e-9 over 6 iterations."""
    geo, bc, phi, rhs = setup(*case, homog=homog, seed=23)
    o = rc.oracle()
    inv = o.invdiags(geo, bc, homog)
    for tol, its, restarts, zero in ((1e-8, 30, 3, 1), (1e-3, 10, 0, 0), (1e-14, 60, 5, 1)):
        want, hist = o.bicgstab1(geo, bc, homog, inv, tol, its, restarts, zero, rhs, phi)
        got, ghist = mg.solve_bicgstab(geo, bc, homog, inv, tol, its, restarts, zero, rhs, phi, mg.ORDER_LEX)
        same(got, want)
        same(ghist, hist)
        for c in checkers()[1:]:
            same(got, c.bicgstab1(geo, bc, homog, inv, tol, its, restarts, zero, rhs, phi)[0])
    want, hist = o.bicgstab1(geo, bc, homog, inv, 1e-30, 6, 0, 1, rhs, phi, order=rc.MC)
    got, ghist = mg.solve_bicgstab(geo, bc, homog, inv, 1e-30, 6, 0, 1, rhs, phi, mg.ORDER_MC)
    assert len(ghist) == len(hist)
    np.testing.assert_allclose(ghist, hist, rtol=1e-9, atol=0)
    scale = np.max(np.abs(want))
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-9 * scale)


@pytest.mark.parametrize("n", [(128, 96), (1024, 768)])
def test_bicgstab2_grid_wide(n):
    """solve_bicgstab2 with the multi-colour preconditioner on grid-wide
    kernels (the full-size path) against the oracle's MC run, 4 iterations,
    1e-10 of the solution's scale."""
    geo, bc, phi, rhs = setup(n[0], n[1], 0.1, 3.0, (0, 0, 0, 0), seed=24)
    o = rc.oracle()
    inv = o.invdiags(geo, bc, 1)
    want = o.bicgstab2(0, geo, bc, 1, inv, 1e-30, 4, 0, 1, rhs, phi, order=rc.MC)
    got = mg.solve_bicgstab2(0, geo, bc, 1, inv, 1e-30, 4, 0, 1, rhs, phi, ordering=mg.ORDER_MC)
    scale = np.max(np.abs(want))
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-10 * scale)


@pytest.mark.parametrize("case", CYCLE_CASES)
@pytest.mark.parametrize("kind,ordering", [(0, 0), (0, 2), (1, 0), (1, 1), (2, 0), (2, 2)])
def test_vcycle_history(case, kind, ordering):
    geo, bc, phi, rhs = setup(*case, seed=7)
    maxit = 6
    o = rc.oracle()
    g = mg.vcycle(kind, geo, bc, 1, 0.0, maxit, phi, rhs, zerophi=True, ordering=ordering)
    c = o.vcycle(kind, geo, bc, 1, 0.0, maxit, phi, rhs, zerophi=True, order=ordering)
    assert g[3] == c[3] and g[4] == c[4]
    np.testing.assert_allclose(g[2], c[2], rtol=1e-12, atol=0)
    scale = np.max(np.abs(c[0][1:-1, 1:-1]))
    np.testing.assert_allclose(g[0][1:-1, 1:-1], c[0][1:-1, 1:-1], rtol=0, atol=1e-12 * scale)
    np.testing.assert_array_equal(g[1], c[1])
    if ordering == 0 and rc.reference_available():
        r = rc.reference().vcycle(kind, geo, bc, 1, 0.0, maxit, phi, rhs, zerophi=True)
        np.testing.assert_allclose(g[0][1:-1, 1:-1], r[0][1:-1, 1:-1], rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("nx,ny", [(8, 8), (20, 7), (300, 260)])
def test_integrate2d_bdry_and_gpnorm(nx, ny):
    rng = np.random.default_rng(nx + ny)
    xf, yf = rng.standard_normal((ny, nx + 1)), rng.standard_normal((ny + 1, nx))
    dx, dy = 1.0 / nx, 0.7 / ny
    got = mg.integrate2d_bdry(nx, ny, dx, dy, xf, yf)
    for c in checkers():
        assert got == c.integrate2d_bdry(nx, ny, dx, dy, xf, yf)
    if nx == ny:
        phi = rng.standard_normal((ny + 2, nx + 2))
        for p in (0, 2, 3):
            arg = np.abs(phi) if p == 3 else phi
            got = mg.gpnorm(0, nx - 1, 0, ny - 1, dx, dy, arg, p)
            for c in checkers():
                assert got == c.gpnorm(nx, ny, dx, dy, arg, p)
    else:
        # the reference's index pattern leaves the array on non-square boxes
        with pytest.raises(mg.MgpError):
            mg.gpnorm(0, nx - 1, 0, ny - 1, dx, dy, rng.standard_normal((ny + 2, nx + 2)), 0)
