"""GPU tests of the pieces of the mixed-precision, fourth-order all-Neumann
solve (8192 x 2048, beta = 8, y-lines), at sizes the CPU checkers finish in
seconds:

* the double-double residual equals the exact residual (tests/exactop.py,
  pinned to the oracle) rounded once, to within an ulp;
* the fourth-order operator equals the oracle's bit for bit on the oracle's
  metric arrays, and the GPU-generated metric set equals the oracle's to
  rounding;
* y-line relaxation (the transposed hierarchy) solves what x-line relaxation
  solves, and rows of 8192 cells are relaxed whole in the FP32 cycle and in
  segments in the FP64 one;
* double-double refinement takes a strongly stretched all-Neumann problem to
  a relative residual below what its FP64 operator can resolve, and the
  solution agrees with a sparse direct solve;
* the fourth-order deferred-correction solve reaches the direct solution of
  the oracle's fourth-order operator, and its error against the manufactured
  solution falls at fourth order.
"""
import math

import numpy as np
import pytest

import exactop
import mms4
import refcall as rc
from cases import probed_matrix, setup

pytestmark = pytest.mark.gpu

mg = pytest.importorskip("mgpoisson2d_amd.capi")
from mgpoisson2d_amd import geometry as G  # noqa: E402


@pytest.mark.parametrize("nx,ny,amp,beta,types", [(10, 9, 0.1, 3.0, (1, 1, 0, 0)), (9, 11, 0.05, 1.0, (1, 1, 1, 1)),
                                                  (12, 8, 0.12, 5.0, (0, 0, 0, 0)), (11, 10, 0.08, 2.0, (0, 1, 1, 0))])
def test_residual_dd_is_the_exact_residual(nx, ny, amp, beta, types):
    geo, bc, phi, rhs = setup(nx, ny, amp, beta, types, homog=True, seed=nx + ny)
    lo = 1e-17 * np.random.default_rng(9).standard_normal(phi.shape) * np.abs(phi)
    got = mg.compute_residual_dd(geo, bc, rhs, phi, lo)
    want = exactop.residual_exact(geo, types, rhs, phi, lo)
    ulp = np.spacing(np.abs(want))
    assert np.all(np.abs(got - want) <= ulp), np.max(np.abs(got - want) / ulp)
    assert np.mean(got == want) > 0.9


@pytest.mark.parametrize("nx,ny", [(5, 5), (13, 11), (24, 20)])
def test_lap4_is_the_oracle_operator(nx, ny):
    o = rc.oracle()
    m = o.metric4(nx, ny, 1.0 / nx, 1.0 / ny, 0.12, 4.0)
    met = [m[k] for k in ("A4", "Bx4", "dA", "dBx", "C4", "By4", "dC", "dBy")]
    phi = np.random.default_rng(nx).standard_normal((ny, nx))
    np.testing.assert_array_equal(mg.lap4(nx, ny, 1.0 / nx, 1.0 / ny, met, phi),
                                  o.lap4(nx, ny, 1.0 / nx, 1.0 / ny, m, phi))


def test_metric4_fill_matches_the_oracle():
    nx, ny, amp, beta = 40, 24, 0.12, 8.0
    m = rc.oracle().metric4(nx, ny, 1.0 / nx, 1.0 / ny, amp, beta)
    got = mg.metric4_fill(mg.Mapping(mg.MAP_WARPED, amp, beta, 0.0, 0.0), nx, ny, 1.0 / nx, 1.0 / ny)
    for k, name in enumerate(("A4", "Bx4", "dA", "dBx", "C4", "By4", "dC", "dBy")):
        want = m[name]
        # last-bit differences of sin, cos and exp; the d* arrays are differences of
        # neighbouring values, so their error is relative to the metric, not to them
        base = np.max(np.abs(m["A4" if k < 4 else "C4"])) * (1.0 if name in ("A4", "Bx4", "C4", "By4") else
                                                          (1.0 / ny if k < 4 else 1.0 / nx) / 12)
        assert np.max(np.abs(got[k] - want)) <= 1e-14 * max(base, np.max(np.abs(want))), name


def mapped(nx, ny, amp, beta):
    from types import SimpleNamespace
    return SimpleNamespace(nx=nx, ny=ny, dx=1.0 / nx, dy=1.0 / ny), mg.Mapping(mg.MAP_WARPED, amp, beta, 0.0, 0.0)


@pytest.mark.parametrize("method", ["KRYLOV_MG", "KRYLOV_MG32"])
def test_y_lines_solve_what_x_lines_solve(method):
    nx, ny, amp, beta, types = 96, 64, 0.1, 5.0, (1, 1, 0, 0)
    geo = G.mapped_geometry(nx, ny, amp, beta)
    bc = G.BoundaryConditions.uniform(types)
    rhs = np.random.default_rng(1).standard_normal((ny, nx)) * geo.J
    sols = []
    for lines in (mg.LINES_X, mg.LINES_Y):
        s = mg.Solver(geo, bc, method=getattr(mg, "METHOD_" + method), lines=lines)
        s.set_rhs(rhs)
        hist, iters, oc = s.solve(1e-10, 300)
        assert oc == mg.CONVERGED, (lines, iters, hist[-3:])
        phi = s.get_phi()
        r, _ = rc.oracle().residual(geo, bc, 1, rhs, phi)
        assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(rhs)
        sols.append(phi[1:-1, 1:-1])
        s.close()
    assert np.max(np.abs(sols[0] - sols[1])) <= 1e-7 * np.max(np.abs(sols[0]))


def test_rows_of_8192_cells():
    """Rows past one 256-thread workgroup (the 8192-cell rows of the
    fourth-order workload): the FP32 cycle solves them whole (512 threads,
    AUTO keeps x-lines), the FP64 cycle with x-lines in two segments of 4096.
    Both reach the reference operator's solution; on a grid whose x couplings
    dominate (dx = 1/8192, dy = 1/16), where the x-line sweep is nearly a
    direct solve when rows are solved whole, the whole-row cycle needs fewer
    iterations than the segmented one."""
    nx, ny, types = 8192, 16, (1, 1, 0, 0)
    geo = G.mapped_geometry(nx, ny, 0.0, 0.0)
    bc = G.BoundaryConditions.uniform(types)
    rhs = np.random.default_rng(5).standard_normal((ny, nx)) * geo.J
    sols, its = [], []
    for method, lines in ((mg.METHOD_KRYLOV_MG32, mg.LINES_AUTO), (mg.METHOD_KRYLOV_MG, mg.LINES_X)):
        s = mg.Solver(geo, bc, method=method, lines=lines)
        s.set_rhs(rhs)
        hist, iters, oc = s.solve(1e-10, 400)
        assert oc == mg.CONVERGED, (method, iters, hist[-3:])
        phi = s.get_phi()
        r, _ = rc.oracle().residual(geo, bc, 1, rhs, phi)
        assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(rhs)
        sols.append(phi[1:-1, 1:-1])
        its.append(s.last_cycles())
        s.close()
    print("V-cycles: whole FP32 rows", its[0], "FP64 segments", its[1])
    assert its[0] < its[1]
    assert np.max(np.abs(sols[0] - sols[1])) <= 1e-7 * np.max(np.abs(sols[1]))
    # the mapped family with stretching: both forms converge
    geo = G.mapped_geometry(nx, 32, 0.1, 3.0)
    rhs = np.random.default_rng(6).standard_normal((32, nx)) * geo.J
    for method in (mg.METHOD_KRYLOV_MG32, mg.METHOD_KRYLOV_MG):
        s = mg.Solver(geo, bc, method=method, lines=mg.LINES_X)
        s.set_rhs(rhs)
        hist, iters, oc = s.solve(1e-10, 400)
        assert oc == mg.CONVERGED, (method, iters, hist[-3:])
        r, _ = rc.oracle().residual(geo, bc, 1, rhs, s.get_phi())
        assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(rhs)
        s.close()


@pytest.mark.parametrize("method", ["KRYLOV_MG", "KRYLOV_MG32"])
def test_fgmres_solves_what_bicgstab_solves(method):
    """The flexible GMRES outer iteration reaches the reference operator's
    solution, as BiCGStab does, and its residual history never rises."""
    nx, ny, amp, beta, types = 128, 96, 0.1, 4.0, (1, 1, 0, 0)
    geo = G.mapped_geometry(nx, ny, amp, beta)
    bc = G.BoundaryConditions.uniform(types)
    rhs = np.random.default_rng(11).standard_normal((ny, nx)) * geo.J
    sols = []
    for krylov, restart in ((mg.KRYLOV_BICGSTAB, 0), (mg.KRYLOV_FGMRES, 0), (mg.KRYLOV_FGMRES, 7)):
        s = mg.Solver(geo, bc, method=getattr(mg, "METHOD_" + method), krylov=krylov, restart=restart)
        s.set_rhs(rhs)
        hist, iters, oc = s.solve(1e-10, 400)
        assert oc == mg.CONVERGED, (krylov, restart, iters, hist[-3:])
        if krylov == mg.KRYLOV_FGMRES:
            assert np.all(np.diff(hist) <= 1e-12), hist
            assert s.last_cycles() == iters
        phi = s.get_phi()
        r, _ = rc.oracle().residual(geo, bc, 1, rhs, phi)
        assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(rhs)
        rn, rh = s.residual()
        # the solver's own residual norm (pnorm 2) is the oracle's
        want = math.sqrt(geo.dx * geo.dy * np.sum(r * r))
        assert abs(rn - want) <= 1e-6 * want + 1e-300
        assert abs(rh - math.sqrt(geo.dx * geo.dy * np.sum(rhs * rhs))) <= 1e-12 * rh
        sols.append(phi[1:-1, 1:-1])
        s.close()
    for other in sols[1:]:
        assert np.max(np.abs(other - sols[0])) <= 1e-7 * np.max(np.abs(sols[0]))


def test_fourth_order_from_the_second_order_solution():
    """Deferred correction: the second-order solution (double-double, FGMRES)
    copied into a fourth-order solver (device copy) leaves a fourth-order
    residual of the size of the truncation-error difference; the fourth-order
    solve from there reaches a target relative to ||rhs|| set through
    residual(), and the solution is the direct solution of the oracle's
    fourth-order operator."""
    nx, ny, amp, beta = 48, 40, 0.12, 4.0
    geo, mapping = mapped(nx, ny, amp, beta)
    bc = G.BoundaryConditions.uniform((0, 0, 0, 0))
    rhs = mms4.rhs_mms(nx, ny, amp, beta)
    s2 = mg.Solver(geo, bc, method=mg.METHOD_KRYLOV_MG32, mapping=mapping, refine=mg.REFINE_DD,
                   krylov=mg.KRYLOV_FGMRES)
    s2.set_rhs(rhs)
    s2.project_mean()
    hist, iters, oc = s2.solve(1e-12, 400)
    assert oc == mg.CONVERGED, (iters, hist[-3:])
    s4 = mg.Solver(geo, bc, method=mg.METHOD_KRYLOV_MG32, mapping=mapping, order=4)
    s4.set_rhs(rhs)
    s4.project_mean()
    s4.copy_phi_from(s2)
    np.testing.assert_array_equal(s4.get_phi()[1:-1, 1:-1], s2.get_phi()[1:-1, 1:-1])
    rn, rh = s4.residual()
    assert 1e-8 * rh < rn < 1e-1 * rh, (rn, rh)
    hist, iters, oc = s4.solve(1e-11 * rh / rn, 400, zerophi=False)
    assert oc == mg.CONVERGED, (iters, hist[-3:])
    rn2, _ = s4.residual()
    assert rn2 <= 1.01e-11 * rh
    phi = s4.get_phi()[1:-1, 1:-1]
    m = rc.oracle().metric4(nx, ny, 1.0 / nx, 1.0 / ny, amp, beta)
    A = mms4.probed_lap4(nx, ny, m)
    ref = mms4.solve_neumann(A, rhs - rhs.mean())
    d = (phi - phi.mean()) - (ref - ref.mean())
    assert np.max(np.abs(d)) <= 1e-8 * np.max(np.abs(ref))
    s2.close()
    s4.close()


def test_second_order_solve_of_the_8192_workload_converges():
    """The 8192 x 2048 all-Neumann workload (beta = 8, warp 0.12) with its
    cell-average rhs, mixed precision, double-double refinement, krylov AUTO:
    BiCGStab breaks down on this rhs and FGMRES carries on from its last good
    iterate; the solve reaches 1e-12."""
    import bench
    nx, ny = 8192, 2048
    geo, mapping = mapped(nx, ny, bench.AMP3, bench.BETA3)
    bc = G.BoundaryConditions.uniform((0, 0, 0, 0))
    rhs = bench.config3_rhs(nx, ny, bench.AMP3, bench.BETA3, bench.SEED + 3, averages=True)
    s = mg.Solver(geo, bc, method=mg.METHOD_KRYLOV_MG32, mapping=mapping, refine=mg.REFINE_DD)
    s.set_rhs(rhs)
    s.project_mean()
    hist, iters, oc = s.solve(1e-12, 1000)
    print("iterations", iters, "V-cycles", s.last_cycles(), "final", hist[-1])
    assert oc == mg.CONVERGED, (iters, hist[-5:])
    assert hist[-1] <= 1e-12
    s.close()


def neumann_direct(A, b):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    n = A.shape[0]
    ones = np.ones((n, 1))
    Ab = sp.bmat([[A, ones], [ones.T, None]]).tocsc()
    return spla.spsolve(Ab, np.concatenate([b.ravel(), [0.0]]))[:n].reshape(b.shape)


def jmean_free(phi, J):
    return phi - np.sum(J * phi) / np.sum(J)


@pytest.mark.parametrize("method", ["KRYLOV_MG", "KRYLOV_MG32"])
def test_dd_refinement_below_the_fp64_floor(method):
    # the configuration's family at 512 x 128: beta = 8, warp 0.12, all Neumann; rows of 512
    # cells take x-lines (AUTO; the y-line orientation is checked above)
    nx, ny, amp, beta = 512, 128, 0.12, 8.0
    geo = G.mapped_geometry(nx, ny, amp, beta)
    bc = G.BoundaryConditions.uniform((0, 0, 0, 0))
    x, y = G.cell_centres(geo, amp, beta)
    f = 5 * math.pi ** 2 * np.cos(math.pi * x) * np.cos(2 * math.pi * y)
    rhs = -geo.J * (f + 1e-2 * np.random.default_rng(3).standard_normal((ny, nx)))
    s = mg.Solver(geo, bc, method=getattr(mg, "METHOD_" + method), refine=mg.REFINE_DD)
    s.set_rhs(rhs)
    m = s.project_mean()
    rhs_p = rhs - geo.J * m
    assert abs(np.sum(rhs_p)) <= 1e-12 * np.sum(np.abs(rhs))
    hist, iters, oc = s.solve(1e-12, 400)
    assert oc == mg.CONVERGED, (iters, hist[-4:])
    hi, lo = s.get_phi(), s.get_phi_lo()
    # the double-double residual (pinned above) of hi + lo meets the tolerance ...
    r = mg.compute_residual_dd(geo, bc, rhs_p, hi, lo)
    rel = np.linalg.norm(r) / np.linalg.norm(rhs_p)
    assert rel <= 1.5e-12, rel
    # ... which the FP64 evaluation of the same operator cannot resolve here
    r64, _ = rc.oracle().residual(geo, bc, 1, rhs_p, hi + lo)
    print(f"{method}: dd relres {rel:.2e}, FP64-evaluated relres {np.linalg.norm(r64) / np.linalg.norm(rhs_p):.2e}")
    assert np.linalg.norm(r64) / np.linalg.norm(rhs_p) > 3 * rel
    # the solution is the direct solution of the operator (up to the constant)
    A = probed_matrix(geo, bc).tocsr()
    ref = neumann_direct(A, rhs_p)
    sol = jmean_free((hi + lo)[1:-1, 1:-1], geo.J)
    ref = jmean_free(ref, geo.J)
    assert np.max(np.abs(sol - ref)) <= 1e-6 * np.max(np.abs(ref))
    s.close()


def fourth_order_solve(nx, ny, amp, beta, method, tol=1e-11):
    geo, mapping = mapped(nx, ny, amp, beta)
    bc = G.BoundaryConditions.uniform((0, 0, 0, 0))
    rhs = mms4.rhs_mms(nx, ny, amp, beta)
    s = mg.Solver(geo, bc, method=getattr(mg, "METHOD_" + method), mapping=mapping, order=4)
    s.set_rhs(rhs)
    s.project_mean()
    hist, iters, oc = s.solve(tol, 400)
    assert oc == mg.CONVERGED, (nx, ny, iters, hist[-4:])
    phi = s.get_phi()[1:-1, 1:-1]
    s.close()
    return phi, rhs, iters


@pytest.mark.parametrize("method", ["KRYLOV_MG", "KRYLOV_MG32"])
def test_fourth_order_solve_is_the_direct_solution(method):
    nx, ny, amp, beta = 48, 40, 0.12, 4.0
    phi, rhs, _ = fourth_order_solve(nx, ny, amp, beta, method)
    m = rc.oracle().metric4(nx, ny, 1.0 / nx, 1.0 / ny, amp, beta)
    A = mms4.probed_lap4(nx, ny, m)
    ref = mms4.solve_neumann(A, rhs - rhs.mean())
    d = (phi - phi.mean()) - (ref - ref.mean())
    assert np.max(np.abs(d)) <= 1e-8 * np.max(np.abs(ref))


def test_fourth_order_rate():
    amp, beta = 0.1, 3.0
    errs = []
    for n in (32, 64, 128):
        phi, _, _ = fourth_order_solve(n, n, amp, beta, "KRYLOV_MG32")
        errs.append(mms4.error_vs_u(phi, n, n, amp, beta)[0])
    rates = [math.log2(errs[k] / errs[k + 1]) for k in range(2)]
    print("fourth-order max errors", errs, "rates", rates)
    assert min(rates) >= 3.5, (errs, rates)
