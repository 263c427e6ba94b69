"""The resident solvers on the GPU.

* KRYLOV_MG (FP64 BiCGStab + semicoarsening line-relaxation V-cycles) must
  reach the requested relative residual of the reference operator, checked
  independently with the oracle's operator (bit-identical to the reference),
  and its solution must equal a sparse direct solve of that operator to
  1e-10 relative (closed check: scipy spsolve on the probed matrix).
  KRYLOV_MG32 (the same with the preconditioner V-cycle in FP32) must meet
  the same bars: the preconditioner's precision changes the iteration
  count, not the FP64 accuracy of the solution.
* The reference drivers on resident data (VCYCLE, DC_VCYCLE) reproduce the
  oracle's relres history.
* Alone-vs-again determinism: solving twice gives identical bits.
"""
import numpy as np
import pytest

import refcall as rc
from cases import probed_matrix, setup

pytestmark = pytest.mark.gpu

mg = pytest.importorskip("mgpoisson2d_amd.capi")


def relres_oracle(geo, bc, rhs, phi):
    o = rc.oracle()
    r, _ = o.residual(geo, bc, 1, rhs, phi)
    r0, _ = o.residual(geo, bc, 1, rhs, np.zeros_like(phi))
    return np.sqrt(np.sum(r * r)) / np.sqrt(np.sum(r0 * r0))


@pytest.mark.parametrize("case", [
    (64, 64, 0.0, 0.0, (1, 1, 1, 1)),
    (128, 128, 0.10, 5.0, (1, 1, 0, 0)),
    (96, 64, 0.12, 3.0, (1, 1, 0, 0)),
    (64, 128, 0.05, 2.0, (1, 0, 1, 0)),
])
@pytest.mark.parametrize("method", [mg.METHOD_KRYLOV_MG, mg.METHOD_KRYLOV_MG32])
def test_krylov_mg_reaches_tolerance(case, method):
    geo, bc, _, rhs = setup(*case, seed=11)
    s = mg.Solver(geo, bc, method=method)
    s.set_rhs(rhs)
    hist, iters, outcome = s.solve(1e-10, 200, zerophi=True)
    assert outcome == mg.CONVERGED, (hist, iters)
    phi = s.get_phi()
    assert relres_oracle(geo, bc, rhs, phi) <= 1.5e-10
    # determinism: a second solve is bit-identical
    s.solve(1e-10, 200, zerophi=True)
    np.testing.assert_array_equal(s.get_phi()[1:-1, 1:-1], phi[1:-1, 1:-1])


@pytest.mark.parametrize("shape", [(48, 45), (40, 90), (36, 4)])
def test_krylov_mg_odd_row_counts(shape):
    """Odd row counts (45 and 90 -> 45: the Neumann y sides let the coarsening
    drop a row and go on; 4: the smallest hierarchy) reach the tolerance."""
    nx, ny = shape
    geo, bc, _, rhs = setup(nx, ny, 0.1, 4.0, (1, 1, 0, 0), seed=17)
    s = mg.Solver(geo, bc, method=mg.METHOD_KRYLOV_MG)
    s.set_rhs(rhs)
    hist, iters, outcome = s.solve(1e-10, 400, zerophi=True)
    assert outcome == mg.CONVERGED, (hist[-3:], iters)
    assert relres_oracle(geo, bc, rhs, s.get_phi()) <= 1.5e-10


@pytest.mark.parametrize("method", [mg.METHOD_KRYLOV_MG, mg.METHOD_KRYLOV_MG32])
def test_krylov_mg_matches_direct_solve(method):
    import scipy.sparse.linalg as spl
    geo, bc, _, rhs = setup(48, 48, 0.1, 4.0, (1, 1, 0, 0), seed=12)
    A = probed_matrix(geo, bc)
    exact = spl.spsolve(A.tocsc(), rhs.ravel()).reshape(geo.ny, geo.nx)
    s = mg.Solver(geo, bc, method=method)
    s.set_rhs(rhs)
    _, _, outcome = s.solve(1e-12, 300, zerophi=True)
    assert outcome == mg.CONVERGED
    phi = s.get_phi()[1:-1, 1:-1]
    np.testing.assert_allclose(phi, exact, rtol=0, atol=1e-10 * np.max(np.abs(exact)))


@pytest.mark.parametrize("method", [mg.METHOD_KRYLOV_MG, mg.METHOD_KRYLOV_MG32])
def test_krylov_mg_all_neumann_compatible(method):
    geo, bc, _, rhs = setup(64, 64, 0.12, 3.0, (0, 0, 0, 0), seed=13)
    s = mg.Solver(geo, bc, method=method)
    s.set_rhs(rhs)
    hist, iters, outcome = s.solve(1e-10, 300, zerophi=True)
    assert outcome == mg.CONVERGED, (hist, iters)
    assert relres_oracle(geo, bc, rhs, s.get_phi()) <= 1.5e-10


# (solver method, oracle cycle kind, ordering)
@pytest.mark.parametrize("method,kind,ordering", [(0, 0, 0), (0, 0, 2), (1, 1, 0), (1, 1, 1), (3, 2, 0), (3, 2, 2)])
@pytest.mark.parametrize("types", [(1, 1, 1, 1), (0, 0, 0, 0)])
def test_resident_reference_drivers(method, kind, ordering, types):
    geo, bc, _, rhs = setup(64, 64, 0.0, 0.0, types, seed=14)
    s = mg.Solver(geo, bc, method=method, ordering=ordering)
    s.set_rhs(rhs)
    hist, iters, outcome = s.solve(1e-10, 30, zerophi=True)
    _, _, h2, it2, oc2 = rc.oracle().vcycle(kind, geo, bc, 1, 1e-10, 30, np.zeros((66, 66)), rhs,
                                            zerophi=True, order=ordering)
    assert (iters, outcome) == (it2, oc2)
    np.testing.assert_allclose(hist, h2, rtol=1e-12, atol=0)


@pytest.mark.parametrize("method", [mg.METHOD_KRYLOV_MG, mg.METHOD_KRYLOV_MG32])
def test_graph_replay_matches_eager(method):
    """The captured-graph iteration must give the bits of the eager one,
    replayed several times with eager work and syncs in between."""
    geo, bc, _, rhs = setup(128, 96, 0.1, 4.0, (1, 1, 0, 0), seed=15)
    eager = mg.Solver(geo, bc, method=method, use_graph=False)
    graph = mg.Solver(geo, bc, method=method, use_graph=True)
    eager.set_rhs(rhs)
    graph.set_rhs(rhs)
    h1, i1, o1 = eager.solve(1e-10, 200, zerophi=True)
    ref = eager.get_phi()
    for _ in range(3):
        h2, i2, o2 = graph.solve(1e-10, 200, zerophi=True)
        assert (i1, o1) == (i2, o2)
        np.testing.assert_array_equal(h1, h2)
        np.testing.assert_array_equal(graph.get_phi()[1:-1, 1:-1], ref[1:-1, 1:-1])
        eager.solve(1e-10, 5, zerophi=True)  # eager work between replays


def zebra_reference(geo, bc, rhs, phi, sweeps, nseg=1):
    """x-line zebra relaxation with scipy on the probed operator: per row j,
    T_j delta = (f - A e)_j with T_j the tridiagonal part of A's (j, j) block,
    rows in the order even interior, odd interior, then 0 and ny-1 (rows of
    one colour do not couple, so a sequential pass equals the parallel one).

    nseg > 1: each row is cut into nseg equal segments, each solved with the
    tridiagonal part of its own diagonal block (the couplings across a cut
    stay in the residual); per colour the even segments of every row go
    first, then the odd ones -- the two launches of the segmented kernel, in
    each of which no segment reads a cell that another one writes."""
    from scipy.linalg import solve_banded
    A = probed_matrix(geo, bc).tocsr()
    nx, ny = geo.nx, geo.ny
    L = nx // nseg
    e = phi[1:-1, 1:-1].copy().ravel()
    f = rhs.ravel()
    colours = [list(range(2, ny - 1, 2)), list(range(1, ny - 1, 2)) + [0, ny - 1]]
    blocks = {}
    for j in range(ny):
        for s in range(nseg):
            a0 = j * nx + s * L
            blk = A[a0:a0 + L, a0:a0 + L]
            ab = np.zeros((3, L))
            ab[0, 1:] = blk.diagonal(1)
            ab[1, :] = blk.diagonal(0)
            ab[2, :-1] = blk.diagonal(-1)
            blocks[j, s] = ab
    parities = [0, 1] if nseg > 1 else [0]
    for _ in range(sweeps):
        for rows in colours:
            for par in parities:
                for j in rows:
                    for s in range(par, nseg, 2 if nseg > 1 else 1):
                        a0 = j * nx + s * L
                        r = f[a0:a0 + L] - A[a0:a0 + L] @ e
                        e[a0:a0 + L] += solve_banded((1, 1), blocks[j, s], r)
    return e.reshape(ny, nx)


@pytest.mark.parametrize("case", [
    (16, 12, 0.0, 0.0, (1, 1, 1, 1)),
    (16, 16, 0.10, 3.0, (1, 1, 0, 0)),
    (20, 12, 0.12, 2.0, (0, 0, 0, 0)),
    (12, 16, 0.08, 1.0, (0, 1, 1, 0)),
    (33, 17, 0.05, 4.0, (1, 0, 0, 1)),
    (257, 6, 0.10, 3.0, (1, 1, 0, 0)),    # one thread owns both row ends
    (515, 5, 0.10, 3.0, (0, 1, 1, 0)),    # chunks of 4 and 5 cells
    (1100, 4, 0.10, 3.0, (1, 1, 0, 0)),   # chunks of 8 and 9 cells
    (2100, 4, 0.10, 3.0, (1, 0, 1, 0)),   # chunks of 16 and 17 cells
])
def test_xline_relax_matches_dense(case):
    geo, bc, phi, rhs = setup(*case, seed=21)
    got = mg.xline_relax(geo, bc, rhs, phi, 2)
    want = zebra_reference(geo, bc, rhs, phi, 2)
    scale = np.max(np.abs(want))
    np.testing.assert_allclose(got[1:-1, 1:-1], want, rtol=0, atol=1e-10 * scale)
    # ghost cells are left as they were
    np.testing.assert_array_equal(got[0, 1:-1], phi[0, 1:-1])


@pytest.mark.parametrize("case,nseg", [
    ((4112, 5, 0.10, 3.0, (1, 1, 0, 0)), 2),   # two segments of 2056 cells
    ((8192, 4, 0.10, 3.0, (0, 1, 1, 0)), 2),   # two of 4096 (the fourth-order workload's rows)
    ((12288, 6, 0.10, 3.0, (1, 0, 0, 1)), 4),  # four of 3072: two per launch
])
def test_xline_relax_segmented_rows(case, nseg):
    """Rows longer than 4111 cells are relaxed in segments (block
    Gauss-Seidel over the segments of a row, even ones first): the result
    matches the dense segmented reference.

    Tolerance: a segment's tridiagonal system has condition number about
    kappa = 2 (dy/dx)^2 here (x couplings 1/dx^2 against the 2/dy^2 the
    y couplings add to the diagonal: 8e6 for 8192 x 4), and both solves
    round at eps kappa relative; so 4 eps kappa of the scale (measured
    3.6e-9 relative at 8192 x 4, where 4 eps kappa = 7.5e-9)."""
    geo, bc, phi, rhs = setup(*case, seed=22)
    got = mg.xline_relax(geo, bc, rhs, phi, 2)
    want = zebra_reference(geo, bc, rhs, phi, 2, nseg=nseg)
    scale = np.max(np.abs(want))
    kappa = 2.0 * (geo.dy / geo.dx) ** 2
    np.testing.assert_allclose(got[1:-1, 1:-1], want, rtol=0, atol=max(1e-10, 4 * np.finfo(float).eps * kappa) * scale)
    # a plain whole-row relaxation differs: the cut is real
    assert np.max(np.abs(zebra_reference(geo, bc, rhs, phi, 2, nseg=1) - want)) > 1e-6 * scale


def projection_fluxes(geo):
    """Smooth face fluxes with zero net boundary flux (the compatibility the
    all-Neumann solve needs), corrected as the reference's projection test
    does: xflux -= (net flux) x (test.f90:2905-2913)."""
    nx, ny, dx, dy = geo.nx, geo.ny, geo.dx, geo.dy
    xn, yc = np.arange(nx + 1) * dx, (np.arange(ny) + 0.5) * dy
    xc, yn = (np.arange(nx) + 0.5) * dx, np.arange(ny + 1) * dy
    xf = np.sin(3 * xn)[None, :] * np.cos(2 * yc)[:, None] + 0.3 * np.cos(5 * yc)[:, None] * xn[None, :] ** 2
    yf = np.cos(4 * xc)[None, :] * np.sin(yn)[:, None] - 0.2 * yn[:, None] * np.sin(2 * xc)[None, :]
    net = np.sum(xf[:, -1] - xf[:, 0]) * dy + np.sum(yf[-1, :] - yf[0, :]) * dx
    xf = xf - net * xn[None, :]
    from mgpoisson2d_amd import geometry as G
    bc = G.BoundaryConditions((0, 0, 0, 0), (1, 1, 1, 1),
                              (xf[:, 0].copy(), xf[:, -1].copy(), yf[0, :].copy(), yf[-1, :].copy()))
    return xf, yf, bc


@pytest.mark.parametrize("amp,beta", [(0.0, 0.0), (0.1, 3.0)])
def test_projection_reference_workflow(amp, beta):
    """The reference's divergence cleaning (test.f90:2817-3024): all-Neumann
    BCs from the flux itself, fmg(tol 1e-30, 10 cycles, smoothing 8/8/8),
    F -= J grad phi.  Against the oracle composing compute_div, fmg,
    compute_grad: max|div| before is exact, the cleaned fluxes agree to 1e-12
    of their scale (the cycles' norms are summed in parallel)."""
    from mgpoisson2d_amd import geometry as G
    geo = G.mapped_geometry(64, 64, amp, beta)
    xf, yf, bc = projection_fluxes(geo)
    o = rc.oracle()
    div0 = o.compute_div(64, 64, geo.dx, geo.dy, xf, yf)
    phi, _, _, it, _ = o.vcycle(2, geo, bc, 1, 1e-30, 10, np.zeros((66, 66)), div0, numcycles=1, sdown=8, sup=8,
                                sbottom=8, zerophi=True)
    _, gx, gy = o.compute_grad(geo, bc, 1, phi)
    want_x, want_y = xf - gx, yf - gy
    div1 = o.compute_div(64, 64, geo.dx, geo.dy, want_x, want_y)

    s = mg.Solver(geo, bc, method=mg.METHOD_FMG, ordering=mg.ORDER_LEX, smooth_down=8, smooth_up=8,
                  smooth_bottom=8, homog=True)
    s.set_flux(xf, yf)
    d0, d1, iters, _ = s.project(1e-30, 10)
    assert iters == it
    assert d0 == np.max(np.abs(div0))
    gxf, gyf = s.get_flux()
    scale = max(np.max(np.abs(xf)), np.max(np.abs(yf)))
    np.testing.assert_allclose(gxf, want_x, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(gyf, want_y, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(d1, np.max(np.abs(div1)), rtol=1e-6, atol=1e-12 * d0)
    assert d1 < 1e-4 * d0


def test_projection_fast_solver():
    """KRYLOV_MG projection on a stretched grid: J div of the cleaned flux is
    the solve's residual, so its 2-norm relative to J div F is within the
    tolerance; recomputed on the host with the oracle's compute_div."""
    from mgpoisson2d_amd import geometry as G
    geo = G.mapped_geometry(256, 192, 0.1, 4.0)
    xf, yf, bc = projection_fluxes(geo)
    s = mg.Solver(geo, bc, method=mg.METHOD_KRYLOV_MG)
    s.set_flux(xf, yf)
    d0, d1, iters, outcome = s.project(1e-10, 200)
    assert outcome == mg.CONVERGED, iters
    o = rc.oracle()
    div0 = o.compute_div(geo.nx, geo.ny, geo.dx, geo.dy, xf, yf)
    gxf, gyf = s.get_flux()
    div1 = o.compute_div(geo.nx, geo.ny, geo.dx, geo.dy, gxf, gyf)
    assert d0 == np.max(np.abs(div0))
    assert d1 == np.max(np.abs(div1))
    assert np.linalg.norm(div1) <= 2e-10 * np.linalg.norm(div0)
    # Neumann faces keep the flux (the correction's boundary fluxes are zero)
    np.testing.assert_array_equal(gxf[:, [0, -1]], xf[:, [0, -1]])
    np.testing.assert_array_equal(gyf[[0, -1], :], yf[[0, -1], :])


@pytest.mark.parametrize("types", [(1, 1, 0, 0), (1, 1, 1, 0), (1, 1, 0, 1), (0, 0, 0, 0)])
@pytest.mark.parametrize("method", [mg.METHOD_KRYLOV_MG, mg.METHOD_KRYLOV_MG32])
def test_odd_row_counts_coarsen_past_a_neumann_side(types, method):
    """ny = 16 * 49: the semicoarsening reaches 49 rows and, with a Neumann y
    side, leaves that side's row out (the top row when both y sides are
    Neumann or only the top one is; the bottom row for (1, 1, 0, 1)) and goes
    on to 24, 12 and 6 rows.  The solve reaches the reference operator's
    residual in a bounded number of iterations at the strong stretching of
    the batched mix's slowest problems: 21 to 53 here, where a hierarchy that
    stops at 49 rows took 105 for (1, 1, 0, 0)."""
    nx, ny = 800, 784
    geo, bc, _, rhs = setup(nx, ny, 0.06, 5.9, types, seed=31)
    s = mg.Solver(geo, bc, method=method)
    s.set_rhs(rhs)
    hist, iters, oc = s.solve(1e-10, 400, zerophi=True)
    print(types, method, "iterations", iters)
    assert oc == mg.CONVERGED, (iters, hist[-3:])
    assert relres_oracle(geo, bc, rhs, s.get_phi()) <= 1.5e-10
    assert iters <= 60
    s.close()


def test_host_copies_larger_than_the_staging_buffer():
    """phi of a 3072 x 3072 grid with ghosts is 3074^2 doubles (75.6 MB): the
    host copies move it through the 64 MiB pinned staging buffer in two
    blocks of rows.  A round trip returns the same bits, and so does the
    one of a grid whose copy fits in one block."""
    from mgpoisson2d_amd import geometry as G
    for n in (3072, 512):
        geo = G.mapped_geometry(n, n, 0.0, 0.0)
        bc = G.BoundaryConditions.uniform((1, 1, 1, 1))
        s = mg.Solver(geo, bc, method=mg.METHOD_KRYLOV_MG)
        phi = np.random.default_rng(n).standard_normal((n + 2, n + 2))
        s.set_phi(phi)
        np.testing.assert_array_equal(s.get_phi(), phi)
        s.close()
