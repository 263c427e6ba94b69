"""The batched solver (mgp_batch) on the GPU.

* Both precisions of the preconditioner cycle (KRYLOV_MG, and KRYLOV_MG32
  with FP32 cycles) meet the same bars below.
* Every problem of a batch of uneven sizes, anisotropies and BCs reaches the
  requested relative residual of the reference operator, checked with the
  oracle's operator (bit-identical to the reference), and a small one equals
  a sparse direct solve of that operator to 1e-10 relative.
* Each result is the same alone and in the mix: a problem solved in a batch
  of one gives the bits (and the iteration count) it gives inside the batch.
* Graph replay and eager launches give the same bits, and so do the launch
  tables rebuilt over the problems still running once the finished ones'
  dispatches have cost about one rebuild.
* Metrics generated on the GPU (mgp_batch_create_mapped) give the solution
  of the single-problem solver built the same way, to the solve tolerance.
"""
from types import SimpleNamespace

import numpy as np
import pytest

import refcall as rc
from cases import probed_matrix, setup

pytestmark = pytest.mark.gpu

mg = pytest.importorskip("mgpoisson2d_amd.capi")
from mgpoisson2d_amd import geometry as G  # noqa: E402

# (nx, ny, amp, beta, bc types) -- uneven sizes, both BC mixes, stretched and not
CASES = [
    (64, 64, 0.10, 4.0, (1, 1, 0, 0)),
    (96, 32, 0.05, 0.0, (1, 1, 1, 1)),
    (48, 128, 0.12, 5.0, (1, 1, 0, 0)),
    (128, 96, 0.08, 2.0, (1, 1, 1, 1)),
    (32, 32, 0.00, 0.0, (1, 1, 0, 0)),
    (200, 16, 0.10, 3.0, (1, 1, 0, 0)),
    (40, 24, 0.15, 6.0, (1, 1, 1, 1)),
    (576, 32, 0.05, 1.0, (1, 1, 0, 0)),  # rows of 513-1024 cells (128 threads) in the same batch
    (48, 90, 0.10, 5.0, (1, 1, 1, 1)),   # 45 rows on the coarsest level (Dirichlet y sides stop there)
    (40, 45, 0.10, 3.0, (1, 1, 1, 1)),   # one level: its own sweep count (22)
    (48, 90, 0.10, 5.0, (1, 1, 0, 0)),   # Neumann y sides: 90 -> 45 -> 22 -> 11 -> 5 rows
]


def relres_oracle(geo, bc, rhs, phi_valid):
    o = rc.oracle()
    phi = np.zeros((geo.ny + 2, geo.nx + 2))
    phi[1:-1, 1:-1] = phi_valid
    r, _ = o.residual(geo, bc, 1, rhs, phi)
    r0, _ = o.residual(geo, bc, 1, rhs, np.zeros_like(phi))
    return np.sqrt(np.sum(r * r)) / np.sqrt(np.sum(r0 * r0))


def problems(cases, seed0=40):
    out = []
    for k, c in enumerate(cases):
        geo, bc, _, rhs = setup(*c, seed=seed0 + k)
        out.append((geo, bc, rhs))
    return out


def run_batch(probs, tol=1e-10, use_graph=True, method=None):
    b = mg.Batch([p[0] for p in probs], [p[1] for p in probs], use_graph=use_graph,
                 method=mg.METHOD_KRYLOV_MG if method is None else method)
    for k, (_, _, rhs) in enumerate(probs):
        b.set_rhs(k, rhs)
    iters, outc = b.solve(tol, 300)
    phis = [b.get_phi(k) for k in range(len(probs))]
    cycles = b.last_cycles()
    b.close()
    return phis, iters, outc, cycles


@pytest.mark.parametrize("method", [mg.METHOD_KRYLOV_MG, mg.METHOD_KRYLOV_MG32])
def test_batch_reaches_tolerance(method):
    probs = problems(CASES)
    phis, iters, outc, cycles = run_batch(probs, method=method)
    assert np.all(outc == mg.CONVERGED), (iters, outc)
    assert cycles == 2 * int(iters.max())
    for (geo, bc, rhs), phi in zip(probs, phis):
        assert relres_oracle(geo, bc, rhs, phi) <= 1.5e-10


def test_batch_matches_direct_solve():
    import scipy.sparse.linalg as spl
    probs = problems([(48, 48, 0.1, 4.0, (1, 1, 0, 0)), (40, 24, 0.12, 3.0, (1, 1, 1, 1))], seed0=60)
    phis, _, outc, _ = run_batch(probs, tol=1e-12)
    assert np.all(outc == mg.CONVERGED)
    for (geo, bc, rhs), phi in zip(probs, phis):
        A = probed_matrix(geo, bc)
        exact = spl.spsolve(A.tocsc(), rhs.ravel()).reshape(geo.ny, geo.nx)
        np.testing.assert_allclose(phi, exact, rtol=0, atol=1e-10 * np.max(np.abs(exact)))


@pytest.mark.parametrize("method", [mg.METHOD_KRYLOV_MG, mg.METHOD_KRYLOV_MG32])
def test_batch_alone_equals_in_mix(method):
    probs = problems(CASES)
    phis, iters, outc, _ = run_batch(probs, method=method)
    for k in (0, 2, 5, 7, 9, 10):
        alone, it1, oc1, _ = run_batch([probs[k]], method=method)
        assert (int(it1[0]), int(oc1[0])) == (int(iters[k]), int(outc[k]))
        np.testing.assert_array_equal(alone[0], phis[k])


@pytest.mark.parametrize("method", [mg.METHOD_KRYLOV_MG, mg.METHOD_KRYLOV_MG32])
def test_batch_graph_matches_eager(method):
    probs = problems(CASES[:4], seed0=70)
    g_phis, g_it, _, _ = run_batch(probs, use_graph=True, method=method)
    e_phis, e_it, _, _ = run_batch(probs, use_graph=False, method=method)
    np.testing.assert_array_equal(g_it, e_it)
    for a, b in zip(g_phis, e_phis):
        np.testing.assert_array_equal(a, b)


def test_batch_resolve_is_deterministic():
    probs = problems(CASES[:3], seed0=80)
    b = mg.Batch([p[0] for p in probs], [p[1] for p in probs])
    for k, (_, _, rhs) in enumerate(probs):
        b.set_rhs(k, rhs)
    it1, _ = b.solve(1e-10, 300)
    first = [b.get_phi(k) for k in range(3)]
    for _ in range(2):
        it2, _ = b.solve(1e-10, 300)
        np.testing.assert_array_equal(it1, it2)
        for k in range(3):
            np.testing.assert_array_equal(b.get_phi(k), first[k])
    b.close()


def test_batch_mapped_matches_single_solver():
    shapes = [(64, 48, 0.1, 5.0), (96, 64, 0.05, 2.0), (32, 128, 0.12, 0.0)]
    geos, maps, bcs, rhss = [], [], [], []
    rng = np.random.default_rng(90)
    for nx, ny, amp, beta in shapes:
        geos.append(SimpleNamespace(nx=nx, ny=ny, dx=1.0 / nx, dy=1.0 / ny))
        maps.append(mg.Mapping(mg.MAP_WARPED, amp, beta, 0.0, 0.0))
        bcs.append(G.BoundaryConditions.uniform((1, 1, 0, 0)))
        rhss.append(rng.standard_normal((ny, nx)))
    b = mg.Batch(geos, bcs, mappings=maps)
    for k, r in enumerate(rhss):
        b.set_rhs(k, r)
    _, outc = b.solve(1e-11, 300)
    assert np.all(outc == mg.CONVERGED)
    for k in range(len(shapes)):
        s = mg.Solver(geos[k], bcs[k], method=mg.METHOD_KRYLOV_MG, mapping=maps[k])
        s.set_rhs(rhss[k])
        _, _, oc = s.solve(1e-11, 300, zerophi=True)
        assert oc == mg.CONVERGED
        ref = s.get_phi()[1:-1, 1:-1]
        np.testing.assert_allclose(b.get_phi(k), ref, rtol=0, atol=1e-9 * np.max(np.abs(ref)))
        s.close()
    b.close()


def test_batch_rejects_bad_input():
    geo, bc, _, _ = setup(32, 32, 0.0, 0.0, (1, 1, 1, 1))
    periodic = G.BoundaryConditions.uniform((2, 2, 1, 1))
    with pytest.raises(mg.MgpError):
        mg.Batch([geo], [periodic])
    b = mg.Batch([geo], [bc])
    with pytest.raises(ValueError):
        b.set_rhs(0, np.zeros((16, 16)))
    rhs = np.zeros((32, 32))
    with pytest.raises(mg.MgpError):
        mg._check(mg.lib.mgp_batch_set_rhs(b._h, 5, rhs.ctypes.data_as(mg._dp)))
    b.close()


@pytest.mark.parametrize("method", [mg.METHOD_KRYLOV_MG, mg.METHOD_KRYLOV_MG32])
def test_rebuilt_tables_keep_the_bits(method):
    """240 easy problems (32 x 32, unstretched) finish within a few
    iterations while 4 stretched Neumann ones run on for 20-40.  The batch
    then charges 7 us per finished problem and iteration, 1680 us per
    iteration here, and rebuilds its tables over the 4 (at 4000 + 60 x 4 us)
    within three iterations, and again as they finish.  The hard problems'
    bits and iteration counts equal those of each solved alone (a batch of
    one never rebuilds), and graph replay equals eager launches."""
    hard = [(96, 368, 0.112, 5.99, (1, 1, 0, 0)), (48, 432, 0.134, 5.89, (1, 1, 0, 0)),
            (64, 208, 0.131, 5.71, (1, 1, 0, 0)), (160, 112, 0.12, 5.5, (1, 1, 0, 0))]
    cases = [(32, 32, 0.0, 0.0, (1, 1, 1, 1))] * 240 + hard
    probs = problems(cases, seed0=300)
    phis, iters, outc, _ = run_batch(probs, method=method)
    assert np.all(outc == mg.CONVERGED)
    assert int(np.max(iters[:240])) + 5 < int(np.min(iters[240:])), (iters[:240].max(), iters[240:])
    e_phis, e_iters, _, _ = run_batch(probs, use_graph=False, method=method)
    np.testing.assert_array_equal(e_iters, iters)
    for k in range(240, 244):
        np.testing.assert_array_equal(e_phis[k], phis[k])
        alone, it1, oc1, _ = run_batch([probs[k]], method=method)
        assert (int(it1[0]), int(oc1[0])) == (int(iters[k]), int(outc[k]))
        np.testing.assert_array_equal(alone[0], phis[k])
        assert relres_oracle(*probs[k], phis[k]) <= 1.5e-10
