"""The serving loop (mgp_server) on the GPU.

* Continuous batching keeps results independent: requests submitted while
  others run (and queued behind a busy slot) give, bit for bit and with the
  same iteration counts, what the same request sequence gives on its slot
  alone.
* A cold request gives the bits of the batched solver on the same problem.
* Warm starts use the slot's last solution: the same rhs again finishes at
  once, a perturbed one in fewer iterations.
* Slots above 2^22 cells run in a batch of their own on a second stream;
  the results of both size classes stay those of each slot alone.
* The HBM budget is enforced at creation.
* FP32 V-cycles inside FP64 refinement (KRYLOV_MG32, the served precision):
  results are still those of each slot alone, and each meets 1e-10 by the
  oracle's operator (bit-identical to the reference's) on the GPU's own
  metrics.
"""
from types import SimpleNamespace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

mg = pytest.importorskip("mgpoisson2d_amd.capi")
from mgpoisson2d_amd import geometry as G  # noqa: E402

SLOTS = [(64, 48, 0.10, 5.0, (1, 1, 0, 0)), (96, 64, 0.05, 2.0, (1, 1, 1, 1)), (48, 112, 0.12, 4.0, (1, 1, 0, 0)),
         (80, 80, 0.00, 0.0, (1, 1, 1, 1))]


def slot_defs(slots):
    geos = [SimpleNamespace(nx=nx, ny=ny, dx=1.0 / nx, dy=1.0 / ny) for nx, ny, *_ in slots]
    maps = [mg.Mapping(mg.MAP_WARPED, a, b, 0.0, 0.0) for _, _, a, b, _ in slots]
    bcs = [G.BoundaryConditions.uniform(t) for *_, t in slots]
    return geos, maps, bcs


def rhs_for(slot, k, slots=SLOTS):
    nx, ny = slots[slot][:2]
    return np.random.default_rng(1000 * slot + k).standard_normal((ny, nx))


# (step at which it is submitted, slot, request index, warm)
SCHEDULE = [(0, 0, 0, False), (0, 1, 0, False), (3, 2, 0, False), (3, 0, 1, True), (5, 3, 0, False),
            (9, 1, 1, True), (12, 2, 1, False)]


def run(schedule, slots=SLOTS, method=mg.METHOD_KRYLOV_MG):
    geos, maps, bcs = slot_defs(slots)
    v = mg.Server(geos, bcs, maps, method=method)
    out, pending, step = {}, {}, 0
    todo = sorted(schedule)
    while todo or pending:
        while todo and todo[0][0] <= step:
            _, slot, k, warm = todo.pop(0)
            pending[v.submit(slot, rhs_for(slot, k, slots), 1e-10, 300, warm)] = (slot, k)
        v.step()
        step += 1
        for t in list(pending):
            r = v.result(t, pending[t][0])
            if r is not None:
                out[pending.pop(t)] = r
        assert step < 5000
    v.close()
    return out


def test_server_results_independent_of_the_mix():
    mixed = run(SCHEDULE)
    assert len(mixed) == len(SCHEDULE)
    for slot in range(len(SLOTS)):
        alone = run([s for s in SCHEDULE if s[1] == slot])
        for key, (phi, it, oc) in alone.items():
            phi2, it2, oc2 = mixed[key]
            assert oc == mg.CONVERGED and oc2 == mg.CONVERGED
            assert it == it2
            np.testing.assert_array_equal(phi, phi2)


def test_server_size_classes_overlap_and_stay_independent():
    """A slot above 2^22 cells runs in a second batch on its own stream; the
    small slots' results and the large one's are those of each alone."""
    slots = SLOTS[:2] + [(2064, 2048, 0.05, 3.0, (1, 1, 0, 0))]
    sched = [(0, 2, 0, False), (0, 0, 0, False), (1, 1, 0, False), (2, 0, 1, True), (4, 1, 1, False)]
    mixed = run(sched, slots)
    assert len(mixed) == len(sched)
    for slot in range(len(slots)):
        alone = run([x for x in sched if x[1] == slot], slots)
        for key, (phi, it, oc) in alone.items():
            phi2, it2, oc2 = mixed[key]
            assert oc == mg.CONVERGED and oc2 == mg.CONVERGED
            assert it == it2
            np.testing.assert_array_equal(phi, phi2)


def test_server_cold_request_matches_batch():
    geos, maps, bcs = slot_defs(SLOTS)
    b = mg.Batch(geos, bcs, mappings=maps)
    for s in range(len(SLOTS)):
        b.set_rhs(s, rhs_for(s, 0))
    it_b, oc_b = b.solve(1e-10, 300)
    res = run([(0, s, 0, False) for s in range(len(SLOTS))])
    for s in range(len(SLOTS)):
        phi, it, oc = res[(s, 0)]
        assert (it, oc) == (int(it_b[s]), int(oc_b[s]))
        np.testing.assert_array_equal(phi, b.get_phi(s))
    b.close()


def test_server_warm_start():
    geos, maps, bcs = slot_defs(SLOTS[:1])
    v = mg.Server(geos, bcs, maps)
    rhs = rhs_for(0, 0)

    def solve(r, warm):
        t = v.submit(0, r, 1e-10, 300, warm)
        while True:
            v.step()
            res = v.result(t, 0)
            if res is not None:
                return res

    phi0, it0, oc0 = solve(rhs, False)
    assert oc0 == mg.CONVERGED and it0 > 0
    phi1, it1, oc1 = solve(rhs, True)
    assert oc1 == mg.CONVERGED and it1 == 0
    np.testing.assert_array_equal(phi0, phi1)
    _, it2, oc2 = solve(rhs + 1e-3 * rhs_for(0, 5), True)
    assert oc2 == mg.CONVERGED and it2 < it0
    # a zero rhs gives phi = 0 whatever the warm start
    phi3, it3, oc3 = solve(np.zeros_like(rhs), True)
    assert oc3 == mg.CONVERGED and it3 == 0 and not np.any(phi3)
    v.close()


def test_server_budget():
    geos, maps, bcs = slot_defs(SLOTS)
    v = mg.Server(geos, bcs, maps)
    need = v.footprint()
    v.close()
    assert need > 0
    with pytest.raises(mg.MgpError):
        mg.Server(geos, bcs, maps, hbm_budget=need / 2)
    mg.Server(geos, bcs, maps, hbm_budget=need * 1.01).close()


def test_server_mixed_precision_independent_and_accurate():
    import refcall as rc
    method = mg.METHOD_KRYLOV_MG32
    mixed = run(SCHEDULE, method=method)
    assert len(mixed) == len(SCHEDULE)
    for slot in range(len(SLOTS)):
        alone = run([s for s in SCHEDULE if s[1] == slot], method=method)
        for key, (phi, it, oc) in alone.items():
            phi2, it2, oc2 = mixed[key]
            assert oc == mg.CONVERGED and oc2 == mg.CONVERGED
            assert it == it2
            np.testing.assert_array_equal(phi, phi2)
    # accuracy against the oracle's operator on the metrics the server built
    o = rc.oracle()
    geos, maps, bcs = slot_defs(SLOTS)
    for (slot, k), (phi, _, _) in mixed.items():
        g = geos[slot]
        geo = mg.metric_fill(maps[slot], g.nx, g.ny, g.dx, g.dy)
        full = G.new_phi(g.nx, g.ny)
        full[1:-1, 1:-1] = phi
        rhs = rhs_for(slot, k)
        r, _ = o.residual(geo, bcs[slot], 1, rhs, full)
        assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(rhs), (slot, k)
