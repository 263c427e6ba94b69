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
* The HBM budget is enforced at creation and when geometries are admitted
  while serving; evicted geometries free their bytes.
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


def drain(v, pending):
    out = {}
    steps = 0
    while pending:
        v.step()
        steps += 1
        for t in list(pending):
            r = v.result(t, pending[t])
            if r is not None:
                out[t] = r
                del pending[t]
        assert steps < 5000
    return out


def test_server_admits_and_evicts_geometries_at_run_time():
    """Geometries join and leave while serving: a geometry whose hierarchy fits the HBM budget
    is admitted and solves what it solves alone; one that does not fit is refused with
    MGP_ERR_NOMEM and the server goes on; an evicted slot frees its bytes (its batch partners
    keep their warm starts) and refuses further requests; the freed room admits the refused
    geometry."""
    geos, maps, bcs = slot_defs(SLOTS[:2])
    probe = mg.Server(geos, bcs, maps)
    base = probe.footprint()
    probe.close()
    extra_small = [(72, 56, 0.08, 3.0, (1, 1, 0, 0))]
    big = [(512, 512, 0.1, 5.0, (1, 1, 0, 0))]
    gs, ms, bs = slot_defs(extra_small)
    gb, mb, bb = slot_defs(big)
    # the budget holds the two slots, the small newcomer and a margin, not the big one
    small_alone = mg.Server(gs, bs, ms)
    small_bytes = small_alone.footprint()
    small_alone.close()
    budget = base + small_bytes + 0.5e6
    v = mg.Server(geos, bcs, maps, hbm_budget=budget)
    t0 = v.submit(0, rhs_for(0, 0), 1e-10, 300, False)
    s_new = v.add_slot(gs[0], bs[0], ms[0])
    assert s_new == 2
    assert v.footprint() <= budget
    t1 = v.submit(s_new, rhs_for(0, 0, extra_small), 1e-10, 300, False)
    with pytest.raises(mg.MgpError, match="out of device memory"):
        v.add_slot(gb[0], bb[0], mb[0])
    # a copy of slot 1's geometry does not fit beside it either
    with pytest.raises(mg.MgpError, match="out of device memory"):
        v.add_slot(geos[1], bcs[1], maps[1])
    res = drain(v, {t0: 0, t1: s_new})
    assert all(r[2] == mg.CONVERGED for r in res.values())
    # the newcomer's result is what it gives alone
    alone = mg.Server(gs, bs, ms)
    ta = alone.submit(0, rhs_for(0, 0, extra_small), 1e-10, 300, False)
    ra = drain(alone, {ta: 0})[ta]
    alone.close()
    np.testing.assert_array_equal(res[t1][0], ra[0])
    assert res[t1][1] == ra[1]
    # eviction frees the slot's bytes and refuses its requests; the other slots keep working
    before = v.footprint()
    v.evict(1)
    assert v.footprint() < before
    with pytest.raises(mg.MgpError):
        v.submit(1, rhs_for(1, 0), 1e-10, 300, False)
    # slot 0 shared its batch with slot 1: its warm start survived the rebuild
    t2 = v.submit(0, rhs_for(0, 0), 1e-10, 300, True)
    r2 = drain(v, {t2: 0})[t2]
    assert r2[2] == mg.CONVERGED and r2[1] == 0  # the same rhs again: already converged
    # the room slot 1 freed admits its copy, which then solves
    s_copy = v.add_slot(geos[1], bcs[1], maps[1])
    t3 = v.submit(s_copy, rhs_for(1, 0), 1e-10, 300, False)
    assert drain(v, {t3: s_copy})[t3][2] == mg.CONVERGED
    for slot in (0, s_new, s_copy):
        v.evict(slot)
    assert v.footprint() == 0
    v.close()
