"""The batch autotuner (mgp_batch_tune / _set_tuning / _tuning) on the GPU.

The tuner picks, per key (nx, BC class, precision), the threads per row of
the batched line solve, and with them the chunk length.  The chunk
partition decides the rounding of a row's solve, so the shapes cannot give
identical bits (unlike the red-black tile shapes); what each must give:

* every shape a key can take solves the line systems: the batch reaches the
  requested residual of the reference operator (oracle), and the solutions
  of all shapes agree to the solve tolerance;
* the choice is a function of the key only: a batch given the tuner's table
  gives the tuned batch's bits, and a problem gives the same bits alone and
  in the mix under one table;
* the tuner returns one entry per key, each a shape with a kernel.
"""
import numpy as np
import pytest

from test_gpu_batch import problems, relres_oracle

pytestmark = pytest.mark.gpu

mg = pytest.importorskip("mgpoisson2d_amd.capi")

# rows of 64-cell, 480-cell (all three widths), 640-cell and 1008-cell classes, both BC mixes
CASES = [
    (64, 48, 0.10, 4.0, (1, 1, 0, 0)),
    (480, 32, 0.05, 2.0, (1, 1, 1, 1)),
    (480, 40, 0.12, 5.0, (1, 1, 0, 0)),
    (640, 24, 0.08, 3.0, (1, 1, 0, 0)),
    (1008, 20, 0.10, 1.0, (1, 1, 1, 1)),
    (96, 64, 0.00, 0.0, (1, 1, 1, 1)),
]


def solve(probs, method, tuning=None, tune_reps=0):
    b = mg.Batch([p[0] for p in probs], [p[1] for p in probs], method=method)
    chosen = None
    if tuning is not None:
        b.set_tuning(tuning)
    if tune_reps:
        chosen = b.tune(tune_reps)
    for k, (_, _, rhs) in enumerate(probs):
        b.set_rhs(k, rhs)
    iters, outc = b.solve(1e-10, 300)
    phis = [b.get_phi(k) for k in range(len(probs))]
    table = b.tuning()
    b.close()
    return phis, iters, outc, table, chosen


@pytest.mark.parametrize("method", [mg.METHOD_KRYLOV_MG, mg.METHOD_KRYLOV_MG32])
def test_every_shape_solves(method):
    probs = problems(CASES, seed0=70)
    fp32 = int(method == mg.METHOD_KRYLOV_MG32)
    base, _, outc, table, _ = solve(probs, method)
    assert np.all(outc == mg.CONVERGED)
    keys = sorted({(e[0], e[1]) for e in table})
    assert len(keys) == len(CASES)
    for threads in (64, 128, 256):
        # every key the width has a kernel for (64 threads: rows up to 519 cells; 128: up to 1031)
        entries = [(nx, bc, fp32, threads) for nx, bc in keys if nx <= {64: 519, 128: 1031, 256: 4111}[threads]]
        phis, iters, outc, now, _ = solve(probs, method, tuning=entries)
        assert np.all(outc == mg.CONVERGED), (threads, iters, outc)
        got = {(e[0], e[1]): e[3] for e in now}
        for nx, bc, _, t in entries:
            assert got[(nx, bc)] == t
        for (geo, bc, rhs), phi, ref in zip(probs, phis, base):
            assert relres_oracle(geo, bc, rhs, phi) <= 1.5e-10
            assert np.max(np.abs(phi - ref)) <= 1e-7 * np.max(np.abs(ref))


def test_shape_without_kernel_is_rejected():
    probs = problems(CASES[:1], seed0=70)
    b = mg.Batch([p[0] for p in probs], [p[1] for p in probs])
    nx = CASES[0][0]
    bc = b.tuning()[0][1]
    with pytest.raises(mg.MgpError):
        b.set_tuning([(nx, bc, 0, 96)])
    b.close()


@pytest.mark.parametrize("method", [mg.METHOD_KRYLOV_MG, mg.METHOD_KRYLOV_MG32])
def test_tuned_table_reproduces_the_tuned_batch(method):
    probs = problems(CASES, seed0=80)
    tuned, it_t, oc_t, table, chosen = solve(probs, method, tune_reps=3)
    assert np.all(oc_t == mg.CONVERGED)
    assert len(chosen) == len(CASES)
    assert sorted(chosen) == sorted(table)
    for nx, bc, fp32, threads in chosen:
        assert threads in (64, 128, 256)
        assert fp32 == int(method == mg.METHOD_KRYLOV_MG32)
    # the table alone, in a fresh batch: the same bits
    again, it_a, _, _, _ = solve(probs, method, tuning=chosen)
    np.testing.assert_array_equal(it_a, it_t)
    for a, b in zip(again, tuned):
        np.testing.assert_array_equal(a, b)
    # and a problem alone under the same table: the same bits as in the mix
    for k in (1, 4):
        alone, it1, _, _, _ = solve([probs[k]], method, tuning=chosen)
        assert int(it1[0]) == int(it_t[k])
        np.testing.assert_array_equal(alone[0], tuned[k])


def test_committed_cache_loads_and_applies():
    entries = mg.load_tuning()
    for e in entries:
        assert len(e) == 4 and e[3] in (64, 128, 256) and e[2] in (0, 1)
    probs = problems(CASES[:3], seed0=90)
    b = mg.Batch([p[0] for p in probs], [p[1] for p in probs], method=mg.METHOD_KRYLOV_MG32)
    b.set_tuning(entries)
    mine = {(e[0], e[1], e[2]): e[3] for e in entries}
    for nx, bc, fp32, threads in b.tuning():
        if (nx, bc, fp32) in mine:
            assert threads == mine[(nx, bc, fp32)]
    b.close()
