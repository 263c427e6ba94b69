"""The bench's contract on the CPU: one GPU only, and the batched mix it
measures is the documented workload.

* `bench.py --gpus N` with N > 1 prints one JSON line naming the reason and
  exits with status 2, before touching a device.
* The 1024-problem mix: nx, ny multiples of 16 in [32, 1024], log-uniform
  (about as many problems below sqrt(32 * 1024) = 181 as above it), beta in
  [0, 6], A in [0, 0.15], the two BC mixes about equally often, one N(0,1)
  rhs per cell of each problem; the same problems from the same seed.
"""
import json
import math
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_several_gpus_are_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_the_mix_is_the_documented_workload():
    probs, rhs = bench.config4_mix(1024)
    assert len(probs) == 1024 and len(rhs) == 1024
    nx = np.array([p[0] for p in probs])
    ny = np.array([p[1] for p in probs])
    amp = np.array([p[2] for p in probs])
    beta = np.array([p[3] for p in probs])
    for n in (nx, ny):
        assert np.all(n % 16 == 0) and n.min() >= 32 and n.max() <= 1024
        mid = math.sqrt(32 * 1024)
        below = int(np.sum(n < mid))
        assert 400 <= below <= 624, below  # log-uniform: half below the geometric mean
    assert beta.min() >= 0.0 and beta.max() <= 6.0 and beta.max() > 5.5
    assert amp.min() >= 0.0 and amp.max() <= 0.15 and amp.max() > 0.14
    types = [p[4] for p in probs]
    assert set(types) == {(1, 1, 1, 1), (1, 1, 0, 0)}
    assert 400 <= types.count((1, 1, 0, 0)) <= 624
    for (x, y, *_), r in zip(probs[:50], rhs[:50]):
        assert r.shape == (y, x)
    allr = np.concatenate([r.ravel() for r in rhs[:200]])
    assert abs(allr.mean()) < 0.01 and abs(allr.std() - 1.0) < 0.01
    # the same problems from the same seed
    probs2, rhs2 = bench.config4_mix(1024)
    assert probs2 == probs
    np.testing.assert_array_equal(rhs2[7], rhs[7])
