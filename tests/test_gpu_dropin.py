"""Drop-in check: one Fortran caller written against the reference's
libMGPoisson interface (tests/fortran/dropin_caller.f90), compiled against the
reference's own modules and against the drop-in layer over the HIP library;
both executables run and their outputs agree.

Operators, inverse diagonals, Gauss-Seidel sweeps, restriction, the
single-level solvers (relax_jacobi, solve_bicgstab), the flux tools
(compute_grad, compute_div, compute_pd, fill_boundary_fluxes) and
extrapolate_ghosts must match bit for bit; the reductions (pnorm, inner_prod,
integrate2d) to 1e-13 (parallel sums; the max norm exactly); the V-cycle and FMG results to 1e-12 (their relres norms are
summed in parallel on the GPU)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "fortran", "build")
N = 64


def run(exe, out):
    subprocess.run([os.path.join(BUILD, exe), out], check=True, timeout=120)
    return np.fromfile(out, dtype=np.float64)


def split(a):
    n2, c2 = N * N, (N // 2) ** 2
    nf = (N + 1) * N
    sizes = [n2, n2, n2, c2, n2, n2, n2, n2, n2, n2, nf, nf, n2, nf, nf, 4, 3, (N + 2) ** 2]
    parts, k = [], 0
    for s in sizes:
        parts.append(a[k:k + s])
        k += s
    assert k == a.size
    return parts


def test_fortran_caller_dropin_matches_reference(tmp_path):
    if not os.path.exists(os.path.join(BUILD, "caller_ref")):
        pytest.skip("reference-linked caller not built (needs oracle/_ref)")
    g = split(run("caller_dropin", str(tmp_path / "dropin.bin")))
    r = split(run("caller_ref", str(tmp_path / "ref.bin")))
    for k in range(4):  # laplacian, invdiags, 3 GS sweeps, restrict
        np.testing.assert_array_equal(g[k], r[k])
    for k in (4, 6, 7):  # DC vcycle x4, plain vcycle x1, fmg x1
        np.testing.assert_allclose(g[k], r[k], rtol=0, atol=1e-12 * np.max(np.abs(r[k])))
    np.testing.assert_array_equal(g[5], r[5])  # rhs restored by the DC driver
    for k in range(8, 15):  # jacobi, bicgstab v1, compute_grad x/y, compute_div, compute_pd, fill_boundary_fluxes
        np.testing.assert_array_equal(g[k], r[k])
    # pnorm p = 0..3, inner_prod, integrate2d with and without J: parallel sums on the GPU
    assert g[15][0] == r[15][0]
    np.testing.assert_allclose(g[15], r[15], rtol=1e-13)
    np.testing.assert_allclose(g[16], r[16], rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(g[17], r[17])  # extrapolate_ghosts order 2
