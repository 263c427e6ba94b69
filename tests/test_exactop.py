"""CPU pin of tests/exactop.py, the restatement of the reference operator
over any number type that checks the double-double residual: evaluated with
Python floats it must equal the oracle's compute_laplacian (itself pinned to
the reference's golden vectors) bit for bit on every homogeneous Dirichlet /
Neumann side mix; evaluated with fractions it is exact, and the exact
residual of a split solution hi + lo differs from the FP64 residual of hi by
about the FP64 operator's rounding, as it must."""
import numpy as np
import pytest

import exactop
import refcall as rc
from cases import OPERATOR_CASES, setup


@pytest.mark.parametrize("case", OPERATOR_CASES)
def test_float_restatement_is_the_oracle_operator(case):
    geo, bc, phi, _ = setup(*case, homog=True, seed=3)
    lap_o, _ = rc.oracle().laplacian(geo, bc, 1, phi)
    lap_p = exactop.lap_float(geo, case[4], phi)
    np.testing.assert_array_equal(lap_p, lap_o)


def test_exact_residual_is_close_to_the_fp64_one():
    nx, ny, amp, beta, types = 10, 9, 0.1, 3.0, (1, 1, 0, 0)
    geo, bc, phi, rhs = setup(nx, ny, amp, beta, types, homog=True, seed=4)
    lo = np.zeros_like(phi)
    exact = exactop.residual_exact(geo, types, rhs, phi, lo)
    fp64, _ = rc.oracle().residual(geo, bc, 1, rhs, phi)
    scale = np.max(np.abs(exactop.lap_float(geo, types, np.abs(phi))))  # ~ |A||phi|
    d = np.max(np.abs(exact - fp64))
    assert 0 < d <= 64 * np.finfo(float).eps * scale
    # a low part below the FP64 resolution of phi changes the exact residual
    lo = 1e-17 * np.random.default_rng(5).standard_normal(phi.shape)
    exact2 = exactop.residual_exact(geo, types, rhs, phi, lo)
    assert np.max(np.abs(exact2 - exact)) > 0
