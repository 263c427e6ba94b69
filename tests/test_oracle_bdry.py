"""CPU pins of the oracle's integrate2d_bdry and gpnorm (arrayutils.f90:721,
802): bit for bit against the reference build where present, and closed
forms -- for F = (a x + b, c y + d) on the unit box the sum is
a ny dx + c nx dy exactly (values are small integers times powers of two; the
reference weights the x faces by dx and the y faces by dy, which is the net
outward flux when dx == dy); the
ghost-layer norm of a field that is 1 on its ghosts (corners excluded) and
arbitrary inside is 1 (p = 0) and (2 (nx dx + ny dy))^(1/p) otherwise."""
import numpy as np
import pytest

import refcall as rc

needs_ref = pytest.mark.skipif(not rc.reference_available(), reason="oracle/_ref not built")


@needs_ref
@pytest.mark.parametrize("nx,ny", [(8, 8), (13, 13), (20, 7)])
@pytest.mark.parametrize("seed", [0, 1])
def test_integrate2d_bdry_oracle_vs_reference(nx, ny, seed):
    rng = np.random.default_rng(seed)
    xf, yf = rng.standard_normal((ny, nx + 1)), rng.standard_normal((ny + 1, nx))
    dx, dy = 1.0 / nx, 0.7 / ny
    a = rc.oracle().integrate2d_bdry(nx, ny, dx, dy, xf, yf)
    b = rc.reference().integrate2d_bdry(nx, ny, dx, dy, xf, yf)
    assert a == b


@needs_ref
@pytest.mark.parametrize("n", [6, 17])
@pytest.mark.parametrize("p", [0, 1, 2, 3])
def test_gpnorm_oracle_vs_reference(n, p):
    rng = np.random.default_rng(n + p)
    phi = rng.standard_normal((n + 2, n + 2))
    if p in (1, 3):
        phi = np.abs(phi)  # the reference takes no absolute value before the root
    a = rc.oracle().gpnorm(n, n, 1.0 / n, 1.0 / n, phi, p)
    b = rc.reference().gpnorm(n, n, 1.0 / n, 1.0 / n, phi, p)
    assert a == b


def test_integrate2d_bdry_closed_form():
    nx, ny = 16, 8
    dx, dy = 1.0 / nx, 1.0 / ny
    a, b, c, d = 3.0, -0.5, 1.25, 2.0
    x = np.arange(nx + 1) * dx
    y = np.arange(ny + 1) * dy
    xf = np.tile(a * x + b, (ny, 1))
    yf = np.tile((c * y + d)[:, None], (1, nx))
    assert rc.oracle().integrate2d_bdry(nx, ny, dx, dy, xf, yf) == a * ny * dx + c * nx * dy


@pytest.mark.parametrize("p", [0, 1, 2, 4])
def test_gpnorm_closed_form(p):
    n = 12
    dx = dy = 1.0 / 16
    phi = np.full((n + 2, n + 2), 1.0)
    phi[1:-1, 1:-1] = np.random.default_rng(3).standard_normal((n, n)) * 100
    phi[0, 0] = phi[0, -1] = phi[-1, 0] = phi[-1, -1] = 1e9  # corner ghosts are skipped
    got = rc.oracle().gpnorm(n, n, dx, dy, phi, p)
    want = 1.0 if p == 0 else (2 * (n * dx + n * dy)) ** (1.0 / p)
    assert abs(got - want) <= 1e-15 * want
