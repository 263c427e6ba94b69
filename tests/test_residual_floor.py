"""CPU evidence for the FP64 residual floor of the stretched, warped 2048^2
problem (x = xi + 0.1 sin 2 pi xi sin 2 pi eta, y = (e^(5 eta) - 1)/(e^5 - 1),
Dirichlet xi faces, Neumann eta faces, rhs = -J (2 pi^2 sin pi x cos pi y +
1e-2 N(0, 1))).

Near the stretched wall the face fluxes are about 1e7 times the J-scaled rhs,
so the residual of any FP64 vector carries the rounding of that vector times
the operator.  Here the exact discrete solution is approached by iterative
refinement with a sparse LU and residuals in x87 extended precision; then:

* ``true``: the residual of that refined FP64 vector, evaluated in extended
  precision -- the representation floor (no FP64 vector does much better:
  it is the rounding of the exact solution);
* ``fp64``: the same residual evaluated in FP64 -- what a solver measures.

Both grow about 4x per halving of h (the operator's scale grows as h^-2
against a fixed rhs).  Measured at 256^2 (2.8e-12 true, 5.1e-12 FP64) and
512^2 (1.1e-11, 2.0e-11) and extrapolated at the measured growth, the 2048^2
floor lies above 1e-10 (about 1.8e-10 true, 3e-10 FP64): no FP64 solution of that
problem has a relative residual of 1e-10, which is why the single-solve
measurement on it states the tolerance it reaches.
"""
import numpy as np
import pytest

from cases import probed_matrix
from mgpoisson2d_amd import geometry as G


def floors(n):
    import scipy.sparse.linalg as spla
    geo = G.mapped_geometry(n, n, 0.1, 5.0)
    bc = G.BoundaryConditions.uniform((1, 1, 0, 0))
    x, y = G.cell_centres(geo, 0.1, 5.0)
    f = 2 * np.pi ** 2 * np.sin(np.pi * x) * np.cos(np.pi * y) + 1e-2 * np.random.default_rng(0).standard_normal((n, n))
    b = (-geo.J * f).ravel()
    A = probed_matrix(geo, bc).tocsr()
    lu = spla.splu(A.tocsc())
    Al = A.astype(np.longdouble)
    bl = b.astype(np.longdouble)
    phi = lu.solve(b)
    for _ in range(4):
        phi = phi + lu.solve((bl - Al @ phi.astype(np.longdouble)).astype(np.float64))
    nb = np.linalg.norm(b)
    true = float(np.linalg.norm((bl - Al @ phi.astype(np.longdouble)).astype(np.float64))) / nb
    fp64 = float(np.linalg.norm(b - A @ phi)) / nb
    return true, fp64


@pytest.mark.skipif(np.finfo(np.longdouble).eps >= np.finfo(np.float64).eps,
                    reason="needs an extended-precision long double (x87)")
def test_fp64_residual_floor_of_the_2048_problem_exceeds_1e10():
    t256, f256 = floors(256)
    t512, f512 = floors(512)
    # the floors grow about 4x per doubling of n
    gt, gf = t512 / t256, f512 / f256
    assert 3.5 <= gt <= 5.0, (t256, t512)
    assert 3.5 <= gf <= 5.0, (f256, f512)
    # extrapolated two doublings further at the measured growth, both lie above
    # 1e-10 (about 1.8e-10 and 3e-10)
    assert t512 * gt ** 2 > 1e-10, (t512, gt)
    assert f512 * gf ** 2 > 1e-10, (f512, gf)
