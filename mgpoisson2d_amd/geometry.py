"""Mapped-grid geometry, boundary data and manufactured problems.

Arrays follow the reference's storage: Fortran column-major ``data(i, j)`` with
i (xi) fastest (arrayutils.f90:107-116).  In numpy that is a C-ordered array of
shape ``(rows, cols) = (ny, nx)`` indexed ``[j, i]``:

* ``J``              cell-centred            ``(ny, nx)``
* ``Jxx``, ``Jxy``   x faces, nodal in xi    ``(ny, nx + 1)``  (Jgup_xx, Jgup_xy)
* ``Jyx``, ``Jyy``   y faces, nodal in eta   ``(ny + 1, nx)``  (Jgup_yx, Jgup_yy)
* solution-like fields carry one ghost layer ``(ny + 2, nx + 2)``, element
  (i, j) at ``[j + 1, i + 1]`` (the reference's phi, MGPoisson2D.f90:524).

Metric terms are J-scaled contravariant metrics, J g^{ij}, evaluated
analytically at the face centres (the reference's test assembles them by
finite differences for x = xi, tests/test.f90:566-716; the closed forms below
reduce to those when x = xi).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# Reference integer codes (arrayutils.f90:27-49).
BCTYPE_NONE = -1
BCTYPE_NEUM = 0
BCTYPE_DIRI = 1
BCTYPE_PERIODIC = 2
BCMODE_UNIFORM = 0
BCMODE_NONUNIFORM = 1
BCMODE_LAH = 2


@dataclasses.dataclass
class Geometry:
    """Grid plus metric arrays; mirrors the reference's ``geo_data``."""

    nx: int
    ny: int
    dx: float
    dy: float
    J: np.ndarray
    Jxx: np.ndarray
    Jxy: np.ndarray
    Jyx: np.ndarray
    Jyy: np.ndarray

    def arrays(self):
        return (self.J, self.Jxx, self.Jxy, self.Jyx, self.Jyy)


@dataclasses.dataclass
class BoundaryConditions:
    """Per-face BC types, modes and data; mirrors ``bdry_data``.

    ``data`` holds four float64 arrays (xlo, xhi, ylo, yhi): one value for
    UNIFORM/LAH faces, ny values for NONUNIFORM x faces, nx for y faces.
    """

    types: tuple
    modes: tuple
    data: tuple

    @staticmethod
    def uniform(types, value=0.0, mode=BCMODE_UNIFORM):
        data = tuple(np.full(1, value, dtype=np.float64) for _ in range(4))
        return BoundaryConditions(tuple(int(t) for t in types), (mode,) * 4, data)


def mapping_derivatives(xi, eta, amp=0.0, beta=0.0):
    """Jacobian entries of x = xi + amp sin(2 pi xi) sin(2 pi eta),
    y = (exp(beta eta) - 1) / (exp(beta) - 1)  (y = eta when beta = 0)."""
    two_pi = 2.0 * math.pi
    x_xi = 1.0 + amp * two_pi * np.cos(two_pi * xi) * np.sin(two_pi * eta)
    x_eta = amp * two_pi * np.sin(two_pi * xi) * np.cos(two_pi * eta)
    y_xi = np.zeros_like(xi * eta)
    if beta == 0.0:
        y_eta = np.ones_like(xi * eta)
    else:
        y_eta = beta * np.exp(beta * eta) / math.expm1(beta) + 0.0 * xi
    return x_xi, x_eta, y_xi, y_eta


def map_points(xi, eta, amp=0.0, beta=0.0):
    x = xi + amp * np.sin(2.0 * math.pi * xi) * np.sin(2.0 * math.pi * eta)
    if beta == 0.0:
        y = eta + 0.0 * xi
    else:
        y = np.expm1(beta * eta) / math.expm1(beta) + 0.0 * xi
    return x, y


def _metrics(xi, eta, amp, beta):
    x_xi, x_eta, y_xi, y_eta = mapping_derivatives(xi, eta, amp, beta)
    J = x_xi * y_eta - x_eta * y_xi
    g11 = (x_eta * x_eta + y_eta * y_eta) / J
    g12 = -(x_xi * x_eta + y_xi * y_eta) / J
    g22 = (x_xi * x_xi + y_xi * y_xi) / J
    return J, g11, g12, g22


def mapped_geometry(nx, ny, amp=0.0, beta=0.0, L=1.0, H=1.0):
    """Analytic J and J g^{ij} for the warped/stretched mapping on
    (xi, eta) in [0, L] x [0, H] (the mapping uses xi/L, eta/H)."""
    dx, dy = L / nx, H / ny
    ic = (np.arange(nx) + 0.5) * dx / L
    jc = (np.arange(ny) + 0.5) * dy / H
    inode = np.arange(nx + 1) * dx / L
    jnode = np.arange(ny + 1) * dy / H
    xi_c, eta_c = np.meshgrid(ic, jc)            # (ny, nx)
    xi_x, eta_x = np.meshgrid(inode, jc)         # (ny, nx+1)
    xi_y, eta_y = np.meshgrid(ic, jnode)         # (ny+1, nx)
    # derivatives w.r.t. the scaled coordinates; chain rule for L, H != 1
    def scaled(xi, eta):
        J, g11, g12, g22 = _metrics(xi, eta, amp, beta)
        # x(s, t) with s = xi/L, t = eta/H: J = J_st/(L H),
        # J g^11 = (L/H) [J g^11]_st, J g^12 = [J g^12]_st, J g^22 = (H/L) [J g^22]_st
        return J / (L * H), g11 * (L / H), g12, g22 * (H / L)
    Jc, _, _, _ = scaled(xi_c, eta_c)
    _, g11x, g12x, _ = scaled(xi_x, eta_x)
    _, _, g12y, g22y = scaled(xi_y, eta_y)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    return Geometry(nx, ny, dx, dy, f64(Jc), f64(g11x), f64(g12x), f64(g12y), f64(g22y))


def cell_centres(geo: Geometry, amp=0.0, beta=0.0, L=1.0, H=1.0):
    ic = (np.arange(geo.nx) + 0.5) * geo.dx / L
    jc = (np.arange(geo.ny) + 0.5) * geo.dy / H
    xi, eta = np.meshgrid(ic, jc)
    return map_points(xi, eta, amp, beta)


def new_phi(nx, ny):
    return np.zeros((ny + 2, nx + 2), dtype=np.float64)


def valid(phi):
    return phi[1:-1, 1:-1]
