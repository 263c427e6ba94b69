"""ctypes front end for the two CPU checkers used by the parity tests.

* ``oracle/liboracle.so``            the C restatement (``mgo_*`` entry points)
* ``oracle/_ref/parity/libmgref.so`` the unmodified reference Fortran plus a
  bind(C) bridge (``ref_*``), built from /root/reference by oracle/Makefile;
  ``oracle/_ref/rb/libmgref_rb.so`` is the same with the DC smoother's
  red-black switch on.

Both expose the same argument order, so one ``Checker`` drives either.  Arrays
use the numpy layouts documented in mgpoisson2d_amd/geometry.py.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "parity", "libmgref.so")
REF_RB_SO = os.path.join(ROOT, "oracle", "_ref", "rb", "libmgref_rb.so")
REF_FAST_SO = os.path.join(ROOT, "oracle", "_ref", "fast", "libmgref_fast.so")

LEX, RB, MC = 0, 1, 2

_i = ctypes.c_int
_d = ctypes.c_double
_p = ctypes.c_void_p


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return _p(a.ctypes.data)


def _iptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return _p(a.ctypes.data)


class Checker:
    """One CPU implementation of the reference API (oracle or reference)."""

    def __init__(self, path, prefix):
        self.path = path
        self.lib = ctypes.CDLL(path)
        self.prefix = prefix
        self.is_oracle = prefix == "mgo_"
        self._keep = []

    def _fn(self, name, restype=None):
        f = getattr(self.lib, self.prefix + name)
        f.restype = restype
        return f

    @staticmethod
    def _geo_args(geo):
        return [_i(geo.nx), _i(geo.ny), _d(geo.dx), _d(geo.dy)] + [_ptr(a) for a in geo.arrays()]

    def _bc_args(self, bc):
        bct = np.ascontiguousarray(bc.types, dtype=np.int32)
        bcm = np.ascontiguousarray(bc.modes, dtype=np.int32)
        data = [np.ascontiguousarray(d, dtype=np.float64) for d in bc.data]
        self._keep = [bct, bcm] + data
        return [_iptr(bct), _iptr(bcm)] + [_ptr(d) for d in data]

    # -- operators -----------------------------------------------------------
    def fill_ghosts(self, geo, bc, homog, do_neum, phi):
        phi = phi.copy()
        self._fn("fill_ghosts")(*self._geo_args(geo), *self._bc_args(bc), _i(homog), _i(do_neum), _ptr(phi))
        return phi

    def extrapolate_ghosts(self, nx, ny, order, phi):
        phi = phi.copy()
        self._fn("extrapolate_ghosts")(_i(nx), _i(ny), _i(order), _ptr(phi))
        return phi

    def laplacian(self, geo, bc, homog, phi, jscale=False):
        phi = phi.copy()
        lap = np.zeros((geo.ny, geo.nx))
        self._fn("laplacian")(*self._geo_args(geo), *self._bc_args(bc), _i(homog), _i(jscale), _ptr(phi), _ptr(lap))
        return lap, phi

    def residual(self, geo, bc, homog, rhs, phi):
        phi = phi.copy()
        res = np.zeros((geo.ny, geo.nx))
        rhs = np.ascontiguousarray(rhs)
        self._fn("residual")(*self._geo_args(geo), *self._bc_args(bc), _i(homog), _ptr(rhs), _ptr(phi), _ptr(res))
        return res, phi

    # -- flux tools ------------------------------------------------------------
    def compute_pd(self, nx, ny, dx, dy, phi, direction, nodedir):
        """pd nodal in x (nodedir 1: shape (ny, nx+1)) or y (2: (ny+1, nx))."""
        pd = np.zeros((ny, nx + 1) if nodedir == 1 else (ny + 1, nx))
        phi = np.ascontiguousarray(phi)
        self._fn("compute_pd")(_i(nx), _i(ny), _d(dx), _d(dy), _ptr(phi), _i(direction), _i(nodedir), _ptr(pd))
        return pd

    def compute_grad(self, geo, bc, homog, phi):
        """Returns (phi with ghosts filled, xflux (ny, nx+1), yflux (ny+1, nx))."""
        phi = phi.copy()
        xf = np.zeros((geo.ny, geo.nx + 1))
        yf = np.zeros((geo.ny + 1, geo.nx))
        self._fn("compute_grad")(*self._geo_args(geo), *self._bc_args(bc), _i(homog), _ptr(phi), _ptr(xf),
                                 _ptr(yf))
        return phi, xf, yf

    def metric_terrain(self, nx, ny, dx, dy, amp):
        """The reference test geometry's metrics: returns a Geometry."""
        from mgpoisson2d_amd import geometry as G
        arrs = [np.zeros((ny, nx)), np.zeros((ny, nx + 1)), np.zeros((ny, nx + 1)), np.zeros((ny + 1, nx)),
                np.zeros((ny + 1, nx))]
        self._fn("metric_terrain")(_i(nx), _i(ny), _d(dx), _d(dy), _d(amp), *[_ptr(a) for a in arrs])
        return G.Geometry(nx, ny, dx, dy, *arrs)

    def fill_boundary_fluxes(self, nx, ny, bc, homog, xflux, yflux):
        xf, yf = np.ascontiguousarray(xflux).copy(), np.ascontiguousarray(yflux).copy()
        self._fn("fill_boundary_fluxes")(_i(nx), _i(ny), *self._bc_args(bc), _i(homog), _ptr(xf), _ptr(yf))
        return xf, yf

    def compute_div(self, nx, ny, dx, dy, xflux, yflux):
        div = np.zeros((ny, nx))
        xflux, yflux = np.ascontiguousarray(xflux), np.ascontiguousarray(yflux)
        self._fn("compute_div")(_i(nx), _i(ny), _d(dx), _d(dy), _ptr(xflux), _ptr(yflux), _ptr(div))
        return div

    def invdiags(self, geo, bc, homog):
        out = np.zeros((geo.ny, geo.nx))
        self._fn("invdiags")(*self._geo_args(geo), *self._bc_args(bc), _i(homog), _ptr(out))
        return out

    def relax_gs(self, geo, bc, homog, invd, omega, iters, rhs, phi, order=LEX):
        phi = phi.copy()
        args = [*self._geo_args(geo), *self._bc_args(bc), _i(homog), _ptr(np.ascontiguousarray(invd)),
                _d(omega), _i(iters), _ptr(np.ascontiguousarray(rhs)), _ptr(phi)]
        if self.is_oracle:
            args.append(_i(order))
        elif order != LEX:
            raise ValueError("the reference relax_gs is lexicographic")
        self._fn("relax_gs")(*args)
        return phi

    def gs_point(self, geo, bc, homog, invd, omega, rhs, phi, i, j):
        """One relax_gs point update at (i, j) in place on phi (ghosts as they
        stand; oracle only)."""
        assert self.is_oracle and phi.flags.c_contiguous
        self._fn("gs_point")(*self._geo_args(geo), *self._bc_args(bc), _i(homog),
                             _ptr(np.ascontiguousarray(invd)), _d(omega), _ptr(np.ascontiguousarray(rhs)),
                             _ptr(phi), _i(i), _i(j))

    # -- fourth-order operator (oracle only) ---------------------------------
    MET4 = ("A4", "Bx4", "dA", "dBx", "C4", "By4", "dC", "dBy")

    def metric4(self, nx, ny, dx, dy, amp, beta, L=0.0, H=0.0):
        """Face-averaged metric set of the warped mapping: dict of x-face
        arrays (ny, nx+1) A4, Bx4, dA, dBx and y-face arrays (ny+1, nx) C4,
        By4, dC, dBy."""
        out = {k: np.zeros((ny, nx + 1)) for k in self.MET4[:4]}
        out.update({k: np.zeros((ny + 1, nx)) for k in self.MET4[4:]})
        self._fn("metric4")(_i(nx), _i(ny), _d(dx), _d(dy), _d(amp), _d(beta), _d(L), _d(H),
                            *[_ptr(out[k]) for k in self.MET4])
        return out

    def lap4(self, nx, ny, dx, dy, met, phi):
        """Fourth-order operator on cell averages phi (ny, nx)."""
        lap = np.zeros((ny, nx))
        arrs = [np.ascontiguousarray(met[k], dtype=np.float64) for k in self.MET4]
        self._fn("lap4")(_i(nx), _i(ny), _d(dx), _d(dy), *[_ptr(a) for a in arrs],
                         _ptr(np.ascontiguousarray(phi, dtype=np.float64)), _ptr(lap))
        return lap

    def dc_laplacian(self, geo, bc, homog, phi):
        phi = phi.copy()
        lap = np.zeros((geo.ny, geo.nx))
        self._fn("dc_laplacian")(*self._geo_args(geo), *self._bc_args(bc), _i(homog), _ptr(phi), _ptr(lap))
        return lap, phi

    def dc_invdiags(self, geo):
        out = np.zeros((geo.ny, geo.nx))
        self._fn("dc_invdiags")(*self._geo_args(geo), _ptr(out))
        return out

    def dc_relax_gs(self, geo, bc, homog, invd, omega, iters, rhs, phi, order=LEX):
        phi = phi.copy()
        args = [*self._geo_args(geo), *self._bc_args(bc), _i(homog), _ptr(np.ascontiguousarray(invd)),
                _d(omega), _i(iters), _ptr(np.ascontiguousarray(rhs)), _ptr(phi)]
        if self.is_oracle:
            args.append(_i(order))
        self._fn("dc_relax_gs")(*args)
        return phi

    def restrict(self, kind, fnx, fny, fine, cnx, cny):
        shape = {0: (cny, cnx), 1: (cny, cnx + 1), 2: (cny + 1, cnx)}[kind]
        crse = np.zeros(shape)
        fine = np.ascontiguousarray(fine)
        self._fn("restrict")(_i(kind), _i(fnx), _i(fny), _ptr(fine), _i(cnx), _i(cny), _ptr(crse))
        return crse

    def prolong(self, fgeo, cgeo, bct, bcm, order, fine, crse):
        fine = fine.copy()
        crse = crse.copy()
        t = np.ascontiguousarray(bct, dtype=np.int32)
        m = np.ascontiguousarray(bcm, dtype=np.int32)
        self._fn("prolong")(_i(fgeo.nx), _i(fgeo.ny), _d(fgeo.dx), _d(fgeo.dy), *[_ptr(a) for a in fgeo.arrays()],
                            _i(cgeo.nx), _i(cgeo.ny), _d(cgeo.dx), _d(cgeo.dy), *[_ptr(a) for a in cgeo.arrays()],
                            _iptr(t), _iptr(m), _i(order), _ptr(fine), _ptr(crse))
        return fine, crse

    def schedule(self, nx, ny, dx, dy, maxdepth=-1):
        rx = np.zeros(64, dtype=np.int32)
        ry = np.zeros(64, dtype=np.int32)
        md = _i(0)
        self._fn("schedule")(_i(nx), _i(ny), _d(dx), _d(dy), _i(maxdepth), ctypes.byref(md), _iptr(rx), _iptr(ry))
        return md.value, rx[: md.value].tolist(), ry[: md.value].tolist()

    def pnorm(self, nx, ny, dx, dy, a, p):
        return self._fn("pnorm", _d)(_i(nx), _i(ny), _d(dx), _d(dy), _ptr(np.ascontiguousarray(a)), _i(p))

    def inner_prod(self, nx, ny, dx, dy, a, b):
        return self._fn("inner_prod", _d)(_i(nx), _i(ny), _d(dx), _d(dy), _ptr(np.ascontiguousarray(a)),
                                          _ptr(np.ascontiguousarray(b)))

    def integrate2d(self, nx, ny, dx, dy, a, J=None):
        a = np.ascontiguousarray(a)
        if self.is_oracle:
            return self._fn("integrate2d", _d)(_i(nx), _i(ny), _d(dx), _d(dy), _ptr(a),
                                               None if J is None else _ptr(np.ascontiguousarray(J)))
        Jarr = np.ones_like(a) if J is None else np.ascontiguousarray(J)
        return self._fn("integrate2d", _d)(_i(nx), _i(ny), _d(dx), _d(dy), _ptr(a), _ptr(Jarr),
                                           _i(0 if J is None else 1))

    def integrate2d_bdry(self, nx, ny, dx, dy, xflux, yflux):
        return self._fn("integrate2d_bdry", _d)(_i(nx), _i(ny), _d(dx), _d(dy), _ptr(np.ascontiguousarray(xflux)),
                                                _ptr(np.ascontiguousarray(yflux)))

    def gpnorm(self, nx, ny, dx, dy, phi, p):
        """gpnorm of a one-ghost field on the box 0..nx-1 x 0..ny-1 (square boxes: the ghost layer)."""
        phi = np.ascontiguousarray(phi)
        if self.is_oracle:
            # the reference's columns jlo-1, jhi+1 and rows ilo-1, ihi+1 relative to the valid cells
            return self._fn("gpnorm", _d)(_i(nx), _i(ny), _d(dx), _d(dy), _ptr(phi), _i(p), _i(-1), _i(ny),
                                          _i(-1), _i(nx))
        return self._fn("gpnorm", _d)(_i(nx), _i(ny), _d(dx), _d(dy), _ptr(phi), _i(p))

    def bicgstab2(self, kind, geo, bc, homog, invd, tol, maxiters, maxrestarts, zerophi, rhs, phi, order=LEX):
        phi = phi.copy()
        args = [_i(kind), *self._geo_args(geo), *self._bc_args(bc), _i(homog), _ptr(np.ascontiguousarray(invd)),
                _d(tol), _i(maxiters), _i(maxrestarts), _i(zerophi), _ptr(np.ascontiguousarray(rhs)), _ptr(phi)]
        if self.is_oracle:
            args.append(_i(order))
        self._fn("bicgstab2")(*args)
        return phi

    def relax_jacobi(self, geo, bc, homog, invd, omega, tol, maxiters, zerophi, rhs, phi):
        """Returns (phi, relres history or None)."""
        phi = phi.copy()
        args = [*self._geo_args(geo), *self._bc_args(bc), _i(homog), _ptr(np.ascontiguousarray(invd)), _d(omega),
                _d(tol), _i(maxiters), _i(zerophi), _ptr(np.ascontiguousarray(rhs)), _ptr(phi)]
        if self.is_oracle:
            relres = np.zeros(maxiters + 1)
            n = self._fn("relax_jacobi", _i)(*args, _ptr(relres))
            return phi, relres[: n + 1]
        self._fn("relax_jacobi")(*args)
        return phi, None

    def bicgstab1(self, geo, bc, homog, invd, tol, maxiters, maxrestarts, zerophi, rhs, phi, order=LEX):
        """Returns (phi, relres history or None); the history is None from the
        oracle where the reference would stop on beta == 0."""
        phi = phi.copy()
        args = [*self._geo_args(geo), *self._bc_args(bc), _i(homog), _ptr(np.ascontiguousarray(invd)), _d(tol),
                _i(maxiters), _i(maxrestarts), _i(zerophi), _ptr(np.ascontiguousarray(rhs)), _ptr(phi)]
        if self.is_oracle:
            relres = np.zeros(maxiters + maxrestarts + 1)
            n = self._fn("bicgstab1", _i)(*args, _i(order), _ptr(relres))
            return phi, (relres[: n + 1] if n >= 0 else None)
        self._fn("bicgstab1")(*args)
        return phi, None

    def vcycle(self, kind, geo, bc, homog, tol, maxiters, phi, rhs, maxdepth=-1, numcycles=1,
               sdown=2, sup=2, sbottom=2, zerophi=False, order=LEX):
        """Returns (phi, rhs_after, relres history or None, iterations or None)."""
        phi = phi.copy()
        rhs = np.ascontiguousarray(rhs).copy()
        args = [_i(kind), *self._geo_args(geo), *self._bc_args(bc), _i(homog), _d(tol), _i(maxiters),
                _i(maxdepth), _i(numcycles), _i(sdown), _i(sup), _i(sbottom), _i(zerophi), _ptr(phi), _ptr(rhs)]
        if self.is_oracle:
            relres = np.zeros(maxiters + 1)
            iters = _i(0)
            args += [_i(order), _ptr(relres), ctypes.byref(iters)]
            status = self._fn("vcycle", _i)(*args)
            return phi, rhs, relres[: iters.value + 1], iters.value, status
        self._fn("vcycle")(*args)
        return phi, rhs, None, None, None


_cache = {}


def oracle():
    if "o" not in _cache:
        _cache["o"] = Checker(ORACLE_SO, "mgo_")
    return _cache["o"]


def reference(rb=False):
    key = "rrb" if rb else "r"
    if key not in _cache:
        path = REF_RB_SO if rb else REF_SO
        _cache[key] = Checker(path, "ref_")
    return _cache[key]


def reference_available(rb=False):
    return os.path.exists(REF_RB_SO if rb else REF_SO)
