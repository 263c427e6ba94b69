"""ctypes binding of include/mgp2d.h (the library's C ABI).

The product path: every function here calls the HIP library
``mgpoisson2d_amd/libmgp2d.so``.  There is no CPU fallback; importing this
module raises if the library has not been built.
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from .geometry import BoundaryConditions, Geometry

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmgp2d.so")

MGP_OK = 0
CONVERGED, DIVERGED, MAXITERS = 0, 1, 2
ORDER_LEX, ORDER_RB, ORDER_MC = 0, 1, 2
METHOD_VCYCLE, METHOD_DC_VCYCLE, METHOD_KRYLOV_MG, METHOD_FMG, METHOD_KRYLOV_MG32 = 0, 1, 2, 3, 4
LINES_AUTO, LINES_X, LINES_Y = 0, 1, 2
REFINE_FP64, REFINE_DD = 0, 1
KRYLOV_AUTO, KRYLOV_BICGSTAB, KRYLOV_FGMRES = 0, 1, 2
MAP_TERRAIN, MAP_WARPED = 0, 1

_d = ctypes.c_double
_i = ctypes.c_int
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class MgpError(RuntimeError):
    pass


class _Geo(ctypes.Structure):
    _fields_ = [("nx", _i), ("ny", _i), ("dx", _d), ("dy", _d),
                ("J", _dp), ("Jxx", _dp), ("Jxy", _dp), ("Jyx", _dp), ("Jyy", _dp)]


class _Bc(ctypes.Structure):
    _fields_ = [("type", _i * 4), ("mode", _i * 4), ("data", _dp * 4)]


class Mapping(ctypes.Structure):
    """mgp_mapping: MAP_TERRAIN (the reference's test geometry) or MAP_WARPED;
    L, H <= 0 mean nx*dx and ny*dy."""
    _fields_ = [("kind", _i), ("amp", _d), ("beta", _d), ("L", _d), ("H", _d)]


class SolverOpts(ctypes.Structure):
    _fields_ = [("method", _i), ("ordering", _i), ("numcycles", _i), ("smooth_down", _i),
                ("smooth_up", _i), ("smooth_bottom", _i), ("maxdepth", _i), ("homog", _i), ("use_graph", _i),
                ("lines", _i), ("refine", _i), ("order", _i), ("krylov", _i), ("restart", _i)]

    def __init__(self, method=2, ordering=2, numcycles=1, smooth_down=2, smooth_up=2, smooth_bottom=2, maxdepth=-1,
                 homog=1, use_graph=1, lines=0, refine=0, order=2, krylov=0, restart=0):
        super().__init__(method, ordering, numcycles, smooth_down, smooth_up, smooth_bottom, maxdepth, homog,
                         use_graph, lines, refine, order, krylov, restart)


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m mgpoisson2d_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    lib.mgp_last_error.restype = ctypes.c_char_p
    lib.mgp_status_string.restype = ctypes.c_char_p
    lib.mgp_solver_stream.restype = ctypes.c_void_p
    lib.mgp_server_footprint.restype = ctypes.c_double
    return lib


lib = _load()

# Every exported symbol of include/mgp2d.h (checked by the CPU test suite).
EXPORTS = (
    "mgp_status_string", "mgp_last_error", "mgp_device_count", "mgp_fill_ghosts",
    "mgp_compute_laplacian", "mgp_compute_residual", "mgp_compute_invdiags", "mgp_relax_gs",
    "mgp_dc_compute_laplacian", "mgp_dc_compute_invdiags", "mgp_dc_relax_gs", "mgp_dc_rb_tiles", "mgp_dc_rb_tile_shapes", "mgp_restrict",
    "mgp_prolong", "mgp_create_mgschedule", "mgp_pnorm", "mgp_solve_bicgstab2", "mgp_vcycle",
    "mgp_solver_default_opts", "mgp_solver_create", "mgp_solver_destroy", "mgp_solver_set_rhs",
    "mgp_solver_set_phi", "mgp_solver_get_phi", "mgp_solver_copy_rhs_from", "mgp_solver_solve",
    "mgp_solver_field", "mgp_solver_stream", "mgp_solver_last_cycles", "mgp_solver_time_kernel",
    "mgp_device_synchronize", "mgp_xline_relax", "mgp_compute_pd", "mgp_compute_grad", "mgp_compute_div",
    "mgp_fill_boundary_fluxes", "mgp_solver_set_flux", "mgp_solver_get_flux", "mgp_solver_project",
    "mgp_relax_jacobi", "mgp_solve_bicgstab", "mgp_metric_fill", "mgp_solver_create_mapped",
    "mgp_inner_prod", "mgp_integrate2d", "mgp_extrapolate_ghosts", "mgp_batch_create",
    "mgp_batch_create_mapped", "mgp_batch_destroy", "mgp_batch_count", "mgp_batch_set_rhs", "mgp_batch_get_phi",
    "mgp_batch_solve", "mgp_batch_last_cycles", "mgp_server_create", "mgp_server_destroy", "mgp_server_footprint",
    "mgp_server_submit", "mgp_server_step", "mgp_server_result", "mgp_integrate2d_bdry", "mgp_gpnorm",
    "mgp_batch_time_kernel", "mgp_compute_residual_dd", "mgp_lap4", "mgp_metric4_fill", "mgp_solver_get_phi_lo",
    "mgp_solver_project_mean", "mgp_pnorm_noalloc", "mgp_server_add_slot", "mgp_server_evict",
    "mgp_solver_copy_phi_from", "mgp_solver_residual", "mgp_batch_tuning", "mgp_batch_set_tuning",
    "mgp_batch_tune",
)


def _check(status):
    if status != MGP_OK:
        msg = lib.mgp_last_error().decode()
        raise MgpError(f"{lib.mgp_status_string(status).decode()}: {msg}")


def _f64(a, shape=None, name="array"):
    """a as a C-contiguous float64 array; with shape given, raises ValueError
    unless a has exactly that shape (the C side sizes every copy from the
    grid, so a wrong shape would read or write past the caller's buffer)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {a.shape}")
    return a


def _cc(nx, ny):
    return (ny, nx)


def _gh(nx, ny):
    return (ny + 2, nx + 2)


def _ptr(a):
    return a.ctypes.data_as(_dp)


class _Args:
    """Keeps numpy buffers alive for the duration of one call."""

    def __init__(self):
        self.keep = []

    def geo(self, g: Geometry):
        nx, ny = g.nx, g.ny
        shapes = ((ny, nx), (ny, nx + 1), (ny, nx + 1), (ny + 1, nx), (ny + 1, nx))
        names = ("J", "Jxx", "Jxy", "Jyx", "Jyy")
        arrs = [_f64(a, sh, nm) for a, sh, nm in zip(g.arrays(), shapes, names)]
        self.keep += arrs
        return ctypes.byref(_Geo(g.nx, g.ny, g.dx, g.dy, *[_ptr(a) for a in arrs]))

    def bc(self, b: BoundaryConditions, nx=None, ny=None):
        data = [_f64(d).ravel() for d in b.data]
        for k in range(4):
            need = 1
            if int(b.modes[k]) == 1 and nx is not None:  # BCMODE_NONUNIFORM: one value per face cell
                need = ny if k < 2 else nx
            if data[k].size < need:
                raise ValueError(f"BC data of side {k} needs {need} values, got {data[k].size}")
        self.keep += data
        s = _Bc()
        for k in range(4):
            s.type[k] = int(b.types[k])
            s.mode[k] = int(b.modes[k])
            s.data[k] = _ptr(data[k])
        return ctypes.byref(s)


def device_count():
    return lib.mgp_device_count()


def fill_ghosts(geo, bc, homog, do_neum, phi):
    a = _Args()
    phi = _f64(phi, _gh(geo.nx, geo.ny), "phi").copy()
    _check(lib.mgp_fill_ghosts(a.geo(geo), a.bc(bc, geo.nx, geo.ny), _i(homog), _i(do_neum), _ptr(phi)))
    return phi


def compute_laplacian(geo, bc, homog, phi, jscale=False):
    a = _Args()
    phi = _f64(phi, _gh(geo.nx, geo.ny), "phi").copy()
    lap = np.zeros((geo.ny, geo.nx))
    _check(lib.mgp_compute_laplacian(a.geo(geo), a.bc(bc, geo.nx, geo.ny), _i(homog), _i(jscale), _ptr(phi), _ptr(lap)))
    return lap, phi


def compute_residual(geo, bc, homog, rhs, phi):
    a = _Args()
    phi = _f64(phi, _gh(geo.nx, geo.ny), "phi").copy()
    rhs = _f64(rhs, _cc(geo.nx, geo.ny), "rhs")
    res = np.zeros((geo.ny, geo.nx))
    _check(lib.mgp_compute_residual(a.geo(geo), a.bc(bc, geo.nx, geo.ny), _i(homog), _ptr(rhs), _ptr(phi), _ptr(res)))
    return res, phi


def compute_invdiags(geo, bc, homog):
    a = _Args()
    out = np.zeros((geo.ny, geo.nx))
    _check(lib.mgp_compute_invdiags(a.geo(geo), a.bc(bc, geo.nx, geo.ny), _i(homog), _ptr(out)))
    return out


def relax_gs(geo, bc, homog, invd, omega, maxiters, rhs, phi, ordering=ORDER_LEX):
    a = _Args()
    phi = _f64(phi, _gh(geo.nx, geo.ny), "phi").copy()
    invd, rhs = _f64(invd, _cc(geo.nx, geo.ny), "invdiags"), _f64(rhs, _cc(geo.nx, geo.ny), "rhs")
    _check(lib.mgp_relax_gs(a.geo(geo), a.bc(bc, geo.nx, geo.ny), _i(homog), _ptr(invd), _d(omega), _i(maxiters),
                            _ptr(rhs), _ptr(phi), _i(ordering)))
    return phi


def xline_relax(geo, bc, rhs, phi, sweeps):
    """x-line zebra sweeps of the homogeneous operator (the fast solver's smoother)."""
    a = _Args()
    phi = _f64(phi, _gh(geo.nx, geo.ny), "phi").copy()
    rhs = _f64(rhs, _cc(geo.nx, geo.ny), "rhs")
    _check(lib.mgp_xline_relax(a.geo(geo), a.bc(bc, geo.nx, geo.ny), _ptr(rhs), _ptr(phi), _i(sweeps)))
    return phi


def dc_compute_laplacian(geo, bc, homog, phi):
    a = _Args()
    phi = _f64(phi, _gh(geo.nx, geo.ny), "phi").copy()
    lap = np.zeros((geo.ny, geo.nx))
    _check(lib.mgp_dc_compute_laplacian(a.geo(geo), a.bc(bc, geo.nx, geo.ny), _i(homog), _ptr(phi), _ptr(lap)))
    return lap, phi


def dc_compute_invdiags(geo):
    a = _Args()
    out = np.zeros((geo.ny, geo.nx))
    _check(lib.mgp_dc_compute_invdiags(a.geo(geo), _ptr(out)))
    return out


def dc_relax_gs(geo, bc, homog, invd, omega, maxiters, rhs, phi, ordering=ORDER_LEX):
    a = _Args()
    phi = _f64(phi, _gh(geo.nx, geo.ny), "phi").copy()
    invd, rhs = _f64(invd, _cc(geo.nx, geo.ny), "invdiags"), _f64(rhs, _cc(geo.nx, geo.ny), "rhs")
    _check(lib.mgp_dc_relax_gs(a.geo(geo), a.bc(bc, geo.nx, geo.ny), _i(homog), _ptr(invd), _d(omega), _i(maxiters),
                               _ptr(rhs), _ptr(phi), _i(ordering)))
    return phi


def dc_rb_tiles(choice):
    """Process-wide switch of the red-black DC sweeps: -1 tile shape timed per
    level (default), 0 one launch per colour, k >= 1 tile shape k.  Returns the
    previous setting; results are bit-identical for every choice."""
    prev = ctypes.c_int(0)
    _check(lib.mgp_dc_rb_tiles(_i(choice), ctypes.byref(prev)))
    return prev.value


def dc_rb_tile_shapes():
    """Number of tile shapes dc_rb_tiles accepts (1 .. n)."""
    return int(lib.mgp_dc_rb_tile_shapes())


def compute_pd(nx, ny, dx, dy, phi, direction, nodedir):
    """Poisson2D::compute_pd: pd is (ny, nx+1) for nodedir 1, (ny+1, nx) for nodedir 2."""
    phi = _f64(phi, _gh(nx, ny), "phi")
    pd = np.zeros((ny, nx + 1) if nodedir == 1 else (ny + 1, nx))
    _check(lib.mgp_compute_pd(_i(nx), _i(ny), _d(dx), _d(dy), _ptr(phi), _i(direction), _i(nodedir), _ptr(pd)))
    return pd


def compute_grad(geo, bc, homog, phi):
    """Poisson2D::compute_grad: returns (phi with ghosts filled, xflux (ny, nx+1), yflux (ny+1, nx))."""
    a = _Args()
    phi = _f64(phi, _gh(geo.nx, geo.ny), "phi").copy()
    xf = np.zeros((geo.ny, geo.nx + 1))
    yf = np.zeros((geo.ny + 1, geo.nx))
    _check(lib.mgp_compute_grad(a.geo(geo), a.bc(bc, geo.nx, geo.ny), _i(homog), _ptr(phi), _ptr(xf), _ptr(yf)))
    return phi, xf, yf


def compute_div(nx, ny, dx, dy, xflux, yflux):
    xflux, yflux = _f64(xflux, (ny, nx + 1), "xflux"), _f64(yflux, (ny + 1, nx), "yflux")
    div = np.zeros((ny, nx))
    _check(lib.mgp_compute_div(_i(nx), _i(ny), _d(dx), _d(dy), _ptr(xflux), _ptr(yflux), _ptr(div)))
    return div


def fill_boundary_fluxes(nx, ny, bc, homog, xflux, yflux):
    a = _Args()
    xf, yf = _f64(xflux, (ny, nx + 1), "xflux").copy(), _f64(yflux, (ny + 1, nx), "yflux").copy()
    _check(lib.mgp_fill_boundary_fluxes(_i(nx), _i(ny), a.bc(bc, nx, ny), _i(homog), _ptr(xf), _ptr(yf)))
    return xf, yf


def metric_fill(mapping, nx, ny, dx, dy):
    """Metrics of a mapping computed on the GPU: returns a Geometry."""
    arrs = [np.zeros((ny, nx)), np.zeros((ny, nx + 1)), np.zeros((ny, nx + 1)), np.zeros((ny + 1, nx)),
            np.zeros((ny + 1, nx))]
    _check(lib.mgp_metric_fill(ctypes.byref(mapping), _i(nx), _i(ny), _d(dx), _d(dy), *[_ptr(a) for a in arrs]))
    return Geometry(nx, ny, dx, dy, *arrs)


def restrict(kind, fnx, fny, fine, cnx, cny):
    shape = {0: (cny, cnx), 1: (cny, cnx + 1), 2: (cny + 1, cnx)}[kind]
    crse = np.zeros(shape)
    fshape = {0: (fny, fnx), 1: (fny, fnx + 1), 2: (fny + 1, fnx)}[kind]
    fine = _f64(fine, fshape, "fine")
    _check(lib.mgp_restrict(_i(kind), _i(fnx), _i(fny), _ptr(fine), _i(cnx), _i(cny), _ptr(crse)))
    return crse


def prolong(fgeo, cgeo, bc_types, bc_modes, order, fine, crse):
    a = _Args()
    fine = _f64(fine, _gh(fgeo.nx, fgeo.ny), "fine").copy()
    crse = _f64(crse, _gh(cgeo.nx, cgeo.ny), "crse").copy()
    t = (_i * 4)(*[int(x) for x in bc_types])
    m = (_i * 4)(*[int(x) for x in bc_modes])
    _check(lib.mgp_prolong(a.geo(fgeo), a.geo(cgeo), t, m, _i(order), _ptr(fine), _ptr(crse)))
    return fine, crse


def create_mgschedule(nx, ny, dx, dy, maxdepth=-1):
    rx = (_i * 64)()
    ry = (_i * 64)()
    md = _i(0)
    _check(lib.mgp_create_mgschedule(_i(nx), _i(ny), _d(dx), _d(dy), _i(maxdepth), ctypes.byref(md), rx, ry))
    return md.value, list(rx[: md.value]), list(ry[: md.value])


def inner_prod(nx, ny, dx, dy, a, b):
    out = _d(0.0)
    a, b = _f64(a, (ny, nx), "a"), _f64(b, (ny, nx), "b")
    _check(lib.mgp_inner_prod(_i(nx), _i(ny), _d(dx), _d(dy), _ptr(a), _ptr(b), ctypes.byref(out)))
    return out.value


def integrate2d(nx, ny, dx, dy, a, J=None):
    """arrayutils integrate2d: sum dx dy J a (J given) or dx dy a."""
    out = _d(0.0)
    a = _f64(a, (ny, nx), "a")
    J = None if J is None else _f64(J, (ny, nx), "J")
    _check(lib.mgp_integrate2d(_i(nx), _i(ny), _d(dx), _d(dy), _ptr(a), None if J is None else _ptr(J),
                               ctypes.byref(out)))
    return out.value


def integrate2d_bdry(nx, ny, dx, dy, xflux, yflux):
    """arrayutils integrate2d_bdry: the net outward flux, xflux (ny, nx+1), yflux (ny+1, nx)."""
    out = _d(0.0)
    xflux, yflux = _f64(xflux, (ny, nx + 1), "xflux"), _f64(yflux, (ny + 1, nx), "yflux")
    _check(lib.mgp_integrate2d_bdry(_i(nx), _i(ny), _d(dx), _d(dy), _ptr(xflux), _ptr(yflux), ctypes.byref(out)))
    return out.value


def gpnorm(ilo, ihi, jlo, jhi, dx, dy, phi, p=0):
    """arrayutils gpnorm of a one-ghost field phi (jhi-jlo+3, ihi-ilo+3) on the box ilo..ihi x jlo..jhi."""
    out = _d(0.0)
    phi = _f64(phi, (jhi - jlo + 3, ihi - ilo + 3), "phi")
    _check(lib.mgp_gpnorm(_i(ilo), _i(ihi), _i(jlo), _i(jhi), _d(dx), _d(dy), _ptr(phi), _i(p), ctypes.byref(out)))
    return out.value


def compute_residual_dd(geo, bc, rhs, phi_hi, phi_lo):
    """rhs - L (phi_hi + phi_lo) in double-double, rounded once (homogeneous Dirichlet / Neumann)."""
    a = _Args()
    rhs = _f64(rhs, _cc(geo.nx, geo.ny), "rhs")
    hi, lo = _f64(phi_hi, _gh(geo.nx, geo.ny), "phi_hi"), _f64(phi_lo, _gh(geo.nx, geo.ny), "phi_lo")
    res = np.zeros((geo.ny, geo.nx))
    _check(lib.mgp_compute_residual_dd(a.geo(geo), a.bc(bc, geo.nx, geo.ny), _ptr(rhs), _ptr(hi), _ptr(lo),
                                       _ptr(res)))
    return res


def _metric4_shapes(nx, ny):
    return [(ny, nx + 1)] * 4 + [(ny + 1, nx)] * 4


def lap4(nx, ny, dx, dy, metrics, phi):
    """Fourth-order operator on cell averages phi (ny, nx); metrics: the 8 arrays of metric4_fill."""
    ms = [_f64(m, sh, f"metrics[{k}]") for k, (m, sh) in enumerate(zip(metrics, _metric4_shapes(nx, ny)))]
    arr = (_dp * 8)(*[_ptr(m) for m in ms])
    phi = _f64(phi, (ny, nx), "phi")
    out = np.zeros((ny, nx))
    _check(lib.mgp_lap4(_i(nx), _i(ny), _d(dx), _d(dy), arr, _ptr(phi), _ptr(out)))
    return out


def metric4_fill(mapping, nx, ny, dx, dy):
    """The 8 face-averaged metric arrays of the fourth-order operator, generated on the GPU."""
    ms = [np.zeros(sh) for sh in _metric4_shapes(nx, ny)]
    arr = (_dp * 8)(*[_ptr(m) for m in ms])
    _check(lib.mgp_metric4_fill(ctypes.byref(mapping), _i(nx), _i(ny), _d(dx), _d(dy), arr))
    return ms


def extrapolate_ghosts(nx, ny, order, phi):
    phi = _f64(phi, _gh(nx, ny), "phi").copy()
    _check(lib.mgp_extrapolate_ghosts(_i(nx), _i(ny), _i(order), _ptr(phi)))
    return phi


def pnorm(nx, ny, dx, dy, arr, p=2):
    out = _d(0.0)
    arr = _f64(arr, (ny, nx), "arr")
    _check(lib.mgp_pnorm(_i(nx), _i(ny), _d(dx), _d(dy), _ptr(arr), _i(p), ctypes.byref(out)))
    return out.value


def relax_jacobi(geo, bc, homog, invd, omega, tol, maxiters, zerophi, rhs, phi, ordering=ORDER_LEX):
    """Poisson2D::relax_jacobi: returns (phi, relres history (when tol > 0), sweeps)."""
    a = _Args()
    phi = _f64(phi, _gh(geo.nx, geo.ny), "phi").copy()
    invd, rhs = _f64(invd, _cc(geo.nx, geo.ny), "invdiags"), _f64(rhs, _cc(geo.nx, geo.ny), "rhs")
    relres = np.zeros(max(maxiters, 0) + 1)
    n = _i(0)
    _check(lib.mgp_relax_jacobi(a.geo(geo), a.bc(bc, geo.nx, geo.ny), _i(homog), _ptr(invd), _d(omega), _d(tol), _i(maxiters),
                                _i(zerophi), _ptr(rhs), _ptr(phi), _i(ordering), _ptr(relres), ctypes.byref(n)))
    return phi, relres[: n.value + 1], n.value


def solve_bicgstab(geo, bc, homog, invd, tol, maxiters, maxrestarts, zerophi, rhs, phi, ordering=ORDER_LEX):
    """Poisson2D::solve_bicgstab (version 1): returns (phi, relres history)."""
    a = _Args()
    phi = _f64(phi, _gh(geo.nx, geo.ny), "phi").copy()
    invd, rhs = _f64(invd, _cc(geo.nx, geo.ny), "invdiags"), _f64(rhs, _cc(geo.nx, geo.ny), "rhs")
    relres = np.zeros(maxiters + maxrestarts + 1)
    last = _i(0)
    _check(lib.mgp_solve_bicgstab(a.geo(geo), a.bc(bc, geo.nx, geo.ny), _i(homog), _ptr(invd), _d(tol), _i(maxiters),
                                  _i(maxrestarts), _i(zerophi), _ptr(rhs), _ptr(phi), _i(ordering), _ptr(relres),
                                  ctypes.byref(last)))
    return phi, relres[: last.value + 1]


def solve_bicgstab2(kind, geo, bc, homog, invd, tol, maxiters, maxrestarts, zerophi, rhs, phi,
                    ordering=ORDER_LEX):
    a = _Args()
    phi = _f64(phi, _gh(geo.nx, geo.ny), "phi").copy()
    invd, rhs = _f64(invd, _cc(geo.nx, geo.ny), "invdiags"), _f64(rhs, _cc(geo.nx, geo.ny), "rhs")
    _check(lib.mgp_solve_bicgstab2(_i(kind), a.geo(geo), a.bc(bc, geo.nx, geo.ny), _i(homog), _ptr(invd), _d(tol),
                                   _i(maxiters), _i(maxrestarts), _i(zerophi), _ptr(rhs), _ptr(phi),
                                   _i(ordering)))
    return phi


def vcycle(kind, geo, bc, homog, tol, maxiters, phi, rhs, maxdepth=-1, numcycles=1, smooth_down=2,
           smooth_up=2, smooth_bottom=2, zerophi=False, ordering=ORDER_LEX):
    """Reference vcycle (kind 0) / DC vcycle (kind 1) / fmg (kind 2) on the GPU.
    Returns (phi, rhs, relres history, iterations, outcome)."""
    a = _Args()
    phi = _f64(phi, _gh(geo.nx, geo.ny), "phi").copy()
    rhs = _f64(rhs, _cc(geo.nx, geo.ny), "rhs").copy()
    relres = np.zeros(maxiters + 1)
    iters, outcome = _i(0), _i(0)
    _check(lib.mgp_vcycle(_i(kind), a.geo(geo), a.bc(bc, geo.nx, geo.ny), _i(homog), _d(tol), _i(maxiters), _i(maxdepth),
                          _i(numcycles), _i(smooth_down), _i(smooth_up), _i(smooth_bottom), _i(zerophi),
                          _ptr(phi), _ptr(rhs), _i(ordering), _ptr(relres), ctypes.byref(iters),
                          ctypes.byref(outcome)))
    return phi, rhs, relres[: iters.value + 1], iters.value, outcome.value


class Solver:
    """A problem resident in HBM (mgp_solver): hierarchy, phi and rhs stay on
    the device between solves."""

    def __init__(self, geo, bc, method=METHOD_KRYLOV_MG, ordering=ORDER_MC, numcycles=1, smooth_down=2,
                 smooth_up=2, smooth_bottom=2, maxdepth=-1, homog=True, use_graph=True, mapping=None,
                 lines=LINES_AUTO, refine=REFINE_FP64, order=2, krylov=KRYLOV_AUTO, restart=0):
        """geo: a Geometry (metrics uploaded), or with mapping given, any object
        with nx, ny, dx, dy (metrics generated on the GPU).  lines, refine,
        order, krylov, restart: the KRYLOV_MG options of mgp_solver_opts
        (numcycles: V-cycles per preconditioner application there)."""
        self.geo = geo
        opts = SolverOpts(method, ordering, numcycles, smooth_down, smooth_up, smooth_bottom, maxdepth,
                          int(homog), int(use_graph), lines, refine, order, krylov, restart)
        a = _Args()
        self._h = ctypes.c_void_p()
        if mapping is None:
            _check(lib.mgp_solver_create(ctypes.byref(self._h), a.geo(geo), a.bc(bc, geo.nx, geo.ny), ctypes.byref(opts)))
        else:
            _check(lib.mgp_solver_create_mapped(ctypes.byref(self._h), ctypes.byref(mapping), _i(geo.nx),
                                                _i(geo.ny), _d(geo.dx), _d(geo.dy), a.bc(bc, geo.nx, geo.ny), ctypes.byref(opts)))

    def close(self):
        if self._h:
            lib.mgp_solver_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_rhs(self, rhs):
        rhs = _f64(rhs, _cc(self.geo.nx, self.geo.ny), "rhs")
        _check(lib.mgp_solver_set_rhs(self._h, _ptr(rhs)))

    def set_phi(self, phi):
        phi = _f64(phi, _gh(self.geo.nx, self.geo.ny), "phi")
        _check(lib.mgp_solver_set_phi(self._h, _ptr(phi)))

    def get_phi(self):
        phi = np.zeros((self.geo.ny + 2, self.geo.nx + 2))
        _check(lib.mgp_solver_get_phi(self._h, _ptr(phi)))
        return phi

    def get_phi_lo(self):
        """Low part of a double-double solution (refine=REFINE_DD): the solution is get_phi() + get_phi_lo()."""
        lo = np.zeros((self.geo.ny + 2, self.geo.nx + 2))
        _check(lib.mgp_solver_get_phi_lo(self._h, _ptr(lo)))
        return lo

    def project_mean(self):
        """rhs -= J sum(rhs)/sum(J) on the resident rhs (all-Neumann compatibility); returns the mean removed."""
        m = _d(0.0)
        _check(lib.mgp_solver_project_mean(self._h, ctypes.byref(m)))
        return m.value

    def copy_phi_from(self, src):
        """The resident solution of another solver of the same shape (device copy)."""
        _check(lib.mgp_solver_copy_phi_from(self._h, src._h))

    def residual(self):
        """(||rhs - A phi||, ||rhs||) on the resident fields, with the solver's refinement residual."""
        rn, rh = _d(0.0), _d(0.0)
        _check(lib.mgp_solver_residual(self._h, ctypes.byref(rn), ctypes.byref(rh)))
        return rn.value, rh.value

    def solve(self, tol=1e-10, maxiters=100, zerophi=True):
        relres = np.zeros(maxiters + 1)
        iters, outcome = _i(0), _i(0)
        _check(lib.mgp_solver_solve(self._h, _d(tol), _i(maxiters), _i(int(zerophi)), _ptr(relres),
                                    ctypes.byref(iters), ctypes.byref(outcome)))
        return relres[: iters.value + 1], iters.value, outcome.value

    def solve_fast(self, tol=1e-10, maxiters=100, zerophi=True):
        """Solve without copying the history out (timed path)."""
        iters, outcome = _i(0), _i(0)
        _check(lib.mgp_solver_solve(self._h, _d(tol), _i(maxiters), _i(int(zerophi)), None,
                                    ctypes.byref(iters), ctypes.byref(outcome)))
        return iters.value, outcome.value

    def set_flux(self, xflux, yflux):
        """Resident face fluxes: xflux (ny, nx+1), yflux (ny+1, nx)."""
        nx, ny = self.geo.nx, self.geo.ny
        xflux, yflux = _f64(xflux, (ny, nx + 1), "xflux"), _f64(yflux, (ny + 1, nx), "yflux")
        _check(lib.mgp_solver_set_flux(self._h, _ptr(xflux), _ptr(yflux)))

    def get_flux(self):
        xf = np.zeros((self.geo.ny, self.geo.nx + 1))
        yf = np.zeros((self.geo.ny + 1, self.geo.nx))
        _check(lib.mgp_solver_get_flux(self._h, _ptr(xf), _ptr(yf)))
        return xf, yf

    def project(self, tol=1e-10, maxiters=100):
        """Divergence cleaning of the resident fluxes: returns (max|J div F|
        before, after, solver iterations, outcome)."""
        dn = np.zeros(2)
        iters, outcome = _i(0), _i(0)
        _check(lib.mgp_solver_project(self._h, _d(tol), _i(maxiters), _ptr(dn), ctypes.byref(iters),
                                      ctypes.byref(outcome)))
        return dn[0], dn[1], iters.value, outcome.value

    def last_cycles(self):
        return lib.mgp_solver_last_cycles(self._h)

    def stream(self):
        return lib.mgp_solver_stream(self._h)

    def time_kernel(self, kernel=0, reps=20, cold=True):
        """Mean ms per launch and algorithmic bytes per launch of a finest-level
        kernel; cold: each launch after a cache flush (as inside a solve)."""
        ms, nbytes = _d(0.0), _d(0.0)
        _check(lib.mgp_solver_time_kernel(self._h, _i(kernel), _i(reps), _i(int(cold)), ctypes.byref(ms), ctypes.byref(nbytes)))
        return ms.value, nbytes.value


def device_synchronize():
    _check(lib.mgp_device_synchronize())


TUNING_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_cache.json")


def load_tuning(path=TUNING_CACHE):
    """The committed autotuner choices: [(nx, bc_class, fp32, threads), ...] ([] if absent)."""
    if not os.path.exists(path):
        return []
    with open(path) as fh:
        return [tuple(e) for e in json.load(fh)["entries"]]


def save_tuning(entries, path=TUNING_CACHE, note=""):
    """Merges entries into the cache at path (a new entry replaces one of the same key)."""
    merged = {(e[0], e[1], e[2]): e[3] for e in load_tuning(path)}
    for e in entries:
        merged[(int(e[0]), int(e[1]), int(e[2]))] = int(e[3])
    out = {"key": ["nx", "bc_class (bit k: side k Dirichlet)", "fp32", "threads"], "note": note,
           "entries": [[k[0], k[1], k[2], v] for k, v in sorted(merged.items())]}
    with open(path, "w") as fh:
        json.dump(out, fh, indent=0)


class Batch:
    """Many independent problems solved together (mgp_batch): each is solved as
    KRYLOV_MG solves it, with one GPU launch per level operation for all.

    geos: a list of Geometry (metrics uploaded), or with mappings given, a list
    of objects with nx, ny, dx, dy (metrics generated on the GPU from
    mappings[k]).  bcs: one BoundaryConditions per problem (Dirichlet or
    Neumann sides; homogeneous)."""

    def __init__(self, geos, bcs, mappings=None, smooth_down=2, smooth_up=2, smooth_bottom=2, maxdepth=-1,
                 use_graph=True, method=METHOD_KRYLOV_MG):
        """method: METHOD_KRYLOV_MG, or METHOD_KRYLOV_MG32 (FP32 preconditioner cycles)."""
        n = len(geos)
        self.shapes = [(g.ny, g.nx) for g in geos]
        opts = SolverOpts(method, ORDER_MC, 1, smooth_down, smooth_up, smooth_bottom, maxdepth, 1,
                          int(use_graph))
        a = _Args()
        bc_arr = (_Bc * n)()
        for k, b in enumerate(bcs):
            bc_arr[k] = a.bc(b, geos[k].nx, geos[k].ny)._obj
        self._h = ctypes.c_void_p()
        if mappings is None:
            geo_arr = (_Geo * n)()
            for k, g in enumerate(geos):
                geo_arr[k] = a.geo(g)._obj
            _check(lib.mgp_batch_create(ctypes.byref(self._h), _i(n), geo_arr, bc_arr, ctypes.byref(opts)))
        else:
            maps = (Mapping * n)(*mappings)
            nx = (_i * n)(*[g.nx for g in geos])
            ny = (_i * n)(*[g.ny for g in geos])
            dx = (_d * n)(*[g.dx for g in geos])
            dy = (_d * n)(*[g.dy for g in geos])
            _check(lib.mgp_batch_create_mapped(ctypes.byref(self._h), _i(n), maps, nx, ny, dx, dy, bc_arr,
                                               ctypes.byref(opts)))

    def __len__(self):
        return len(self.shapes)

    def close(self):
        if self._h:
            lib.mgp_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_rhs(self, k, rhs):
        rhs = _f64(rhs)
        if rhs.shape != self.shapes[k]:
            raise ValueError(f"rhs of problem {k} must have shape {self.shapes[k]}")
        _check(lib.mgp_batch_set_rhs(self._h, _i(k), _ptr(rhs)))

    def get_phi(self, k):
        phi = np.zeros(self.shapes[k])
        _check(lib.mgp_batch_get_phi(self._h, _i(k), _ptr(phi)))
        return phi

    def solve(self, tol=1e-10, maxiters=200, zerophi=True):
        """Returns (iterations, outcomes) per problem."""
        n = len(self)
        iters = np.zeros(n, dtype=np.int32)
        outc = np.zeros(n, dtype=np.int32)
        _check(lib.mgp_batch_solve(self._h, _d(tol), _i(maxiters), _i(int(zerophi)),
                                   iters.ctypes.data_as(ctypes.POINTER(_i)), outc.ctypes.data_as(ctypes.POINTER(_i))))
        return iters, outc

    def last_cycles(self):
        return lib.mgp_batch_last_cycles(self._h)

    def time_kernel(self, kernel=0, reps=10):
        """Mean ms and algorithmic bytes of one launch set of the finest-level
        line relaxation (even rows) of every problem (between solves only)."""
        ms, nbytes = _d(0.0), _d(0.0)
        _check(lib.mgp_batch_time_kernel(self._h, _i(kernel), _i(reps), ctypes.byref(ms), ctypes.byref(nbytes)))
        return ms.value, nbytes.value

    def tuning(self):
        """The line-solve shape per key present: [(nx, bc_class, fp32, threads), ...]."""
        n = _i(0)
        _check(lib.mgp_batch_tuning(self._h, None, _i(0), ctypes.byref(n)))
        tab = (_i * (4 * n.value))()
        _check(lib.mgp_batch_tuning(self._h, tab, n, ctypes.byref(n)))
        return [tuple(tab[4 * k:4 * k + 4]) for k in range(n.value)]

    def set_tuning(self, entries):
        """Applies (nx, bc_class, fp32, threads) entries matching problems of the
        batch; returns the number of problems whose shape changed."""
        entries = [tuple(int(v) for v in e) for e in entries]
        tab = (_i * (4 * len(entries)))(*[v for e in entries for v in e])
        changed = _i(0)
        _check(lib.mgp_batch_set_tuning(self._h, tab, _i(len(entries)), ctypes.byref(changed)))
        return changed.value

    def tune(self, reps=10):
        """Times the candidate line-solve shapes per key, applies the fastest,
        returns the choices (between solves only)."""
        n = len(self)
        tab = (_i * (4 * n))()
        cnt = _i(0)
        _check(lib.mgp_batch_tune(self._h, _i(reps), tab, _i(n), ctypes.byref(cnt)))
        return [tuple(tab[4 * k:4 * k + 4]) for k in range(min(cnt.value, n))]


class Server:
    """Serving loop over resident geometries (mgp_server): submit(slot, rhs)
    queues a solve; step() admits queued requests to idle slots and advances
    every running solve by one iteration; result(ticket) returns
    (phi, iterations, outcome) once finished, else None."""

    def __init__(self, geos, bcs, mappings, hbm_budget=0.0, smooth_down=2, smooth_up=2, smooth_bottom=2,
                 use_graph=True, method=METHOD_KRYLOV_MG):
        n = len(geos)
        self.shapes = [(g.ny, g.nx) for g in geos]
        opts = SolverOpts(method, ORDER_MC, 1, smooth_down, smooth_up, smooth_bottom, -1, 1,
                          int(use_graph))
        a = _Args()
        bc_arr = (_Bc * n)()
        for k, b in enumerate(bcs):
            bc_arr[k] = a.bc(b, geos[k].nx, geos[k].ny)._obj
        maps = (Mapping * n)(*mappings)
        nx = (_i * n)(*[g.nx for g in geos])
        ny = (_i * n)(*[g.ny for g in geos])
        dx = (_d * n)(*[g.dx for g in geos])
        dy = (_d * n)(*[g.dy for g in geos])
        self._h = ctypes.c_void_p()
        _check(lib.mgp_server_create(ctypes.byref(self._h), _i(n), maps, nx, ny, dx, dy, bc_arr, ctypes.byref(opts),
                                     _d(hbm_budget)))

    def close(self):
        if self._h:
            lib.mgp_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def footprint(self):
        return lib.mgp_server_footprint(self._h)

    def add_slot(self, geo, bc, mapping):
        """Admits a geometry while serving; raises MgpError (out of device memory) if its
        hierarchy does not fit the budget.  Returns the new slot's index."""
        a = _Args()
        slot = _i(0)
        _check(lib.mgp_server_add_slot(self._h, ctypes.byref(mapping), _i(geo.nx), _i(geo.ny), _d(geo.dx),
                                       _d(geo.dy), a.bc(bc, geo.nx, geo.ny), ctypes.byref(slot)))
        self.shapes.append((geo.ny, geo.nx))
        return slot.value

    def evict(self, slot):
        """Frees an idle slot's hierarchy and warm start."""
        _check(lib.mgp_server_evict(self._h, _i(slot)))

    def submit(self, slot, rhs, tol=1e-10, maxiters=400, warm=False):
        rhs = _f64(rhs)
        if rhs.shape != self.shapes[slot]:
            raise ValueError(f"rhs of slot {slot} must have shape {self.shapes[slot]}")
        t = ctypes.c_longlong(0)
        _check(lib.mgp_server_submit(self._h, _i(slot), _ptr(rhs), _d(tol), _i(maxiters), _i(int(warm)),
                                     ctypes.byref(t)))
        return t.value

    def step(self):
        """Returns (solves running on the device, requests waiting)."""
        r, q = _i(0), _i(0)
        _check(lib.mgp_server_step(self._h, ctypes.byref(r), ctypes.byref(q)))
        return r.value, q.value

    def result(self, ticket, slot):
        phi = np.zeros(self.shapes[slot])
        it, oc, done = _i(0), _i(0), _i(0)
        _check(lib.mgp_server_result(self._h, ctypes.c_longlong(ticket), _ptr(phi), ctypes.byref(it),
                                     ctypes.byref(oc), ctypes.byref(done)))
        return (phi, it.value, oc.value) if done.value else None
