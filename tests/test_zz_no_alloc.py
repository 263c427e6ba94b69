"""mgp_pnorm_noalloc (no device allocation; partial sums in the library's
static scratch array) against mgp_pnorm on the same field: two calls in a row,
the second on a smaller grid, for p = 0, 1, 2, all through the C ABI.

The input is a resident solver's rhs field (mgp_solver_field, the padded
layout the entry point takes); the output is written into a ghost corner of
that solver's phi field and read back with mgp_solver_get_phi.  The norms must
equal mgp_pnorm's bit for bit: the kernels and the row blocks are the same."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_pnorm_noalloc_matches_pnorm_twice():
    from mgpoisson2d_amd import capi
    from mgpoisson2d_amd import geometry as G
    fn = capi.lib.mgp_pnorm_noalloc
    fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_int,
                   ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    fn.restype = ctypes.c_int
    rng = np.random.default_rng(11)
    bc = G.BoundaryConditions.uniform((1, 1, 1, 1))
    for nx, ny in ((1000, 700), (37, 19)):  # the second call is smaller: stale partials must not leak in
        geo = G.mapped_geometry(nx, ny, 0.0, 0.0)
        s = capi.Solver(geo, bc, method=capi.METHOD_VCYCLE, ordering=capi.ORDER_LEX)
        a = rng.standard_normal((ny, nx))
        s.set_rhs(a)
        ptr, pitch = ctypes.POINTER(ctypes.c_double)(), ctypes.c_int(0)
        assert capi.lib.mgp_solver_field(s._h, 1, ctypes.byref(ptr), ctypes.byref(pitch)) == 0
        phi, phi_pitch = ctypes.POINTER(ctypes.c_double)(), ctypes.c_int(0)
        assert capi.lib.mgp_solver_field(s._h, 0, ctypes.byref(phi), ctypes.byref(phi_pitch)) == 0
        rhs_addr = ctypes.cast(ptr, ctypes.c_void_p).value
        out_addr = ctypes.cast(phi, ctypes.c_void_p).value + 8 * 31  # phi's ghost corner (-1, -1)
        for p in (0, 1, 2):
            print(f"pnorm_noalloc nx={nx} ny={ny} p={p}", flush=True)
            status = fn(nx, ny, geo.dx, geo.dy, rhs_addr, pitch.value, p, out_addr, s.stream())
            assert status == 0, capi.lib.mgp_last_error().decode()
            capi.device_synchronize()
            got = float(s.get_phi()[0, 0])
            want = capi.pnorm(nx, ny, geo.dx, geo.dy, a, p)
            assert got == want, (nx, ny, p, got, want)
        # arguments outside the documented range are rejected before any device work
        assert fn(nx, ny, geo.dx, geo.dy, rhs_addr, nx, 0, out_addr, s.stream()) == 1  # pitch too small
        assert fn(nx, ny, geo.dx, geo.dy, rhs_addr, pitch.value, 3, out_addr, s.stream()) == 1  # p > 2
        s.close()
