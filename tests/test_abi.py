"""The C-ABI library loads and exports every entry point include/mgp2d.h
declares; host-only entry points answer without a GPU."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "mgp2d.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mgp_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from mgpoisson2d_amd import capi
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(capi.lib, n)]
    assert not missing, missing
    assert set(capi.EXPORTS) == set(names)


def test_host_only_entry_points():
    from mgpoisson2d_amd import capi
    assert capi.lib.mgp_status_string(0) == b"ok"
    # the schedule is host code (MGPoisson2D.f90:38-129)
    assert capi.create_mgschedule(64, 64, 1 / 64, 1 / 64) == (4, [2, 2, 2, 2], [2, 2, 2, 2])
    md, rx, ry = capi.create_mgschedule(8192, 2048, 1 / 8192, 1 / 2048)
    assert (rx[:3], ry[:3]) == ([2, 2, 2], [1, 1, 2])


def test_dc_rb_tile_switch_is_host_only():
    import pytest
    from mgpoisson2d_amd import capi
    n = capi.dc_rb_tile_shapes()
    assert n >= 2
    prev = capi.dc_rb_tiles(0)
    try:
        assert capi.dc_rb_tiles(n) == 0
        assert capi.dc_rb_tiles(-1) == n
        for bad in (-2, n + 1):
            with pytest.raises(capi.MgpError):
                capi.dc_rb_tiles(bad)
        assert capi.dc_rb_tiles(-1) == -1  # a rejected choice leaves the setting alone
    finally:
        capi.dc_rb_tiles(prev)


def test_bad_arguments_rejected_without_device_work():
    import numpy as np
    import pytest
    from mgpoisson2d_amd import capi, geometry as G
    geo = G.mapped_geometry(2, 2)  # too small: the reference stencils need 3 cells
    with pytest.raises(capi.MgpError):
        capi.compute_invdiags(geo, G.BoundaryConditions.uniform((1, 1, 1, 1)), 1)
    with pytest.raises(capi.MgpError):
        capi.restrict(0, 10, 10, np.zeros((10, 10)), 3, 5)
    with pytest.raises(capi.MgpError):
        capi.compute_pd(2, 8, 0.5, 0.125, np.zeros((10, 4)), 1, 1)
    with pytest.raises(capi.MgpError):
        capi.compute_pd(8, 8, 0.125, 0.125, np.zeros((10, 10)), 3, 1)
    with pytest.raises(capi.MgpError):
        capi.compute_div(8, 8, 0.0, 0.125, np.zeros((8, 9)), np.zeros((9, 8)))
    with pytest.raises(capi.MgpError):
        capi.extrapolate_ghosts(8, 8, 4, np.zeros((10, 10)))
    with pytest.raises(capi.MgpError):
        capi.pnorm(8, 8, 0.125, 0.125, np.zeros((8, 8)), -1)
    with pytest.raises(capi.MgpError):
        capi.metric_fill(capi.Mapping(7, 0.0, 0.0, 0.0, 0.0), 8, 8, 0.125, 0.125)


def test_batch_and_server_arguments_rejected_without_device_work():
    import ctypes
    import pytest
    from mgpoisson2d_amd import capi
    h = ctypes.c_void_p()
    # no problems, or no BCs: rejected before any HIP call
    with pytest.raises(capi.MgpError):
        capi._check(capi.lib.mgp_batch_create(ctypes.byref(h), 0, None, None, None))
    with pytest.raises(capi.MgpError):
        capi._check(capi.lib.mgp_server_create(ctypes.byref(h), 0, None, None, None, None, None, None, None,
                                               ctypes.c_double(0.0)))
    with pytest.raises(capi.MgpError):
        capi._check(capi.lib.mgp_batch_solve(None, ctypes.c_double(1e-10), 10, 1, None, None))
    with pytest.raises(capi.MgpError):
        capi._check(capi.lib.mgp_server_step(None, None, None))
    done = ctypes.c_int(7)
    with pytest.raises(capi.MgpError):
        capi._check(capi.lib.mgp_server_result(None, ctypes.c_longlong(1), None, None, None, ctypes.byref(done)))
    assert capi.lib.mgp_batch_count(None) == 0 and capi.lib.mgp_batch_last_cycles(None) == 0
    assert capi.lib.mgp_server_footprint(None) == 0.0


def test_wrong_shapes_raise_before_the_library_is_called():
    """Every wrapper checks its arrays against the grid: the C side sizes its
    copies from nx, ny, so a wrong shape must never reach it."""
    import numpy as np
    import pytest
    from mgpoisson2d_amd import capi, geometry as G
    nx, ny = 12, 8
    geo = G.mapped_geometry(nx, ny, 0.1, 2.0)
    bc = G.BoundaryConditions.uniform((1, 1, 0, 0))
    good_phi, good_cc = np.zeros((ny + 2, nx + 2)), np.zeros((ny, nx))
    bad = [
        lambda: capi.fill_ghosts(geo, bc, 1, 0, np.zeros((ny, nx))),
        lambda: capi.compute_laplacian(geo, bc, 1, np.zeros((nx + 2, ny + 2))),
        lambda: capi.compute_residual(geo, bc, 1, np.zeros((nx, ny)), good_phi),
        lambda: capi.relax_gs(geo, bc, 1, good_cc, 1.0, 1, good_cc, np.zeros((ny + 1, nx + 2))),
        lambda: capi.relax_gs(geo, bc, 1, np.zeros(nx * ny), 1.0, 1, good_cc, good_phi),
        lambda: capi.dc_relax_gs(geo, bc, 1, good_cc, 1.0, 1, np.zeros((ny, nx + 1)), good_phi),
        lambda: capi.xline_relax(geo, bc, good_cc, good_phi[:-1], 1),
        lambda: capi.compute_grad(geo, bc, 1, good_cc),
        lambda: capi.compute_div(nx, ny, 1.0, 1.0, np.zeros((ny + 1, nx)), np.zeros((ny + 1, nx))),
        lambda: capi.fill_boundary_fluxes(nx, ny, bc, 1, np.zeros((ny, nx)), np.zeros((ny + 1, nx))),
        lambda: capi.restrict(0, nx, ny, np.zeros((ny, nx + 1)), nx // 2, ny // 2),
        lambda: capi.pnorm(nx, ny, 1.0, 1.0, np.zeros((nx, ny)), 2),
        lambda: capi.inner_prod(nx, ny, 1.0, 1.0, good_cc, np.zeros((ny + 2, nx + 2))),
        lambda: capi.integrate2d(nx, ny, 1.0, 1.0, good_cc, np.zeros((ny, nx + 1))),
        lambda: capi.extrapolate_ghosts(nx, ny, 2, good_cc),
        lambda: capi.vcycle(0, geo, bc, 1, 1e-10, 2, good_cc, good_cc),
        lambda: capi.compute_pd(nx, ny, 1.0, 1.0, good_cc, 1, 1),
    ]
    for k, call in enumerate(bad):
        with pytest.raises(ValueError):
            call()
    # a metric array of the wrong shape, and NONUNIFORM BC data that is too short
    geo_bad = G.Geometry(nx, ny, geo.dx, geo.dy, geo.J, geo.Jxy[:, :-1], geo.Jxy, geo.Jyx, geo.Jyy)
    with pytest.raises(ValueError):
        capi.compute_invdiags(geo_bad, bc, 1)
    bc_short = G.BoundaryConditions((1, 1, 0, 0), (1, 0, 0, 0), tuple(np.zeros(3) for _ in range(4)))
    with pytest.raises(ValueError):
        capi.compute_invdiags(geo, bc_short, 1)


def test_gpnorm_rejects_out_of_array_index_pattern_without_device_work():
    import numpy as np
    import pytest
    from mgpoisson2d_amd import capi
    # columns jlo-1 and jhi+1 of a 10 x 4 box: jhi+1 = 4 lies inside, but the rows ilo-1 and
    # ihi+1 = 10 lie past the 4 + 2 rows of the array
    with pytest.raises(capi.MgpError):
        capi.gpnorm(0, 9, 0, 3, 0.1, 0.25, np.zeros((6, 12)), 0)
    with pytest.raises(ValueError):
        capi.gpnorm(0, 9, 0, 9, 0.1, 0.1, np.zeros((11, 12)), 0)
