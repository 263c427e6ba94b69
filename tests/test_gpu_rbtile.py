"""Red-black DC sweeps with several sweeps per LDS tile load.

The tiled smoother must give, bit for bit, what the reference's red-black
relax_gs gives (MGPoisson2D_DD.f90:227-285: modrhs once, then per sweep a
ghost fill, colour 0 and colour 1), for every tile shape and sweep count:
the oracle's dc_relax_gs (bit-identical to the reference built with the RB
branch) is the check, and so is the one-launch-per-colour GPU path.  Cases
cover windows that cross tile edges inside the domain, every Neumann /
Dirichlet side mix of the LAH level BCs, tiny grids smaller than one tile,
odd sweep counts (a 2-sweep and a 1-sweep launch, ping-ponged and copied
back), and periodic sides (which keep the one-launch-per-colour path).  A DC
V-cycle solve with RB ordering gives the same history and phi under every
choice.
"""
import numpy as np
import pytest

import refcall as rc
from cases import setup

pytestmark = pytest.mark.gpu

mg = pytest.importorskip("mgpoisson2d_amd.capi")
from mgpoisson2d_amd import geometry as G  # noqa: E402

NSHAPES = mg.dc_rb_tile_shapes()
CHOICES = [0, -1] + list(range(1, NSHAPES + 1))


@pytest.fixture(autouse=True)
def restore_choice():
    prev = mg.dc_rb_tiles(-1)
    yield
    mg.dc_rb_tiles(prev)


def relax_all(geo, lah, inv, rhs, phi, sweeps, omega=0.8):
    out = {}
    for ch in CHOICES:
        mg.dc_rb_tiles(ch)
        out[ch] = mg.dc_relax_gs(geo, lah, 1, inv, omega, sweeps, rhs, phi, mg.ORDER_RB)
    return out


# (nx, ny, amp, beta, types): tiles cross inside the domain in both
# directions; all side mixes; grids smaller than a tile
CASES = [
    (300, 70, 0.10, 3.0, (1, 1, 0, 0)),
    (260, 90, 0.05, 2.0, (0, 0, 1, 1)),
    (129, 31, 0.12, 1.0, (0, 1, 1, 0)),
    (121, 57, 0.08, 4.0, (0, 0, 0, 0)),
    (97, 41, 0.00, 0.0, (1, 1, 1, 1)),
    (7, 5, 0.10, 1.0, (1, 0, 0, 1)),
    (3, 4, 0.00, 0.0, (0, 1, 0, 1)),
]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("sweeps", [1, 2, 3, 5])
def test_rb_tiles_match_oracle(case, sweeps):
    geo, _, phi, rhs = setup(*case, seed=21)
    lah = G.BoundaryConditions.uniform(case[4], mode=G.BCMODE_LAH)
    inv = mg.dc_compute_invdiags(geo)
    ref = rc.oracle().dc_relax_gs(geo, lah, 1, inv, 0.8, sweeps, rhs, phi, rc.RB)
    for ch, got in relax_all(geo, lah, inv, rhs, phi, sweeps).items():
        np.testing.assert_array_equal(got, ref, err_msg=f"choice {ch}")


def test_rb_tiles_large_grid():
    """Many tiles per direction on a 1500 x 700 mapped grid, the V(2,2)
    sweep count: every shape equals the per-colour launches and the oracle."""
    case = (1500, 700, 0.10, 5.0, (1, 1, 0, 0))
    geo, _, phi, rhs = setup(*case, seed=22)
    lah = G.BoundaryConditions.uniform(case[4], mode=G.BCMODE_LAH)
    inv = mg.dc_compute_invdiags(geo)
    out = relax_all(geo, lah, inv, rhs, phi, 2)
    ref = rc.oracle().dc_relax_gs(geo, lah, 1, inv, 0.8, 2, rhs, phi, rc.RB)
    for ch, got in out.items():
        np.testing.assert_array_equal(got, ref, err_msg=f"choice {ch}")


def test_rb_periodic_keeps_per_colour_path():
    case = (64, 32, 0.05, 1.0, (2, 2, 1, 1))
    geo, _, phi, rhs = setup(*case, seed=23)
    lah = G.BoundaryConditions.uniform(case[4], mode=G.BCMODE_LAH)
    inv = mg.dc_compute_invdiags(geo)
    ref = rc.oracle().dc_relax_gs(geo, lah, 1, inv, 0.8, 3, rhs, phi, rc.RB)
    for ch, got in relax_all(geo, lah, inv, rhs, phi, 3).items():
        np.testing.assert_array_equal(got, ref, err_msg=f"choice {ch}")


@pytest.mark.parametrize("sweeps", [2, 3])
def test_dc_vcycle_rb_same_under_every_choice(sweeps):
    geo, bc, _, rhs = setup(512, 512, 0.1, 3.0, (1, 1, 0, 0), seed=24)
    runs = {}
    for ch in (0, -1, 2):
        mg.dc_rb_tiles(ch)
        s = mg.Solver(geo, bc, method=mg.METHOD_DC_VCYCLE, ordering=mg.ORDER_RB, smooth_down=sweeps,
                      smooth_up=sweeps, smooth_bottom=sweeps)
        s.set_rhs(rhs)
        hist, iters, outcome = s.solve(1e-9, 12, zerophi=True)
        runs[ch] = (np.asarray(hist).copy(), iters, outcome, s.get_phi().copy())
        s.close()
    h0, i0, o0, p0 = runs[0]
    for ch, (h, i, o, p) in runs.items():
        assert (i, o) == (i0, o0), ch
        np.testing.assert_array_equal(h, h0)
        np.testing.assert_array_equal(p, p0)


def test_rb_tiles_choice_validated():
    with pytest.raises(mg.MgpError):
        mg.dc_rb_tiles(NSHAPES + 1)
    with pytest.raises(mg.MgpError):
        mg.dc_rb_tiles(-2)
