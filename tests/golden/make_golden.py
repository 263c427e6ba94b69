"""Write tests/golden/reference_vectors.npz from the reference Fortran.

Runs the unmodified reference (EdSantilli/MGPoisson2D, built from
/root/reference into oracle/_ref by `make -C oracle ref`) through its bind(C)
bridge on small seeded cases and stores inputs and outputs.  The fixture lets
the oracle and the GPU library be checked against the reference where the
reference build is absent (the GPU box has no /root/reference).

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import refcall as rc  # noqa: E402
from cases import CYCLE_CASES, OPERATOR_CASES, setup  # noqa: E402
from mgpoisson2d_amd import geometry as G  # noqa: E402

OUT = os.path.join(HERE, "reference_vectors.npz")


def main():
    if not rc.reference_available(rb=True):
        raise SystemExit("build the reference first: make -C oracle ref")
    ref, ref_rb = rc.reference(), rc.reference(rb=True)
    data = {}

    def put(key, arr):
        data[key] = np.asarray(arr)

    for k, case in enumerate(OPERATOR_CASES):
        for homog in (1, 0):
            geo, bc, phi, rhs = setup(*case, homog=homog, seed=100 + k)
            p = f"op{k}_h{homog}_"
            for name, a in zip(("J", "Jxx", "Jxy", "Jyx", "Jyy"), geo.arrays()):
                put(p + name, a)
            put(p + "shape", [geo.nx, geo.ny, geo.dx, geo.dy, *case[4]])
            put(p + "phi", phi)
            put(p + "rhs", rhs)
            lap, phig = ref.laplacian(geo, bc, homog, phi)
            put(p + "lap", lap)
            put(p + "phi_filled", phig)
            put(p + "res", ref.residual(geo, bc, homog, rhs, phi)[0])
            inv = ref.invdiags(geo, bc, homog)
            put(p + "invd", inv)
            put(p + "fill_neum", ref.fill_ghosts(geo, bc, homog, 1, phi))
            put(p + "gs3", ref.relax_gs(geo, bc, homog, inv, 1.0, 3, rhs, phi))
            lah = G.BoundaryConditions.uniform(case[4], mode=G.BCMODE_LAH)
            put(p + "dc_lap", ref.dc_laplacian(geo, lah, 1, phi)[0])
            dinv = ref.dc_invdiags(geo)
            put(p + "dc_invd", dinv)
            put(p + "dc_lex3", ref.dc_relax_gs(geo, lah, 1, dinv, 0.8, 3, rhs, phi))
            put(p + "dc_rb3", ref_rb.dc_relax_gs(geo, lah, 1, dinv, 0.8, 3, rhs, phi))
            # reductions and ghost extrapolation (arrayutils.f90:637-796, 1497-1566)
            v = np.ascontiguousarray(phi[1:-1, 1:-1])
            nx_, ny_, dx_, dy_ = geo.nx, geo.ny, geo.dx, geo.dy
            put(p + "reductions", [ref.pnorm(nx_, ny_, dx_, dy_, v, q) for q in range(5)]
                + [ref.inner_prod(nx_, ny_, dx_, dy_, v, rhs), ref.integrate2d(nx_, ny_, dx_, dy_, v, geo.J),
                   ref.integrate2d(nx_, ny_, dx_, dy_, v)])
            for order in range(4):
                put(p + f"extrap{order}", ref.extrapolate_ghosts(nx_, ny_, order, phi))
            # single-level solvers (poisson2d.f90:1358-2248)
            put(p + "jacobi", ref.relax_jacobi(geo, bc, homog, inv, 0.8, 1e-3, 12, 0, rhs, phi)[0])
            put(p + "bicgstab1", ref.bicgstab1(geo, bc, homog, inv, 1e-8, 20, 2, 0, rhs, phi)[0])
            # flux tools (poisson2d.f90:404-879): J grad phi on the faces, J div of it
            _, xf, yf = ref.compute_grad(geo, bc, homog, phi)
            put(p + "grad_x", xf)
            put(p + "grad_y", yf)
            put(p + "div", ref.compute_div(geo.nx, geo.ny, geo.dx, geo.dy, xf, yf))
            for nodedir in (1, 2):
                for d in (1, 2):
                    put(p + f"pd{nodedir}{d}", ref.compute_pd(geo.nx, geo.ny, geo.dx, geo.dy, phi, d, nodedir))
    for k, case in enumerate(CYCLE_CASES):
        geo, bc, phi, rhs = setup(*case, seed=200 + k)
        p = f"cyc{k}_"
        for name, a in zip(("J", "Jxx", "Jxy", "Jyx", "Jyy"), geo.arrays()):
            put(p + name, a)
        put(p + "shape", [geo.nx, geo.ny, geo.dx, geo.dy, *case[4]])
        put(p + "rhs", rhs)
        for kind in (0, 1, 2):  # MGPoisson2D::vcycle, DeferredMGPoisson2D::vcycle, MGPoisson2D::fmg
            out = ref.vcycle(kind, geo, bc, 1, 0.0, 4, np.zeros_like(phi), rhs, zerophi=True)
            put(p + f"vcycle{kind}_phi4", out[0])
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({os.path.getsize(OUT) / 1024:.0f} KiB, {len(data)} arrays)")


if __name__ == "__main__":
    main()
