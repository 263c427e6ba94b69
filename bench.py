"""Benchmark: FP64 solves/s to 1e-10 relative residual on one MI355X.

Workload (the FP64 drop-in single-solve configuration): 2048 x 2048 cells on
(xi, eta) in [0,1]^2, mapping x = xi + 0.1 sin(2 pi xi) sin(2 pi eta),
y = (exp(5 eta) - 1)/(exp(5) - 1), analytic J and J g^ij at faces, Dirichlet
on the xi faces, Neumann on the eta faces, rhs = -J (f + 1e-2 N(0,1)) with f
the manufactured source of u = sin(pi x) cos(pi y) (the operator is +Laplacian),
zero initial guess, solved with the resident fast solver (FP64 BiCGStab on
the reference operator, one V(2,2) semicoarsening x-line cycle per
preconditioner application) to ||rhs - L phi|| <= tol ||rhs||.

tol defaults to 1e-9 at 2048^2: the FP64 evaluation of this operator has a
rounding floor near 2e-10 relative at that size (the flux terms near the
stretched wall are ~1e7 times the J-scaled rhs), so no FP64 solution vector
reaches 1e-10 there; 1e-10 is reachable up to 1024^2 (--n 1024 --tol 1e-10).

A step is one complete solve from phi = 0 with rhs already resident in HBM.
The same solves with the preconditioner V-cycle in FP32 (KRYLOV_MG32; the
Krylov vectors, operator, residual and solution stay FP64) are timed beside
them and reported under "mixed", with their own relative residual and their
distance from the FP64 solution.
Prints one JSON line.  --gpus N with N > 1 exits: this build measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NX = NY = 2048
AMP, BETA = 0.1, 5.0
TOL = 1e-9
SEED = 20261202


def config2_problem(nx=NX, ny=NY, seed=SEED):
    from mgpoisson2d_amd import geometry as G
    geo = G.mapped_geometry(nx, ny, AMP, BETA)
    x, y = G.cell_centres(geo, AMP, BETA)
    f = 2.0 * math.pi ** 2 * np.sin(math.pi * x) * np.cos(math.pi * y)
    rng = np.random.default_rng(seed)
    rhs = -geo.J * (f + 1e-2 * rng.standard_normal((ny, nx)))
    bc = G.BoundaryConditions.uniform((G.BCTYPE_DIRI, G.BCTYPE_DIRI, G.BCTYPE_NEUM, G.BCTYPE_NEUM))
    return geo, bc, np.ascontiguousarray(rhs)


def config4_mix(n=1024, seed=SEED):
    """The batched problem mix: n problems, nx, ny = 16 round(exp(U[ln 32, ln 1024]) / 16)
    each, mapping x = xi + A sin(2 pi xi) sin(2 pi eta), y = (exp(beta eta) - 1)/(exp(beta) - 1)
    on [0,1]^2 with beta ~ U[0, 6], A ~ U[0, 0.15]; BCs all-Dirichlet or Dirichlet xi /
    Neumann eta with equal odds; rhs N(0,1) per cell."""
    rng = np.random.default_rng(seed)
    lo, hi = math.log(32.0), math.log(1024.0)
    probs = []
    for _ in range(n):
        nx = int(16 * round(math.exp(rng.uniform(lo, hi)) / 16))
        ny = int(16 * round(math.exp(rng.uniform(lo, hi)) / 16))
        beta, amp = float(rng.uniform(0.0, 6.0)), float(rng.uniform(0.0, 0.15))
        types = (1, 1, 1, 1) if rng.uniform() < 0.5 else (1, 1, 0, 0)
        probs.append((nx, ny, amp, beta, types))
    rhs = [rng.standard_normal((ny, nx)) for nx, ny, _, _, _ in probs]
    return probs, rhs


def batched_bench(capi, n=1024, tol=1e-10, steps=2, sample=48, sd=2, su=2, sb=2, use_graph=True, mixed=True):
    """Solves/s of the batched solver on the problem mix, against the plain FP64
    single-problem solver (KRYLOV_MG, one problem at a time) timed on a sample of
    the same mix; per-solve latencies of the plain path give p50/p99."""
    from types import SimpleNamespace
    from mgpoisson2d_amd import geometry as G
    probs, rhs = config4_mix(n)
    geos = [SimpleNamespace(nx=nx, ny=ny, dx=1.0 / nx, dy=1.0 / ny) for nx, ny, _, _, _ in probs]
    maps = [capi.Mapping(capi.MAP_WARPED, amp, beta, 0.0, 0.0) for _, _, amp, beta, _ in probs]
    bcs = [G.BoundaryConditions.uniform(t) for *_, t in probs]
    t0 = time.perf_counter()
    b = capi.Batch(geos, bcs, mappings=maps, smooth_down=sd, smooth_up=su, smooth_bottom=sb, use_graph=use_graph)
    for k in range(n):
        b.set_rhs(k, rhs[k])
    capi.device_synchronize()
    setup_s = time.perf_counter() - t0
    iters, outc = b.solve(tol, 400)  # warm-up
    if not np.all(outc == capi.CONVERGED):
        raise SystemExit(f"batched warm-up: {int(np.sum(outc != capi.CONVERGED))} problems did not converge")
    times = []
    for _ in range(steps):
        capi.device_synchronize()
        t1 = time.perf_counter()
        iters, outc = b.solve(tol, 400)
        capi.device_synchronize()
        times.append(time.perf_counter() - t1)
    if not np.all(outc == capi.CONVERGED):
        raise SystemExit("batched solve did not converge")
    cycles = b.last_cycles()
    mixed_line = None
    if mixed:
        # the same mix with FP32 preconditioner cycles (FP64 Krylov, residual and answer)
        b32 = capi.Batch(geos, bcs, mappings=maps, smooth_down=sd, smooth_up=su, smooth_bottom=sb,
                         use_graph=use_graph, method=capi.METHOD_KRYLOV_MG32)
        for k in range(n):
            b32.set_rhs(k, rhs[k])
        b32.solve(tol, 400)
        times32 = []
        for _ in range(steps):
            capi.device_synchronize()
            t1 = time.perf_counter()
            it32, oc32 = b32.solve(tol, 400)
            capi.device_synchronize()
            times32.append(time.perf_counter() - t1)
        if not np.all(oc32 == capi.CONVERGED):
            raise SystemExit("batched mixed-precision solve did not converge")
        d32 = max(float(np.max(np.abs(b32.get_phi(k) - b.get_phi(k))) / max(np.max(np.abs(b.get_phi(k))), 1e-300))
                  for k in range(0, n, max(1, n // 64)))
        mixed_line = {"solves_per_s": n / min(times32), "batch_s": min(times32),
                      "v_cycles_mean": 2 * float(np.mean(it32)), "v_cycles_max": 2 * int(np.max(it32)),
                      "max_rel_diff_vs_fp64_batch": d32}
        b32.close()
        print(f"# batched mixed: solve {min(times32):.3f} s", file=sys.stderr, flush=True)
    print(f"# batched: {n} problems, setup {setup_s:.2f} s, solve {min(times):.3f} s, "
          f"{2 * float(np.mean(iters)):.1f} V-cycles mean, {cycles} max", file=sys.stderr, flush=True)
    # plain FP64 path on a sample, one problem at a time (hierarchy set-up not timed)
    lat = []
    diffs = []
    for k in range(min(sample, n)):
        s1 = capi.Solver(geos[k], bcs[k], method=capi.METHOD_KRYLOV_MG, mapping=maps[k], smooth_down=sd,
                         smooth_up=su, smooth_bottom=sb, use_graph=use_graph)
        s1.set_rhs(rhs[k])
        s1.solve_fast(tol, 400, zerophi=True)
        capi.device_synchronize()
        t1 = time.perf_counter()
        _, oc = s1.solve_fast(tol, 400, zerophi=True)
        capi.device_synchronize()
        lat.append(time.perf_counter() - t1)
        if oc != capi.CONVERGED:
            raise SystemExit(f"plain FP64 solve of problem {k} did not converge")
        ref = s1.get_phi()[1:-1, 1:-1]
        diffs.append(float(np.max(np.abs(b.get_phi(k) - ref)) / max(np.max(np.abs(ref)), 1e-300)))
        s1.close()
    b.close()
    lat = np.array(lat)
    batched = n / min(times)
    plain = len(lat) / float(np.sum(lat))
    cells = int(sum(nx * ny for nx, ny, *_ in probs))
    return {
        "workload": f"{n} mapped-grid problems, nx, ny log-uniform in [32, 1024] (multiples of 16), "
                    f"beta U[0,6], A U[0,0.15], BCs all-Dirichlet or Dirichlet xi / Neumann eta, rhs N(0,1), "
                    f"FP64 to {tol:g} relative residual each",
        "cells": cells,
        "batched_solves_per_s": batched,
        "batched_batch_s": min(times),
        "batched_setup_s": setup_s,
        "plain_fp64_solves_per_s": plain,
        "plain_sample": len(lat),
        "batched_over_plain": batched / plain,
        "plain_latency_p50_ms": 1e3 * float(np.percentile(lat, 50)),
        "plain_latency_p99_ms": 1e3 * float(np.percentile(lat, 99)),
        "v_cycles_mean": 2 * float(np.mean(iters)),
        "v_cycles_max": cycles,
        "max_rel_diff_batched_vs_plain": max(diffs),
        "batched_mixed": mixed_line,
        "batched_mixed_over_plain": (mixed_line["solves_per_s"] / plain) if mixed_line else None,
    }


def serving_bench(capi, n_small=16, big_n=4096, big_share=0.1, load=0.7, n_requests=200, tol=1e-10,
                  budget=200e9, seed=SEED + 5, mixed=True):
    """Continuous batching: n_small resident geometries from the batched mix plus one
    big_n^2 (beta = 6) geometry that takes big_share of the requests; each request is the
    slot's base rhs plus a 5 % perturbation, warm-started from the slot's last solution.
    Capacity is measured closed-loop (one request in flight per slot); then requests
    arrive as a Poisson process at `load` x capacity and their latency (arrival to result)
    is recorded."""
    from types import SimpleNamespace
    from mgpoisson2d_amd import geometry as G
    probs, base = config4_mix(n_small, seed)
    probs.append((big_n, big_n, 0.1, 6.0, (1, 1, 0, 0)))
    rng = np.random.default_rng(seed + 1)
    base.append(rng.standard_normal((big_n, big_n)))
    pert = [[rng.standard_normal(b.shape) for _ in range(3)] for b in base]
    geos = [SimpleNamespace(nx=nx, ny=ny, dx=1.0 / nx, dy=1.0 / ny) for nx, ny, *_ in probs]
    maps = [capi.Mapping(capi.MAP_WARPED, amp, beta, 0.0, 0.0) for _, _, amp, beta, _ in probs]
    bcs = [G.BoundaryConditions.uniform(t) for *_, t in probs]
    v = capi.Server(geos, bcs, maps, hbm_budget=budget,
                    method=capi.METHOD_KRYLOV_MG32 if mixed else capi.METHOD_KRYLOV_MG)
    big = len(probs) - 1

    def draw():
        slot = big if rng.uniform() < big_share else int(rng.integers(n_small))
        return slot, base[slot] + 0.05 * pert[slot][int(rng.integers(3))]

    # cold solves of every slot (hierarchies and phi become resident)
    pend = {v.submit(k, base[k], tol, 400, False): k for k in range(len(probs))}
    while pend:
        v.step()
        for t in list(pend):
            if v.result(t, pend[t]) is not None:
                del pend[t]
    # capacity, closed loop: as many requests in flight as slots
    t0 = time.perf_counter()
    pend = {}
    for _ in range(len(probs)):
        slot, r = draw()
        pend[v.submit(slot, r, tol, 400, True)] = slot
    done = 0
    while done < n_requests // 2:
        v.step()
        for t in list(pend):
            if v.result(t, pend[t]) is not None:
                del pend[t]
                done += 1
                slot, r = draw()
                pend[v.submit(slot, r, tol, 400, True)] = slot
    while pend:
        v.step()
        for t in list(pend):
            if v.result(t, pend[t]) is not None:
                del pend[t]
                done += 1
    capacity = done / (time.perf_counter() - t0)
    # open loop at load x capacity
    rate = load * capacity
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    reqs = [draw() for _ in range(n_requests)]
    lat, big_lat, pend, nxt = [], [], {}, 0
    t0 = time.perf_counter()
    while nxt < n_requests or pend:
        now = time.perf_counter() - t0
        while nxt < n_requests and arrivals[nxt] <= now:
            slot, r = reqs[nxt]
            pend[v.submit(slot, r, tol, 400, True)] = (slot, arrivals[nxt])
            nxt += 1
        running, queued = v.step()
        now = time.perf_counter() - t0
        for t in list(pend):
            slot, arr = pend[t]
            res = v.result(t, slot)
            if res is not None:
                if res[2] != capi.CONVERGED:
                    raise SystemExit(f"served solve on slot {slot} did not converge")
                (big_lat if slot == big else lat).append(now - arr)
                del pend[t]
        if running == 0 and queued == 0 and not pend and nxt < n_requests:
            time.sleep(max(0.0, arrivals[nxt] - (time.perf_counter() - t0)))
    span = time.perf_counter() - t0
    footprint = v.footprint()
    v.close()
    all_lat = np.array(lat + big_lat)
    return {
        "workload": f"{n_small} resident geometries from the batched mix plus one {big_n}^2 beta=6 geometry "
                    f"({100 * big_share:.0f} % of requests); rhs = base + 5 % perturbation, warm start from the "
                    f"slot's last solution, {'FP32 cycles in FP64 BiCGStab' if mixed else 'FP64'} to {tol:g} "
                    f"relative residual; Poisson arrivals at {load} x the closed-loop capacity",
        "capacity_solves_per_s": capacity,
        "offered_per_s": rate,
        "served_per_s": n_requests / span,
        "requests": n_requests,
        "latency_p50_ms": 1e3 * float(np.percentile(all_lat, 50)),
        "latency_p99_ms": 1e3 * float(np.percentile(all_lat, 99)),
        "small_latency_p50_ms": 1e3 * float(np.percentile(lat, 50)) if lat else None,
        "small_latency_p99_ms": 1e3 * float(np.percentile(lat, 99)) if lat else None,
        "big_latency_p50_ms": 1e3 * float(np.percentile(big_lat, 50)) if big_lat else None,
        "hbm_footprint_GB": footprint / 1e9,
        "hbm_budget_GB": budget / 1e9,
    }


def dc_rb_bench(capi, geo, bc, rhs=None, reps=20, cycles=5):
    """The DC red-black smoother (DeferredMGPoisson2D, RB ordering) on the
    finest level: one call of 2 sweeps with LDS tiles (several sweeps per
    load) against one launch per colour, and the tile kernel alone against
    the HBM roofline (algorithmic bytes: phi, rhs, invd and the four face
    metrics read once, phi written once)."""
    s = capi.Solver(geo, bc, method=capi.METHOD_DC_VCYCLE, ordering=capi.ORDER_RB)
    tiled_ms, nbytes = s.time_kernel(4, reps=reps)
    colour_ms, _ = s.time_kernel(5, reps=reps)
    kern_ms, _ = s.time_kernel(6, reps=reps)
    # whole DC V(2,2) cycles (every level's smoother, residuals, transfers, bottom solve)
    cycle_ms = {}
    if rhs is not None:
        s.set_rhs(rhs)
        for name, choice in (("tiled", -1), ("per_colour", 0)):
            prev = capi.dc_rb_tiles(choice)
            s.solve(1e-30, 1, zerophi=True)  # first use of each level shape picks its tile shape
            t0 = time.perf_counter()
            _, iters, _ = s.solve(1e-30, cycles, zerophi=True)
            cycle_ms[name] = (time.perf_counter() - t0) * 1e3 / max(iters, 1)
            capi.dc_rb_tiles(prev)
    s.close()
    gbps = nbytes / (kern_ms * 1e-3) / 1e9
    traffic = None
    path = os.path.join(ROOT, "profiles", "traffic_dc_rb_2048.json")
    if geo.nx == NX and geo.ny == NX and os.path.exists(path):
        with open(path) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    return {
        "workload": f"DC red-black relax_gs, 2 sweeps (modrhs fused), {geo.nx}x{geo.ny} finest level",
        "tiled_ms": tiled_ms,
        "per_colour_ms": colour_ms,
        "speedup": colour_ms / tiled_ms,
        "tile_kernel_ms": kern_ms,
        "tile_kernel_GBps": gbps,
        "tile_kernel_frac": gbps / HBM_PEAK_GBS,
        "alg_bytes_per_call": nbytes,
        "tile_kernel_traffic": traffic,
        "timing": f"HIP events around {reps} back-to-back calls on the solver stream",
        "dc_vcycle_ms": cycle_ms.get("tiled"),
        "dc_vcycle_ms_per_colour": cycle_ms.get("per_colour"),
        "dc_vcycle_timing": f"host clock around {cycles} DC V(2,2) cycles (MGPoisson2D_DD.f90 vcycle, RB ordering) "
                            "with their residual checks",
    }


def cpu_baseline(geo, bc, rhs, gpu_cycles, budget_s=20.0):
    """Time the reference Fortran (oracle/_ref/fast, -O3 -march=native, one
    core) on the same 2048^2 problem: DeferredMGPoisson2D::vcycle, one V(2,2)
    cycle per call, as many calls as fit in the budget.  The reference cannot
    reach 1e-10 here (its cycles stall and then diverge on this mapping), so
    the solves/s figure is for equal work: 1 / (seconds per reference V-cycle
    x V-cycle applications of one GPU solve)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import refcall
    if not os.path.exists(refcall.REF_FAST_SO):
        return None
    ref = refcall.Checker(refcall.REF_FAST_SO, "ref_")
    phi = np.zeros((geo.ny + 2, geo.nx + 2))
    cycles, t0 = 0, time.perf_counter()
    while True:
        phi, _, _, _, _ = ref.vcycle(1, geo, bc, 1, 0.0, 1, phi, rhs, sdown=2, sup=2, sbottom=2)
        cycles += 1
        if time.perf_counter() - t0 > budget_s or cycles >= 8:
            break
    sec_per_cycle = (time.perf_counter() - t0) / cycles
    return {
        "value": 1.0 / (sec_per_cycle * max(gpu_cycles, 1)),
        "unit": "solves/s",
        "cores": 1,
        "kind": "reference",
        "sample": (f"reference DeferredMGPoisson2D::vcycle V(2,2) on the same 2048^2 problem, "
                   f"{cycles} cycles timed, {sec_per_cycle:.3f} s/cycle; equal-work solves/s = "
                   f"1/(s_per_cycle x {gpu_cycles} V-cycles of the GPU solve) -- the reference "
                   f"itself does not reach 1e-10 on this mapping"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=NX)
    ap.add_argument("--tol", type=float, default=TOL)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sd", type=int, default=2, help="pre-smoothing sweeps per level")
    ap.add_argument("--su", type=int, default=2, help="post-smoothing sweeps per level")
    ap.add_argument("--sb", type=int, default=2, help="sweeps on the coarsest level")
    ap.add_argument("--method", choices=("fp64", "mixed"), default="fp64",
                    help="solver of the headline value: FP64 cycles (default) or FP32 cycles in FP64 BiCGStab")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision comparison")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched problem-mix comparison")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched problem-mix comparison")
    ap.add_argument("--batch-n", type=int, default=1024, help="problems in the batched mix")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop measurement")
    ap.add_argument("--serving-only", action="store_true", help="run only the serving-loop measurement")
    ap.add_argument("--no-dc-rb", action="store_true", help="skip the DC red-black smoother measurement")
    ap.add_argument("--dc-rb-only", action="store_true", help="run only the DC red-black smoother measurement")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch each iteration's kernels eagerly instead of replaying the captured hipGraph")
    ap.add_argument("--kernels-only", action="store_true",
                    help="set up the solver and time only the x-line, residual and copy kernels "
                         "(for hardware-counter passes)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic_xline_2048.json"),
                    help="measured HBM bytes per x-line launch (from FETCH_SIZE/WRITE_SIZE passes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1"}))
        sys.exit(2)

    from mgpoisson2d_amd import capi
    if args.batch_only:
        print(json.dumps(batched_bench(capi, args.batch_n, use_graph=not args.no_graph)), flush=True)
        return
    if args.serving_only:
        print(json.dumps({"mixed": serving_bench(capi), "fp64": serving_bench(capi, mixed=False)}), flush=True)
        return
    geo, bc, rhs = config2_problem(args.n, args.n)
    if args.dc_rb_only:
        print(json.dumps(dc_rb_bench(capi, geo, bc, rhs)), flush=True)
        return
    main_method = capi.METHOD_KRYLOV_MG32 if args.method == "mixed" else capi.METHOD_KRYLOV_MG
    other_method = capi.METHOD_KRYLOV_MG if args.method == "mixed" else capi.METHOD_KRYLOV_MG32

    def make(method):
        s = capi.Solver(geo, bc, method=method, smooth_down=args.sd, smooth_up=args.su, smooth_bottom=args.sb,
                        use_graph=not args.no_graph)
        s.set_rhs(rhs)
        return s

    solver = make(main_method)
    print(f"# setup done: {args.n}x{args.n}", file=sys.stderr, flush=True)
    if args.kernels_only:
        mixed = make(capi.METHOD_KRYLOV_MG32) if not args.no_mixed else None
        for cold in (True, False):
            for s, k, name in ((solver, 0, "x-line even rows"), (solver, 1, "residual"), (solver, 2, "copy"),
                               (mixed, 3, "x-line even rows FP32")):
                if s is None:
                    continue
                ms, nbytes = s.time_kernel(k, reps=5, cold=cold)
                print(f"# {'cold' if cold else 'warm'} {name}: {ms:.4f} ms/launch, {nbytes / 1e6:.1f} MB "
                      f"algorithmic, {nbytes / (ms * 1e-3) / 1e9:.0f} GB/s", flush=True)
        if mixed is not None:
            mixed.close()
        solver.close()
        return

    def timed(s, label):
        for _ in range(args.warmup):
            iters, outcome = s.solve_fast(args.tol, 400, zerophi=True)
            if outcome != capi.CONVERGED:
                raise SystemExit(f"{label} warmup solve did not converge (outcome {outcome}, {iters} iterations)")
        capi.device_synchronize()
        t0 = time.perf_counter()
        cycles = 0
        for _ in range(args.steps):
            iters, outcome = s.solve_fast(args.tol, 400, zerophi=True)
            cycles += s.last_cycles()
            if outcome != capi.CONVERGED:
                raise SystemExit(f"{label} solve did not converge (outcome {outcome})")
        capi.device_synchronize()
        dt = time.perf_counter() - t0
        print(f"# {label}: {args.steps} solves in {dt:.3f} s, {cycles / args.steps:.1f} V-cycles per solve",
              file=sys.stderr, flush=True)
        # verify the last solution against the operator itself
        relres, _, _ = s.solve(args.tol, 400, zerophi=True)
        return dt, cycles / args.steps, float(relres[-1])

    dt, cycles_per_solve, final_relres = timed(solver, args.method)
    ms_per_step = 1e3 * dt / args.steps

    compare = None
    if not args.no_mixed:
        other = make(other_method)
        odt, ocycles, orelres = timed(other, "fp64" if args.method == "mixed" else "mixed")
        a, b = solver.get_phi()[1:-1, 1:-1], other.get_phi()[1:-1, 1:-1]
        diff = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        mixed_s, fp64_s = (solver, other) if args.method == "mixed" else (other, solver)
        mixed_dt, fp64_dt = (dt, odt) if args.method == "mixed" else (odt, dt)
        mk_ms, mk_bytes = mixed_s.time_kernel(3, reps=20, cold=True)
        compare = {
            "mixed_solves_per_s": args.steps / mixed_dt,
            "fp64_solves_per_s": args.steps / fp64_dt,
            "mixed_over_fp64": fp64_dt / mixed_dt,
            "mixed_v_cycles_per_solve": ocycles if args.method == "fp64" else cycles_per_solve,
            "mixed_final_relres": orelres if args.method == "fp64" else final_relres,
            "max_rel_diff_mixed_vs_fp64_phi": diff,
            "xline_fp32_cold_ms": mk_ms,
            "xline_fp32_GBps": mk_bytes / (mk_ms * 1e-3) / 1e9,
        }
        other.close()

    # dominant kernel: finest-level x-line relaxation (even rows), HIP events
    # around back-to-back launches whose operands are cache-cold (as inside a solve)
    mixed_main = args.method == "mixed"
    rk = 3 if mixed_main else 0
    k_ms, k_bytes = solver.time_kernel(rk, reps=20, cold=True)
    achieved = k_bytes / (k_ms * 1e-3) / 1e9
    kw_ms, _ = solver.time_kernel(rk, reps=20, cold=False)
    r_ms, r_bytes = solver.time_kernel(1, reps=20, cold=True)
    c_ms, c_bytes = solver.time_kernel(2, reps=20, cold=True)

    traffic = None
    if args.n == NX and not mixed_main and os.path.exists(args.traffic):
        with open(args.traffic) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")

    batched = None
    if not args.no_batched:
        batched = batched_bench(capi, args.batch_n, use_graph=not args.no_graph)
    serving = None
    if not args.no_serving:
        serving = serving_bench(capi)
    dc_rb = None
    if not args.no_dc_rb:
        dc_rb = dc_rb_bench(capi, geo, bc, rhs)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(geo, bc, rhs, int(round(cycles_per_solve)))

    line = {
        "metric": "solves/s @1e-10 resid, mixed/batched vs plain FP64; p99 latency; HBM GB/s",
        "value": args.steps / dt,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if mixed_main else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{'mixed-precision' if mixed_main else 'FP64'} single solve, {args.n}x{args.n} "
                        f"warped/stretched mapped grid "
                        f"(A=0.1, beta=5), Dirichlet xi / Neumann eta, to {args.tol:g} relative residual",
            "tol": args.tol,
            "method": f"FP64 BiCGStab + {'FP32 ' if mixed_main else ''}V({args.sd},{args.su}) y-semicoarsening "
                      f"x-line multigrid preconditioner ({args.sb} sweeps on the coarsest level)",
            "v_cycles_per_solve": cycles_per_solve,
            "final_relres": final_relres,
            "residual_kernel_ms": r_ms,
            "residual_kernel_GBps": r_bytes / (r_ms * 1e-3) / 1e9,
            "copy_kernel_GBps": c_bytes / (c_ms * 1e-3) / 1e9,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_xline{' FP32' if mixed_main else ''} (finest level, even rows)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "avg_launch_ms": k_ms,
            "timing": "HIP events around 20 back-to-back launches rotating over 4 copies of the operands (cache-cold)",
            "warm_avg_launch_ms": kw_ms,
            "alg_bytes_per_launch": k_bytes,
        },
        "cpu_baseline": cpu,
        "mixed": compare,
        "batched": batched,
        "serving": serving,
        "dc_rb_smoother": dc_rb,
    }
    print(json.dumps(line), flush=True)
    solver.close()


if __name__ == "__main__":
    main()
