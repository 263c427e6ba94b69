"""Benchmark: solves/s at 1e-10 relative residual on one MI355X, mixed/batched
against plain FP64.

Headline workload (the batched uneven problem mix): 1024 independent mapped-
grid Poisson problems, nx, ny = 16 round(exp(U[ln 32, ln 1024]) / 16),
mapping x = xi + A sin(2 pi xi) sin(2 pi eta), y = (exp(beta eta) - 1)/(exp(beta) - 1)
on [0,1]^2 with beta ~ U[0,6], A ~ U[0,0.15], all-Dirichlet or Dirichlet xi /
Neumann eta BCs (50/50), rhs N(0,1) per cell, all from one seed.  Each problem
is solved to ||rhs - L phi|| <= 1e-10 ||rhs|| from phi = 0 by the batched
solver with FP32 V-cycles inside FP64 BiCGStab with iterative refinement (the
operator, residual and solution stay FP64), one launch per level operation
for all problems.  A step is one batched solve of the whole mix with the rhs
resident in HBM; value = problems solved per second.

Beside it, in the same run:
* plain FP64: the same 1024 problems solved one at a time by the
  single-problem FP64 solver (hierarchy set-up untimed), with p50/p99
  latency; the ratio value / plain is the metric's "mixed/batched vs plain";
* the batched solver with FP64 cycles;
* roofline: the batch's dominant kernel, the finest-level line relaxation of
  every problem (one launch per chunk class, side by side), cache-cold;
* single_2048: the FP64 drop-in single solve (2048^2 stretched, warped map);
* serving: continuous batching of an open-loop request stream, against the
  plain FP64 one-at-a-time path on the same arrivals;
* dc_rb_smoother: the DC red-black smoother with LDS tiles;
* cpu_baseline: the reference Fortran (one core) on a sample of the mix,
  equal work.
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
TOL = 1e-10
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


def mix_problems(capi, n):
    from types import SimpleNamespace
    from mgpoisson2d_amd import geometry as G
    probs, rhs = config4_mix(n)
    geos = [SimpleNamespace(nx=nx, ny=ny, dx=1.0 / nx, dy=1.0 / ny) for nx, ny, _, _, _ in probs]
    maps = [capi.Mapping(capi.MAP_WARPED, amp, beta, 0.0, 0.0) for _, _, amp, beta, _ in probs]
    bcs = [G.BoundaryConditions.uniform(t) for *_, t in probs]
    return probs, rhs, geos, maps, bcs


def make_batch(capi, geos, bcs, maps, rhs, method, use_graph=True, sd=2, su=2, sb=2, tuning=None):
    """tuning: autotuner entries (nx, bc_class, fp32, threads) to apply, or None for the default shapes."""
    b = capi.Batch(geos, bcs, mappings=maps, smooth_down=sd, smooth_up=su, smooth_bottom=sb, use_graph=use_graph,
                   method=method)
    if tuning:
        b.set_tuning(tuning)
    for k in range(len(geos)):
        b.set_rhs(k, rhs[k])
    capi.device_synchronize()
    return b


def timed_batch(capi, b, tol, steps, warmup, label):
    """warmup untimed batched solves, then `steps` timed ones; returns (seconds, iterations, outcomes)."""
    for _ in range(warmup):
        iters, outc = b.solve(tol, 400)
        if not np.all(outc == capi.CONVERGED):
            raise SystemExit(f"{label} warm-up: {int(np.sum(outc != capi.CONVERGED))} problems did not converge")
    capi.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        iters, outc = b.solve(tol, 400)
    capi.device_synchronize()
    dt = time.perf_counter() - t0
    if not np.all(outc == capi.CONVERGED):
        raise SystemExit(f"{label}: {int(np.sum(outc != capi.CONVERGED))} problems did not converge")
    print(f"# {label}: {steps} batched solves of {len(iters)} problems in {dt:.3f} s, "
          f"{2 * float(np.mean(iters)):.1f} V-cycles mean, {2 * int(np.max(iters))} max", file=sys.stderr, flush=True)
    return dt, iters, outc


def plain_fp64_mix(capi, geos, bcs, maps, rhs, tol, use_graph=True):
    """The plain FP64 single-problem solver, one problem at a time over the whole
    mix: per problem a resident solver (set-up untimed), one untimed solve (graph
    capture), then one timed solve.  Returns (latencies in s, solutions)."""
    lat, sols = [], []
    for k in range(len(geos)):
        s1 = capi.Solver(geos[k], bcs[k], method=capi.METHOD_KRYLOV_MG, mapping=maps[k], use_graph=use_graph)
        s1.set_rhs(rhs[k])
        s1.solve_fast(tol, 400, zerophi=True)
        capi.device_synchronize()
        t1 = time.perf_counter()
        _, oc = s1.solve_fast(tol, 400, zerophi=True)
        capi.device_synchronize()
        lat.append(time.perf_counter() - t1)
        if oc != capi.CONVERGED:
            raise SystemExit(f"plain FP64 solve of problem {k} did not converge")
        sols.append(s1.get_phi()[1:-1, 1:-1])
        s1.close()
    print(f"# plain fp64: {len(lat)} problems one at a time in {sum(lat):.3f} s", file=sys.stderr, flush=True)
    return np.array(lat), sols


def rel_diff(a, b):
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


def serving_bench(capi, n_small=16, big_n=4096, big_share=0.1, load=0.7, n_requests=200, tol=1e-10,
                  budget=200e9, seed=SEED + 5, pool=6):
    """Continuous batching: n_small resident geometries from the batched mix plus one
    big_n^2 (beta = 6) geometry that takes big_share of the requests.  Every request
    carries an N(0,1) rhs drawn for it from a seeded pool of `pool` independent draws
    per geometry (a slot never receives the same rhs twice in a row), warm-started
    from the slot's last solution; FP32 cycles in FP64 BiCGStab to tol.  Capacity is
    measured closed-loop (one request in flight per slot); then requests arrive as a
    Poisson process at `load` x capacity and their latency (arrival to result) is
    recorded.  The same arrivals are then served by the plain FP64 single-problem
    solver one request at a time (resident per-slot solvers).  Its tolerance is
    relative to its initial residual (the reference's relres), so it solves from
    phi = 0, where that initial residual is ||rhs||, the server's reference."""
    from types import SimpleNamespace
    from mgpoisson2d_amd import geometry as G
    probs, _ = config4_mix(n_small, seed)
    probs.append((big_n, big_n, 0.1, 6.0, (1, 1, 0, 0)))
    rng = np.random.default_rng(seed + 1)
    pools = [[rng.standard_normal((ny, nx)) for _ in range(pool)] for nx, ny, *_ in probs]
    geos = [SimpleNamespace(nx=nx, ny=ny, dx=1.0 / nx, dy=1.0 / ny) for nx, ny, *_ in probs]
    maps = [capi.Mapping(capi.MAP_WARPED, amp, beta, 0.0, 0.0) for _, _, amp, beta, _ in probs]
    bcs = [G.BoundaryConditions.uniform(t) for *_, t in probs]
    big = len(probs) - 1
    last = [-1] * len(probs)

    def draw():
        slot = big if rng.uniform() < big_share else int(rng.integers(n_small))
        k = int(rng.integers(pool - 1))
        k = k if k < last[slot] else k + 1  # never the slot's previous rhs
        last[slot] = k
        return slot, pools[slot][k]

    v = capi.Server(geos, bcs, maps, hbm_budget=budget, method=capi.METHOD_KRYLOV_MG32)
    # cold solves of every slot (hierarchies and phi become resident)
    pend = {v.submit(k, pools[k][0], tol, 400, False): k for k in range(len(probs))}
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
    print(f"# serving: {n_requests} requests in {span:.2f} s", file=sys.stderr, flush=True)

    # the plain FP64 path, one request at a time, on the same arrivals
    solvers = [capi.Solver(geos[k], bcs[k], method=capi.METHOD_KRYLOV_MG, mapping=maps[k]) for k in range(len(probs))]
    for k, s1 in enumerate(solvers):  # cold solves (resident phi, captured graphs)
        s1.set_rhs(pools[k][0])
        s1.solve_fast(tol, 400, zerophi=True)
    capi.device_synchronize()
    plat, pbig, service = [], [], []
    t0 = time.perf_counter()
    for (slot, r), arr in zip(reqs, arrivals):
        now = time.perf_counter() - t0
        if now < arr:
            time.sleep(arr - now)
        t1 = time.perf_counter()
        solvers[slot].set_rhs(r)
        _, oc = solvers[slot].solve_fast(tol, 400, zerophi=True)
        solvers[slot].get_phi()
        if oc != capi.CONVERGED:
            raise SystemExit(f"plain FP64 served solve on slot {slot} did not converge")
        end = time.perf_counter()
        service.append(end - t1)
        (pbig if slot == big else plat).append(end - t0 - arr)
    for s1 in solvers:
        s1.close()
    print(f"# serving plain fp64: {n_requests} requests in {time.perf_counter() - t0:.2f} s", file=sys.stderr,
          flush=True)
    all_lat = np.array(lat + big_lat)
    all_plain = np.array(plat + pbig)
    return {
        "workload": f"{n_small} resident geometries from the batched mix plus one {big_n}^2 beta=6 geometry "
                    f"({100 * big_share:.0f} % of requests); each request an N(0,1) rhs (seeded pool of {pool} "
                    f"per geometry, never the slot's previous one), warm start from the slot's last solution, "
                    f"FP32 cycles in FP64 BiCGStab to {tol:g} relative residual; Poisson arrivals at {load} x the "
                    f"closed-loop capacity",
        "capacity_solves_per_s": capacity,
        "offered_per_s": rate,
        "served_per_s": n_requests / span,
        "requests": n_requests,
        "latency_p50_ms": 1e3 * float(np.percentile(all_lat, 50)),
        "latency_p99_ms": 1e3 * float(np.percentile(all_lat, 99)),
        "small_latency_p50_ms": 1e3 * float(np.percentile(lat, 50)) if lat else None,
        "small_latency_p99_ms": 1e3 * float(np.percentile(lat, 99)) if lat else None,
        "big_latency_p50_ms": 1e3 * float(np.percentile(big_lat, 50)) if big_lat else None,
        "big_latency_p99_ms": 1e3 * float(np.percentile(big_lat, 99)) if big_lat else None,
        "hbm_footprint_GB": footprint / 1e9,
        "hbm_budget_GB": budget / 1e9,
        "plain_fp64_one_at_a_time": {
            "capacity_solves_per_s": len(service) / float(np.sum(service)),
            "latency_p50_ms": 1e3 * float(np.percentile(all_plain, 50)),
            "latency_p99_ms": 1e3 * float(np.percentile(all_plain, 99)),
            "small_latency_p50_ms": 1e3 * float(np.percentile(plat, 50)) if plat else None,
            "small_latency_p99_ms": 1e3 * float(np.percentile(plat, 99)) if plat else None,
            "big_latency_p50_ms": 1e3 * float(np.percentile(pbig, 50)) if pbig else None,
            "big_latency_p99_ms": 1e3 * float(np.percentile(pbig, 99)) if pbig else None,
        },
    }


def mc_bench(capi, geo, bc, reps=20):
    """The multi-colour GS sweep of the plain operator (Poisson2D relax_gs, MC
    ordering) on the finest level: the four interior colours in one LDS-tiled
    launch against one launch per colour, each followed by the ring phase
    (algorithmic bytes: phi, rhs, invd and the four face metrics read once,
    phi written once)."""
    s = capi.Solver(geo, bc, method=capi.METHOD_VCYCLE, ordering=capi.ORDER_MC)
    tiled_ms, nbytes = s.time_kernel(7, reps=reps)
    colour_ms, _ = s.time_kernel(8, reps=reps)
    s.close()
    gbps = nbytes / (tiled_ms * 1e-3) / 1e9
    return {
        "workload": f"Poisson2D relax_gs, MC ordering, 1 sweep, {geo.nx}x{geo.ny} finest level",
        "tiled_ms": tiled_ms,
        "per_colour_ms": colour_ms,
        "speedup": colour_ms / tiled_ms,
        "tiled_GBps": gbps,
        "tiled_frac": gbps / HBM_PEAK_GBS,
        "alg_bytes_per_sweep": nbytes,
        "timing": f"HIP events around {reps} back-to-back sweeps on the solver stream (interior and ring phase)",
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


def single_bench(capi, args, steps=5, warmup=1):
    """The FP64 drop-in single solve (2048^2 stretched, warped map; Dirichlet xi,
    Neumann eta) with the resident solver, FP64 and FP32 cycles, to args.single_tol;
    the finest-level x-line kernel against the HBM roofline; the reference on one
    core for equal work.  With args.kernels_only: only the kernel timings."""
    geo, bc, rhs = config2_problem(args.n, args.n)

    def make(method, refine=capi.REFINE_DD):
        s = capi.Solver(geo, bc, method=method, use_graph=not args.no_graph, refine=refine)
        s.set_rhs(rhs)
        return s

    solver = make(capi.METHOD_KRYLOV_MG)
    print(f"# single: set-up done, {args.n}x{args.n}", file=sys.stderr, flush=True)
    if args.kernels_only:
        mixed = make(capi.METHOD_KRYLOV_MG32)
        out = {}
        for cold in (True, False):
            for s, k, name in ((solver, 0, "x-line even rows"), (solver, 1, "residual"), (solver, 2, "copy"),
                               (mixed, 3, "x-line even rows FP32")):
                ms, nbytes = s.time_kernel(k, reps=5, cold=cold)
                key = f"{'cold' if cold else 'warm'} {name}"
                out[key] = {"ms": ms, "GBps": nbytes / (ms * 1e-3) / 1e9}
                print(f"# {key}: {ms:.4f} ms/launch, {nbytes / 1e6:.1f} MB algorithmic, "
                      f"{nbytes / (ms * 1e-3) / 1e9:.0f} GB/s", flush=True)
        mixed.close()
        solver.close()
        return out

    def timed(s, label):
        for _ in range(warmup):
            iters, outcome = s.solve_fast(args.single_tol, 400, zerophi=True)
            if outcome != capi.CONVERGED:
                raise SystemExit(f"{label} warmup solve did not converge (outcome {outcome}, {iters} iterations)")
        capi.device_synchronize()
        t0 = time.perf_counter()
        cycles = 0
        for _ in range(steps):
            iters, outcome = s.solve_fast(args.single_tol, 400, zerophi=True)
            cycles += s.last_cycles()
            if outcome != capi.CONVERGED:
                raise SystemExit(f"{label} solve did not converge (outcome {outcome})")
        capi.device_synchronize()
        dt = time.perf_counter() - t0
        print(f"# single {label}: {steps} solves in {dt:.3f} s, {cycles / steps:.1f} V-cycles per solve",
              file=sys.stderr, flush=True)
        relres, _, _ = s.solve(args.single_tol, 400, zerophi=True)
        return dt, cycles / steps, float(relres[-1])

    dt, cyc, rr = timed(solver, "fp64")
    mixed = make(capi.METHOD_KRYLOV_MG32)
    mdt, mcyc, mrr = timed(mixed, "mixed")
    diff = rel_diff(mixed.get_phi()[1:-1, 1:-1], solver.get_phi()[1:-1, 1:-1])
    k_ms, k_bytes = solver.time_kernel(0, reps=20, cold=True)
    kw_ms, _ = solver.time_kernel(0, reps=20, cold=False)
    m_ms, m_bytes = mixed.time_kernel(3, reps=20, cold=True)
    r_ms, r_bytes = solver.time_kernel(1, reps=20, cold=True)
    c_ms, c_bytes = solver.time_kernel(2, reps=20, cold=True)
    mixed.close()
    solver.close()
    traffic = None
    if args.n == NX and os.path.exists(args.traffic):
        with open(args.traffic) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    out = {
        "workload": f"FP64 single solve, {args.n}x{args.n} warped/stretched mapped grid (A=0.1, beta=5), "
                    f"Dirichlet xi / Neumann eta, to {args.single_tol:g} relative residual (solution and "
                    f"refinement residual in double-double: the FP64 operator's rounding alone is about "
                    f"2e-10 of the rhs at 2048^2)",
        "tol": args.single_tol,
        "fp64_solves_per_s": steps / dt,
        "fp64_v_cycles_per_solve": cyc,
        "fp64_final_relres": rr,
        "mixed_solves_per_s": steps / mdt,
        "mixed_v_cycles_per_solve": mcyc,
        "mixed_final_relres": mrr,
        "max_rel_diff_mixed_vs_fp64_phi": diff,
        "xline_fp64": {"kernel": "k_xline (finest level, even rows)", "avg_launch_ms": k_ms,
                       "warm_avg_launch_ms": kw_ms, "alg_bytes_per_launch": k_bytes,
                       "GBps": k_bytes / (k_ms * 1e-3) / 1e9, "frac": k_bytes / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                       "traffic": traffic,
                       "timing": "HIP events around 20 back-to-back launches rotating over 4 copies of the "
                                 "operands (cache-cold)"},
        "xline_fp32_cold_ms": m_ms,
        "xline_fp32_GBps": m_bytes / (m_ms * 1e-3) / 1e9,
        "residual_kernel_GBps": r_bytes / (r_ms * 1e-3) / 1e9,
        "copy_kernel_GBps": c_bytes / (c_ms * 1e-3) / 1e9,
    }
    if not args.no_cpu_baseline:
        out["cpu_reference"] = cpu_baseline(geo, bc, rhs, int(round(cyc)))
    return out


NX3, NY3, AMP3, BETA3 = 8192, 2048, 0.12, 8.0


def warped_points(xi, eta, amp, beta):
    x = xi + amp * np.sin(2 * math.pi * xi) * np.sin(2 * math.pi * eta)
    y = np.expm1(beta * eta) / math.expm1(beta)
    return x, y


def warped_jacobian(xi, eta, amp, beta):
    x_xi = 1 + amp * 2 * math.pi * np.cos(2 * math.pi * xi) * np.sin(2 * math.pi * eta)
    return x_xi * (beta * np.exp(beta * eta) / math.expm1(beta))


def _rows_parallel(ny, fill):
    """fill(j0, j1) for row blocks of 64 on a thread pool (numpy releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    blocks = [(j, min(j + 64, ny)) for j in range(0, ny, 64)]
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        list(ex.map(lambda b: fill(*b), blocks))


def config3_rhs(nx, ny, amp, beta, seed, averages):
    """-J (f + 1e-2 N(0,1)) for u = cos(pi x) cos(2 pi y), f = 5 pi^2 u (L = +J Laplacian), at cell
    centres (second order) or, for the fourth-order operator, -(<J f> + 1e-2 J N(0,1)) with <J f>
    the cell average by 2 x 2 Gauss-Legendre points (exact to cubics: an O(h^4) average);
    float64, (ny, nx)."""
    noise = np.random.default_rng(seed).standard_normal((ny, nx))
    out = np.empty((ny, nx))
    t, w = np.polynomial.legendre.leggauss(2)
    t, w = (t + 1) / 2, w / 2

    def fill(j0, j1):
        Xc, Ec = np.meshgrid((np.arange(nx) + 0.5) / nx, (np.arange(j0, j1) + 0.5) / ny)
        Jc = warped_jacobian(Xc, Ec, amp, beta)
        if not averages:
            x, y = warped_points(Xc, Ec, amp, beta)
            out[j0:j1] = -Jc * (5 * math.pi ** 2 * np.cos(math.pi * x) * np.cos(2 * math.pi * y) + 1e-2 * noise[j0:j1])
            return
        acc = np.zeros((j1 - j0, nx))
        for a, wa in zip(t, w):
            for b, wb in zip(t, w):
                X, E = np.meshgrid((np.arange(nx) + a) / nx, (np.arange(j0, j1) + b) / ny)
                x, y = warped_points(X, E, amp, beta)
                acc += wa * wb * warped_jacobian(X, E, amp, beta) * (5 * math.pi ** 2) * np.cos(math.pi * x) * np.cos(
                    2 * math.pi * y)
        out[j0:j1] = -(acc + 1e-2 * Jc * noise[j0:j1])

    _rows_parallel(ny, fill)
    return out


def u_cells(nx, ny, amp, beta, averages):
    """u at cell centres, or its cell averages (2 x 2 Gauss-Legendre points)."""
    out = np.empty((ny, nx))
    t, w = np.polynomial.legendre.leggauss(2)
    t, w = (t + 1) / 2, w / 2

    def fill(j0, j1):
        if not averages:
            x, y = warped_points(*np.meshgrid((np.arange(nx) + 0.5) / nx, (np.arange(j0, j1) + 0.5) / ny), amp, beta)
            out[j0:j1] = np.cos(math.pi * x) * np.cos(2 * math.pi * y)
            return
        acc = np.zeros((j1 - j0, nx))
        for a, wa in zip(t, w):
            for b, wb in zip(t, w):
                x, y = warped_points(*np.meshgrid((np.arange(nx) + a) / nx, (np.arange(j0, j1) + b) / ny), amp, beta)
                acc += wa * wb * np.cos(math.pi * x) * np.cos(2 * math.pi * y)
        out[j0:j1] = acc

    _rows_parallel(ny, fill)
    return out


def cfg3_bench(capi, nx=NX3, ny=NY3, tol=1e-12, tol4=1e-7, seed=SEED + 3):
    """Mixed precision + deferred correction on the strongly stretched all-Neumann
    grid: 8192 x 2048 cells, x = xi + 0.12 sin(2 pi xi) sin(2 pi eta),
    y = (exp(8 eta) - 1)/(exp(8) - 1), u = cos(pi x) cos(2 pi y), rhs with 1e-2
    N(0,1) noise, projected to J-weighted zero mean in FP64 (compensated sums).

    The preconditioner is the y-semicoarsening V(2,2) cycle with x-line
    relaxation: the FP32 cycle solves the 8192-cell rows whole (512-thread
    line solve); the FP64 cycle cannot hold such a row in LDS and relaxes
    y-lines on the transposed hierarchy (AUTO).  Near the stretched wall the
    warp skews the cells (the strong coupling runs about 18 degrees off the
    grid lines), the cycle converges slowly and BiCGStab's recurrences break
    down, so the outer iteration is flexible GMRES(30).

    * second order: FP32 cycles inside FP64 refinement, solution and
      refinement residual in double-double, to `tol`; the same with FP64
      cycles is the FP64 reference solution;
    * fourth order (deferred correction): the second-order solve of the
      cell-average rhs, copied on the device into a solver of the
      fourth-order finite-volume operator (same second-order cycles inside),
      which iterates to |f - L4 u| <= tol4 |f|.  The FP64 evaluation of L4
      rounds at about 3e-8 of |f| on this grid (|L4 (phi (1 + eps N(0,1)))
      - L4 phi| measured), so 1e-7.
    Reported per solve: time (one untimed solve first), V-cycles, final relative
    residual, max |phi - u| / max |u| (both with the J-weighted mean removed), and
    the distance of the FP32-cycle solution from the FP64-cycle one."""
    from types import SimpleNamespace
    from mgpoisson2d_amd import geometry as G
    geo = SimpleNamespace(nx=nx, ny=ny, dx=1.0 / nx, dy=1.0 / ny)
    mp = capi.Mapping(capi.MAP_WARPED, AMP3, BETA3, 0.0, 0.0)
    bc = G.BoundaryConditions.uniform((0, 0, 0, 0))
    J = warped_jacobian(*np.meshgrid((np.arange(nx) + 0.5) / nx, (np.arange(ny) + 0.5) / ny), AMP3, BETA3)
    outcome = lambda oc: {capi.CONVERGED: "converged", capi.DIVERGED: "diverged"}.get(oc, "maxiters")

    def mean_free(a):
        return a - np.sum(J * a) / np.sum(J)

    def second_order(method, rhs):
        s = capi.Solver(geo, bc, method=method, mapping=mp, refine=capi.REFINE_DD, krylov=capi.KRYLOV_FGMRES)
        s.set_rhs(rhs)
        s.project_mean()
        return s

    out = {"workload": f"{nx}x{ny} cells, warp {AMP3}, stretching beta={BETA3}, all-Neumann, "
                       f"u = cos(pi x) cos(2 pi y) + 1e-2 N(0,1) rhs noise, J-weighted zero-mean projection",
           "outer_iteration": "FGMRES(30), preconditioner one V(2,2) cycle"}
    # second order
    rhs = config3_rhs(nx, ny, AMP3, BETA3, seed, averages=False)
    u = mean_free(u_cells(nx, ny, AMP3, BETA3, averages=False))
    sols = {}
    for method, label in ((capi.METHOD_KRYLOV_MG, "fp64"), (capi.METHOD_KRYLOV_MG32, "mixed")):
        s = second_order(method, rhs)
        s.solve(tol, 1000)  # first touch
        capi.device_synchronize()
        t0 = time.perf_counter()
        hist, iters, oc = s.solve(tol, 1000)
        capi.device_synchronize()
        dt = time.perf_counter() - t0
        phi = mean_free(s.get_phi()[1:-1, 1:-1] + s.get_phi_lo()[1:-1, 1:-1])
        sols[label] = phi
        out[f"order2_{label}"] = {
            "outcome": outcome(oc), "solve_s": dt, "v_cycles": s.last_cycles(), "final_relres": float(hist[-1]),
            "residual": "double-double", "max_err_vs_u_rel": float(np.max(np.abs(phi - u)) / np.max(np.abs(u)))}
        s.close()
        print(f"# cfg3 order 2 {label}: {dt:.3f} s, {out[f'order2_{label}']['v_cycles']} V-cycles, "
              f"relres {hist[-1]:.2e}", file=sys.stderr, flush=True)
    out["order2_max_rel_diff_mixed_vs_fp64"] = rel_diff(sols["mixed"], sols["fp64"])
    # fourth order: deferred correction from the second-order solution
    rhs = config3_rhs(nx, ny, AMP3, BETA3, seed, averages=True)
    u = mean_free(u_cells(nx, ny, AMP3, BETA3, averages=True))
    sols = {}
    for method, label in ((capi.METHOD_KRYLOV_MG, "fp64"), (capi.METHOD_KRYLOV_MG32, "mixed")):
        s2 = second_order(method, rhs)
        s4 = capi.Solver(geo, bc, method=method, mapping=mp, order=4)
        s4.set_rhs(rhs)
        s4.project_mean()
        for timed in (False, True):
            capi.device_synchronize()
            t0 = time.perf_counter()
            h2, i2, o2 = s2.solve(tol, 1000)
            s4.copy_phi_from(s2)
            rn, rh = s4.residual()
            h4, i4, o4 = s4.solve(tol4 * rh / rn, 1000, zerophi=False)
            capi.device_synchronize()
            dt = time.perf_counter() - t0
        rn4, _ = s4.residual()
        phi = mean_free(s4.get_phi()[1:-1, 1:-1])
        sols[label] = phi
        out[f"order4_{label}"] = {
            "outcome": outcome(o4) if o2 == capi.CONVERGED else "second-order step " + outcome(o2),
            "solve_s": dt, "v_cycles": s2.last_cycles() + s4.last_cycles(),
            "v_cycles_second_order_step": s2.last_cycles(), "relres_L4_after_second_order_step": rn / rh,
            "final_relres_L4": rn4 / rh, "residual": "FP64",
            "max_err_vs_u_rel": float(np.max(np.abs(phi - u)) / np.max(np.abs(u)))}
        s2.close()
        s4.close()
        print(f"# cfg3 order 4 {label}: {dt:.3f} s, {out[f'order4_{label}']['v_cycles']} V-cycles, "
              f"|f - L4 u|/|f| {rn / rh:.2e} -> {rn4 / rh:.2e}", file=sys.stderr, flush=True)
    out["order4_max_rel_diff_mixed_vs_fp64"] = rel_diff(sols["mixed"], sols["fp64"])
    for k in ("order2_fp64", "order4_fp64"):
        if out[k]["outcome"] != "converged":
            raise SystemExit(f"cfg3 {k} did not converge")
    return out


def tune_bench(capi, args, probs, rhs, geos, maps, bcs, reps=5):
    """The batch autotuner on the mix: per precision, timed batched solves with the
    default line-solve shapes, the tuner (per (nx, BC class) key: `reps` launches of
    the finest-level even-row relaxation of the key's problems per candidate width),
    timed solves with its choices.  The choices of both precisions go to
    gpurun_out/tuning_cache.json, the cache format of capi.save_tuning."""
    out = {"workload": f"batched mix, {len(probs)} problems", "tol": args.tol}
    entries = []
    for method, label in ((capi.METHOD_KRYLOV_MG32, "mixed"), (capi.METHOD_KRYLOV_MG, "fp64")):
        b = make_batch(capi, geos, bcs, maps, rhs, method, not args.no_graph)
        dt0, _, _ = timed_batch(capi, b, args.tol, 2, 1, f"{label} default shapes")
        t0 = time.perf_counter()
        chosen = b.tune(reps)
        tune_s = time.perf_counter() - t0
        for k in range(len(geos)):
            b.set_rhs(k, rhs[k])
        dt1, _, _ = timed_batch(capi, b, args.tol, 2, 1, f"{label} tuned shapes")
        b.close()
        entries += chosen
        counts = {}
        for e in chosen:
            counts[e[3]] = counts.get(e[3], 0) + 1
        out[label] = {"default_solves_per_s": 2 * len(probs) / dt0, "tuned_solves_per_s": 2 * len(probs) / dt1,
                      "tune_s": tune_s, "keys": len(chosen), "keys_per_width": counts}
    os.makedirs(os.path.join(ROOT, "gpurun_out"), exist_ok=True)
    capi.save_tuning(entries, os.path.join(ROOT, "gpurun_out", "tuning_cache.json"),
                     note=f"bench.py --tune on the {len(probs)}-problem mix, {reps} launches per candidate")
    return out


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


def cpu_baseline_mix(probs, rhs, gpu_cycles, sample=24, cycles=2, budget_s=25.0):
    """The reference Fortran (oracle/_ref/fast, one core) on a sample of the batched
    mix: DeferredMGPoisson2D::vcycle V(2,2), `cycles` cycles per problem.  The mix's
    grids are not square, and the reference's schedule then semicoarsens and its
    prolongation stops (MGPoisson2D.f90:385-388), so each sampled problem is timed
    with equal spacing in xi and eta (the same cells, the same mapping family on
    [0, nx h] x [0, ny h]), which keeps the reference's 2x2 schedule.  Equal work:
    seconds per reference V-cycle on each problem x the V-cycles the batched GPU
    solve spent on it."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import refcall
    from mgpoisson2d_amd import geometry as G
    if not os.path.exists(refcall.REF_FAST_SO):
        return None
    ref = refcall.Checker(refcall.REF_FAST_SO, "ref_")
    idx = np.random.default_rng(SEED + 9).choice(len(probs), size=min(sample, len(probs)), replace=False)
    spent, work, used = 0.0, 0.0, 0
    for k in idx:
        nx, ny, amp, beta, types = probs[k]
        h = 1.0 / max(nx, ny)
        geo = G.mapped_geometry(nx, ny, amp, beta, L=nx * h, H=ny * h)
        bc = G.BoundaryConditions.uniform(types)
        phi = np.zeros((ny + 2, nx + 2))
        t0 = time.perf_counter()
        for _ in range(cycles):
            phi, _, _, _, _ = ref.vcycle(1, geo, bc, 1, 0.0, 1, phi, rhs[k], sdown=2, sup=2, sbottom=2)
        dt = time.perf_counter() - t0
        spent += dt
        work += dt / cycles * max(int(gpu_cycles[k]), 1)
        used += 1
        if spent > budget_s:
            break
    return {
        "value": used / work,
        "unit": "solves/s",
        "cores": 1,
        "kind": "reference",
        "sample": (f"reference DeferredMGPoisson2D::vcycle V(2,2) on {used} problems of the batched mix "
                   f"({cycles} cycles each, {spent:.1f} s; equal spacing in xi and eta so the reference's "
                   f"2x2 schedule runs); equal-work solves/s = problems / sum over them of (s per reference "
                   f"V-cycle x V-cycles the batched GPU solve spent on that problem)"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed batched solves of the whole mix")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--batch-n", type=int, default=1024, help="problems in the batched mix")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 one-at-a-time comparison")
    ap.add_argument("--no-batched-fp64", action="store_true", help="skip the batched mix with FP64 cycles")
    ap.add_argument("--no-single", action="store_true", help="skip the 2048^2 single-solve section")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop measurement")
    ap.add_argument("--serving-only", action="store_true", help="run only the serving-loop measurement")
    ap.add_argument("--no-dc-rb", action="store_true", help="skip the DC red-black smoother measurement")
    ap.add_argument("--dc-rb-only", action="store_true", help="run only the DC red-black smoother measurement")
    ap.add_argument("--single-only", action="store_true", help="run only the 2048^2 single-solve section")
    ap.add_argument("--no-cfg3", action="store_true", help="skip the 8192x2048 mixed-precision / fourth-order section")
    ap.add_argument("--cfg3-only", action="store_true", help="run only the 8192x2048 section")
    ap.add_argument("--cfg3-n", type=int, default=NX3, help="nx of the cfg3 section (ny = nx / 4)")
    ap.add_argument("--n", type=int, default=NX, help="grid of the single-solve section")
    ap.add_argument("--single-tol", type=float, default=TOL, help="tolerance of the single-solve section")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch each iteration's kernels eagerly instead of replaying the captured hipGraph")
    ap.add_argument("--kernels-only", action="store_true",
                    help="time only the finest-level kernels (single solver: x-line, residual, copy; the "
                         "batch: its line relaxation) for hardware-counter passes")
    ap.add_argument("--batch-kernels-only", action="store_true",
                    help="set up the mixed-precision batch and time only its finest-level line relaxation "
                         "(and the calibration copy), for hardware-counter passes")
    ap.add_argument("--with-solve", action="store_true",
                    help="with --batch-kernels-only: one batched solve (captured graphs) before the timings")
    ap.add_argument("--tune", action="store_true",
                    help="autotune the batched line-solve shapes on the mix (both precisions), report the "
                         "batched solves/s before and after, and write the choices to gpurun_out/tuning_cache.json")
    ap.add_argument("--no-tuning", action="store_true",
                    help="batched sections with the default line-solve shapes instead of the committed cache")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic_xline_2048.json"),
                    help="measured HBM bytes per single-solver x-line launch (FETCH_SIZE/WRITE_SIZE passes)")
    ap.add_argument("--batch-traffic", default=os.path.join(ROOT, "profiles", "traffic_batch_xline32.json"),
                    help="measured HBM bytes per batched FP32 x-line launch set (FETCH_SIZE/WRITE_SIZE passes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1"}))
        sys.exit(2)

    from mgpoisson2d_amd import capi
    use_graph = not args.no_graph
    if args.serving_only:
        print(json.dumps(serving_bench(capi)), flush=True)
        return
    if args.dc_rb_only:
        geo, bc, rhs = config2_problem(args.n, args.n)
        out = dc_rb_bench(capi, geo, bc, rhs)
        out["mc_sweep"] = mc_bench(capi, geo, bc)
        print(json.dumps(out), flush=True)
        return
    if args.cfg3_only:
        print(json.dumps(cfg3_bench(capi, args.cfg3_n, args.cfg3_n // 4)), flush=True)
        return
    if args.single_only or args.kernels_only:
        print(json.dumps(single_bench(capi, args)), flush=True)
        return

    probs, rhs, geos, maps, bcs = mix_problems(capi, args.batch_n)
    cells = int(sum(nx * ny for nx, ny, *_ in probs))
    if args.tune:
        print(json.dumps(tune_bench(capi, args, probs, rhs, geos, maps, bcs)), flush=True)
        return
    tuning = [] if args.no_tuning else capi.load_tuning()
    t0 = time.perf_counter()
    b32 = make_batch(capi, geos, bcs, maps, rhs, capi.METHOD_KRYLOV_MG32, use_graph, tuning=tuning)
    setup_s = time.perf_counter() - t0
    print(f"# batch set-up: {len(probs)} problems, {cells} cells, {setup_s:.2f} s", file=sys.stderr, flush=True)
    if args.batch_kernels_only:
        if args.with_solve:
            iters, _ = b32.solve(args.tol, 400)
            print(f"# batched solve: {2 * int(np.max(iters))} V-cycles for the slowest problem", flush=True)
        ms, nbytes = b32.time_kernel(0, reps=5)
        print(f"# batched FP32 x-line set: {ms:.4f} ms, {nbytes / 1e6:.1f} MB algorithmic, "
              f"{nbytes / (ms * 1e-3) / 1e9:.0f} GB/s", flush=True)
        b32.close()
        # the counters' calibration: a plain copy of 32 MiB (2048^2 doubles) read and written
        geo, bc, _ = config2_problem(NX, NX)
        s = capi.Solver(geo, bc, method=capi.METHOD_KRYLOV_MG)
        cms, cbytes = s.time_kernel(2, reps=5, cold=True)
        print(f"# copy (calibration): {cms:.4f} ms, {cbytes / 1e6:.1f} MB", flush=True)
        s.close()
        return
    dt, iters32, _ = timed_batch(capi, b32, args.tol, args.steps, args.warmup, "batched mixed")
    value = len(probs) * args.steps / dt
    # dominant kernel: finest-level line relaxation (even rows) of every problem, FP32
    k_ms, k_bytes = b32.time_kernel(0, reps=10)
    achieved = k_bytes / (k_ms * 1e-3) / 1e9
    traffic = None
    if len(probs) == 1024 and os.path.exists(args.batch_traffic):
        with open(args.batch_traffic) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    mixed_phi = [b32.get_phi(k) for k in range(len(probs))]
    b32.close()

    # the same mix with FP64 cycles
    batched_fp64 = None
    if not args.no_batched_fp64:
        b64 = make_batch(capi, geos, bcs, maps, rhs, capi.METHOD_KRYLOV_MG, use_graph, tuning=tuning)
        dt64, iters64, _ = timed_batch(capi, b64, args.tol, max(1, min(args.steps, 2)), 1, "batched fp64")
        b64.close()
        batched_fp64 = {
            "solves_per_s": len(probs) * max(1, min(args.steps, 2)) / dt64,
            "v_cycles_mean": 2 * float(np.mean(iters64)),
        }

    plain = None
    if not args.no_plain:
        lat, sols = plain_fp64_mix(capi, geos, bcs, maps, rhs, args.tol, use_graph)
        plain_rate = len(lat) / float(np.sum(lat))
        plain = {
            "solves_per_s": plain_rate,
            "problems": len(lat),
            "latency_p50_ms": 1e3 * float(np.percentile(lat, 50)),
            "latency_p99_ms": 1e3 * float(np.percentile(lat, 99)),
            "max_rel_diff_mixed_batched_vs_plain": max(rel_diff(mixed_phi[k], sols[k]) for k in range(len(sols))),
        }

    single = None if args.no_single else single_bench(capi, args)
    cfg3 = None if args.no_cfg3 else cfg3_bench(capi, args.cfg3_n, args.cfg3_n // 4)
    serving = None if args.no_serving else serving_bench(capi)
    dc_rb = None
    if not args.no_dc_rb:
        geo, bc, rhs2 = config2_problem(NX, NX)
        dc_rb = dc_rb_bench(capi, geo, bc, rhs2)
        dc_rb["mc_sweep"] = mc_bench(capi, geo, bc)
    cpu = None if args.no_cpu_baseline else cpu_baseline_mix(probs, rhs, 2 * iters32)

    line = {
        "metric": "solves/s @1e-10 resid, mixed/batched vs plain FP64; p99 latency; HBM GB/s",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"batched uneven problem mix: {len(probs)} mapped-grid Poisson problems, nx, ny "
                        f"log-uniform in [32, 1024] (multiples of 16), beta U[0,6], A U[0,0.15], BCs all-Dirichlet "
                        f"or Dirichlet xi / Neumann eta, rhs N(0,1); each to {args.tol:g} relative residual",
            "tol": args.tol,
            "cells": cells,
            "method": "batched: FP64 BiCGStab with iterative refinement on the reference 9-point operator, "
                      "preconditioned by FP32 V(2,2) y-semicoarsening x-line cycles; one launch per level "
                      "operation for all problems",
            "step": "one batched solve of the whole mix from phi = 0, rhs resident in HBM",
            "v_cycles_mean": 2 * float(np.mean(iters32)),
            "v_cycles_max": 2 * int(np.max(iters32)),
            "setup_s": setup_s,
        },
        "plain_fp64": plain,
        "mixed_batched_over_plain_fp64": (value / plain["solves_per_s"]) if plain else None,
        "batched_fp64": batched_fp64,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_xline_batch_* FP32 (finest-level line relaxation, even rows, of all 1024 problems: "
                      "one launch per chunk class, side by side)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "avg_launch_ms": k_ms,
            "alg_bytes_per_launch": k_bytes,
            "timing": "HIP events on the batch stream around 10 back-to-back launch sets (1.7 GB of operands "
                      "per set, so every set starts cache-cold)",
        },
        "cpu_baseline": cpu,
        "single_2048": single,
        "cfg3_8192x2048": cfg3,
        "serving": serving,
        "dc_rb_smoother": dc_rb,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
