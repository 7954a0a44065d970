"""Benchmark: FP64 point-updates/s of the 2+1D Teukolsky evolution on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one full classic-RK4 time step of the single solve (--integrator
rk3: the reference's SSP-RK3): the particle source deposit on its 128x128
patch, then one fused kernel per RK stage (the last one also extracts Psi
and integrates h) over N=16384 rho points on
[-400, 1200] M and M=64 theta points, a=0.9, m=2, dt = 0.5 drho, zero
initial fields, source on a seeded plunge (mass ratio 1e-3, inclination
30 degrees, from 1.02 x the ISCO).  One point-update advances (Psi, Pi) at
one grid point by one full step.

Prints one JSON line with the throughput, the dominant kernel's roofline
position (HIP events around every stage launch) and the host-core baseline
(the CPU oracle's OpenMP build on a bounded sample of the same workload).
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
# Algorithmic bytes per point-update (state traffic of one sweep per stage):
# SSP-RK3 64 + 96 + 96; classic RK4 512 = the accumulator form's
# 96 + 160 + 160 + 96 (the kernels use the stage-value form, which moves
# 64 + 128 + 96 + 96 = 384; the roofline keeps the 512 figure, not
# renormalised to the cheaper form);
# the staggered second-order scheme 64 (half step: fields in, half-point
# fields out) + 96 (full step: half-point fields in, fields in and out).
B_ALG = {"rk3": 256.0, "rk4": 512.0, "so2": 160.0}
# State bytes the implemented kernels move per point-update (RK4 in its
# stage-value form: 64 + 128 + 96 + 96), without the 8 B per stage of the dbdx
# plane.
B_MOVED = {"rk3": 256.0, "rk4": 384.0, "so2": 160.0}
STAGES = {"rk3": 3, "rk4": 4, "so2": 2}
INTEGRATORS = {"rk3": 0, "rk4": 1, "so2": 2}
INTEGRATOR_NAMES = {"rk4": "classic RK4", "rk3": "SSP-RK3",
                    "so2": "staggered second-order scheme (TD_GPU_code/Old)"}


# Timed single solves: name -> (maker in bh_ringdown_amd.workloads, its grid size).
WORKLOADS = {
    "single_16384x64_plunge": ("single_16384x64", None),
    "high_spin_16384x32_plunge": ("high_spin", (16384, 32)),
    "high_spin_32768x64_plunge": ("high_spin", (32768, 64)),
    "high_spin_65536x128_plunge": ("high_spin", (65536, 128)),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def pmc_traffic(integrator, N, M):
    """HBM bytes per stage launch (FETCH_SIZE + WRITE_SIZE, scaled by the copy
    calibration) from the committed rocprofv3 --pmc summary of this
    integrator and grid (profiles/k_stage_<integrator>_<N>x<M>_plunge_pmc.txt),
    or None."""
    path = os.path.join(ROOT, "profiles", f"k_stage_{integrator}_{N}x{M}_plunge_pmc.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(p, rows, seconds=12.0, threads=None):
    """The oracle (g++ -O3 OpenMP build) on the first steps of the same
    workload: source, the RK stages, boundaries, extraction and projection
    per step.  The one-off coefficient set-up is timed separately and
    subtracted (a one-step run)."""
    from bh_ringdown_amd import teuk
    from oracle import oracle as O
    nth = threads or int(os.environ.get("OMP_NUM_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    _, r_c, th = teuk.grid(p)
    op = O.make_params(p.N, p.M, p.xmin, p.xmax, p.a, p.m, dt=p.dt, j_extract=p.j_extract, nthreads=nth,
                       integrator=p.integrator)
    name = {0: "SSP-RK3", 1: "RK4", 2: "staggered-scheme"}[p.integrator]
    stag = p.integrator == 2
    r_h = teuk.grid_half(p)[1] if stag else None

    def timed(n):
        sched = (teuk.trajectory_schedule_staggered if stag else teuk.trajectory_schedule)(rows, n, p.dt)
        ost = [O.SrcStep(*[getattr(x, f) for f, _ in teuk.SourceStep._fields_]) for x in sched]
        u = np.zeros((4, p.M, p.N))
        t0 = time.perf_counter()
        if stag:
            rc, *_ = O.run_so2(op, r_h, r_c, th, u, n, src_steps=ost, timing=True)
        else:
            rc, *_ = O.run(op, r_c, th, u, n, src_steps=ost, timing=True)
        if rc != 0:
            raise RuntimeError(f"oracle run failed ({rc})")
        return time.perf_counter() - t0

    t1 = timed(1)
    t3 = timed(3)
    per = max((t3 - t1) / 2, 1e-6)
    n = int(max(4, min(2000, seconds / per)))
    el = timed(n + 1) - t1
    return {"value": p.N * p.M * n / el, "unit": "point-updates/s", "cores": nth, "kind": "port",
            "sample": f"first {n} {name} steps (with source, extraction, projection) of the "
                      f"{p.N}x{p.M} plunge solve; oracle g++ -O3 OpenMP, set-up excluded",
            "cpu": cpu_model()}


def batch_section(scale, grid=None, plain=True):
    """The 72 uneven solves (9 m x 8 spins, grids by spin, or all on `grid`)
    at `scale` of their final time: the plain one-solve-at-a-time loop
    against one batch (plain=False: the batch alone, for profiler runs of
    the batched kernels)."""
    from bh_ringdown_amd import teuk, tune, workloads
    items = workloads.batch_72(seed=0, t_final=1200.0 * scale, grid=grid)
    total = sum(p.N * p.M * p.nsteps() for p, _ in items)
    log(f"bench: batch of {len(items)} solves, {total:.3e} point-updates")
    probe = {0, 31, 71} if plain else set()
    t0 = time.perf_counter()
    plain_out = {}
    for i, (p, rows) in enumerate(items if plain else []):
        with teuk.Solver(p) as s:
            s.set_trajectory(rows, p.nsteps())
            s.step(p.nsteps())
            s.sync()
            if i in probe:
                plain_out[i] = s.output()["h"]
    t_plain = time.perf_counter() - t0 if plain else None
    if plain:
        log(f"bench: plain loop {t_plain:.2f} s")
    shapes = [(p.N, p.M) for p, _ in items]
    tiling = tune.best_batch_tiling(shapes) or {"waves_per_block": 4, "waves_per_eu": 2, "order": 0, "chunk": 32}
    groups = tune.best_batch_groups(shapes)
    t0 = time.perf_counter()
    solvers = []
    with teuk.Batch(integrator=1) as b:
        b.set_tiling(**tiling)
        b.set_cache_groups(groups)
        for p, rows in items:
            s = teuk.Solver(p)
            s.set_trajectory(rows, p.nsteps())
            b.add(s)
            solvers.append(s)
        b.step(max(p.nsteps() for p, _ in items))
        b.sync()
        same = all(np.array_equal(solvers[i].output()["h"], plain_out[i]) for i in probe)
    for s in solvers:
        s.close()
    t_batch = time.perf_counter() - t0
    log(f"bench: batched {t_batch:.2f} s")
    grids = f"all {grid[0]}x{grid[1]}" if grid else "N x M by spin in 4096x32, 8192x32, 16384x64"
    return {"workload": f"72 solves (m -4..4 x 8 seeded spins; {grids}) to t_final {1200.0 * scale:g} "
                        f"({scale:g} of 1200), RK4, plunge source",
            "solves_per_hour": 3600.0 * len(items) / t_batch,
            "plain_solves_per_hour": 3600.0 * len(items) / t_plain if plain else None,
            "speedup_vs_plain": t_plain / t_batch if plain else None, "point_updates_per_s": total / t_batch,
            "plain_point_updates_per_s": total / t_plain if plain else None, "tiling": tiling,
            "cache_group_bytes": groups, "members_bitwise_equal_alone": same if plain else None}


def serve_section(scale, seconds, policy="sjf"):
    """A seeded Poisson stream (20 requests/s for `seconds`, request lengths
    scaled by `scale`) through the continuous-batching server, and the same
    requests through the plain loop."""
    from bh_ringdown_amd import serve
    reqs = serve.poisson_requests(20.0, seconds, seed=0, t_range=(500.0 * scale, 2000.0 * scale))
    log(f"bench: serving {len(reqs)} requests")
    srv = serve.Server(hbm_budget=64 << 30, chunk=128, target_points=1 << 20, shortest_first=(policy == "sjf"))
    t0 = time.perf_counter()
    res_s = srv.run(reqs)
    wall_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    res_p = serve.run_plain(reqs)
    wall_p = time.perf_counter() - t0
    same = all(np.array_equal(res_s[r.id].records["h"], res_p[r.id].records["h"]) for r in reqs)
    ss, sp = serve.summary(res_s, wall_s), serve.summary(res_p, wall_p)
    log(f"bench: served {wall_s:.2f} s, plain {wall_p:.2f} s")
    return {"workload": f"Poisson 20 req/s for {seconds:g} s (seed 0, {len(reqs)} requests), a~U[0,0.99), "
                        f"m in 1..4, iota~U[0,pi/2) capped 85 deg, q log-uniform [1e-4,1e-2], N in 4096/8192/16384, "
                        f"M in 32/64, t_final~U[{500 * scale:g},{2000 * scale:g}) ({scale:g} of the full lengths), RK4",
            "policy": "continuous batching, " + ("shortest work first" if policy == "sjf" else "arrival order")
                      + ", occupancy target 1M points, 128-step chunks",
            "served": ss, "server_time_s": srv.stats, "plain": sp, "solves_per_hour_speedup": ss["solves_per_hour"] / sp["solves_per_hour"],
            "p99_ratio_plain_over_served": sp["p99_s"] / ss["p99_s"], "results_bitwise_equal_alone": same}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--timing-steps", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--integrator", choices=list(INTEGRATORS), default="rk4")
    ap.add_argument("--workload", choices=list(WORKLOADS), default="single_16384x64_plunge",
                    help="the single solve timed (the default is the headline configuration; the a=0.99 "
                         "ladder grids are larger single solves of the same path)")
    ap.add_argument("--wpb", type=int, default=0, help="wavefronts per block (0 = default)")
    ap.add_argument("--wpe", type=int, default=0, help="occupancy target, waves per SIMD (0 = default)")
    ap.add_argument("--rows", type=int, default=0, help="theta rows per wavefront, 2, 4 or 8 (0 = default)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--batch-sections", choices=["both", "mixed", "small"], default="both",
                    help="which batch sections run: the mixed 72 solves, the 72 at 4096x32, or both")
    ap.add_argument("--no-plain", action="store_true",
                    help="batch sections without the one-at-a-time comparison (profiler runs of the batched kernels)")
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the batch/serve solve lengths")
    ap.add_argument("--serve-seconds", type=float, default=1.0)
    ap.add_argument("--serve-policy", choices=["fifo", "sjf"], default="sjf",
                    help="admission order of waiting requests: arrival order or shortest work first")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 2 GiB device to device 4 times (a kernel of known traffic, larger than "
                         "the last-level cache, for scaling the HBM counters of a profiler run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(independent solves shard one per GPU with no communication)", file=sys.stderr)
        return 2

    from bh_ringdown_amd import teuk, tune, workloads
    if args.copy_calibration:
        import torch
        x = torch.ones(256 << 20, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(4):
            y.copy_(x)
        torch.cuda.synchronize()
        log("bench: copy calibration, 4 x (2 GiB read + 2 GiB written)")
        del x, y
    integ = INTEGRATORS[args.integrator]
    maker, size = WORKLOADS[args.workload]
    p = getattr(workloads, maker)(*(size or ()), integrator=integ)
    need = args.warmup + args.steps + args.timing_steps
    p.T = max(p.T, need * p.dt * (1.0 + 1e-12))
    nt = int(p.T / p.dt)
    _, r_c, th = teuk.grid(p)
    rows = workloads.plunge_rows(p, nt, eta=1e-3, iota_deg=30.0, seed=0)
    pts = p.N * p.M
    log(f"bench: workload {args.workload} N={p.N} M={p.M} warmup={args.warmup} steps={args.steps}")
    # launch shape: the tuned one for this grid and integrator (committed
    # table), unless given on the command line; None = the library default
    tiling = tune.best_solver_tiling(p.N, p.M, integ)
    if args.wpb or args.wpe or args.rows:
        tiling = {"waves_per_block": args.wpb or 1, "waves_per_eu": args.wpe or 2, "theta_rows": args.rows or 2}
    with teuk.Solver(p, r_c, th) as s:
        if tiling:
            s.set_tiling(**tiling)
        s.set_trajectory(rows, nt)
        s.step(args.warmup)
        s.sync()
        t0 = time.perf_counter()
        s.step(args.steps)
        s.sync()
        el = time.perf_counter() - t0
        log(f"bench: timed {args.steps} steps in {el:.3f} s")
        # Kernel timing pass: HIP events around each stage launch, same stream.
        s.enable_timing(True)
        s.step(args.timing_steps)
        s.sync()
        launches, ms = s.timing()
    value = pts * args.steps / el
    per_launch_s = ms / 1e3 / launches
    # B_alg is per full step; a step is STAGES stage launches
    b_alg = B_ALG[args.integrator]
    achieved = (b_alg / STAGES[args.integrator]) * pts / per_launch_s / 1e9
    moved = (B_MOVED[args.integrator] / STAGES[args.integrator]) * pts / per_launch_s / 1e9
    traffic = pmc_traffic(args.integrator, p.N, p.M)
    out = {
        "metric": "FP64 point-updates/s (%HBM roofline)",
        "value": value,
        "unit": "point-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "N": p.N, "M": p.M, "rho": [p.xmin, p.xmax],
                   "a": p.a, "m": p.m, "dt_over_drho": 0.5,
                   "integrator": INTEGRATOR_NAMES[args.integrator],
                   "t_final": p.T,
                   "source": "seeded plunge, mass ratio 1e-3, inclination 30 deg, from 1.02 r_isco",
                   "extraction_rho": p.xmin + p.j_extract * p.dx,
                   "launch": tiling or "library default"},
        # headline: algorithmic bytes (RK4 in its accumulator form 512 B,
        # SSP-RK3 256 B per point-update) over the average stage launch; beside it the
        # state bytes the kernels move (RK4 stage-value form 384 B) and the
        # counter-measured HBM bytes per launch (committed rocprofv3 --pmc
        # summary of this grid), each over the same launch time
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_so2_half + k_so2_full (edges fused)" if args.integrator == "so2" else "k_stage (fused RK stage)", "bytes_per_point_update": b_alg,
                     "avg_launch_us": per_launch_s * 1e6,
                     "moved_bytes_per_point_update": B_MOVED[args.integrator], "achieved_moved": moved,
                     "frac_moved": moved / HBM_PEAK_GBS,
                     "achieved_counter": traffic / per_launch_s / 1e9 if traffic else None,
                     "frac_counter": traffic / per_launch_s / 1e9 / HBM_PEAK_GBS if traffic else None},
    }
    if not args.no_batch:
        if args.batch_sections in ("both", "mixed"):
            out["batch"] = batch_section(args.scale, plain=not args.no_plain)
        if args.batch_sections in ("both", "small"):
            out["batch_small_grids"] = batch_section(args.scale, grid=(4096, 32), plain=not args.no_plain)
    if not args.no_serve:
        out["serving"] = serve_section(args.scale, args.serve_seconds, args.serve_policy)
    if not args.no_cpu:
        log("bench: cpu baseline")
        out["cpu_baseline"] = cpu_baseline(p, rows)
        one = cpu_baseline(p, rows, seconds=6.0, threads=1)
        out["cpu_baseline"]["value_1_core"] = one["value"]
        out["cpu_baseline"]["sample_1_core"] = one["sample"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
