"""Autotuner for the stage-kernel launch shape and the batch schedule.

Candidates (all give bitwise-identical results; only speed differs):
  waves_per_block  1, 2, 4  -- wavefronts (64 rho columns x a chunk of theta
                               rows each) stacked along theta in one block
  waves_per_eu     2, 3, 4  -- occupancy the kernel is compiled for (3 fits
                               2-row wavefronts with a few spilled registers;
                               4 forces heavy spilling, kept as a candidate
                               for single solves because the trade-off is
                               measured, not assumed)
  theta_rows       2, 4, 8  -- theta rows per wavefront (short chunks: more
                               wavefronts and fewer registers; long ones:
                               fewer theta halo re-reads)
  order            0, 1     -- batch tile table: contiguous run per XCD, or
                               dealt in list order
  chunk            16, 32, 64 -- steps per captured hipGraph
  group_bytes      0, 256 MB  -- batch cache groups: off, or members stepped
                               a chunk at a time in groups whose field state
                               fits the last-level cache

tune_solver() times single-solve stage launches with HIP events;
tune_batch() times whole batch steps.  Winners are kept in a JSON cache
(TEUK_TUNE_CACHE, default ~/.cache/bh_ringdown_amd/tuning.json) keyed by
grid shape and integrator, and looked up by best_solver_tiling() /
best_batch_tiling() / best_batch_groups().
"""
import itertools
import json
import os
import time

import numpy as np

from . import teuk

CACHE = os.environ.get("TEUK_TUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "bh_ringdown_amd",
                                                          "tuning.json")
# Results of a tuning run on an MI355X, used where the local cache has no entry.
DEFAULTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_defaults.json")


def _read(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _load():
    data = _read(DEFAULTS)
    data.update(_read(CACHE))
    return data


def _store(key, value):
    data = _read(CACHE)
    data[key] = value
    os.makedirs(os.path.dirname(CACHE), exist_ok=True)
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def solver_key(N, M, integrator):
    return f"solver/{N}x{M}/i{integrator}"


def batch_key(shapes, integrator):
    """Shapes with their counts: "batch/4096x32*24,8192x32*32,16384x64*16/i1"."""
    counts = {}
    for sh in shapes:
        counts[sh] = counts.get(sh, 0) + 1
    return "batch/" + ",".join(f"{n}x{m}*{c}" for (n, m), c in sorted(counts.items())) + f"/i{integrator}"


def _probe_params(N, M, integrator, steps):
    # a stable, source-free probe grid of the right shape (dt / dtheta <= 2)
    p = teuk.default_params(N, M, -0.04 * N, 0.06 * N, 0.7, 2, dt_over_dx=0.5)
    p.T = (steps + 1) * p.dt
    p.integrator = integrator
    return p


def tune_solver(N, M, integrator=1, steps=60, repeats=2, store=True, rows=(2, 4, 8)):
    """Time every (waves_per_block, waves_per_eu, theta_rows) on an N x M
    grid; returns (best, {candidate: us per stage launch})."""
    times = {}
    for wpb, wpe, tr in itertools.product((1, 2, 4), (2, 3, 4), rows):
        p = _probe_params(N, M, integrator, steps * repeats + 10)
        with teuk.Solver(p) as s:
            s.set_tiling(wpb, wpe, tr)
            s.step(5)
            s.sync()
            s.enable_timing(True)
            for _ in range(repeats):
                s.step(steps)
            s.sync()
            n, ms = s.timing()
        times[(wpb, wpe, tr)] = 1e3 * ms / n
    best = min(times, key=times.get)
    if store:
        _store(solver_key(N, M, integrator), {"waves_per_block": best[0], "waves_per_eu": best[1],
                                              "theta_rows": best[2], "us_per_launch": times[best]})
    return best, times


# Cache-group sizes tried for batches: off, and groups whose field state fits
# the 256 MB last-level cache.
GROUP_BYTES = (0, 256 << 20)


def tune_batch(shapes, integrator=1, steps=64, store=True, chunks=(16, 32, 64), group_bytes=GROUP_BYTES,
               rows=(2, 4, 8), wpes=(2, 3)):
    """Time whole batch steps of probe solvers of the given (N, M) shapes
    (one probe per entry, so repeat a shape to tune a batch of many) for
    every (waves_per_block, waves_per_eu, order, chunk, cache-group bytes,
    theta_rows); returns (best, {candidate: seconds per step})."""
    times = {}
    for wpb, wpe, order, chunk, gb, tr in itertools.product((1, 2, 4), wpes, (0, 1), chunks, group_bytes, rows):
        solvers = [teuk.Solver(_probe_params(n, m, integrator, 3 * steps + 10)) for n, m in shapes]
        try:
            with teuk.Batch(integrator=integrator) as b:
                b.set_tiling(wpb, wpe, order, chunk, tr)
                b.set_cache_groups(gb)
                for s in solvers:
                    b.add(s)
                b.step(chunk)
                b.sync()
                t0 = time.perf_counter()
                b.step(2 * steps)
                b.sync()
                times[(wpb, wpe, order, chunk, gb, tr)] = (time.perf_counter() - t0) / (2 * steps)
        finally:
            for s in solvers:
                s.close()
    best = min(times, key=times.get)
    if store:
        _store(batch_key(shapes, integrator), {"waves_per_block": best[0], "waves_per_eu": best[1],
                                               "order": best[2], "chunk": best[3], "group_bytes": best[4],
                                               "theta_rows": best[5], "s_per_step": times[best]})
    return best, times


def best_solver_tiling(N, M, integrator=1):
    """Tuned launch shape for a grid as Solver.set_tiling keyword arguments,
    or None."""
    e = _load().get(solver_key(N, M, integrator))
    if not e:
        return None
    return {"waves_per_block": e["waves_per_block"], "waves_per_eu": e["waves_per_eu"],
            "theta_rows": e.get("theta_rows", 8)}


def best_batch_tiling(shapes, integrator=1):
    """Tuned batch schedule as set_tiling keyword arguments, or None.  Falls
    back to the entry for the same set of shapes (any counts), then to one
    for the same set of row counts M, if the exact batch was not tuned."""
    data = _load()
    e = data.get(batch_key(shapes, integrator))

    def kshapes(k):
        return {tuple(map(int, x.split("*")[0].split("x"))) for x in k.split("/")[1].split(",")}

    if e is None:
        cands = [(k, v) for k, v in sorted(data.items()) if k.startswith("batch/") and k.endswith(f"/i{integrator}")]
        e = next((v for k, v in cands if kshapes(k) == set(shapes)), None)
        if e is None:
            ms = {m for _, m in shapes}
            e = next((v for k, v in cands if {m for _, m in kshapes(k)} == ms), None)
    if not e:
        return None
    return {"waves_per_block": e["waves_per_block"], "waves_per_eu": e["waves_per_eu"], "order": e["order"],
            "chunk": e["chunk"], "theta_rows": e.get("theta_rows", 8)}


def best_batch_groups(shapes, integrator=1):
    """Tuned cache-group bytes for a batch (0 = off, also when not tuned)."""
    data = _load()
    e = data.get(batch_key(shapes, integrator))
    return int(e.get("group_bytes", 0)) if e else 0


def main(argv=None):
    """python -m bh_ringdown_amd.tune [--solver NxM ...] [--batch NxM,NxM,...]"""
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--solver", nargs="*", default=[])
    ap.add_argument("--batch", nargs="*", default=[])
    ap.add_argument("--integrator", type=int, default=1)
    ap.add_argument("--chunks", default="16,32,64", help="steps per captured graph tried for batches")
    args = ap.parse_args(argv)
    for s in args.solver:
        N, M = map(int, s.split("x"))
        best, t = tune_solver(N, M, args.integrator)
        print(json.dumps({"solver": s, "best": best, "us": {",".join(map(str, k)): round(v, 3) for k, v in t.items()}}),
              flush=True)
    for spec in args.batch:
        shapes = []
        for x in spec.split(","):   # NxM or NxM*count
            sh, _, cnt = x.partition("*")
            shapes += [tuple(map(int, sh.split("x")))] * int(cnt or 1)
        best, t = tune_batch(shapes, args.integrator, chunks=tuple(int(c) for c in args.chunks.split(",")))
        print(json.dumps({"batch": spec, "best": best,
                          "ms_per_step": {",".join(map(str, k)): round(1e3 * v, 4) for k, v in t.items()}}),
              flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
