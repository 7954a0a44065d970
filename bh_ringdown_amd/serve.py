"""Serving many Teukolsky solves on one GPU with continuous batching.

Requests (spin a, mode m, inclination, mass ratio, grid, length) arrive over
time.  The server keeps one libteuk batch: every loop iteration it enqueues
a chunk of steps for all resident solves (one launch per RK stage for the
whole batch), and while that chunk runs on the GPU it prepares the requests
that have arrived (grid, plunge trajectory, device set-up) and admits them
as HBM budget frees up.  After each chunk it retires the solves that
reached their final time (with all their psi4 / h_lm records) and, every
`stream_every` chunks, streams the new records of the other resident
solves back to the host (each read is host time the GPU waits for).  A result does not depend on what else was
in the batch: batched steps are bitwise the steps of a solve run alone.
A solve that fails (the device sentinel found a non-finite or runaway value
in its fields, teuk_status) is retired on its own with `failed` set; the
other resident solves go on.

The plain loop it is compared with runs the same requests one at a time,
in arrival order, each as its own solver (one solve per launch).
"""
import math
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import teuk


@dataclass
class Request:
    id: int
    arrival: float          # seconds after the stream starts
    a: float
    m: int
    iota: float             # radians
    eta: float              # mass ratio
    N: int
    M: int
    t_final: float
    seed: int
    xmin: float = -400.0    # rho domain (the single-solve domain by default)
    xmax: float = 1200.0


@dataclass
class Result:
    id: int
    arrival: float
    admitted: float = math.nan
    done: float = math.nan
    steps: int = 0
    points: int = 0
    records: dict = field(default_factory=dict)     # output(): k, t, dedt, h
    psi: np.ndarray = None
    failed: bool = False    # retired by the divergence sentinel (records up to the failure)

    @property
    def latency(self):
        return self.done - self.arrival


MAX_IOTA = math.radians(85.0)   # the trajectory generator's limit (polar passes excluded)


def poisson_requests(rate, duration, seed, t_range=(500.0, 2000.0), grids=((4096, 8192, 16384), (32, 64)),
                     m_range=(1, 4), eta_range=(1e-4, 1e-2), a_max=0.99, max_requests=None):
    """A seeded Poisson request stream: exponential inter-arrival times at
    `rate` per second for `duration` seconds; a ~ U[0, a_max), m ~ U{m_range},
    inclination ~ U[0, pi/2) (capped at 85 degrees), mass ratio log-uniform,
    N and M uniform over `grids`, final time ~ U[t_range)."""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration or (max_requests is not None and len(out) >= max_requests):
            break
        out.append(Request(
            id=len(out), arrival=t, a=float(rng.uniform(0.0, a_max)),
            m=int(rng.integers(m_range[0], m_range[1] + 1)),
            iota=min(float(rng.uniform(0.0, 0.5 * math.pi)), MAX_IOTA),
            eta=float(math.exp(rng.uniform(math.log(eta_range[0]), math.log(eta_range[1])))),
            N=int(rng.choice(grids[0])), M=int(rng.choice(grids[1])),
            t_final=float(rng.uniform(*t_range)), seed=int(rng.integers(1 << 31))))
    return out


def request_params(req, integrator=1, Fr=2):
    """Solver parameters of a request: dt = 0.5 drho, psi4 at the outer
    boundary (scri, the reference's extraction column)."""
    p = teuk.default_params(req.N, req.M, req.xmin, req.xmax, req.a, req.m, dt_over_dx=0.5, T=req.t_final)
    p.integrator = integrator
    p.Fr = Fr
    return p


def prepare_host(req, integrator=1):
    """Host set-up of one request (no GPU calls): parameters, grid and the
    seeded plunge's trajectory rows."""
    p = request_params(req, integrator)
    _, r_c, th = teuk.grid(p)
    nt = p.nsteps()
    rows = teuk.plunge_trajectory(req.a, req.eta, req.iota, p.dt / 2, (2 * nt + 6) * p.dt / 2, seed=req.seed)
    return p, r_c, th, rows


def prepare_device(p, r_c, th, rows):
    """Device set-up of a prepared request: the solver and its source schedule."""
    s = teuk.Solver(p, r_c, th)
    s.set_trajectory(rows, p.nsteps())
    return p, s


def prepare(req, integrator=1):
    """Host and device set-up of one request: grid, seeded plunge, solver."""
    return prepare_device(*prepare_host(req, integrator))


class Server:
    """Continuous batching over a libteuk batch with an HBM budget.

    Admission is limited by the HBM budget and by an occupancy target: once
    the resident solves together update `target_points` grid points per step
    (enough wavefronts to keep every compute unit streaming), admitting more
    would not raise throughput and would only stretch every resident solve's
    latency, so later requests wait for a resident one to finish.  Waiting
    requests are admitted shortest first (grid points x steps) when
    `shortest_first`, else in arrival order."""

    def __init__(self, hbm_budget, chunk=128, integrator=1, tiling=None, max_resident=64, target_points=1 << 20,
                 shortest_first=True, stream_every=8):
        self.hbm_budget = int(hbm_budget)
        self.chunk = chunk
        self.integrator = integrator
        self.tiling = tiling or {}
        self.max_resident = max_resident
        self.target_points = int(target_points)
        self.shortest_first = shortest_first
        self.stream_every = max(1, int(stream_every))   # chunks between record reads of a resident solve
        self.stats = {}

    def run(self, requests, clock=time.perf_counter, on_records=None, deadline=None, sample_s=None, on_sample=None):
        """Serve `requests` (sorted by arrival); returns {id: Result}.

        deadline: seconds after the start at which serving stops (after the
        chunk in flight); solves still resident then are closed with their
        records so far and `done` unset, waiting requests are left unserved.
        sample_s: every that many seconds append (t, arrived, admitted,
        completed, waiting, resident solves, resident points, point-updates
        done) to stats["timeline"] and pass it to on_sample."""
        pending = sorted(requests, key=lambda r: r.arrival)
        results = {r.id: Result(r.id, r.arrival) for r in pending}
        queue = []            # arrived, not admitted
        live = {}             # id -> (req, params, solver, records read)
        t0 = clock()
        now = lambda: clock() - t0
        # host time spent preparing requests, enqueuing chunks, waiting for
        # the GPU and streaming records back
        st = self.stats = {"prepare_s": 0.0, "setup_s": 0.0, "enqueue_s": 0.0, "sync_s": 0.0, "stream_s": 0.0, "chunks": 0,
                           "failed": 0, "unfinished": 0, "unserved": 0, "point_updates": 0}
        if sample_s:
            st["timeline"] = []
        next_sample = sample_s or math.inf
        n_admitted = n_done = 0

        def collect(rid):
            # the records streamed so far, joined once (not per chunk)
            outs, psis = live[rid][4], live[rid][5]
            if outs:
                res = results[rid]
                res.records = {key: np.concatenate([o[key] for o in outs]) for key in outs[0]}
                res.psi = np.concatenate(psis)

        def pu_done():
            # completed steps of finished and resident solves
            return st["point_updates"] + sum(v[1].N * v[1].M * (v[1].nsteps() - v[2].remaining())
                                             for v in live.values())
        worker = ThreadPoolExecutor(max_workers=1)
        prepared = {}   # request id -> future of its host set-up

        def timed_prepare(req):
            # host work only: every GPU call stays on this thread (a graph
            # capture on the batch stream must not see another thread's
            # allocations or copies)
            ts = time.perf_counter()
            try:
                return prepare_host(req, self.integrator)
            finally:
                st["prepare_s"] += time.perf_counter() - ts

        batch = teuk.Batch(integrator=self.integrator, hbm_budget=self.hbm_budget)
        if self.tiling:
            batch.set_tiling(chunk=self.chunk, **self.tiling)
        else:
            # measured on an MI355X over the 16-request window of the
            # Poisson stream at full lengths (N 4096-16384, M 32/64, about
            # 1 M resident points): 2-row wavefronts, 1 per block, built for
            # 3 waves per SIMD, 11.07 s; 4-row 11.21 s; 8-row 2 per block
            # 11.35 s
            batch.set_tiling(1, 3, 0, self.chunk, theta_rows=2)
        try:
            while pending or queue or live:
                if deadline is not None and now() >= deadline:
                    break
                # 1. enqueue a chunk for everything resident (asynchronous)
                if live:
                    ts = clock()
                    batch.step(self.chunk)
                    st["enqueue_s"] += clock() - ts
                    st["chunks"] += 1
                # 2. meanwhile: arrivals and admission by footprint
                while pending and pending[0].arrival <= now():
                    queue.append(pending.pop(0))
                if self.shortest_first:
                    queue.sort(key=lambda r: (r.N * r.M * r.t_final / ((r.xmax - r.xmin) / r.N), r.arrival))
                while queue and len(live) < self.max_resident:
                    resident = sum(v[1].N * v[1].M for v in live.values())
                    if live and resident >= self.target_points:
                        break
                    req = queue[0]
                    p = request_params(req, self.integrator)
                    if not live and teuk.footprint(p) > self.hbm_budget:
                        raise ValueError(f"request {req.id} alone exceeds the HBM budget")
                    if batch.info()["bytes"] + teuk.footprint(p) > self.hbm_budget:
                        break
                    # Host set-up runs on the worker thread while the chunk
                    # runs; admit when it is ready (wait for it only with
                    # nothing resident to step).
                    if req.id not in prepared:
                        # keep the host set-ups of the next few waiting
                        # requests only (shortest-first admission can move
                        # a prepared request back in the queue)
                        if len(prepared) >= 8:
                            near = {r.id for r in queue[:8]}
                            for rid in [rid for rid in prepared if rid not in near]:
                                del prepared[rid]
                        prepared[req.id] = worker.submit(timed_prepare, req)
                    if live and not prepared[req.id].done():
                        break
                    fut = prepared.pop(req.id)
                    try:
                        ts = clock()
                        p, s = prepare_device(*fut.result())
                        st["setup_s"] += clock() - ts
                    except teuk.TeukError:
                        # its set-up failed (e.g. no trajectory for its
                        # parameters): retire it, serve the others
                        queue.pop(0)
                        results[req.id].failed = True
                        results[req.id].done = now()
                        st["failed"] += 1
                        continue
                    if not batch.add(s):
                        s.close()
                        break
                    queue.pop(0)
                    live[req.id] = [req, p, s, 0, [], []]   # records read, their chunks (records, psi)
                    results[req.id].admitted = now()
                    n_admitted += 1
                if not live:
                    if pending:
                        time.sleep(max(0.0, min(0.05, pending[0].arrival - now())))
                    continue
                # 3. chunk done: stream records, retire finished and failed solves
                ts = clock()
                try:
                    batch.sync()
                    any_failed = False
                except teuk.TeukError as e:
                    if e.code != teuk.TEUK_EDIVERGED:
                        raise
                    any_failed = True
                st["sync_s"] += clock() - ts
                ts = clock()
                for rid in list(live):
                    req, p, s, nread, outs, psis = live[rid]
                    if any_failed and s.status() == teuk.TEUK_EDIVERGED:
                        res = results[rid]
                        res.failed = True
                        res.done = now()
                        out = s.output(first=nread)     # its records up to the failure
                        if len(out["k"]):
                            outs.append(out)
                            psis.append(s.psi(first=nread))
                        collect(rid)
                        batch.remove(s)
                        s.close()
                        del live[rid]
                        st["failed"] += 1
                        continue
                    finished = s.remaining() == 0
                    if not finished and st["chunks"] % self.stream_every:
                        continue   # read its records at a later chunk (each read costs host time the GPU waits for)
                    out = s.output(first=nread)
                    n_new = len(out["k"])
                    if n_new:
                        psi = s.psi(first=nread)
                        outs.append(out)
                        psis.append(psi)
                        if on_records is not None:
                            on_records(rid, out, psi)
                        live[rid][3] = nread + n_new
                    if finished:
                        res = results[rid]
                        res.done = now()
                        res.steps = p.nsteps()
                        res.points = p.N * p.M * res.steps
                        st["point_updates"] += res.points
                        n_done += 1
                        collect(rid)
                        batch.remove(s)
                        s.close()
                        del live[rid]
                st["stream_s"] += clock() - ts
                if now() >= next_sample:
                    n_arrived = len(requests) - len(pending)
                    row = (round(now(), 3), n_arrived, n_admitted, n_done, len(queue), len(live),
                           sum(v[1].N * v[1].M for v in live.values()), pu_done())
                    st["timeline"].append(row)
                    if on_sample is not None:
                        on_sample(row)
                    next_sample += sample_s
        finally:
            st["unfinished"] = len(live)
            st["unserved"] = len(pending) + len(queue)
            if live:
                try:
                    batch.sync()
                except teuk.TeukError:
                    pass
                st["point_updates"] = pu_done()
            for rid, (req, p, s, *_) in live.items():
                collect(rid)
                s.close()
            worker.shutdown(wait=True)       # host set-ups never admitted are dropped
            batch.close()
        return results


def run_plain(requests, clock=time.perf_counter, integrator=1, deadline=None, stats=None):
    """The same requests one at a time in arrival order, each as its own
    solver stepped alone; returns {id: Result}.  With a deadline the loop
    stops there (a solve is stepped in pieces of 4096 steps so that it can
    stop in time; its records so far are kept, `done` unset); `stats` then
    receives the point-updates done and the unfinished / unserved counts."""
    results = {}
    t0 = clock()
    now = lambda: clock() - t0
    pu = 0
    unfinished = 0
    ordered = sorted(requests, key=lambda r: r.arrival)
    served = 0
    for req in ordered:
        if deadline is not None and now() >= deadline:
            break
        wait = req.arrival - now()
        if wait > 0:
            time.sleep(wait if deadline is None else min(wait, max(0.0, deadline - now())))
            if deadline is not None and now() >= deadline:
                break
        served += 1
        res = Result(req.id, req.arrival, admitted=now())
        p, s = prepare(req, integrator)
        n = p.nsteps()
        done_steps = 0
        try:
            piece = n if deadline is None else 4096
            while done_steps < n:
                k = min(piece, n - done_steps)
                s.step(k)
                try:
                    s.sync()
                except teuk.TeukError as e:
                    if e.code != teuk.TEUK_EDIVERGED:
                        raise
                    res.failed = True
                done_steps += k
                if res.failed or (deadline is not None and now() >= deadline):
                    break
            res.records = s.output()
            res.psi = s.psi()
        finally:
            s.close()
        pu += p.N * p.M * done_steps
        if done_steps == n or res.failed:
            res.done = now()
            res.steps = n
            res.points = p.N * p.M * n
        else:
            unfinished += 1
        results[req.id] = res
    if stats is not None:
        stats.update({"point_updates": pu, "unfinished": unfinished, "unserved": len(ordered) - served})
    return results


def summary(results, wall=None):
    """p50 / p99 latency, solves per hour and point-updates per second of
    the completed solves (failed ones are counted apart)."""
    done = [r for r in results.values() if not math.isnan(r.done) and not r.failed]
    failed = sum(1 for r in results.values() if r.failed)
    lat = np.array([r.latency for r in done])
    span = wall if wall is not None else max(r.done for r in done)
    pts = sum(r.points for r in done)
    pct = lambda q: float(np.percentile(lat, q)) if len(lat) else None
    return {"solves": len(done), "failed": failed, "p50_s": pct(50),
            "p99_s": pct(99), "solves_per_hour": 3600.0 * len(done) / span,
            "point_updates_per_s": pts / span, "wall_s": span}


def request_work(req):
    """Point-updates of a request (grid points x steps)."""
    return req.N * req.M * request_params(req).nsteps()


def main(argv=None):
    """A sustained-load run: a seeded Poisson stream through the server for
    a fixed wall time, reporting the offered load against what was served.

        python -m bh_ringdown_amd.serve [--rate 20] [--seconds 60] [--deadline 75]
                                        [--scale 1.0] [--policy sjf|fifo] [--plain]
    """
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=float, default=20.0, help="requests per second")
    ap.add_argument("--seconds", type=float, default=60.0, help="length of the arrival stream")
    ap.add_argument("--deadline", type=float, default=None, help="stop serving at this time (default seconds + 15)")
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the request lengths t_final ~ U[500, 2000)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--policy", choices=["sjf", "fifo"], default="sjf")
    ap.add_argument("--sample", type=float, default=5.0, help="seconds between timeline rows")
    ap.add_argument("--plain", action="store_true", help="also run the one-at-a-time loop to the same deadline")
    args = ap.parse_args(argv)
    deadline = args.deadline if args.deadline is not None else args.seconds + 15.0
    reqs = poisson_requests(args.rate, args.seconds, args.seed, t_range=(500.0 * args.scale, 2000.0 * args.scale))
    work = sum(request_work(r) for r in reqs)
    head = {"requests": len(reqs), "rate_per_s": args.rate, "arrival_window_s": args.seconds, "deadline_s": deadline,
            "offered_point_updates": work, "offered_point_updates_per_s": work / args.seconds,
            "t_final_range": [500.0 * args.scale, 2000.0 * args.scale], "policy": args.policy}
    print(json.dumps({"stream": head}), flush=True)
    cols = ("t", "arrived", "admitted", "completed", "waiting", "resident", "resident_points", "point_updates")
    srv = Server(hbm_budget=64 << 30, chunk=128, target_points=1 << 20, shortest_first=(args.policy == "sjf"))
    res = srv.run(reqs, deadline=deadline, sample_s=args.sample,
                  on_sample=lambda row: print(json.dumps({"served_at": dict(zip(cols, row))}), flush=True))
    st = srv.stats
    by_id = {r.id: r for r in reqs}
    failed = [{k: getattr(by_id[i], k) for k in ("id", "a", "m", "iota", "eta", "N", "M", "t_final", "seed")}
              for i, r in sorted(res.items()) if r.failed]
    out = {"stream": head,
           "served": {**summary(res, deadline), "failed_requests": failed,
                      "sustained_point_updates_per_s": st["point_updates"] / deadline,
                      "unfinished": st["unfinished"], "unserved": st["unserved"],
                      "host_s": {k: st[k] for k in ("prepare_s", "setup_s", "enqueue_s", "sync_s", "stream_s", "chunks")}}}
    if args.plain:
        pst = {}
        resp = run_plain(reqs, deadline=deadline, stats=pst)
        out["plain"] = {**summary(resp, deadline), "sustained_point_updates_per_s": pst["point_updates"] / deadline,
                        "unfinished": pst["unfinished"], "unserved": pst["unserved"]}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
