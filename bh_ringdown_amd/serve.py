"""Serving many Teukolsky solves on one GPU with continuous batching.

Requests (spin a, mode m, inclination, mass ratio, grid, length) arrive over
time.  The server keeps one libteuk batch: every loop iteration it enqueues
a chunk of steps for all resident solves (one launch per RK stage for the
whole batch), and while that chunk runs on the GPU it prepares the requests
that have arrived (grid, plunge trajectory, device set-up) and admits them
as HBM budget frees up.  After each chunk it streams the new psi4 / h_lm
records of every resident solve back to the host and retires the solves
that reached their final time.  A result does not depend on what else was
in the batch: batched steps are bitwise the steps of a solve run alone.
A solve that fails (the device sentinel found a non-finite or runaway value
in its fields, teuk_status) is retired on its own with `failed` set; the
other resident solves go on.

The plain loop it is compared with runs the same requests one at a time,
in arrival order, each as its own solver (one solve per launch).
"""
import math
import time
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


def prepare(req, integrator=1):
    """Host and device set-up of one request: grid, seeded plunge, solver."""
    p = request_params(req, integrator)
    _, r_c, th = teuk.grid(p)
    nt = p.nsteps()
    rows = teuk.plunge_trajectory(req.a, req.eta, req.iota, p.dt / 2, (2 * nt + 6) * p.dt / 2, seed=req.seed)
    s = teuk.Solver(p, r_c, th)
    s.set_trajectory(rows, nt)
    return p, s


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
                 shortest_first=True):
        self.hbm_budget = int(hbm_budget)
        self.chunk = chunk
        self.integrator = integrator
        self.tiling = tiling or {}
        self.max_resident = max_resident
        self.target_points = int(target_points)
        self.shortest_first = shortest_first
        self.stats = {}

    def run(self, requests, clock=time.perf_counter, on_records=None):
        """Serve `requests` (sorted by arrival); returns {id: Result}."""
        pending = sorted(requests, key=lambda r: r.arrival)
        results = {r.id: Result(r.id, r.arrival) for r in pending}
        queue = []            # arrived, not admitted
        live = {}             # id -> (req, params, solver, records read)
        t0 = clock()
        now = lambda: clock() - t0
        # host time spent preparing requests, enqueuing chunks, waiting for
        # the GPU and streaming records back
        st = self.stats = {"prepare_s": 0.0, "enqueue_s": 0.0, "sync_s": 0.0, "stream_s": 0.0, "chunks": 0,
                           "failed": 0}
        batch = teuk.Batch(integrator=self.integrator, hbm_budget=self.hbm_budget)
        if self.tiling:
            batch.set_tiling(chunk=self.chunk, **self.tiling)
        else:
            batch.set_tiling(4, 2, 0, self.chunk)
        try:
            while pending or queue or live:
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
                    ts = clock()
                    p, s = prepare(req, self.integrator)
                    st["prepare_s"] += clock() - ts
                    if not batch.add(s):
                        s.close()
                        break
                    queue.pop(0)
                    live[req.id] = [req, p, s, 0]
                    results[req.id].admitted = now()
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
                    req, p, s, nread = live[rid]
                    if any_failed and s.status() == teuk.TEUK_EDIVERGED:
                        res = results[rid]
                        res.failed = True
                        res.done = now()
                        batch.remove(s)
                        s.close()
                        del live[rid]
                        st["failed"] += 1
                        continue
                    out = s.output(first=nread)
                    n_new = len(out["k"])
                    if n_new:
                        res = results[rid]
                        psi = s.psi(first=nread)
                        for key, v in out.items():
                            res.records[key] = np.concatenate([res.records[key], v]) if key in res.records else v
                        res.psi = psi if res.psi is None else np.concatenate([res.psi, psi])
                        if on_records is not None:
                            on_records(rid, out, psi)
                        live[rid][3] = nread + n_new
                    if s.remaining() == 0:
                        res = results[rid]
                        res.done = now()
                        res.steps = p.nsteps()
                        res.points = p.N * p.M * res.steps
                        batch.remove(s)
                        s.close()
                        del live[rid]
                st["stream_s"] += clock() - ts
        finally:
            for req, p, s, _ in live.values():
                s.close()
            batch.close()
        return results


def run_plain(requests, clock=time.perf_counter, integrator=1):
    """The same requests one at a time in arrival order, each as its own
    solver stepped alone; returns {id: Result}."""
    results = {}
    t0 = clock()
    now = lambda: clock() - t0
    for req in sorted(requests, key=lambda r: r.arrival):
        wait = req.arrival - now()
        if wait > 0:
            time.sleep(wait)
        res = Result(req.id, req.arrival, admitted=now())
        p, s = prepare(req, integrator)
        try:
            n = p.nsteps()
            s.step(n)
            try:
                s.sync()
            except teuk.TeukError as e:
                if e.code != teuk.TEUK_EDIVERGED:
                    raise
                res.failed = True
            res.records = s.output()
            res.psi = s.psi()
        finally:
            s.close()
        res.done = now()
        res.steps = n
        res.points = p.N * p.M * n
        results[req.id] = res
    return results


def summary(results, wall=None):
    """p50 / p99 latency, solves per hour and point-updates per second of
    the completed solves (failed ones are counted apart)."""
    done = [r for r in results.values() if not math.isnan(r.done) and not r.failed]
    failed = sum(1 for r in results.values() if r.failed)
    lat = np.array([r.latency for r in done])
    span = wall if wall is not None else max(r.done for r in done)
    pts = sum(r.points for r in done)
    return {"solves": len(done), "failed": failed, "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)), "solves_per_hour": 3600.0 * len(done) / span,
            "point_updates_per_s": pts / span, "wall_s": span}
