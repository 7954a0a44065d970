"""Host-side logic of the serving loop and the launch-shape tuner, without a
GPU: the seeded request stream, the solver parameters of a request, the
latency/throughput summary, and the tuner's committed defaults under a local
cache."""
import json
import math

import numpy as np
import pytest

from bh_ringdown_amd import serve, tune


def test_poisson_stream_is_seeded_and_within_its_ranges():
    a = serve.poisson_requests(20.0, 30.0, seed=3)
    b = serve.poisson_requests(20.0, 30.0, seed=3)
    c = serve.poisson_requests(20.0, 30.0, seed=4)
    assert [r.arrival for r in a] == [r.arrival for r in b]
    assert [r.arrival for r in a] != [r.arrival for r in c]
    t = np.array([r.arrival for r in a])
    assert np.all(np.diff(t) > 0) and t[-1] < 30.0
    # Poisson count: mean 600, standard deviation ~24.5
    assert abs(len(a) - 600) < 5 * math.sqrt(600)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 1 / 20.0) < 0.2 / 20.0
    for r in a:
        assert 0.0 <= r.a < 0.99 and 1 <= r.m <= 4
        assert 0.0 <= r.iota <= serve.MAX_IOTA
        assert 1e-4 <= r.eta <= 1e-2
        assert r.N in (4096, 8192, 16384) and r.M in (32, 64)
        assert 500.0 <= r.t_final < 2000.0
    assert len(serve.poisson_requests(20.0, 30.0, seed=3, max_requests=7)) == 7


def test_request_parameters():
    r = serve.Request(id=0, arrival=0.0, a=0.3, m=2, iota=0.2, eta=1e-3, N=4096, M=32, t_final=800.0, seed=1)
    p = serve.request_params(r, integrator=1, Fr=2)
    assert p.dx == 1600.0 / 4096 and p.dt == 0.5 * p.dx
    assert p.T == 800.0 and p.integrator == 1 and p.Fr == 2 and p.j_extract == -1
    assert p.nsteps() == int(800.0 / p.dt)


def test_summary_percentiles_and_rates():
    res = {}
    lat = [1.0, 2.0, 3.0, 4.0, 10.0]
    for i, l in enumerate(lat):
        res[i] = serve.Result(i, arrival=float(i), admitted=float(i), done=float(i) + l, steps=100, points=1000)
    res[99] = serve.Result(99, arrival=0.0)          # not finished: ignored
    s = serve.summary(res, wall=20.0)
    assert s["solves"] == 5
    assert s["p50_s"] == np.percentile(lat, 50) and s["p99_s"] == np.percentile(lat, 99)
    assert s["solves_per_hour"] == 3600.0 * 5 / 20.0
    assert s["point_updates_per_s"] == 5000 / 20.0


def test_tuner_defaults_and_local_cache(tmp_path, monkeypatch):
    """With no local measurements the committed table answers; a local cache
    entry overrides it for its key only."""
    cache = tmp_path / "tuning.json"
    monkeypatch.setattr(tune, "CACHE", str(cache))
    defaults = json.load(open(tune.DEFAULTS))
    key = tune.solver_key(16384, 64, 1)
    assert key in defaults
    def shape(e):
        return {"waves_per_block": e["waves_per_block"], "waves_per_eu": e["waves_per_eu"],
                "theta_rows": e.get("theta_rows", 8)}

    assert tune.best_solver_tiling(16384, 64, 1) == shape(defaults[key])
    other = tune.solver_key(4096, 32, 1)
    cache.write_text(json.dumps({key: {"waves_per_block": 1, "waves_per_eu": 4, "theta_rows": 4,
                                       "us_per_launch": 1.0}}))
    assert tune.best_solver_tiling(16384, 64, 1) == {"waves_per_block": 1, "waves_per_eu": 4, "theta_rows": 4}
    assert tune.best_solver_tiling(4096, 32, 1) == shape(defaults[other])
    assert tune.best_solver_tiling(1024, 16, 1) is None


class _FakeSolver:
    """Stands in for teuk.Solver in the server's host logic: steps advance a
    counter, every step writes one record, and a planted failure step makes
    the solve report TEUK_EDIVERGED from then on."""

    def __init__(self, p, fail_at=None):
        self.p, self.k, self.fail_at, self.closed = p, 0, fail_at, False

    def advance(self, n):
        self.k = min(self.p.nsteps(), self.k + n)

    def status(self):
        from bh_ringdown_amd import teuk
        return teuk.TEUK_EDIVERGED if self.fail_at is not None and self.k >= self.fail_at else teuk.TEUK_OK

    def output(self, first=0):
        k = np.arange(first, self.k)
        return {"k": k, "write": k * 0, "t": k * 1.0, "dedt": k * 0.0, "h": np.zeros((len(k), 14))}

    def psi(self, first=0):
        return np.zeros((max(0, self.k - first), 2, self.p.M))

    def remaining(self):
        return self.p.nsteps() - self.k

    def close(self):
        self.closed = True


def _fake_host(req, integrator=1):
    """prepare_host's stand-in: the request's parameters, and the request
    itself in place of the grid (the fake device set-up reads its id)."""
    return serve.request_params(req, integrator), req, None, None


class _FakeBatch:
    def __init__(self, integrator=1, hbm_budget=0):
        self.members = []

    def set_tiling(self, *a, **k):
        pass

    def add(self, s):
        self.members.append(s)
        return True

    def remove(self, s):
        self.members.remove(s)

    def step(self, n):
        for s in self.members:
            s.advance(n)

    def sync(self):
        from bh_ringdown_amd import teuk
        if any(s.status() == teuk.TEUK_EDIVERGED for s in self.members):
            raise teuk.TeukError(teuk.TEUK_EDIVERGED, "teuk_batch_sync")

    def info(self):
        return {"members": len(self.members), "active": len(self.members), "bytes": 0}

    def close(self):
        pass


def test_server_retires_a_failed_solve_and_keeps_the_others(monkeypatch):
    """One resident solve diverges part-way: the server marks it failed and
    closes it, and every other request still runs to its end with all of
    its records (the batch's own failure report is per member)."""
    from bh_ringdown_amd import teuk
    reqs = [serve.Request(id=i, arrival=0.0, a=0.5, m=2, iota=0.1, eta=1e-3, N=4096, M=32,
                          t_final=20.0 + 10.0 * i, seed=i) for i in range(4)]
    made = {}

    def fake_device(p, req, th, rows):
        s = _FakeSolver(p, fail_at=60 if req.id == 1 else None)
        made[req.id] = s
        return p, s

    monkeypatch.setattr(serve, "prepare_host", _fake_host)
    monkeypatch.setattr(serve, "prepare_device", fake_device)
    monkeypatch.setattr(serve.teuk, "Batch", _FakeBatch)
    srv = serve.Server(hbm_budget=1 << 40, chunk=16, target_points=1 << 30)
    res = srv.run(reqs)
    assert res[1].failed and not any(res[i].failed for i in (0, 2, 3))
    assert srv.stats["failed"] == 1
    assert made[1].closed and made[1].k < made[1].p.nsteps()
    for i in (0, 2, 3):
        p = made[i].p
        assert made[i].closed and made[i].k == p.nsteps()
        assert np.array_equal(res[i].records["k"], np.arange(p.nsteps()))
    s = serve.summary(res, wall=1.0)
    assert s["solves"] == 3 and s["failed"] == 1


def test_server_deadline_accounts_for_served_and_unserved_work(monkeypatch):
    """A stream longer than the server can finish by its deadline: serving
    stops there, resident solves are closed with their partial work counted,
    waiting requests are counted unserved, and the load timeline is
    consistent (counts never decrease, arrived >= admitted >= completed)."""
    reqs = [serve.Request(id=i, arrival=0.5 * i, a=0.5, m=2, iota=0.1, eta=1e-3, N=4096, M=32,
                          t_final=40.0, seed=i) for i in range(12)]
    made = {}

    def fake_device(p, req, th, rows):
        made[req.id] = _FakeSolver(p)
        return p, made[req.id]

    ticks = iter(np.arange(0.0, 1e6, 0.25))
    monkeypatch.setattr(serve, "prepare_host", _fake_host)
    monkeypatch.setattr(serve, "prepare_device", fake_device)
    monkeypatch.setattr(serve.teuk, "Batch", _FakeBatch)
    monkeypatch.setattr(serve.time, "sleep", lambda s: None)
    # one resident solve at a time, 64 steps per chunk, clock +0.25 s per reading
    srv = serve.Server(hbm_budget=1 << 40, chunk=64, target_points=1, max_resident=1)
    rows = []
    res = srv.run(reqs, clock=lambda: float(next(ticks)), deadline=30.0, sample_s=2.0, on_sample=rows.append)
    st = srv.stats
    done = [r for r in res.values() if not math.isnan(r.done)]
    assert 0 < len(done) < len(reqs)
    assert st["unfinished"] == 1 and st["unserved"] == len(reqs) - len(done) - 1
    partial = [s for i, s in made.items() if math.isnan(res[i].done)]
    assert len(partial) == 1 and 0 < partial[0].k < partial[0].p.nsteps() and partial[0].closed
    want = sum(r.points for r in done) + partial[0].p.N * partial[0].p.M * partial[0].k
    assert st["point_updates"] == want
    assert rows == st["timeline"] and len(rows) >= 3
    t = np.array(rows)
    assert np.all(np.diff(t, axis=0)[:, [0, 1, 2, 3, 7]] >= 0)
    assert np.all(t[:, 1] >= t[:, 2]) and np.all(t[:, 2] >= t[:, 3])
    s = serve.summary(res, wall=30.0)
    assert s["solves"] == len(done)


def test_server_retires_a_request_whose_setup_fails(monkeypatch):
    """A request whose set-up raises (its trajectory cannot be generated) is
    retired as failed; the requests around it are served."""
    from bh_ringdown_amd import teuk
    reqs = [serve.Request(id=i, arrival=0.0, a=0.5, m=2, iota=0.1, eta=1e-3, N=4096, M=32,
                          t_final=20.0, seed=i) for i in range(3)]

    def fake_host(req, integrator=1):
        if req.id == 1:
            raise teuk.TeukError(teuk.TEUK_EINVAL, "teuk_traj_plunge")
        return _fake_host(req, integrator)

    monkeypatch.setattr(serve, "prepare_host", fake_host)
    monkeypatch.setattr(serve, "prepare_device", lambda p, req, th, rows: (p, _FakeSolver(p)))
    monkeypatch.setattr(serve.teuk, "Batch", _FakeBatch)
    srv = serve.Server(hbm_budget=1 << 40, chunk=16, target_points=1 << 30)
    res = srv.run(reqs)
    assert res[1].failed and srv.stats["failed"] == 1
    assert not res[0].failed and not res[2].failed
    assert not math.isnan(res[0].done) and not math.isnan(res[2].done)


def test_batch_schedule_lookup(tmp_path, monkeypatch):
    """Batch keys count each shape; an exact entry gives the tiling and the
    cache-group size, the same shapes in other counts fall back to it, and
    an untuned batch gets no groups."""
    monkeypatch.setattr(tune, "CACHE", str(tmp_path / "tuning.json"))
    small = [(4096, 32)] * 72
    assert tune.batch_key(small, 1) == "batch/4096x32*72/i1"
    assert tune.batch_key([(8192, 32), (4096, 32), (8192, 32)], 1) == "batch/4096x32*1,8192x32*2/i1"
    defaults = json.load(open(tune.DEFAULTS))
    e = defaults[tune.batch_key(small, 1)]
    assert tune.best_batch_groups(small) == e["group_bytes"]
    assert tune.best_batch_tiling([(4096, 32)] * 8)["chunk"] == e["chunk"]   # same shapes, other count
    assert tune.best_batch_groups([(4096, 32)] * 8) == 0                     # groups only for the tuned batch
    assert tune.best_batch_tiling([(2048, 16)]) is None
