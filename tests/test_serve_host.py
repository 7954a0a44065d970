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
    want = (defaults[key]["waves_per_block"], defaults[key]["waves_per_eu"])
    assert tune.best_solver_tiling(16384, 64, 1) == want
    other = tune.solver_key(4096, 32, 1)
    cache.write_text(json.dumps({key: {"waves_per_block": 1, "waves_per_eu": 4, "us_per_launch": 1.0}}))
    assert tune.best_solver_tiling(16384, 64, 1) == (1, 4)
    assert tune.best_solver_tiling(4096, 32, 1) == (defaults[other]["waves_per_block"],
                                                     defaults[other]["waves_per_eu"])
    assert tune.best_solver_tiling(1024, 16, 1) is None
