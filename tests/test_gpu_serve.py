"""Continuous-batching server: requests arriving over time are admitted by
HBM footprint into one batch; each result (every h_lm record and psi4
snapshot) is bitwise the result of the same request solved alone."""
import math

import numpy as np
import pytest

from bh_ringdown_amd import serve, teuk

pytestmark = pytest.mark.gpu


def small_requests(n, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        reqs.append(serve.Request(
            id=i, arrival=0.03 * i, a=float(rng.uniform(0.0, 0.95)), m=int(rng.integers(1, 5)),
            iota=float(rng.uniform(0.0, 1.2)), eta=float(10 ** rng.uniform(-4, -2)),
            N=int(rng.choice([256, 512])), M=int(rng.choice([16, 32])),
            t_final=float(rng.uniform(20.0, 40.0)), seed=int(rng.integers(1 << 30)),
            xmin=-40.0, xmax=60.0))
    return reqs


def test_requests_same_alone_and_in_the_mix():
    reqs = small_requests(8, seed=11)
    print("serve: 8 small requests", flush=True)
    plain = serve.run_plain(reqs)
    fps = [teuk.footprint(serve.request_params(r)) for r in reqs]
    budget = int(2.5 * max(fps))               # at most a few resident at once
    seen = []
    srv = serve.Server(hbm_budget=budget, chunk=16, tiling={"waves_per_block": 2, "waves_per_eu": 2,
                                                             "order": 0})
    served = srv.run(reqs, on_records=lambda rid, out, psi: seen.append((rid, len(out["k"]))))
    assert set(served) == set(plain) == {r.id for r in reqs}
    for r in reqs:
        a, b = served[r.id], plain[r.id]
        assert not math.isnan(a.done) and a.done >= a.admitted >= a.arrival
        assert np.array_equal(a.records["k"], b.records["k"])
        assert np.array_equal(a.records["h"], b.records["h"])
        assert np.array_equal(a.records["dedt"], b.records["dedt"])
        assert np.array_equal(a.psi, b.psi)
    # records were streamed back in pieces while solves were resident
    streamed = {}
    for rid, n in seen:
        streamed[rid] = streamed.get(rid, 0) + n
    assert all(streamed[r.id] == len(plain[r.id].records["k"]) for r in reqs)
    assert max(sum(1 for rid, _ in seen if rid == r.id) for r in reqs) > 1
    s = serve.summary(served)
    assert s["solves"] == 8 and s["p99_s"] >= s["p50_s"] > 0


def test_request_larger_than_budget_is_refused():
    reqs = small_requests(1, seed=3)
    srv = serve.Server(hbm_budget=1024)
    with pytest.raises(ValueError):
        srv.run(reqs)
