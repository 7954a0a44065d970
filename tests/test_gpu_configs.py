"""The benchmark configurations at their own grid sizes, every solve checked
against the oracle (not against another run of the HIP path):

* config 2 -- N = 16384, M = 64, rho in [-400, 1200], a = 0.9, m = 2, zero
  fields and the seeded plunge (mass ratio 1e-3, inclination 30 degrees):
  the first 256 steps live against the oracle (RK4 and SSP-RK3), and the
  whole run (t = 1500, 30720 steps, RK4 and SSP-RK3) against the oracle's records
  (tests/golden/config2_{rk4,rk3}_records.npz, tests/golden/make_config2_golden.py);
* config 3 -- the 72 batched solves (m = -4..4 x 8 seeded spins, grids
  4096x32 / 8192x32 / 16384x64 by spin, one inclined plunge per spin),
  40 steps each in one batch, every member against the oracle;
* config 4 -- requests drawn from the serving distribution (N in 4096 /
  8192 / 16384, M in 32 / 64, random a, m, inclination, mass ratio), lengths
  cut to a few M, served by the continuous-batching server, every result
  against the oracle.

Equality is bit for bit (the HIP path and the oracle share the point
arithmetic; see test_gpu_parity.py), which implies the stated tolerances:
fields 1e-11, psi4 and h_lm 1e-8, QNM amplitudes 1e-8 relative.
"""
import hashlib
import math
import os

import numpy as np
import pytest

from bh_ringdown_amd import qnm, serve, teuk, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
QNM_TOL = 1e-8


def oracle_solve(p, rows, nsteps, u0=None):
    _, r_c, th = teuk.grid(p)
    ost = None
    if rows is not None:
        sched = teuk.trajectory_schedule(rows, nsteps, p.dt)
        ost = [O.SrcStep(*[getattr(x, f) for f, _ in teuk.SourceStep._fields_]) for x in sched]
    op = O.make_params(p.N, p.M, p.xmin, p.xmax, p.a, p.m, dt=p.dt, integrator=p.integrator,
                       j_extract=p.j_extract, ptsx=p.ptsx, ptsy=p.ptsy)
    u = np.zeros((4, p.M, p.N)) if u0 is None else u0.copy()
    rc, psi, h, dedt = O.run(op, r_c, th, u, nsteps, src_steps=ost)
    assert rc == 0
    return u, psi, h, dedt


def gpu_solve(p, rows, nsteps):
    with teuk.Solver(p) as s:
        s.set_trajectory(rows, nsteps)
        s.step(nsteps)
        s.sync()
        return s.get_fields(), s.output(), s.psi()


def same_as_oracle(ug, out, psig, uo, psio, ho, dedto):
    ks = out["k"]
    return (np.array_equal(ug, uo) and np.array_equal(psig, psio[ks]) and np.array_equal(out["h"], ho[ks])
            and np.array_equal(out["dedt"], dedto[ks]))


@pytest.mark.parametrize("integrator", [1, 0])
def test_config2_first_256_steps(integrator):
    p = workloads.single_16384x64(integrator=integrator)
    n = 256
    rows = workloads.plunge_rows(p, p.nsteps())
    print(f"config 2, integrator {integrator}: {n} steps", flush=True)
    ug, out, psig = gpu_solve(p, rows, n)
    uo, psio, ho, dedto = oracle_solve(p, rows, n)
    assert np.abs(uo).max() > 0
    assert same_as_oracle(ug, out, psig, uo, psio, ho, dedto)


@pytest.mark.parametrize("name,integrator", [("rk4", 1), ("rk3", 0)])
def test_config2_full_length_against_oracle_records(name, integrator):
    path = os.path.join(HERE, "golden", f"config2_{name}_records.npz")
    if not os.path.exists(path):
        pytest.skip(f"oracle records not generated (tests/golden/make_config2_golden.py --integrator {name})")
    gold = np.load(path)
    p = workloads.single_16384x64(integrator=integrator)
    n = p.nsteps()
    assert n == int(gold["nsteps"]) == 30720
    rows = workloads.plunge_rows(p, n)
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_config2_golden as G
    if "traj_sha256" not in gold.files or not np.array_equal(G.traj_digest(rows), gold["traj_sha256"]):
        pytest.skip("the records were made from another trajectory (the generator changed since); "
                    "regenerate them with tests/golden/make_config2_golden.py")
    print(f"config 2 full length: {n} {name} steps", flush=True)
    ug, out, psig = gpu_solve(p, rows, n)
    ks = out["k"]
    assert np.array_equal(ks, np.arange(0, n, p.Fr))
    pos_h = np.searchsorted(ks, gold["h_k"])
    assert np.array_equal(out["h"][pos_h], gold["h"])
    assert np.array_equal(out["dedt"][pos_h], gold["dedt"])
    assert np.array_equal(psig[np.searchsorted(ks, gold["psi_k"])], gold["psi"])
    assert hashlib.sha256(np.ascontiguousarray(ug).tobytes()).digest() == gold["fields_sha256"].tobytes()
    # QNM amplitudes of the whole h record, fitted as the oracle's were
    amp, phase = G.fit_qnm(ks * p.dt, out["h"], p.a, int(p.m), G.tables_a09())
    assert np.all(np.abs(amp - gold["qnm_amp"]) <= QNM_TOL * np.abs(gold["qnm_amp"]))
    dph = np.angle(np.exp(1j * (phase - gold["qnm_phase"])))
    assert np.all(np.abs(dph) <= QNM_TOL * np.maximum(1.0, np.abs(gold["qnm_phase"])))


def test_config3_batch_every_member_against_oracle():
    steps = 40
    items = workloads.batch_72(seed=0)
    for p, _ in items:
        p.T = steps * p.dt * (1 + 1e-12)
    grids = sorted({(p.N, p.M) for p, _ in items})
    assert grids == [(4096, 32), (8192, 32), (16384, 64)]
    print(f"config 3: 72 members, {steps} steps each, grids {grids}", flush=True)
    solvers = []
    with teuk.Batch(integrator=1) as b:
        for p, rows in items:
            s = teuk.Solver(p)
            s.set_trajectory(rows, steps)
            assert b.add(s)
            solvers.append(s)
        b.step(steps)
        b.sync()
        for i, ((p, rows), s) in enumerate(zip(items, solvers)):
            uo, psio, ho, dedto = oracle_solve(p, rows, steps)
            assert same_as_oracle(s.get_fields(), s.output(), s.psi(), uo, psio, ho, dedto), (i, p.N, p.M, p.a, p.m)
    for s in solvers:
        s.close()


def test_config4_served_requests_against_oracle():
    reqs = serve.poisson_requests(20.0, 10.0, seed=4, t_range=(3.0, 6.0), max_requests=6)
    assert len({(r.N, r.M) for r in reqs}) >= 3
    print("config 4: " + ", ".join(f"{r.N}x{r.M} m={r.m} a={r.a:.2f}" for r in reqs), flush=True)
    srv = serve.Server(hbm_budget=64 << 30, chunk=16, target_points=1 << 23)
    served = srv.run(reqs)
    for r in reqs:
        res = served[r.id]
        assert not res.failed and not math.isnan(res.done)
        p = serve.request_params(r)
        nt = p.nsteps()
        rows = teuk.plunge_trajectory(r.a, r.eta, r.iota, p.dt / 2, (2 * nt + 6) * p.dt / 2, seed=r.seed)
        uo, psio, ho, dedto = oracle_solve(p, rows, nt)
        ks = res.records["k"]
        assert np.array_equal(ks, np.arange(0, nt, p.Fr))
        assert np.array_equal(res.records["h"], ho[ks]), r.id
        assert np.array_equal(res.records["dedt"], dedto[ks]), r.id
        assert np.array_equal(res.psi, psio[ks]), r.id
