"""Batched stepping (teuk_batch_*): every member of a batch of uneven solves
must end bit for bit where stepping it alone ends, through the C ABI."""
import math

import numpy as np
import pytest

from bh_ringdown_amd import teuk, workloads

pytestmark = pytest.mark.gpu


def member(N, M, xmin, xmax, a, m, steps, integrator, seed, source):
    p = teuk.default_params(N, M, xmin, xmax, a, m, dt_over_dx=0.5)
    p.T = steps * p.dt * (1 + 1e-12)
    p.integrator = integrator
    p.Fr = 3
    p.j_extract = N - 40
    u0 = workloads.gaussian_pulse(p, seed=seed, center=0.5 * (xmin + xmax), width=4.0)
    rows = None
    if source:
        rows = teuk.plunge_trajectory(a, 1e-2, 0.4, p.dt / 2, (2 * steps + 6) * p.dt / 2, seed=seed, r0_over_isco=1.3)
    return p, u0, rows


def run_alone(p, u0, rows, steps):
    _, r_c, th = teuk.grid(p)
    with teuk.Solver(p, r_c, th) as s:
        s.set_fields(u0)
        if rows is not None:
            s.set_trajectory(rows, steps)
        s.step(steps)
        s.sync()
        return s.get_fields(), s.output(), s.psi()


CASES = [  # N, M, xmin, xmax, a, m, steps, seed, source
    (256, 32, -40.0, 60.0, 0.5, 2, 70, 1, True),
    (512, 16, -60.0, 90.0, 0.9, -3, 45, 2, False),
    (128, 64, -30.0, 50.0, 0.0, 1, 101, 3, True),
    (384, 32, -50.0, 70.0, 0.7, 0, 64, 4, False),
    (256, 24, -40.0, 60.0, 0.6, 2, 50, 5, False),   # 3 and 5 theta chunks: partial blocks
    (128, 40, -30.0, 50.0, 0.8, -2, 38, 6, True),
]


@pytest.mark.parametrize("integrator", [1, 0])
@pytest.mark.parametrize("wpb,wpe,order,chunk,groups,rows", [(2, 2, 0, 32, 0, 2), (1, 4, 1, 7, 0, 2),
                                                             (4, 3, 0, 16, 0, 4), (4, 2, 0, 32, 2 << 20, 2),
                                                             (2, 2, 1, 16, 1, 8)])
def test_batch_members_equal_alone(integrator, wpb, wpe, order, chunk, groups, rows):
    """groups: cache-group budget in bytes (0 = one group; 1 = a group per
    member; 2 MiB = two or three members per group here); rows: theta rows
    per wavefront (members run alone with the default 2; with 2-row
    wavefronts the tile table puts the edge chunks first, within each XCD's
    run for order 0 and over the whole list for order 1)."""
    specs = [member(N, M, x0, x1, a, m, st, integrator, seed, src) for N, M, x0, x1, a, m, st, seed, src in CASES]
    steps = [c[6] for c in CASES]
    print(f"batch integ={integrator} wpb={wpb} wpe={wpe} order={order} chunk={chunk} groups={groups} "
          f"rows={rows}", flush=True)
    alone = [run_alone(p, u0, rows, st) for (p, u0, rows), st in zip(specs, steps)]
    solvers = []
    with teuk.Batch(integrator=integrator) as b:
        b.set_tiling(wpb, wpe, order, chunk, rows)
        b.set_cache_groups(groups)
        for (p, u0, rows), st in zip(specs, steps):
            _, r_c, th = teuk.grid(p)
            s = teuk.Solver(p, r_c, th)
            s.set_fields(u0)
            if rows is not None:
                s.set_trajectory(rows, st)
            assert b.add(s)
            solvers.append(s)
        # uneven calls: the longest member needs max(steps); others stop at their own end
        b.step(13)
        b.step(max(steps))
        b.sync()
        assert b.info()["active"] == 0
        for s, (u, out, psi), st in zip(solvers, alone, steps):
            assert s.remaining() == 0
            assert s.progress()[0] == st
            assert np.array_equal(s.get_fields(), u)
            o = s.output()
            assert np.array_equal(o["h"], out["h"]) and np.array_equal(o["t"], out["t"])
            assert np.array_equal(s.psi(), psi)
    for s in solvers:
        s.close()


def test_batch_admission_by_footprint():
    p, _, _ = member(256, 32, -40.0, 60.0, 0.5, 2, 20, 1, 1, False)
    fp = teuk.footprint(p)
    _, r_c, th = teuk.grid(p)
    with teuk.Batch(integrator=1, hbm_budget=int(2.5 * fp)) as b:
        ss = [teuk.Solver(p, r_c, th) for _ in range(3)]
        assert b.add(ss[0]) and b.add(ss[1])
        assert not b.add(ss[2])            # over budget: refused, nothing changed
        assert b.info()["members"] == 2
        b.remove(ss[0])
        assert b.add(ss[2])
        b.step(20)
        b.sync()
    for s in ss:
        s.close()


def test_solver_in_batch_refuses_own_stepping():
    p, _, _ = member(256, 32, -40.0, 60.0, 0.5, 2, 20, 1, 1, False)
    with teuk.Batch(integrator=1) as b, teuk.Solver(p) as s:
        assert b.add(s)
        with pytest.raises(teuk.TeukError):
            s.step(1)
        b.remove(s)
        s.step(1)
        s.sync()


@pytest.mark.parametrize("M", [32, 40])
def test_single_solver_tiling_is_bitwise_neutral(M):
    """Every launch shape gives the same bits, also where the waves per block
    do not divide the theta chunks (M = 40: 5 chunks of 8 rows, 9 of 4 rows,
    17 of 2 rows)."""
    p, u0, rows = member(512, M, -40.0, 60.0, 0.6, 2, 40, 1, 5, True)
    ref = run_alone(p, u0, rows, 40)[0]
    _, r_c, th = teuk.grid(p)
    for wpb, wpe, rows_ in [(1, 2, 8), (2, 4, 2), (4, 4, 4), (4, 2, 8), (2, 2, 4), (1, 3, 2), (2, 3, 4)]:
        with teuk.Solver(p, r_c, th) as s:
            s.set_tiling(wpb, wpe, rows_)
            s.set_fields(u0)
            s.set_trajectory(rows, 40)
            s.step(40)
            s.sync()
            assert np.array_equal(s.get_fields(), ref), (wpb, wpe, rows_)
