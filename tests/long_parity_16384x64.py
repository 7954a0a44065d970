"""Long full-size parity run (not collected by pytest: it takes minutes).

The single-solve workload -- 16384 x 64, rho in [-400, 1200], a = 0.9,
m = 2, classic RK4, the seeded plunge source from 1.02 r_isco -- stepped
`--steps` steps from zero fields on the GPU through the C ABI and by the
oracle on the host cores, then compared bit for bit: the fields, psi at the
extraction column and every h_lm record.

    python tests/long_parity_16384x64.py --steps 6000
"""
import argparse
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from bh_ringdown_amd import teuk, workloads  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--threads", type=int, default=O.DEFAULT_THREADS)
    args = ap.parse_args()
    steps = args.steps
    p = workloads.single_16384x64(integrator=1)
    p.T = steps * p.dt * (1.0 + 1e-12)
    rows = workloads.plunge_rows(p, steps)
    _, r_c, th = teuk.grid(p)
    print(f"16384x64 RK4 plunge, {steps} steps", flush=True)
    t0 = time.perf_counter()
    with teuk.Solver(p, r_c, th) as s:
        s.set_trajectory(rows, steps)
        s.step(steps)
        s.sync()
        ug, out, psig = s.get_fields(), s.output(), s.psi()
    print(f"device run: {time.perf_counter() - t0:.1f} s", flush=True)
    sched = teuk.trajectory_schedule(rows, steps, p.dt)
    op = O.make_params(p.N, p.M, p.xmin, p.xmax, p.a, p.m, dt=p.dt, integrator=p.integrator,
                       j_extract=p.j_extract, nthreads=args.threads)
    ost = [O.SrcStep(*[getattr(x, f) for f, _ in teuk.SourceStep._fields_]) for x in sched]
    uo = np.zeros((4, p.M, p.N))
    t0 = time.perf_counter()
    rc, psio, ho, dedto = O.run(op, r_c, th, uo, steps, src_steps=ost)
    print(f"oracle run ({args.threads} threads): {time.perf_counter() - t0:.1f} s, rc {rc}", flush=True)
    ks = out["k"]
    same_u = np.array_equal(ug, uo)
    same_psi = np.array_equal(psig, psio[ks])
    same_h = np.array_equal(out["h"], ho[ks])
    same_e = np.array_equal(out["dedt"], dedto[ks])
    nz = float(np.abs(uo).max())
    print(f"fields bitwise equal: {same_u} (max |u| {nz:.3e}, {ug.size} values)", flush=True)
    print(f"psi bitwise equal: {same_psi}; h_lm records ({len(ks)}) bitwise equal: {same_h}; dE/dt: {same_e}",
          flush=True)
    if not (rc == 0 and same_u and same_psi and same_h and same_e and nz > 0):
        rel = np.linalg.norm(ug - uo) / max(np.linalg.norm(uo), 1e-300)
        print(f"MISMATCH: relative L2 field difference {rel:.3e}", flush=True)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
