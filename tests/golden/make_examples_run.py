"""Runs the reference's plunge example through libteuk on the GPU and writes
tests/golden/examples_a07_run.npz: the h_lm records (l = 2..6 of m = 2, the
modes the QNM fit uses) from 60 M before the light-ring time to the end of the
reference's waveform, in the reference's time frame (body.cu:207), for

  * "ref":  the grid of the shipped parm.h:1-8 and main.c:46 (N = 16000,
            M = 256, rho in [-50, 50], dt = 0.002, Fr = 50), SSP-RK3;
  * "half": N = 8000, M = 128, dt = 0.004, Fr = 25;
  * other grids named N:M:dt:Fr; the committed file also holds
    16000:512:0.002:50 (the reference's production theta resolution, which
    reproduces its waveform), 16000:128:0.002:50 and 8000:256:0.004:25;

all on the trajectory examples_pin.solver_rows builds (source switched on at
t = 2140).  GPU time: ~85 s for "ref", ~170 s for 16000:512, ~10 s for
"half":

    python tests/golden/make_examples_run.py ref half 16000:512:0.002:50 ...

tests/test_examples_pin.py checks the records against the reference's own
waveform and QNM amplitudes; tests/test_gpu_examples.py reruns "half".
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

GRIDS = {"ref": (16000, 256, 0.002, 50), "half": (8000, 128, 0.004, 25)}
T0, T_END, XMAX = 2140.0, 2800.0, 50.0
T_KEEP = 2429.3        # light-ring time 2489.3 less 60 M (the span qnm_a07.npz keeps)


def run(N, M, dt, Fr, t0=T0, t_end=T_END, approach="ot", log=print):
    from bh_ringdown_amd import teuk
    import examples_pin as E
    p = teuk.default_params(N, M, -XMAX, XMAX, E.A, 2.0, dt_over_dx=0.32)
    p.dt, p.Fr, p.integrator, p.j_extract = dt, Fr, 0, -1
    nsteps = int(round((t_end - t0) / dt))
    p.T = nsteps * dt * (1 + 1e-12)
    rows = E.solver_rows(t0, t_end + 4 * dt, dt / 2, approach=approach)
    t1 = time.time()
    with teuk.Solver(p) as s:
        s.set_trajectory(rows, nsteps)
        done = 0
        while done < nsteps:
            n = min(40000, nsteps - done)
            s.step(n)
            s.sync()
            done += n
            log(f"  N={N} M={M}: {done}/{nsteps} steps, {time.time() - t1:.1f} s")
        out = s.output()
    t = t0 + out["t"] - XMAX            # the reference's waveform time, timer - Xmax
    keep = t >= T_KEEP - 1e-9
    return t[keep], out["h"][keep, :10]


def main(argv):
    names = argv or list(GRIDS)
    path = os.path.join(HERE, "examples_a07_run.npz")
    data = dict(np.load(path)) if os.path.exists(path) else {}
    for name in names:
        N, M, dt, Fr = GRIDS[name] if name in GRIDS else (int(name.split(":")[0]), int(name.split(":")[1]),
                                                           float(name.split(":")[2]), int(name.split(":")[3]))
        print(f"examples run {name}: N={N} M={M} dt={dt} Fr={Fr}", flush=True)
        t, h = run(N, M, dt, Fr, log=lambda s: print(s, flush=True))
        tag = name.replace(":", "_").replace(".", "p")
        data[f"{tag}_t"] = t
        data[f"{tag}_h"] = h.astype(np.float32)
        data[f"{tag}_grid"] = np.array([N, M, dt, Fr])
        out = os.environ.get("EXAMPLES_RUN_OUT", path)
        np.savez_compressed(out, **data)
        print("wrote", out, flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
