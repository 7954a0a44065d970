"""Self-convergence of psi4 over three resolutions on one GPU.

    python -m bh_ringdown_amd.convergence [--a 0.99] [--t-final 2000] [--base 16384x32]

Runs the same seeded plunge (mass ratio 1e-3, inclination 30 degrees, from
1.02 r_isco) on rho in [-600, 1600] at (N, M), (2N, 2M) and (4N, 4M) with
dt = 0.5 drho and classic RK4, extracts psi4 at the same rho (a column that
lies on all three grids) and the theta rows the three grids share, and
reports the convergence ratio |psi_1 - psi_2| / |psi_2 - psi_4| over time
windows (2^p for order p).  --no-source evolves Gaussian initial data
instead, which isolates the order of the evolution scheme from the source's
resolution-scaled delta widths and the 1.5 dt time offset of the reference's
two-rows-per-step trajectory reading (body.cu:112-129).
"""
import argparse
import json
import math

import numpy as np

from . import teuk, workloads


def run(N, M, a, t_final, j_extract, seed=0, xmin=-600.0, xmax=1600.0, source=True):
    p = teuk.default_params(N, M, xmin, xmax, a, 2, dt_over_dx=0.5, T=t_final)
    p.integrator = 1
    p.Fr = 1
    p.j_extract = j_extract
    nt = p.nsteps()
    with teuk.Solver(p) as s:
        if source:
            s.set_trajectory(workloads.plunge_rows(p, nt, eta=1e-3, iota_deg=30.0, seed=seed), nt)
        else:
            s.set_fields(workloads.gaussian_pulse(p, seed=seed, center=0.0, width=8.0))
        s.step(nt)
        s.sync()
        out = s.output()
        psi = s.psi()
    return p, out, psi


def study(a=0.99, t_final=2000.0, base="16384x32", rho=1000.0, windows=8, refine="both", source=True,
          log=print):
    """One three-resolution study; returns the report (per time window: rms
    psi, |psi_1 - psi_2|, |psi_2 - psi_4| and the order log2 of their ratio)."""
    N, M = map(int, base.split("x"))
    xmin, xmax = -600.0, 1600.0
    j0 = int(round((rho - xmin) / ((xmax - xmin) / N)))
    res = []
    for k in range(3):
        f = 2 ** k
        fr = f if refine in ("both", "rho") else 1
        ft = f if refine in ("both", "theta") else 1
        p, out, psi = run(N * fr, M * ft, a, t_final, j0 * fr, source=source)
        # Record k holds psi after step k, at time (k + 1) dt: the coarse
        # record k matches the refined record fr (k + 1) - 1.  Shared theta
        # rows: every ft-th row.
        psi_c = psi[fr - 1::fr, :, ::ft]
        res.append((p, out["t"][fr - 1::fr] + p.dt, psi_c[:, 0] + 1j * psi_c[:, 1]))
        log(json.dumps({"N": p.N, "M": p.M, "steps": p.nsteps(), "rho": xmin + j0 * fr * p.dx}))
    n = min(len(r[1]) for r in res)
    t = res[0][1][:n]
    d12 = np.abs(res[0][2][:n] - res[1][2][:n])
    d24 = np.abs(res[1][2][:n] - res[2][2][:n])
    amp = np.abs(res[2][2][:n])
    edges = np.linspace(0, n, windows + 1).astype(int)
    report = []
    for lo, hi in zip(edges[:-1], edges[1:]):
        e12 = float(np.sqrt(np.mean(d12[lo:hi] ** 2)))
        e24 = float(np.sqrt(np.mean(d24[lo:hi] ** 2)))
        report.append({"t": [float(t[lo]), float(t[hi - 1])], "rms_psi": float(np.sqrt(np.mean(amp[lo:hi] ** 2))),
                       "diff_12": e12, "diff_24": e24,
                       "order": (math.log2(e12 / e24) if e24 > 0 and e12 > 0 else None)})
    return {"a": a, "t_final": t_final, "base": base, "refine": refine, "source": source, "rho": rho,
            "grids": [[r[0].N, r[0].M, r[0].nsteps()] for r in res], "windows": report}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--a", type=float, default=0.99)
    ap.add_argument("--t-final", type=float, default=2000.0)
    ap.add_argument("--base", default="16384x32")
    ap.add_argument("--rho", type=float, default=1000.0)
    ap.add_argument("--windows", type=int, default=8)
    ap.add_argument("--refine", choices=["both", "rho", "theta"], default="both",
                    help="which directions are refined (dt follows drho)")
    ap.add_argument("--no-source", action="store_true",
                    help="Gaussian initial data (width 8 M at rho = 0) instead of the plunge source")
    args = ap.parse_args(argv)
    rep = study(args.a, args.t_final, args.base, args.rho, args.windows, args.refine, not args.no_source,
                log=lambda s: print(s, flush=True))
    print(json.dumps(rep), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
