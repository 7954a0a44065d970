"""Writes tests/golden/convergence_a099.json: three three-resolution
self-convergence studies of psi4 (bh_ringdown_amd.convergence.study), a = 0.99,
rho in [-600, 1600], dt = 0.5 drho, classic RK4, psi at rho = 1000:

  * "plunge":   the high-resolution configuration -- the seeded plunge (mass
                ratio 1e-3, inclination 30 degrees) from zero fields to
                t = 2000 at 16384x32, 32768x64, 65536x128 (1.1e12
                point-updates, about two minutes of GPU time);
  * "rho":      Gaussian initial data, rho (and dt) refined, theta fixed:
                8192x32, 16384x32, 32768x32 to t = 2000;
  * "theta":    the same data, theta refined alone: 16384x32, 16384x64,
                16384x128 (dt / dtheta stays below the scheme's stability
                limit, about 3.6, only with the finer rho spacing).

    python tests/golden/make_convergence_record.py [--out path]

tests/test_convergence_record.py checks the orders the record shows.
"""
import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from bh_ringdown_amd import convergence  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "convergence_a099.json"))
    args = ap.parse_args()
    log = lambda s: print(s, flush=True)
    rec = {}
    for name, base, refine, source in (("rho", "8192x32", "rho", False), ("theta", "16384x32", "theta", False),
                                       ("plunge", "16384x32", "both", True)):
        rec[name] = convergence.study(0.99, 2000.0, base, 1000.0, 8, refine, source, log=log)
        with open(args.out, "w") as f:
            json.dump(rec, f, indent=1)
        print("wrote", name, "to", args.out, flush=True)


if __name__ == "__main__":
    main()
