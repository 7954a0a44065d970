"""Write tests/golden/source_t4.json: the canonical coefficients of the
point-particle T_4 at random particle states, from the symbolic derivation
in source_derivation.py, evaluated at 30 significant digits.

    python tests/golden/make_source_golden.py      (about 5 minutes)

The C++ jet evaluator (teuk_t4.h, used by libteuk and by the oracle) is a
separate implementation of the same equations; tests/test_source_math.py
checks it against these vectors.  The vectors carry the standard
normalisation (Teukolsky's T_4 with delta(phi) -> e^{-i m phi_p}/pi); the
evaluator returns the reference's, half of it.
"""
import json
import os
import random
import sys

import mpmath
import sympy as sp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import source_derivation as S  # noqa: E402

ORDER = [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2)]   # teuk_t4.h output order


def main():
    mpmath.mp.dps = 30
    c = S.derive()
    jets = S.field_jets(S.used_jets(c))
    U = S.u_series()
    rng = random.Random(20261202)
    cases = []
    for n in range(10):
        a = rng.choice([1e-8, 0.3, 0.7, 0.9, 0.99]) if n else 0.5
        M = 1.0
        rh = 1 + (1 - a * a) ** 0.5
        st = {
            "mass": M, "a": a, "E": rng.uniform(0.8, 1.0), "lz": rng.uniform(-3.5, 3.5), "nmu": 1.0,
            "mm": float(rng.choice([-4, -2, -1, 0, 1, 2, 3, 4])),
            "rp": rng.uniform(rh + 0.3, 12.0), "tp": rng.uniform(0.3, 2.8), "phip": rng.uniform(0, 6.2),
            "drdt": rng.uniform(-0.6, 0.1), "d2rdt2": rng.uniform(-0.05, 0.01),
            "dthdt": rng.uniform(-0.05, 0.05), "d2thdt2": rng.uniform(-0.01, 0.01),
            "dphidt": rng.uniform(0.0, 0.3), "d2phidt2": rng.uniform(-0.01, 0.01),
        }
        base = {getattr(S, k) if k not in ("mass", "a", "E", "lz", "nmu", "mm") else
                {"mass": S.M, "a": S.a, "E": S.E, "lz": S.Lz, "nmu": S.mu, "mm": S.m}[k]: sp.Float(v, 30)
                for k, v in st.items()}
        vals = dict(base)
        for sym, e in jets.items():
            vals[sym] = sp.N(e.xreplace(base), 30)
        for sym, e in zip((S.U0, S.U1, S.U2), U):
            vals[sym] = sp.N(e.xreplace(base), 30)
        mode = sp.N((sp.cos(S.m * S.phip) - sp.I * sp.sin(S.m * S.phip)).xreplace(base) / sp.pi, 30)
        out = []
        for key in ORDER:
            z = complex(sp.N(c[key].xreplace(vals) * mode, 30))
            out += [z.real, z.imag]
        cases.append({"state": st, "coeffs": out})
        print(n, out[:2], flush=True)
    with open(os.path.join(HERE, "source_t4.json"), "w") as f:
        json.dump({"order": ORDER, "cases": cases}, f, indent=1)


if __name__ == "__main__":
    main()
