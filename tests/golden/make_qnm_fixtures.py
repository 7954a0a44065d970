"""Writes tests/golden/qnm_a07.npz from the reference's QNM_Extraction
Examples (a = 0.7, equatorial plunge, m = 2, -2, 0): the waveform text rows the
fit uses, the trajectory rows near the light ring, the QNM-frequency and
overlap table rows that bracket a = 0.7, and the reference's own fit outputs
(every 20th row of spheroidal_coefficients_*.out, all of
excitation_coefficients_*.dat).  Data only; run in a tree that has the
reference at /root/reference:

    python tests/golden/make_qnm_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from bh_ringdown_amd import qnm  # noqa: E402  (table names only)

REF = "/root/reference/QNM_Extraction"
EX = os.path.join(REF, "Examples")
A, MMAX, MODES = 0.7, 4, (2, -2, 0)


def main():
    out = {}
    traj = np.loadtxt(os.path.join(EX, "a0.7_thi000_thf090.0_n.traj"))
    near = np.abs(traj[:, 1] - qnm.light_ring_radius(A)) < 0.2
    out["traj_t"], out["traj_r"] = traj[near, 0], traj[near, 1]
    t_lr = qnm.light_ring_time(traj[:, 0], traj[:, 1], A)
    out["t_lr"] = np.array(t_lr)
    names = set()
    for m in MODES:
        tag = f"hm{m}" if m >= 0 else f"hmm{-m}"
        # the waveform text rows from t_lr - 60 on, kept verbatim (compressed)
        with open(os.path.join(EX, f"{tag}_a0.7_thi000_thf090.0_n.dat")) as f:
            lines = [ln for ln in f if float(ln.split()[0]) >= t_lr - 60.0]
        out[f"wave_{m}"] = np.frombuffer("".join(lines).encode(), dtype=np.uint8)
        lab = f"m{m}" if m >= 0 else f"mm{-m}"
        s = np.loadtxt(os.path.join(EX, f"spheroidal_coefficients_{lab}_a70_thinc000_thf090.0_n.out"))
        out[f"spheroidal_{m}"] = s[::20]
        out[f"spheroidal_{m}_rows"] = np.array(s.shape[0])
        out[f"excitation_{m}"] = np.atleast_2d(np.loadtxt(os.path.join(EX, f"excitation_coefficients_a70_thinc000_h{lab}.dat")))
        for lv in qnm.mode_list(m, MMAX):
            for lp, mp in qnm.spheroidal_modes(m, MMAX):
                names.add(f"swsh/l{lv}m{mp}lp{abs(lp)}np0.dat" if mp >= 0 else f"swsh/l{lv}mm{-mp}lp{abs(lp)}np0.dat")
                mm_, aa = mp, A
                while mm_ < 0:
                    aa, mm_ = -aa, -mm_
                names.add(f"kerr/n1l{abs(lp)}m{mm_}.dat" if aa >= 0 else f"kerr/n1l{abs(lp)}mm{mm_}.dat")
    tables = {}
    for name in sorted(names):
        d = np.atleast_2d(np.loadtxt(os.path.join(REF, "Tables", name)))
        keep = np.abs(d[:, 0] - A) <= (2e-4 if name.startswith("kerr") else 0.06)
        tables[name] = d[keep][:, :3]
    out["table_names"] = np.array(sorted(tables))
    for i, name in enumerate(sorted(tables)):
        out[f"table_{i}"] = tables[name]
    np.savez_compressed(os.path.join(HERE, "qnm_a07.npz"), **out)
    print("wrote", os.path.join(HERE, "qnm_a07.npz"), len(tables), "tables")
    # table rows for fitting the a = 0.5, m = 2 plain-method waveform
    write_tables(0.5, [2], os.path.join(HERE, "qnm_tables_a05_m2.npz"))


def table_names(a, modes):
    names = set()
    for m in modes:
        for lv in qnm.mode_list(m, MMAX):
            for lp, mp in qnm.spheroidal_modes(m, MMAX):
                names.add(f"swsh/l{lv}m{mp}lp{abs(lp)}np0.dat" if mp >= 0 else f"swsh/l{lv}mm{-mp}lp{abs(lp)}np0.dat")
                mm_, aa = mp, a
                while mm_ < 0:
                    aa, mm_ = -aa, -mm_
                names.add(f"kerr/n1l{abs(lp)}m{mm_}.dat" if aa >= 0 else f"kerr/n1l{abs(lp)}mm{mm_}.dat")
    return sorted(names)


def write_tables(a, modes, path):
    out = {}
    names = table_names(a, modes)
    for i, name in enumerate(names):
        d = np.atleast_2d(np.loadtxt(os.path.join(REF, "Tables", name)))
        keep = np.abs(d[:, 0] - a) <= (2e-4 if name.startswith("kerr") else 0.06)
        out[f"table_{i}"] = d[keep][:, :3]
    out["table_names"] = np.array(names)
    np.savez_compressed(path, **out)
    print("wrote", path, len(names), "tables")


if __name__ == "__main__":
    main()
