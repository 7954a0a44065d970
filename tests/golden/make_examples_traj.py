"""Writes tests/golden/examples_a07_traj.npz: the trajectory of the
reference's QNM_Extraction Examples (a0.7_thi000_thf090.0_n.traj, an a = 0.7
prograde equatorial plunge sampled every 0.02 M from 2440.06 to 2690.04,
10 columns t r theta dr/dt dtheta/dt dphi/dt phi E Lz Q), the columns that
vary, and the constant ones checked constant.  The reference solver's own
waveform for this plunge is Examples/hm2_a0.7_thi000_thf090.0_n.dat (kept in
qnm_a07.npz).  Data only; run in a tree that has the reference at
/root/reference:

    python tests/golden/make_examples_traj.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EX = "/root/reference/QNM_Extraction/Examples"


def main():
    d = np.loadtxt(os.path.join(EX, "a0.7_thi000_thf090.0_n.traj"))
    assert d.shape[1] == 10
    # equatorial: theta, dtheta/dt and Q are constant over the file
    for col in (2, 4, 9):
        assert np.all(d[:, col] == d[0, col]), col
    out = {
        "t": d[:, 0], "r": d[:, 1], "rdot": d[:, 3], "phidot": d[:, 5], "phi": d[:, 6],
        "E": d[:, 7], "Lz": d[:, 8],
        "theta": np.array(d[0, 2]), "thetadot": np.array(d[0, 4]), "Q": np.array(d[0, 9]),
    }
    path = os.path.join(HERE, "examples_a07_traj.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, d.shape[0], "rows")


if __name__ == "__main__":
    main()
