"""Writes tests/golden/ot_x_vs_l.npz: every 100th row of the reference's
GenOT/OT_X_vs_L.dat (L, X(L): the universal Ori-Thorne transition function,
Apte & Hughes 2019, Fig. 2).  Data only; run in a tree that has the
reference at /root/reference:

    python tests/golden/make_ot_fixture.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    d = np.loadtxt("/root/reference/GenOT/OT_X_vs_L.dat")
    np.savez_compressed(os.path.join(HERE, "ot_x_vs_l.npz"), L=d[::100, 0], X=d[::100, 1])
    print("wrote ot_x_vs_l.npz", d[::100].shape)


if __name__ == "__main__":
    main()
