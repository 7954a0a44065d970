"""Self-convergence of psi4 at a = 0.99 (the high-resolution configuration),
from the committed record tests/golden/convergence_a099.json
(tests/golden/make_convergence_record.py: bh_ringdown_amd.convergence on the
GPU, grids doubling, psi at rho = 1000 compared at shared times and theta
rows).  Orders are log2(|psi_1 - psi_2| / |psi_2 - psi_4|) over 250 M
windows; only windows where the wave is present (rms psi above 1e-2, or
1e-3 for the plunge) are assessed -- before the pulse arrives the differences
are zero and after it has passed they are a 1e-6 late-time residue.

* rho and dt refined, theta fixed, Gaussian data: 4th order (measured 4.3
  and 4.9) -- RK4's time error, with dt proportional to drho, dominates the
  6th-order rho stencils;
* theta refined alone: about 1.7 -- the reference's axis treatment (the first
  and last three theta rows extrapolated, kern.cu:320-409) limits the theta
  direction to about second order;
* the plunge from zero fields at 16384x32, 32768x64, 65536x128 to t = 2000:
  converging (every difference shrinks), at 0.4-1.5 -- the source's
  Gaussian delta widths scale with the grid (kern.cu:1191-1196) and the
  trajectory is read 1.5 dt ahead (body.cu:112-129), so 6th-order
  self-convergence does not hold for the sourced problem."""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
REC = json.load(open(os.path.join(HERE, "golden", "convergence_a099.json")))


def signal_windows(study, floor):
    return [w for w in REC[study]["windows"] if w["rms_psi"] > floor]


def test_record_grids():
    assert REC["rho"]["grids"] == [[8192, 32, 14894], [16384, 32, 29789], [32768, 32, 59578]]
    assert REC["theta"]["grids"] == [[16384, 32, 29789], [16384, 64, 29789], [16384, 128, 29789]]
    assert REC["plunge"]["grids"] == [[16384, 32, 29789], [32768, 64, 59578], [65536, 128, 119156]]
    assert all(REC[k]["a"] == 0.99 and REC[k]["t_final"] == 2000.0 for k in REC)


def test_rho_refinement_is_fourth_order():
    ws = signal_windows("rho", 1e-2)
    assert len(ws) >= 2
    assert all(w["order"] > 3.8 for w in ws), [w["order"] for w in ws]


def test_theta_refinement_is_about_second_order():
    ws = signal_windows("theta", 1e-2)
    assert len(ws) >= 2
    assert all(1.4 < w["order"] < 2.3 for w in ws), [w["order"] for w in ws]


def test_plunge_converges():
    ws = signal_windows("plunge", 1e-3)
    assert len(ws) >= 2
    assert all(w["diff_24"] < w["diff_12"] and w["order"] > 0.3 for w in ws), [w["order"] for w in ws]
    # where the signal is largest the finest pair still differs by 81 % of
    # it (the coarser pair by 108 %): the sourced problem converges slowly
    top = max(ws, key=lambda w: w["rms_psi"])
    assert top["diff_24"] < 0.85 * top["rms_psi"] < top["diff_12"]
