"""The reference-side binding: TD_GPU_code/HighOrder/main.c and proj.c,
compiled with g++ where they lie (build.sh:7), linked with body() from
bh_ringdown_amd/driver/body_teuk.cpp and libteuk in place of body.cu and
kern.cu (oracle/_ref/body_main, built by `make -C oracle body` from the
reference tree with oracle/body_parm/parm.h: 1024 x 32, T = 260).  Its
stdout and every waveform file hm0<m>_/hmm<|m|>_<name>.dat of main.c's m-loop
(-5..5, main.c:50-55) must equal those of bin/teuk_run run with the same
parameters, byte for byte: the binding adds nothing and loses nothing.
teuk_run's rows are checked against the oracle in test_gpu_driver.py."""
import os
import subprocess

import numpy as np
import pytest

from bh_ringdown_amd import teuk

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BODY_MAIN = os.path.join(ROOT, "oracle", "_ref", "body_main")
DRIVER = os.path.join(ROOT, "bh_ringdown_amd", "bin", "teuk_run")
NAME = "a0.0_thi000_thf090.0_n"  # the trajectory name body.cu:68,77 hard-codes

# oracle/body_parm/parm.h and main.c:40-46
PARAMS = {"N": 1024, "M": 32, "Xmax": 50.0, "Xmin": -50.0, "Fr": 50, "T": 260.0, "PTSX": 128, "PTSY": 128,
          "mass": 1.0, "aa": 1e-8, "dt": 0.002, "mmin": -5, "mmax": 5, "integrator": 0, "j_extract": -1}


def wave_files(m):
    return ("hmm%d" if m < 0 else "hm0%d") % abs(m) + "_" + NAME + ".dat"


@pytest.mark.timeout(600)
@pytest.mark.skipif(not os.path.exists(BODY_MAIN),
                    reason="oracle/_ref/body_main is built only where the reference tree is present")
def test_reference_main_with_body_binding_matches_teuk_run(tmp_path):
    assert os.path.exists(DRIVER), "driver not built (bh_ringdown_amd.build)"
    nt = int(PARAMS["T"] / PARAMS["dt"])
    rows = teuk.plunge_trajectory(PARAMS["aa"], 1e-2, 0.3, PARAMS["dt"] / 2, (2 * nt + 20) * PARAMS["dt"] / 2,
                                  seed=3, r0_over_isco=1.3)
    (tmp_path / "Trajectories").mkdir()
    teuk.write_traj(tmp_path / "Trajectories" / (NAME + ".traj"), rows)
    ref_dir = tmp_path / "main_c"
    ref_dir.mkdir()
    run_dir = tmp_path / "teuk_run"
    run_dir.mkdir()
    params = tmp_path / "params.txt"
    params.write_text("".join(f"{k} {v!r}\n" for k, v in PARAMS.items()))

    print("binding: reference main.c + body_teuk.cpp, 1024x32, m = -5..5, 130000 steps each", flush=True)
    ref = subprocess.run([BODY_MAIN], cwd=ref_dir, capture_output=True, text=True, timeout=300)
    assert ref.returncode == 0, ref.stderr
    print("binding: teuk_run with the same parameters", flush=True)
    run = subprocess.run([DRIVER, "-p", str(params), NAME], cwd=run_dir, capture_output=True, text=True,
                         timeout=300)
    assert run.returncode == 0, run.stderr

    assert ref.stdout.startswith("\nWELCOME TO TEUKOLSKY CODE!\n\n")
    assert ref.stdout.endswith("All done.\nBye bye!\n")
    assert ref.stdout.count(" procs 1 my_rank 0 \n") == 11
    assert ref.stdout == run.stdout
    for m in range(-5, 6):
        got = (ref_dir / wave_files(m)).read_bytes()
        want = (run_dir / wave_files(m)).read_bytes()
        assert got == want, f"m = {m}"
        table = np.loadtxt(ref_dir / wave_files(m))
        # rows every Fr steps from k dt >= Xmax + 200 (body.cu:205-208)
        assert table.shape[1] == 15 and 95 <= table.shape[0] <= 100, (m, table.shape)
        assert table[0, 0] == pytest.approx(200.0) and np.all(np.diff(table[:, 0]) > 0)
        assert np.all(np.isfinite(table)) and np.any(table[:, 1:] != 0.0), f"m = {m}"
