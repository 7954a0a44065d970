"""The committed launch-shape table (bh_ringdown_amd/tuning_defaults.json)
holds only shapes the library accepts: wavefronts per block 1, 2 or 4,
kernels built for 2-4 waves per SIMD, 2, 4 or 8 theta rows per wavefront,
batch tile order 0 or 1, graph chunks of 1-256 steps and cache groups off or
a positive byte budget; and its keys name grids the solver accepts (N a
multiple of 64, M a multiple of 8 from 16).  A bad entry would reach the GPU
only as a refused set_tiling call, so it is caught here."""
import json
import re

from bh_ringdown_amd import tune


def grids(key):
    return [tuple(map(int, re.match(r"(\d+)x(\d+)", part).groups())) for part in key.split("/")[1].split(",")]


def test_committed_launch_shapes_are_accepted_values():
    table = json.load(open(tune.DEFAULTS))
    assert table
    for key, e in table.items():
        assert key.startswith(("solver/", "batch/")) and key.endswith(("/i0", "/i1")), key
        for N, M in grids(key):
            assert N % 64 == 0 and N >= 128 and M % 8 == 0 and M >= 16, key
        assert e["waves_per_block"] in (1, 2, 4), key
        assert 2 <= e["waves_per_eu"] <= 4, key
        assert e.get("theta_rows", 8) in (2, 4, 8), key
        if key.startswith("batch/"):
            assert e["order"] in (0, 1) and 1 <= e["chunk"] <= 256 and e["group_bytes"] >= 0, key
        else:
            assert e["us_per_launch"] > 0, key
