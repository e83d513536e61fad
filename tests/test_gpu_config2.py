"""The benchmark's 3D shot (256x256x128 model + 20-cell C-PML, 296x296x168
padded, radius 4, 128x128 receivers, the seeded dipping-layer model) with the
time axis cut to 500 steps per pass, migrated through liberm.so with the
benchmark's launch shape (32x8 tiles, z sweep 6, XCD y-band tile order,
region-split C-PML tile lists) and compared with the FP64 oracle's fixture
(tests/golden/config2.py).

Tolerances: images rel L2 <= 1e-4 over excited cells (ex_amp > 1e-5 max) whose
excitation step agrees, at most 2% of excited cells on another step (an argmax
of FP32 samples can pick a neighbouring step where two are within rounding);
the receiver-pass vz slice <= 1e-5.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import config2  # noqa: E402

pytestmark = pytest.mark.gpu


def test_gpu_config2_shot_matches_fp64_fixture(erm):
    fx = np.load(config2.FIXTURE)
    print(f"config-2 shot at nt={int(fx['nt'])} vs the FP64 fixture", flush=True)
    p, et, ea, imgs, vz = config2.gpu_shot(erm, int(fx["nt"]), tile_x=32, tile_y=8, zchunk=6, xcd_map=2)
    err = config2.compare(fx, p, et, ea, imgs, vz)
    print(err, flush=True)
    assert err["ex_time_mismatch_frac"] < 0.02, err
    assert err["cells_compared"] > 5000, err
    for k in ("img_tp_rel_l2", "img_pp_rel_l2", "img_ps_rel_l2"):
        assert err[k] < 1e-4, (k, err)
    assert err["vz_rel_l2"] < 1e-5, err
