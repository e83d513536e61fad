"""The batched 3D workload at its real layout: 24 shots on a random surface
layout over the 600x600x250 model, windows of 192..320 cells laterally and
full depth, 20-cell C-PML, radius 4 (bench.py --workload batch), with the time
axis cut to 200 steps per pass.

* the records, excitation times and images of six windows of the batched
  run (the smallest, the largest and four others) are bit-identical to the
  same window migrated alone through the one-shot path (each window is its
  own context with the same per-cell arithmetic; tests/test_gpu_batch.py
  shows it for every window and launch shape on smaller layouts);
* the stacked images equal the sum of all 24 window images at their offsets;
* the smallest window (shot 7) matches the FP64 oracle's fixture
  (tests/golden/config3.py) as the single-shot benchmark does
  (tests/test_gpu_config2.py): images rel L2 <= 1e-4 over excited cells
  with the same excitation step, at most 2% of them on another step, the
  receiver-pass vz slice <= 1e-5.
"""
import os
import sys
import time

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import config2  # noqa: E402
import config3  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(600)
def test_gpu_config3_batch_layout(erm):
    t0 = time.time()
    p, true, mig, wins = config3.batch_inputs(erm)
    print(f"config-3 batch: {len(wins)} windows over {p.nx}x{p.ny}x{p.nz}, nt={p.lt}", flush=True)
    expect = [np.zeros((p.nx, p.ny, p.nz), np.float32) for _ in range(3)]
    fx = np.load(config3.FIXTURE)
    with erm.Batch3D(p) as b:
        b.load_model(0, *true)
        b.load_model(1, *mig)
        b.set_shots(wins)
        b.set_launch(8, 6, 0, tile_x=32)
        print(f"  [{time.time() - t0:.0f} s] models loaded, shots set", flush=True)
        stacked = b.run()
        print(f"  [{time.time() - t0:.0f} s] batched run done", flush=True)
        largest = max(range(len(wins)), key=lambda s: wins[s][5] * wins[s][6])
        alone = {config3.WINDOW, largest, 0, 8, 15, 23}
        for s, w in enumerate(wins):
            v = b.shot(s)
            imgs_b = v.get_images()
            sx, sy, sz, x0, y0, wx, wy = w
            for k in range(3):
                expect[k][x0:x0 + wx, y0:y0 + wy, :] += imgs_b[k]
            if s not in alone:
                continue
            rec_b, et_b = v.get_record(), v.get_field("ex_time")
            q = config3.window_params(erm, p, w)
            if s == config3.WINDOW:
                err = config2.compare(fx, q, config2.model_region(et_b, q),
                                      config2.model_region(v.get_field("ex_amp"), q), imgs_b,
                                      v.get_field("vz")[:, et_b.shape[1] // 2, :])
                print(f"  [{time.time() - t0:.0f} s] window {s} vs FP64: {err}", flush=True)
                assert err["ex_time_mismatch_frac"] < 0.02, err
                assert err["cells_compared"] > 2000, err
                for k in ("img_tp_rel_l2", "img_pp_rel_l2", "img_ps_rel_l2"):
                    assert err[k] < 1e-4, (k, err)
                assert err["vz_rel_l2"] < 1e-5, err
            with erm.RTM3D(q) as c:
                c.load_model(0, *[config3.cut(a, w) for a in true])
                c.load_model(1, *[config3.cut(a, w) for a in mig])
                c.set_shot(0, 0, 0)
                rec = c.forward_real()
                c.forward_smoothed()
                c.backward_smoothed()
                imgs = c.get_images()
                et = c.get_field("ex_time")
            np.testing.assert_array_equal(rec_b, rec)
            np.testing.assert_array_equal(et_b, et)
            for k in range(3):
                np.testing.assert_array_equal(imgs_b[k], imgs[k])
            print(f"  [{time.time() - t0:.0f} s] window {s} {w}: bit-identical to its one-shot migration", flush=True)
    for k in range(3):
        np.testing.assert_allclose(stacked[k], expect[k], rtol=1e-6, atol=1e-6 * np.abs(expect[k]).max())
