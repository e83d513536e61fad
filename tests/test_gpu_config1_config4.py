"""The two 2D configurations at their full grids on the GPU, against the
FP64 oracle (the same helpers as bench.py's rtm2d and visco workloads).

* Config 1: 2D elastic RTM of one shot, 400x200 model + 40-cell C-PML
  (480x280), d = 10 m, nt = 2000, three flat layers + a Gaussian anomaly,
  20 Hz Ricker, 400 surface receivers: the GPU's tp / PP / PS excitation
  images against the FP64 oracle's, rel L2 <= 1e-4 over excited cells with
  the same excitation step (at most 2% of them on another step), and the
  receiver record of the model pass <= 1e-5.
* Config 4: 2D viscoelastic LSRTM, 600x300 model (d = 8 m, L = 3 SLS, smooth
  random Q, 40-cell damping): two CG iterations of one shot with nt cut to
  500 on the full grid, GPU CG (batched operators) against CGLS on the FP64
  oracle's Born / adjoint operators: misfits and model <= 1e-4.
"""
import importlib
import os
import sys
import time

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(400)
def test_gpu_config1_rtm_shot_matches_fp64(erm, oracle2d_mod):
    import bench
    p, true, mig = bench.rtm2d_case(erm)
    t0 = time.time()
    with erm.Elastic2D(p) as g:
        g.load_model(0, *true)
        g.load_model(1, *mig)
        rec = g.forward()
        imgs = g.rtm(rec)
        et = g.get_field("ex_time")
    print(f"config 1: GPU shot in {time.time() - t0:.1f} s", flush=True)
    o = oracle2d_mod.Oracle2D(p, 64)
    o.load_model(0, *true)
    o.load_model(1, *mig)
    o.step(0, 0, p.lt)
    ref_rec = o.record()
    o.zero(7)
    o.step(1, 0, p.lt)
    o.zero(1)
    o.step(2, p.lt - 1, p.lt)
    print(f"config 1: FP64 oracle shot done at {time.time() - t0:.1f} s", flush=True)
    err = bench.rtm2d_errors(p, o.images(), o.get_field("ex_time"), o.get_field("ex_amp"), imgs, et)
    e_rec = float(np.linalg.norm(rec - ref_rec) / np.linalg.norm(ref_rec))
    print(f"config 1 vs FP64: {err}, record {e_rec:.2e}", flush=True)
    assert e_rec < 1e-5
    assert err["ex_time_mismatch_frac"] < 0.02
    for k in ("img_tp_rel_l2", "img_pp_rel_l2", "img_ps_rel_l2"):
        assert err[k] < 1e-4, (k, err)


@pytest.mark.timeout(400)
def test_gpu_config4_cg_matches_fp64_cgls(erm):
    import bench
    learned = importlib.import_module(bench.PKG + ".learned")
    nx, nz = 600, 300
    p = learned.survey_params(nx, nz, 3000, freq=20.0, d=8.0, dt=0.8, b=40)
    p.nsls = 3
    p.nrec, p.rec_x0 = 600, 0
    bvp, bvs, brho, qp, qs, dm = bench.visco_model(nx, nz, 4)
    chk = bench.visco_oracle_check(erm, learned, p, (bvp, bvs, brho, qp, qs), dm)
    print(f"config 4: {chk}", flush=True)
    assert chk["model_rel_diff_gpu_vs_fp64"] < 1e-4, chk
    for a, b in zip(chk["misfit_gpu"], chk["misfit_fp64"]):
        assert abs(a - b) <= 1e-4 * max(abs(b), 1e-30), chk
