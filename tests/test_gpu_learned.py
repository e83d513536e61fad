"""What the trained U-Net preconditioner gains over plain CG-LSRTM and over
the illumination diagonal alone, on the GPU.

A reduced version of the learned workflow (learned.py): 48 random 64x48
training models (4 shots each) run 30 CG iterations preconditioned by the
illumination diagonal w; pairs (w * gradient_k, m_30 - m_k) at 9 iterations k,
each target normalised by its own RMS, train the U-Net on the GPU.  On a
held-out 128x48 model with 8 shots, flexible CG alternating the net
(z = net(w g)) and the diagonal (z = w g) must reach the metric's data-misfit
target 1e-2 in at most half the iterations and time of plain CG, and in at most 0.8x the
iterations and 0.85x the time of CG with the diagonal alone, so the net's
own contribution is tested, not only the diagonal's.  All runs use the same
operators and the same CG code.
"""
import importlib

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(600)
def test_learned_preconditioner_beats_plain_cg_and_the_diagonal(erm):
    L = importlib.import_module("2d_and_3d_elastic_reverse_time_migration_amd.learned")
    print("learned preconditioner: training pairs from 48 models", flush=True)
    pairs = L.generate(train_models=48, train_shape=(64, 48), train_shots=4, train_lt=700, seed0=1000,
                       illum_eps=1e-3, log=lambda s: None)
    net, st = L.train(pairs, epochs=30, batch=16, lr=1e-3, seed=1, log=print)
    assert st["loss_first_last"][1] < 0.7 * st["loss_first_last"][0]
    res = L.evaluate(net, test_shape=(128, 48), test_shots=8, test_lt=700, test_iters=500, tol=1e-2,
                     test_seed=5000, log=print, every=50, warm=0, illum_eps=1e-3, schedules=("alt2",))
    net.close()
    p, d, w = res["plain"], res["illumination"], res["schedule_alt2"]
    print(f"plain: {p['iters_to_tol']} it / {p['s_to_tol']} s; illumination diagonal: {d['iters_to_tol']} it /"
          f" {d['s_to_tol']} s; learned (net alternating with the diagonal): {w['iters_to_tol']} it /"
          f" {w['s_to_tol']} s; model error vs true {res['model_err_vs_true']}", flush=True)
    assert p["iters_to_tol"] is not None and d["iters_to_tol"] is not None and w["iters_to_tol"] is not None
    assert w["iters_to_tol"] <= 0.5 * p["iters_to_tol"] and w["s_to_tol"] <= 0.5 * p["s_to_tol"]
    assert w["iters_to_tol"] <= 0.8 * d["iters_to_tol"] and w["s_to_tol"] <= 0.85 * d["s_to_tol"]
