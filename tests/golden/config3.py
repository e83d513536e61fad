"""FP64 oracle fixture of one window of the batched 3D workload, and the
script that writes it (python tests/golden/config3.py; a few minutes on 8
cores).

The batched workload (bench.py --workload batch: 24 shots on a random
surface layout over a 600x600x250 model, seed 3, windows of 192..320 cells
laterally and full depth, 20-cell C-PML, radius 4, dt 0.8 ms, 15 Hz,
receivers every 2 cells over each window) with the time axis cut to NT = 400
steps per pass.  The fixture is the smallest window (shot 7: 207 x 193 cells
at (393, 367), source at (566, 463)) migrated by the FP64 oracle on its own
window grid: model pass in the true model, source pass with excitation,
receiver pass with imaging.  Kept, as for the single-shot fixture
(tests/golden/config2.py): the inline slice y = wy/2 and a lattice every
8 / 8 / 4 cells of ex_time, ex_amp and the three images, the images' masked
norms and the receiver-pass vz slice y = nny/2.  The GPU test
(tests/test_gpu_config3.py) migrates all 24 windows in one batch and checks
this window against these arrays.
"""
import ctypes
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

import config2  # noqa: E402

FIXTURE = os.path.join(HERE, "config3_window7_nt400.npz")
NT = 400
WINDOW = 7


def batch_inputs(erm, nt=NT):
    """(full params, true model, migration model, the 24 windows) of the batched workload."""
    import bench
    return bench.batch_setup(erm, nt, erm.BOUNDARY_CPML)


def window_params(erm, p, w):
    """The parameters the batch gives window w = (sx, sy, sz, x0, y0, wx, wy):
    its own lateral grid, the source in window coordinates, receivers every
    receiver_interval cells over the whole window."""
    q = erm.default_params()
    ctypes.memmove(ctypes.addressof(q), ctypes.addressof(p), ctypes.sizeof(q))
    sx, sy, sz, x0, y0, wx, wy = w
    q.nx, q.ny = wx, wy
    q.shot_start_x, q.shot_start_y, q.shot_start_z = sx - x0, sy - y0, sz
    q.receiver_start_x = q.receiver_start_y = 0
    q.receiver_num_x = (wx + p.receiver_interval_x - 1) // p.receiver_interval_x
    q.receiver_num_y = (wy + p.receiver_interval_y - 1) // p.receiver_interval_y
    return q


def cut(a, w):
    sx, sy, sz, x0, y0, wx, wy = w
    return np.ascontiguousarray(a[x0:x0 + wx, y0:y0 + wy, :])


def main():
    import importlib
    erm = importlib.import_module("2d_and_3d_elastic_reverse_time_migration_amd")
    import oracle3d
    oracle3d.set_threads(os.cpu_count() or 1)
    p, true, mig, wins = batch_inputs(erm)
    w = wins[WINDOW]
    q = window_params(erm, p, w)
    t0 = time.time()
    o = oracle3d.Oracle3D(q, 64)
    o.load_model(0, *[cut(a, w) for a in true])
    o.load_model(1, *[cut(a, w) for a in mig])
    del true, mig
    o.rtm_shot()
    print(f"oracle window {WINDOW} {w}: {time.time() - t0:.0f} s", flush=True)
    out = {}
    et = config2.model_region(o.get_field("ex_time"), q)
    ea = config2.model_region(o.get_field("ex_amp"), q)
    imgs = o.images()
    for name, a in (("ex_time", et), ("ex_amp", ea), ("img_tp", imgs[0]), ("img_pp", imgs[1]), ("img_ps", imgs[2])):
        s, g = config2.sample(a, q)
        dt = np.int16 if name == "ex_time" else np.float32
        out["slice_" + name] = s.astype(dt)
        out["grid_" + name] = g.astype(dt)
    mask = ea > 1e-3 * ea.max()
    for name, a in zip(("tp", "pp", "ps"), imgs):
        out["norm_" + name] = np.float64(np.linalg.norm(a[mask]))
    vz = o.get_field("vz")
    out["vz_slice"] = vz[:, vz.shape[1] // 2, :].astype(np.float32)
    out["nt"] = np.int32(NT)
    out["window"] = np.asarray(w, np.int32)
    np.savez_compressed(FIXTURE, **out)
    print(f"wrote {FIXTURE} ({os.path.getsize(FIXTURE) / 1e6:.2f} MB)", flush=True)


if __name__ == "__main__":
    main()
