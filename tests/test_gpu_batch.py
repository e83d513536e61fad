"""Batched multi-shot 3D RTM (erm_batch_*) on the GPU.

Each window of a batch is an erm_ctx, so the reference result for shot s is
the same window migrated alone through the one-shot path (erm_forward_real,
erm_forward_smoothed, erm_backward_smoothed).  The batched kernels run the
same per-cell arithmetic (the library is built with expression-level FP
contraction only, so every instantiation of a tile body rounds alike), so
records and images must be bit-identical; the stacked image must equal the
sum of the window images placed at their offsets; graph replay, plain
launches, persistent grids and every autotuner candidate give identical
images.
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

WINDOWS = [  # (sx, sy, sz, x0, y0, wx, wy): uneven apertures
    (20, 18, 3, 4, 2, 34, 30),
    (40, 30, 2, 22, 14, 40, 34),
    (12, 40, 4, 0, 26, 28, 26),
]


def _full(boundary, R=4, lt=70, cut=0):
    p = cases.params(nx=64, ny=52, nz=30, b=8 if R == 4 else 10, R=R, lt=lt, rec=(1, 1, 0, 0, 3, 3, 2))
    p.boundary = boundary
    p.cut_direct_wave = cut
    p.receiver_interval_x, p.receiver_interval_y = 3, 2
    p.receiver_start_z = 2
    vp, vs, rho = cases.layered(p, seed=11)
    sm = [cases.smooth(a, 4) for a in (vp, vs, rho)]
    return p, (vp, vs, rho), sm


def _window_params(p, w):
    import ctypes
    q = cases.Params()
    ctypes.memmove(ctypes.addressof(q), ctypes.addressof(p), ctypes.sizeof(q))
    sx, sy, sz, x0, y0, wx, wy = w
    q.nx, q.ny = wx, wy
    q.shot_start_x, q.shot_start_y, q.shot_start_z = sx - x0, sy - y0, sz
    q.receiver_start_x = q.receiver_start_y = 0
    q.receiver_num_x = (wx + p.receiver_interval_x - 1) // p.receiver_interval_x
    q.receiver_num_y = (wy + p.receiver_interval_y - 1) // p.receiver_interval_y
    return q


def _cut(a, w):
    sx, sy, sz, x0, y0, wx, wy = w
    return np.ascontiguousarray(a[x0:x0 + wx, y0:y0 + wy, :])


@pytest.mark.parametrize("boundary,R,cut", [(0, 4, 0), (1, 4, 3), (1, 6, 0)])
def test_batch_matches_single_shots(erm, boundary, R, cut):
    """cut > 0: the records are direct-wave muted, per window, in both paths
    (lt long enough for data to survive the 87-sample wavelet window)."""
    p, true, sm = _full(boundary, R, lt=170 if cut else 70, cut=cut)
    with erm.Batch3D(p) as b:
        b.load_model(0, *true)
        b.load_model(1, *sm)
        b.set_shots(WINDOWS)
        b.set_launch(8, 5, 0)
        stacked = b.run()
        per = []
        for s in range(len(WINDOWS)):
            v = b.shot(s)
            per.append((v.get_images(), v.get_record(), v.get_field("ex_time")))
        t = b.timing()
    assert t["steps_model"] == p.lt and t["cells_per_step"] > 0
    expect = [np.zeros((p.nx, p.ny, p.nz), np.float32) for _ in range(3)]
    for s, w in enumerate(WINDOWS):
        q = _window_params(p, w)
        with erm.RTM3D(q) as c:
            c.load_model(0, *[_cut(a, w) for a in true])
            c.load_model(1, *[_cut(a, w) for a in sm])
            c.set_shot(0, 0, 0)
            rec = c.forward_real()
            c.forward_smoothed()
            c.backward_smoothed()
            imgs = c.get_images()
            ext = c.get_field("ex_time")
        np.testing.assert_array_equal(per[s][1], rec)
        np.testing.assert_array_equal(per[s][2], ext)
        for q_ in range(3):
            np.testing.assert_array_equal(per[s][0][q_], imgs[q_])
            sx, sy, sz, x0, y0, wx, wy = w
            expect[q_][x0:x0 + wx, y0:y0 + wy, :] += per[s][0][q_]
        assert np.abs(imgs[1]).max() > 0
    for q_ in range(3):
        np.testing.assert_allclose(stacked[q_], expect[q_], rtol=1e-6, atol=1e-6 * np.abs(expect[q_]).max())


def test_graph_persistent_and_autotune_identical(erm):
    p, true, sm = _full(1, 4, lt=50)
    with erm.Batch3D(p) as b:
        b.load_model(0, *true)
        b.load_model(1, *sm)
        b.set_shots(WINDOWS)
        b.set_launch(4, 8, 0)
        b.use_graph(False)
        ref = b.run()
        b.use_graph(True)
        b.set_launch(8, 4, 64)  # persistent grid: 64 workgroups stride over the table
        g1 = b.run()
        b.set_launch(16, 6, 0, tile_x=32)  # 32-wide tiles
        g3 = b.run()
        best = b.autotune(steps=2)
        g2 = b.run()
    assert (best["tile_x"], best["tile_y"]) in ((64, 4), (64, 8), (32, 8), (32, 16))
    assert best["zchunk"] in (4, 6, 8, 16) and best["ms_per_step"] > 0
    for q in range(3):
        np.testing.assert_array_equal(g1[q], ref[q])
        np.testing.assert_array_equal(g2[q], ref[q])
        np.testing.assert_array_equal(g3[q], ref[q])


def test_batch_errors(erm):
    p, true, sm = _full(0, 4, lt=10)
    with erm.Batch3D(p) as b:
        with pytest.raises(erm.ErmError):
            b.set_shots([(5, 5, 2, 40, 0, 30, 20)])   # window beyond the model
        with pytest.raises(erm.ErmError):
            b.set_shots([(50, 5, 2, 0, 0, 30, 20)])   # source outside its window
        b.set_shots(WINDOWS[:1])
        with pytest.raises(erm.ErmError):
            b.run()                                    # no model
        with pytest.raises(erm.ErmError):
            b.set_launch(5, 4, 0)
