"""erm2d_angle_gathers_noalloc (the allocation-free variant of
erm2d_angle_gathers, staging in a fixed per-GPU device array) against
erm2d_angle_gathers, through the C ABI: two calls in a row, the second on a
smaller grid with fewer bins, must give bit-identical outputs (same kernel,
same inputs; the second call must not see the first call's staging).  Sizes
above the array's capacity are rejected with ERM_E_SIZE before any device
work."""
import ctypes

import numpy as np
import pytest

import cases2d

pytestmark = pytest.mark.gpu


def _shot(erm, nx, nz, zi, src):
    p = cases2d.params(nx=nx, nz=nz, b=12, R=4, L=0, lt=500, freq=25.0, coe=250.0, src=src, rec_dx=1, rec_z=2)
    z = np.arange(nz)[None, :]
    vp = (np.where(z < zi, 2000.0, 3000.0) * np.ones((nx, 1))).astype(np.float32)
    rho = (np.where(z < zi, 2000.0, 2400.0) * np.ones((nx, 1))).astype(np.float32)
    s = np.full_like(vp, 2000.0)
    g = erm.Elastic2D(p)
    g.load_model(0, vp, (vp / 1.732).astype(np.float32), rho)
    g.load_model(1, s, (s / 1.732).astype(np.float32), np.full_like(vp, 2000.0))
    g.forward()
    g.rtm()
    return p, g


def _gathers(erm, g, p, nang, dang, fn):
    fp = ctypes.POINTER(ctypes.c_float)
    out = [np.full((nang, p.nx, p.nz), np.nan, np.float32), np.full((nang, p.nx, p.nz), np.nan, np.float32)]
    out += [np.full((p.nx, p.nz), np.nan, np.float32) for _ in range(3)]
    rc = fn(g._h, nang, ctypes.c_float(dang), *[o.ctypes.data_as(fp) for o in out])
    return rc, out


def test_angle_gathers_noalloc_matches_allocating_entry_point(erm):
    lib = erm.lib
    fn_new = lib.erm2d_angle_gathers_noalloc
    fn_new.restype = ctypes.c_int
    fn_new.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_float] + [ctypes.POINTER(ctypes.c_float)] * 5
    fn_ref = lib.erm2d_angle_gathers
    for nx, nz, zi, src, nang, dang in ((90, 60, 36, (45, 3), 30, 3.0), (50, 40, 24, (25, 3), 12, 5.0)):
        print(f"no-alloc angle gathers {nx}x{nz}, {nang} bins", flush=True)
        p, g = _shot(erm, nx, nz, zi, src)
        try:
            rc_ref, ref = _gathers(erm, g, p, nang, dang, fn_ref)
            rc_new, new = _gathers(erm, g, p, nang, dang, fn_new)
            assert rc_ref == 0 and rc_new == 0, (rc_ref, rc_new, erm.lib.erm2d_last_error(g._h))
            for a, b in zip(new, ref):
                np.testing.assert_array_equal(a, b)
            assert np.count_nonzero(new[0]) > 0
        finally:
            g.close()


def test_angle_gathers_noalloc_rejects_oversize(erm):
    lib = erm.lib
    fn_new = lib.erm2d_angle_gathers_noalloc
    fn_new.restype = ctypes.c_int
    fn_new.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_float] + [ctypes.POINTER(ctypes.c_float)] * 5
    p = cases2d.params(nx=700, nz=300, b=10, R=4, L=0, lt=10, freq=25.0, src=(350, 3))
    g = erm.Elastic2D(p)
    try:
        # 2 * 20 * 210000 + 3 * 210000 floats > 2^23: rejected before any device work
        nul = ctypes.POINTER(ctypes.c_float)()
        rc = fn_new(g._h, 20, ctypes.c_float(3.0), nul, nul, nul, nul, nul)
        assert rc == -5, rc  # ERM_E_SIZE
    finally:
        g.close()
