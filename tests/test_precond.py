"""Host-side preconditioner adaptors of the 2D LSRTM (no GPU needed): the
diagonal preconditioner erm2d_precond_diag through the C ABI."""
import ctypes

import numpy as np


def test_precond_diag_scales_every_plane(erm):
    rng = np.random.default_rng(0)
    nx, nz = 13, 7
    w = rng.random((nx, nz)).astype(np.float32)
    g = rng.standard_normal((3, nx, nz)).astype(np.float32)
    z = np.zeros_like(g)
    d = erm.DiagPrecond(w)
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    rc = erm.lib.erm2d_precond_diag(fp(g), fp(z), ctypes.addressof(d._s))
    assert rc == 0
    np.testing.assert_array_equal(z, (w[None] * g).astype(np.float32))


def test_precond_diag_rejects_missing_weights(erm):
    g = np.zeros((3, 4, 4), np.float32)
    z = np.zeros_like(g)
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    assert erm.lib.erm2d_precond_diag(fp(g), fp(z), None) != 0
