"""Boundary-strip saving and backward source reconstruction (erm_forward_smoothed_strips,
erm_reconstruct, erm_backward_smoothed_xcorr) on the GPU.

* Reconstruction: the source pass saved with strips and stepped back to the
  start of step k must reproduce, in the model region, the wavefield of a
  plain source pass stopped after k steps (FP32 forward-then-reverse
  rounding: rel L2 <= 1e-4 for v, sigma, tp, vP).
* Cross-correlation images from the reconstructed source wavefield against
  the FP64 oracle, which keeps the whole source history in memory:
  rel L2 <= 1e-4 (PP, PS, tp).
* The excitation images of the xcorr receiver pass equal erm_backward_smoothed's.
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu
PASS_SOURCE = 1
FIELDS = ["vx", "vy", "vz", "txx", "tyy", "tzz", "txy", "txz", "tyz", "tp", "vxp", "vyp", "vzp"]


def _case(boundary, R=4, lt=140):
    b = 10 if R == 6 else 8
    p = cases.params(nx=36, ny=30, nz=28, b=b, R=R, lt=lt, dt=1.0, freq=25.0, src=(18, 15, 3),
                     rec=(12, 10, 1, 1, 3, 3, 2))
    p.lt_forward = p.lt
    p.boundary = boundary
    vp, vs, rho = cases.layered(p, seed=21)
    sm = [cases.smooth(a, 3) for a in (vp, vs, rho)]
    return p, (vp, vs, rho), sm


def _region(p, a):
    return a[p.bu:p.bu + p.nz, p.bf:p.bf + p.ny, p.bl:p.bl + p.nx]


@pytest.mark.parametrize("boundary,R", [(0, 4), (1, 4), (1, 6)])
def test_reconstruction_matches_forward(erm, boundary, R):
    p, true, sm = _case(boundary, R)
    with erm.RTM3D(p) as g, erm.RTM3D(p) as h:
        g.load_model(1, *sm)
        h.load_model(1, *sm)
        assert g.strip_bytes() > 0
        g.forward_smoothed_strips()
        k = 40
        g.reconstruct(k)
        h.zero_wavefield()
        h.zero_excitation()
        h.step(PASS_SOURCE, 0, k)
        for f in FIELDS:
            a, b = _region(p, g.get_reconstructed_field(f)), _region(p, h.get_field(f))
            e = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            assert e < 1e-4, (f, e)


@pytest.mark.parametrize("boundary", [0, 1])
def test_xcorr_images_match_oracle(erm, oracle, boundary):
    p, true, sm = _case(boundary)
    with erm.RTM3D(p) as g:
        g.load_model(0, *true)
        g.load_model(1, *sm)
        rec = g.forward_real()
        g.forward_smoothed()
        g.backward_smoothed()
        ex_ref = g.get_images()
        g.forward_smoothed_strips()
        g.backward_smoothed_xcorr()
        ex = g.get_images()
        xc = g.get_xcorr_images()
    for a, b in zip(ex, ex_ref):
        np.testing.assert_array_equal(a, b)
    o = oracle.Oracle3D(p, 64)
    o.load_model(0, *true)
    o.load_model(1, *sm)
    o.set_record(rec)
    xo = o.xcorr_shot(use_record=True)
    for name, a, b in zip(("tp", "pp", "ps"), xc, xo):
        e = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"xcorr {name}: rel L2 vs FP64 oracle {e:.2e}")
        assert e < 1e-4, (name, e)


@pytest.mark.parametrize("boundary,decim", [(0, 1), (1, 1), (1, 3)])
def test_xcorr_history_matches_strips_and_oracle(erm, oracle, boundary, decim):
    """The decimated-history strategy: with decim = 1 its images equal the
    boundary-strip reconstruction's (<= 1e-5) and, for any decim, the FP64
    oracle's decimated cross-correlation (<= 1e-4); the excitation images of
    its receiver pass are those of erm_backward_smoothed (bit-identical)."""
    p, true, sm = _case(boundary)
    with erm.RTM3D(p) as g:
        g.load_model(0, *true)
        g.load_model(1, *sm)
        rec = g.forward_real()
        g.forward_smoothed()
        g.backward_smoothed()
        ex_ref = g.get_images()
        if decim == 1:
            g.forward_smoothed_strips()
            g.backward_smoothed_xcorr()
            xs = g.get_xcorr_images()
        assert g.history_bytes(decim) == 16.0 * p.nx * p.ny * p.nz * (p.lt // decim)
        g.forward_smoothed_history(decim)
        g.backward_smoothed_xcorr_history()
        ex = g.get_images()
        xh = g.get_xcorr_images()
    for a, b in zip(ex, ex_ref):
        np.testing.assert_array_equal(a, b)
    if decim == 1:
        for a, b in zip(xh, xs):
            assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b)
    o = oracle.Oracle3D(p, 64)
    o.load_model(0, *true)
    o.load_model(1, *sm)
    o.set_record(rec)
    xo = o.xcorr_shot(use_record=True, decim=decim)
    for name, a, b in zip(("tp", "pp", "ps"), xh, xo):
        e = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"history decim {decim} {name}: rel L2 vs FP64 oracle {e:.2e}")
        assert e < 1e-4, (name, e)


def test_strip_errors(erm):
    p, true, sm = _case(0)
    with erm.RTM3D(p) as g:
        g.load_model(1, *sm)
        with pytest.raises(erm.ErmError):
            g.backward_smoothed_xcorr()          # no strips yet
        with pytest.raises(erm.ErmError):
            g.backward_smoothed_xcorr_history()  # no history yet
        with pytest.raises(erm.ErmError):
            g.forward_smoothed_history(0)
    p.lt_forward = p.lt - 10
    with erm.RTM3D(p) as g:
        g.load_model(1, *sm)
        with pytest.raises(erm.ErmError):
            g.forward_smoothed_strips()          # needs lt_forward == lt
    p, true, sm = _case(0)
    p.bu = 2
    with erm.RTM3D(p) as g:
        g.load_model(1, *[a for a in sm])
        with pytest.raises(erm.ErmError):
            g.forward_smoothed_strips()          # boundary narrower than the stencil
