"""GPU parity of the 3D Poynting vectors (erm_params.poynting = 1; the
cuda_cal_source_poyn_3D / cuda_cal_receiver_poyn_3D calls that the reference
comments out at 3D_forward_or_back_together.cu:283,432) against the FP64
oracle on a full RTM shot: the receiver P and S vectors at each cell's
imaging step and the source P vector at its excitation step, compared on the
cells whose excitation step agrees with the oracle's (as the images), each
component <= 1e-4 rel L2.  The same shot with poynting = 0 gives the same
images, and the receiver fields are refused there.
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

COMPS = ("poyn_sx", "poyn_sy", "poyn_sz", "poyn_rpx", "poyn_rpy", "poyn_rpz", "poyn_rsx", "poyn_rsy", "poyn_rsz")


def test_gpu_poynting_vectors_match_fp64_oracle(erm, oracle):
    import test_gpu3d as T
    p, model, smodel, zi = T._reflector_case()
    p.poynting = 1
    print(f"Poynting parity: padded {tuple(p.padded_shape)}, lt={p.lt}", flush=True)
    with erm.RTM3D(p) as g:
        g.load_model(0, *model)
        g.load_model(1, *smodel)
        g.forward_real()
        g.forward_smoothed()
        gt = g.get_field("ex_time")
        g.backward_smoothed()
        gimg = g.get_field("img_pp")
        gv = {k: g.get_field(k) for k in COMPS}
    o = oracle.Oracle3D(p, 64)
    o.load_model(0, *model)
    o.load_model(1, *smodel)
    o.rtm_shot()
    ot, oa = o.get_field("ex_time"), o.get_field("ex_amp")
    mask = (oa > 1e-3 * oa.max()) & (gt == ot)
    assert mask.sum() > 1000
    for k in COMPS:
        ov = o.get_field(k)
        e = cases.rel_l2(gv[k][mask], ov[mask])
        print(f"  {k}: rel L2 {e:.2e}", flush=True)
        assert e < 1e-4, (k, e)
    # the receiver S vector is not a copy of the P one, nor zero
    assert cases.rel_l2(gv["poyn_rsz"][mask], gv["poyn_rpz"][mask]) > 0.1
    o.close()
    p.poynting = 0
    with erm.RTM3D(p) as g:
        g.load_model(0, *model)
        g.load_model(1, *smodel)
        g.forward_real()
        g.forward_smoothed()
        g.backward_smoothed()
        np.testing.assert_array_equal(g.get_field("img_pp"), gimg)
        with pytest.raises(erm.ErmError):
            g.get_field("poyn_rpx")
