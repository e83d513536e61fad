"""Direct-wave mute of the modelled record on the GPU (erm_forward_real with
cut_direct_wave > 0): bit-identical to the oracle's definition
(oracle3d.mute_direct) applied to the unmuted GPU record of the same run."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("boundary", [0, 1])
def test_gpu_mute_matches_oracle_definition(erm, oracle, boundary):
    p = cases.params(nx=30, ny=26, nz=20, b=8, R=4, lt=150, dt=1.0, freq=25.0, src=(13, 11, 3),
                     rec=(7, 6, 1, 1, 4, 4, 2))
    p.boundary = boundary
    vp, vs, rho = cases.layered(p, seed=8)
    print(f"mute boundary={boundary}", flush=True)
    with erm.RTM3D(p) as g:
        g.load_model(0, vp, vs, rho)
        g.set_shot(0, 0, 0)
        ru = g.forward_real()
    p.cut_direct_wave = 4
    with erm.RTM3D(p) as g:
        g.load_model(0, vp, vs, rho)
        g.set_shot(0, 0, 0)
        rm = g.forward_real()
    ref = oracle.mute_direct(p, ru, vp)
    np.testing.assert_array_equal(rm, ref.astype(np.float32))
    assert np.count_nonzero(rm) < np.count_nonzero(ru)
    assert np.abs(rm).max() > 0
