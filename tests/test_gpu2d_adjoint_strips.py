"""The exact-transpose LSRTM adjoint with the background from boundary strips
(erm2d_params.adjoint_source = ERM2D_SRC_STRIPS; no reference counterpart):
the background pass keeps a ring of width R + 1 around the model region per
half step instead of the E/F history, and the background is stepped back
before each transposed step, the reversed kernels recomputing E and F.

Against the FP64 oracle's history adjoint (which the strips oracle equals to
rounding, tests/test_oracle2d.py): <= 1e-4 rel L2 per parameter, as the
history adjoint on the GPU; the dot-product test against Born <= 1e-4
(FP32); the batched operator (5 shots in groups of 3, a ragged last group)
equals the one-shot strips operator summed over shots; CG runs on it.
"""
import numpy as np
import pytest

import cases
import cases2d

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("R,para", [(4, 2), (6, 2), (4, 3)])
def test_gpu2d_strips_adjoint_matches_oracle(erm, oracle2d_mod, R, para):
    import test_oracle2d as T
    p, bg = T._lsrtm_case(0, R=R, lt=200)
    p.inversion_para = para
    p.adjoint_source = erm.SRC_STRIPS
    rng = np.random.default_rng(21)
    dm = (rng.standard_normal((3, p.nx, p.nz)) * cases2d.taper_mask(p)).astype(np.float32)
    r = rng.standard_normal((2, p.lt, p.nrec)).astype(np.float32)
    print(f"2D strips adjoint parity R={R} inversion_para={para}", flush=True)
    with erm.Elastic2D(p) as g:
        g.load_model(1, *bg)
        d = g.born(dm)
        gr = g.adjoint(r)
    o = oracle2d_mod.Oracle2D(p)
    o.load_model(1, *bg)
    og = o.adjoint(r)
    for c in range(3):
        e = cases.rel_l2(gr[c], og[c])
        assert e < 1e-4, (c, e)
    lhs = float(np.sum(d.astype(np.float64) * r))
    rhs = float(np.sum(dm.astype(np.float64) * gr))
    assert abs(lhs - rhs) / abs(lhs) < 1e-4, (lhs, rhs)


def test_gpu2d_strips_adjoint_batched_and_cg(erm):
    import test_oracle2d as T
    p, bg = T._lsrtm_case(0, lt=160)
    p.adjoint_source = erm.SRC_STRIPS
    rng = np.random.default_rng(22)
    shots = [(6, 2), (14, 3), (20, 2), (27, 4), (33, 2)]
    r = rng.standard_normal((len(shots), 2, p.lt, p.nrec)).astype(np.float32)
    print("2D strips adjoint: batched vs one-shot, CG", flush=True)
    with erm.Elastic2D(p) as g:
        g.load_model(1, *bg)
        one = g.adjoint_shots(shots, r)
        g.set_batch(3)
        bat = g.adjoint_shots(shots, r)
        for c in range(3):
            e = cases.rel_l2(bat[c], one[c])
            assert e < 1e-6, (c, e)
        dm = (rng.standard_normal((3, p.nx, p.nz)) * cases2d.taper_mask(p)).astype(np.float32)
        d = g.born_shots(shots, dm)
        m, misfit = g.lsrtm_cg(shots, d, 6)[:2]
    assert np.all(np.diff(misfit) <= 1e-6) and misfit[-1] < 0.9 * misfit[0], misfit
    q = cases2d.params(nx=40, nz=30, b=10, R=4, L=0, lt=160, freq=25.0, rec_dx=2)
    q.adjoint_source = erm.SRC_HISTORY
    with erm.Elastic2D(q) as h:
        h.load_model(1, *bg)
        ref = h.adjoint_shots(shots, r)
    for c in range(3):
        e = cases.rel_l2(one[c], ref[c])
        assert e < 1e-4, ("strips vs history", c, e)


def test_gpu2d_strips_adjoint_rejects_bad_setups(erm):
    import test_oracle2d as T
    p, bg = T._lsrtm_case(1, lt=20)
    p.adjoint_source = erm.SRC_STRIPS
    with pytest.raises(erm.ErmError):
        erm.Elastic2D(p)  # viscoelastic
    p, bg = T._lsrtm_case(0, lt=20)
    p.adjoint_source = erm.SRC_STRIPS
    p.adjoint_mode = 1
    with pytest.raises(erm.ErmError):
        erm.Elastic2D(p)  # reference-mode adjoint
    p.adjoint_mode = 0
    p.adjoint_source = 1
    with pytest.raises(erm.ErmError):
        erm.Elastic2D(p)  # reversed propagation alone is not offered for the adjoint
    p.adjoint_source = erm.SRC_STRIPS
    with erm.Elastic2D(p) as g:
        g.load_model(1, *bg)
        with pytest.raises(erm.ErmError):
            g.adjoint_ref(np.zeros((2, p.lt, p.nrec), np.float32))
