"""Batched 2D LSRTM operators (erm2d_set_batch, erm2d_born_shots,
erm2d_adjoint_shots; no reference counterpart): several shots per launch,
along the grid's z dimension, must give the one-shot operators' results.

Born records are bit-identical shot by shot (the same arithmetic per cell);
the summed adjoint gradient agrees to 1e-6 rel L2 (summation order over
shots differs); CG with batched operators follows the one-shot CG iterates
(misfits 1e-5).  Cases: elastic and 3-mechanism viscoelastic, radius 4 and
6, every inversion_para, 5 shots in groups of 3 (a ragged last group).
"""
import numpy as np
import pytest

import cases
import cases2d

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L,R,para", [(0, 4, 2), (3, 4, 1), (0, 6, 3)])
def test_gpu2d_batched_born_adjoint_equal_one_shot(erm, L, R, para):
    import test_oracle2d as T
    p, bg = T._lsrtm_case(L, R=R, lt=160)
    p.inversion_para = para
    rng = np.random.default_rng(7)
    dm = (rng.standard_normal((3, p.nx, p.nz)) * cases2d.taper_mask(p)).astype(np.float32)
    shots = [(6, 2), (14, 3), (20, 2), (27, 4), (33, 2)]
    r = rng.standard_normal((len(shots), 2, p.lt, p.nrec)).astype(np.float32)
    print(f"batched 2D operators L={L} R={R} inversion_para={para}, 5 shots in groups of 3", flush=True)
    with erm.Elastic2D(p) as g:
        g.load_model(1, *bg)
        d1, g1 = [], np.zeros((3, p.nx, p.nz), np.float64)
        for k, (sx, sz) in enumerate(shots):
            g.set_shot(sx, sz)
            d1.append(g.born(dm))
            g1 += g.adjoint(r[k])
        d1 = np.stack(d1)
        seq = g.born_shots(shots, dm)  # no batch set: one shot per launch
        np.testing.assert_array_equal(seq, d1)
        g.set_batch(3)
        db = g.born_shots(shots, dm)
        gb = g.adjoint_shots(shots, r)
        np.testing.assert_array_equal(db, d1)
        for c in range(3):
            e = cases.rel_l2(gb[c], g1[c])
            assert e < 1e-6, (c, e)
        # the single-shot operators still work after batching, on the set shot
        g.set_shot(*shots[1])
        np.testing.assert_array_equal(g.born(dm), d1[1])
        g.set_batch(1)


@pytest.mark.parametrize("L,R,mode", [(0, 4, 0), (3, 6, 0), (3, 4, 1)])
def test_gpu2d_tall_workgroups_give_the_same_operators(erm, L, R, mode):
    """tile_rows = 8 (the taller stencil workgroups that large batched
    launches use by default) against tile_rows = 4: Born records bit-identical
    (the same arithmetic per cell), adjoint gradients equal to the last bits,
    batched and one shot, for the exact and the reference-mode adjoint."""
    import test_oracle2d as T
    p, bg = T._lsrtm_case(L, R=R, lt=120)
    p.adjoint_mode = mode
    rng = np.random.default_rng(17)
    dm = (rng.standard_normal((3, p.nx, p.nz)) * cases2d.taper_mask(p)).astype(np.float32)
    shots = [(6, 2), (14, 3), (20, 2), (27, 4)]
    r = rng.standard_normal((len(shots), 2, p.lt, p.nrec)).astype(np.float32)
    print(f"tile_rows 8 vs 4: L={L} R={R} adjoint_mode={mode}", flush=True)
    out = {}
    for rows in (4, 8):
        p.tile_rows = rows
        with erm.Elastic2D(p) as g:
            g.load_model(1, *bg)
            g.set_shot(*shots[0])
            one = (g.born(dm), g.adjoint(r[0]))
            g.set_batch(len(shots))
            out[rows] = one + (g.born_shots(shots, dm), g.adjoint_shots(shots, r))
    for k in (0, 2):
        np.testing.assert_array_equal(out[8][k], out[4][k])
    for k in (1, 3):
        for c in range(3):
            e = cases.rel_l2(out[8][k][c], out[4][k][c])
            assert e < 1e-6, (k, c, e)


def test_gpu2d_batched_cg_follows_one_shot_cg(erm):
    p = cases2d.params(nx=40, nz=24, b=10, R=4, L=0, lt=160, freq=25.0, src=(20, 1), rec_dx=1, rec_z=1, nrec=38,
                       rec_x0=1)
    vp, vs, rho, _, _ = cases2d.model(p, seed=4, q=False)
    bg = (cases2d.smooth2(vp), cases2d.smooth2(vs), cases2d.smooth2(rho))
    rng = np.random.default_rng(2)
    dm = np.zeros((3, p.nx, p.nz), np.float32)
    dm[:2, 4:-4, 6:-4] = 0.05 * rng.standard_normal((2, p.nx - 8, p.nz - 10))
    shots = [(6, 1), (14, 1), (26, 1), (34, 1), (20, 1)]
    print("batched CG vs one-shot CG, 5 shots, batch 4", flush=True)
    with erm.Elastic2D(p) as g:
        g.load_model(1, *bg)
        d = g.born_shots(shots, dm)
        m1, mis1, _, it1 = g.lsrtm_cg(shots, d, 6, pmask=3)
        g.set_batch(4)
        mb, misb, secb, itb = g.lsrtm_cg(shots, d, 6, pmask=3)
        assert it1 == itb == 6
        np.testing.assert_allclose(misb, mis1, rtol=1e-5)
        np.testing.assert_allclose(mb, m1, rtol=1e-4, atol=1e-4 * np.abs(m1).max())
        g.set_batch(1)
