"""2D C-PML (no reference counterpart; the configured 2D workload asks for a
40-cell CPML): pinned on the CPU oracle by the same closed properties as the
3D one, then GPU parity against it.

* Rc = 1 makes d0 = 0: the C-PML run equals the run with no absorption
  (damping coefficient 0) exactly.
* Against an enlarged domain (40 more cells per face, so no boundary
  reflection reaches the receivers in the window) the C-PML receiver traces
  differ by < 5e-3 (L2 over the whole record; measured 3.1e-4 at radius 4
  and 6 with 12 updated absorbing cells) while the run with no absorption
  differs by > 50x more (0.70); the reference's cubic damping sits in between
  (0.075 / 0.15).
* GPU: wavefields and record <= 1e-5 vs the FP64 oracle (vector P <= 1e-4 or
  10x the FP32 oracle's own error, as in tests/test_gpu2d.py),
  elastic and viscoelastic; the LSRTM operators refuse the C-PML boundary.
"""
import numpy as np
import pytest

import cases
import cases2d


def _case(b=16, n=40, lt=260, boundary=1, coe=0.0, Rc=1e-3, R=4, L=0):
    p = cases2d.params(nx=n, nz=n, b=b, R=R, L=L, lt=lt, dt=1.0, freq=25.0, coe=coe, src=(n // 2, n // 2),
                       nrec=8, rec_dx=4, rec_z=3)
    p.rec_x0 = 4
    p.boundary = boundary
    p.cpml_r = Rc
    p.cpml_vmax = 2500.0
    return p


def _run(oracle2d_mod, p, L=0):
    o = oracle2d_mod.Oracle2D(p)
    shape = (p.nx, p.nz)
    vp = np.full(shape, 2500.0, np.float32)
    q = np.full(shape, 50.0, np.float32) if L else None
    o.load_model(0, vp, (vp / 1.732).astype(np.float32), np.full(shape, 2000.0, np.float32), q,
                 None if q is None else (0.6 * q).astype(np.float32))
    o.step(0, 0, p.lt)
    return o


def test_cpml2d_with_unit_reflection_is_exactly_no_absorption(oracle2d_mod):
    oc = _run(oracle2d_mod, _case(boundary=1, Rc=1.0, lt=80))
    od = _run(oracle2d_mod, _case(boundary=0, coe=0.0, lt=80))
    for k in ["vx", "vz", "txx", "tzz", "txz", "tp", "vxp", "vzp"]:
        np.testing.assert_array_equal(oc.get_field(k), od.get_field(k))


@pytest.mark.parametrize("R", [4, 6])
def test_cpml2d_reflection_against_enlarged_domain(oracle2d_mod, R):
    b = R + 12
    ref = _run(oracle2d_mod, _case(b=b + 40, boundary=0, R=R)).record()
    errs = {name: cases.rel_l2(_run(oracle2d_mod, p).record(), ref)
            for name, p in [("cpml", _case(b=b, boundary=1, R=R)), ("none", _case(b=b, boundary=0, R=R)),
                            ("damping", _case(b=b, boundary=0, coe=200.0, R=R))]}
    assert errs["cpml"] < 5e-3, errs
    assert errs["none"] > 50 * errs["cpml"], errs
    assert errs["damping"] > errs["cpml"], errs


@pytest.mark.gpu
@pytest.mark.parametrize("L,R", [(0, 4), (3, 4), (0, 6)])
def test_gpu_cpml2d_matches_fp64_oracle(erm, oracle2d_mod, L, R):
    p = cases2d.params(nx=80, nz=60, b=20, R=R, L=L, lt=300, freq=20.0, boundary=1, cpml_r=1e-3)
    vp, vs, rho, qp, qs = cases2d.model(p, seed=6, q=L > 0)
    print(f"2D C-PML parity L={L} R={R}", flush=True)
    g = erm.Elastic2D(p)
    g.load_model(0, vp, vs, rho, qp, qs)
    rec = g.forward()
    o = oracle2d_mod.Oracle2D(p)
    o.load_model(0, vp, vs, rho, qp, qs)
    o.step(0, 0, p.lt)
    assert cases.rel_l2(rec, o.record()) < 1e-5
    for k in ["vx", "vz", "txx", "tzz", "txz", "tp"]:
        e = cases.rel_l2(g.get_field(k), o.get_field(k))
        assert e < 1e-5, (k, e)
    o32 = oracle2d_mod.Oracle2D(p, 32)
    o32.load_model(0, vp, vs, rho, qp, qs)
    o32.step(0, 0, p.lt)
    for k in ["vxp", "vzp"]:   # vector P integrates grad tp: as the FP32 oracle itself (see test_gpu2d)
        ref = o.get_field(k)
        e, e32 = cases.rel_l2(g.get_field(k), ref), cases.rel_l2(o32.get_field(k), ref)
        assert e < max(1e-4, 10 * e32), (k, e, e32)
    g.load_model(1, *[cases2d.smooth2(a) for a in (vp, vs, rho)], qp, qs)
    with pytest.raises(erm.ErmError):
        g.born(np.zeros((3, p.nx, p.nz), np.float32))
    g.close()
