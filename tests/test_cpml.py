"""C-PML (no reference counterpart): pinned by closed properties on the CPU
oracle, then GPU parity against it.

* Rc = 1 makes d0 = 0, so every memory variable stays zero: the CPML run must
  equal the run without any absorbing layer (damping coefficient 0) exactly.
* Against an enlarged domain (the same model padded by 34 more cells, so no
  boundary reflection reaches the receivers in the window), the CPML
  receiver traces (whole time history, 4x4 receivers in the model) differ by
  < 3e-3 (L2; measured 8.5e-4 with 6 updated absorbing cells) while the same
  run with no absorption differs by > 100x more (measured 0.61).
"""
import numpy as np
import pytest

import cases

PASS_MODEL = 0


def _case(R=4, b=10, n=24, lt=110, boundary=1, coe=0.0, Rc=1e-3):
    p = cases.params(nx=n, ny=n, nz=n, b=b, R=R, lt=lt, dt=1.0, freq=25.0, coe=coe,
                     src=(n // 2, n // 2, n // 2), rec=(4, 4, 1, 1, 5, 5, 2))
    p.boundary = boundary
    p.cpml_r = Rc
    p.cpml_vmax = 2500.0
    return p


def _run(oracle, p, prec=64, steps=None):
    o = oracle.Oracle3D(p, prec)
    o.load_model(0, *cases.homogeneous(p, vp=2500.0, rho=2000.0))
    o.step(PASS_MODEL, 0, p.lt if steps is None else steps)
    return o


def _model_region(p, a):
    return a[p.bu:p.bu + p.nz, p.bf:p.bf + p.ny, p.bl:p.bl + p.nx]


def test_cpml_with_unit_reflection_is_exactly_no_absorption(oracle):
    pc = _case(boundary=1, Rc=1.0, lt=60)
    pd = _case(boundary=0, coe=0.0, lt=60)
    oc, od = _run(oracle, pc), _run(oracle, pd)
    for k in ["vx", "vz", "txx", "txz", "tp", "vzp"]:
        np.testing.assert_array_equal(oc.get_field(k), od.get_field(k))


@pytest.mark.parametrize("R", [4, 6])
def test_cpml_reflection_against_enlarged_domain(oracle, R):
    # 25 Hz Ricker peaks at step 43; at 0.25 cell/step the wave has travelled
    # 32 cells by step 170: out to the grid edge (20-22 cells) and back to the
    # receivers without absorption; the enlarged domain's edge is >50 cells away
    n, b, lt = 20, R + 6, 170   # 6 updated absorbing cells per face
    big = _case(R=R, b=b + 34, n=n, lt=lt, boundary=0, coe=0.0)
    ob = _run(oracle, big)
    ref = ob.get_record()
    errs = {}
    for name, p in [("cpml", _case(R=R, b=b, n=n, lt=lt, boundary=1)),
                    ("none", _case(R=R, b=b, n=n, lt=lt, boundary=0, coe=0.0))]:
        o = _run(oracle, p)
        errs[name] = cases.rel_l2(o.get_record(), ref)
    assert errs["cpml"] < 3e-3, errs
    assert errs["none"] > 100 * errs["cpml"], errs


@pytest.mark.gpu
@pytest.mark.parametrize("R", [4, 6])
def test_gpu_cpml_matches_fp64_oracle(erm, oracle, R):
    p = cases.params(nx=40, ny=34, nz=30, b=10, R=R, lt=120, dt=1.0, freq=25.0)
    p.boundary = 1
    vp, vs, rho = cases.layered(p, seed=3)
    print(f"CPML parity R={R}: padded {tuple(p.padded_shape)}", flush=True)
    g = erm.RTM3D(p)
    g.load_model(0, vp, vs, rho)
    g.zero_wavefield()
    g.step(PASS_MODEL, 0, p.lt)
    o = oracle.Oracle3D(p, 64)
    o.load_model(0, vp, vs, rho)
    o.step(PASS_MODEL, 0, p.lt)
    for k in ["vx", "vy", "vz", "txx", "tyy", "tzz", "txy", "txz", "tyz", "tp"]:
        e = cases.rel_l2(g.get_field(k), o.get_field(k))
        assert e < 1e-5, (k, e)
    for k in ["vxp", "vyp", "vzp"]:
        e = cases.rel_l2(g.get_field(k), o.get_field(k))
        assert e < 1e-4, (k, e)
    assert cases.rel_l2(g.get_record(), o.get_record()) < 1e-5
    # a second shot after zeroing starts from a clean memory-variable state
    g.zero_wavefield()
    g.step(PASS_MODEL, 0, 30)
    o.zero()
    o.step(PASS_MODEL, 0, 30)
    assert cases.rel_l2(g.get_field("vz"), o.get_field("vz")) < 1e-5
    g.close()


@pytest.mark.gpu
def test_gpu_cpml_region_split_and_autotune_are_exact(erm):
    """The region-split launches (plain kernel on tiles without C-PML cells),
    run on one stream or concurrently on two (region_streams), in the
    XCD-aware tile orders (xcd_map 1: contiguous chunks, 2: y bands), tiles
    32 x 8 (default), 32 x 16, 64 x 4, 64 x 8 and every autotuned launch shape
    give bit-identical shots to one C-PML kernel over all tiles with the
    default shape."""
    p = cases.params(nx=150, ny=40, nz=44, b=12, R=4, lt=80, dt=1.0, freq=25.0, src=(75, 20, 3))
    p.boundary = 1
    vp, vs, rho = cases.layered(p, seed=5)
    sm = [cases.smooth(a, 3) for a in (vp, vs, rho)]
    out = []
    for split_off, tune, streams, xcd, tx, ty in ((1, False, 0, 0, 0, 0), (0, False, 0, 0, 0, 0), (0, True, 0, 0, 0, 0),
                                                  (0, False, 1, 0, 0, 0), (0, True, 1, 0, 0, 0), (0, False, 0, 1, 0, 0),
                                                  (0, False, 0, 2, 0, 0), (0, True, 0, 2, 0, 0), (0, False, 0, 2, 32, 16),
                                                  (0, False, 0, 2, 64, 4), (0, False, 0, 0, 64, 8), (1, False, 0, 0, 64, 4)):
        p.cpml_split = split_off
        p.region_streams = streams
        p.xcd_map = xcd
        p.tile_x, p.tile_y = tx, ty
        print(f"C-PML shot: split_off={split_off} tune={tune} streams={streams} xcd_map={xcd} tile={tx}x{ty}",
              flush=True)
        with erm.RTM3D(p) as g:
            g.load_model(0, vp, vs, rho)
            g.load_model(1, *sm)
            rec = g.forward_real()
            if tune:
                t = g.autotune(steps=2)
                assert (t["tile_x"], t["tile_y"]) in ((64, 4), (64, 8), (32, 8), (32, 16))
                assert t["zchunk"] in (4, 5, 6, 7, 8, 12, 16, 32)
            g.forward_smoothed()
            g.backward_smoothed()
            out.append((rec, g.get_images(), g.get_field("ex_time")))
    for r in out[1:]:
        np.testing.assert_array_equal(r[0], out[0][0])
        np.testing.assert_array_equal(r[2], out[0][2])
        for q in range(3):
            np.testing.assert_array_equal(r[1][q], out[0][1][q])
