"""Pins the 2D CPU oracle (oracle/oracle2d.cpp) before it checks the GPU.

* single-step closed forms on quadratic fields for the 8 staggered 2D
  derivatives (the positions of viscoelastic_equation.cu:137-176);
* SLS: the reference's L=1 parameters (viscoelastic_equation.cu:43-56) give
  Q_P(w0) = Qp and Q_S(w0) = Qs for the standard-linear-solid modulus
  M(w) = M_R (1 + i w tau_eps)/(1 + i w tau_sig); the L=3 tau-method hits
  Q exactly at w0 and stays within 20% over [f/2, 2f];
* the viscoelastic step tends to the elastic one as Q -> infinity;
* Born vs the centred finite difference of forward modelling (O(h^2));
* the adjoint is the exact transpose: <L dm, r> = <dm, L^T r> to 1e-12 (FP64).
"""
import numpy as np
import pytest

import cases2d


def _gsls_q(s, w, L):
    """Q(w) of the discrete SLS parameters (per unit relaxed modulus)."""
    # M(w)/M_R = fM - sum_l g_l tau_l * (something): rebuild tau_eps/tau_sig from fM, g
    # For L=1: fM = te/ts, g = (1 - te/ts)/ts.  For L mechanisms: fM = 1 + L tau,
    # g_l = -tau/ts_l.  Both: M(w)/M_R = fM + sum_l g_l ts_l * (1/(1 + i w ts_l) - 1) ... closed form below.
    Mw = s["fM"] + 0j
    for l in range(L):
        ts = s["taus"][l]
        # relaxation of mechanism l: -g_l ts_l * (i w ts_l)/(1 + i w ts_l) subtracted from the unrelaxed value
        Mw += s["gP"][l] * ts * (1j * w * ts) / (1 + 1j * w * ts) - s["gP"][l] * ts
    return Mw.real / Mw.imag


@pytest.mark.parametrize("qp", [30.0, 80.0, 200.0])
def test_sls_reference_l1_hits_q_at_reference_frequency(oracle2d_mod, qp):
    f0 = 20.0
    w = 2 * 3.1415926 * f0
    s = oracle2d_mod.sls(1, f0, qp, 0.6 * qp)
    te_ts = s["fM"]
    ts = s["taus"][0]
    # reference closed form (A.4): M(w)/M_R = (1 + i w te)/(1 + i w ts)
    te = te_ts * ts
    Mw = (1 + 1j * w * te) / (1 + 1j * w * ts)
    assert abs(Mw.real / Mw.imag - qp) / qp < 1e-9
    teS = s["fU"] * ts
    Mws = (1 + 1j * w * teS) / (1 + 1j * w * ts)
    assert abs(Mws.real / Mws.imag - 0.6 * qp) / (0.6 * qp) < 1e-9
    # memory coefficient (1 - te/ts)/ts as in the reference kernels
    assert abs(s["gP"][0] - (1 - te_ts) / ts) < 1e-9 * abs(s["gP"][0])


def test_sls_three_mechanisms_flat_q(oracle2d_mod):
    f0, qp = 20.0, 50.0
    s = oracle2d_mod.sls(3, f0, qp, 30.0)
    tau = (s["fM"] - 1) / 3
    for f in np.linspace(f0 / 2, 2 * f0, 25):
        w = 2 * np.pi * f
        Mw = 1 + sum(tau * (1j * w * t) / (1 + 1j * w * t) for t in s["taus"])
        q = Mw.real / Mw.imag
        if abs(f - f0) < 1e-9:
            assert abs(q - qp) / qp < 1e-6
        assert abs(q - qp) / qp < 0.2, (f, q)
    np.testing.assert_allclose(s["gP"], -tau / s["taus"], rtol=1e-12)
    w0 = 2 * 3.1415926 * f0
    Mw0 = 1 + sum(tau * (1j * w0 * t) / (1 + 1j * w0 * t) for t in s["taus"])
    assert abs(Mw0.real / Mw0.imag - qp) / qp < 1e-3


# ---------------------------------------------------------------- closed forms
CASES = [  # (field set, axis 0=x 1=z, {output: (modulus key, half shift)})
    ("vx", 0, {"txx": ("M", +0.5), "tzz": ("L", +0.5), "tp": ("M", +0.5)}),
    ("vx", 1, {"txz": ("U", +0.5)}),
    ("vz", 0, {"txz": ("U", -0.5)}),
    ("vz", 1, {"tzz": ("M", -0.5), "txx": ("L", -0.5), "tp": ("M", -0.5)}),
    ("txx", 0, {"vx": ("B", -0.5)}),
    ("txz", 1, {"vx": ("B", -0.5)}),
    ("txz", 0, {"vz": ("B", +0.5)}),
    ("tzz", 1, {"vz": ("B", +0.5)}),
    ("tp", 0, {"vxp": ("B", -0.5)}),
    ("tp", 1, {"vzp": ("B", +0.5)}),
]
OUTS = ["vx", "vz", "txx", "tzz", "txz", "tp", "vxp", "vzp"]


def closed_form_2d(factory, R):
    p = cases2d.params(nx=26, nz=22, b=0, R=R, L=0, lt=1, src=(13, 11), nrec=2, rec_z=11)
    vp = np.full((p.nx, p.nz), 2500.0, np.float32)
    vs = np.full_like(vp, 1300.0)
    rho = np.full_like(vp, 1800.0)
    M, U = 1800.0 * 2500.0 ** 2, 1800.0 * 1300.0 ** 2
    mod = {"M": M, "U": U, "L": M - 2 * U, "B": 1 / 1800.0}
    delta = p.dt / 1000.0
    for comp, axis, expect in CASES:
        e = factory(p)
        e.load_model(1, vp, vs, rho)
        nnx, nnz = p.padded_shape
        idx = np.indices((nnx, nnz))[axis].astype(np.float64)
        e.set_field(comp, (idx * p.dx) ** 2)
        e.step(1, 100000, 1)
        sl = (slice(R, nnx - R), slice(R, nnz - R))
        # a velocity input is checked on the stresses it drives (the velocities
        # then respond to those stresses near the never-updated margins) and
        # a stress input on the velocities
        outs = ["txx", "tzz", "txz", "tp"] if comp.startswith("v") else ["vx", "vz", "vxp", "vzp"]
        for o in outs:
            got = np.asarray(e.get_field(o), np.float64)[sl]
            if o in expect:
                key, sh = expect[o]
                want = (mod[key] * delta * 2.0 * (idx + sh) * p.dx)[sl]
                err = np.max(np.abs(got - want)) / np.max(np.abs(want))
            elif o == comp:
                continue
            else:
                ref = mod["M"] if comp.startswith("v") else mod["B"]
                err = np.max(np.abs(got)) / (ref * delta * p.dx * nnx)
            yield comp, axis, o, err


@pytest.mark.parametrize("R", [4, 6])
def test_closed_form_2d_derivatives(oracle2d_mod, R):
    tol = 1e-12 if R == 4 else 1e-9
    for comp, axis, o, err in closed_form_2d(lambda p: oracle2d_mod.Oracle2D(p), R):
        assert err < tol, (comp, axis, o, err)


def test_visco_tends_to_elastic_for_large_q(oracle2d_mod):
    p0 = cases2d.params(L=0, lt=150)
    p1 = cases2d.params(L=1, lt=150)
    vp, vs, rho, _, _ = cases2d.model(p0)
    big = np.full(vp.shape, 1e7, np.float32)
    o0, o1 = oracle2d_mod.Oracle2D(p0), oracle2d_mod.Oracle2D(p1)
    o0.load_model(0, vp, vs, rho)
    o1.load_model(0, vp, vs, rho, big, big)
    o0.step(0, 0, p0.lt)
    o1.step(0, 0, p1.lt)
    r0, r1 = o0.record(), o1.record()
    assert np.linalg.norm(r1 - r0) / np.linalg.norm(r0) < 1e-5


def test_attenuation_reduces_amplitude(oracle2d_mod):
    p0 = cases2d.params(L=0, lt=300)
    p1 = cases2d.params(L=1, lt=300)
    vp, vs, rho, _, _ = cases2d.model(p0)
    q = np.full(vp.shape, 20.0, np.float32)
    o0, o1 = oracle2d_mod.Oracle2D(p0), oracle2d_mod.Oracle2D(p1)
    o0.load_model(0, vp, vs, rho)
    o1.load_model(0, vp, vs, rho, q, 0.6 * q)
    o0.step(0, 0, p0.lt)
    o1.step(0, 0, p1.lt)
    late = slice(150, 300)
    assert np.linalg.norm(o1.record()[:, late]) < 0.8 * np.linalg.norm(o0.record()[:, late])


# ---------------------------------------------------------------- Born / adjoint
def _lsrtm_case(L, R=4, lt=160):
    p = cases2d.params(nx=40, nz=30, b=10, R=R, L=L, lt=lt, freq=25.0, rec_dx=2)
    vp, vs, rho, qp, qs = cases2d.model(p, seed=3)
    return p, (cases2d.smooth2(vp), cases2d.smooth2(vs), cases2d.smooth2(rho), qp, qs)


def perturbed_model(bg, dm, eps, para):
    """(vp, vs, rho) of the background with the relative perturbation eps*dm of
    the inversion_para parameters applied to the parameters themselves:
    1 lambda, mu, rho; 2 vp, vs, rho; 3 Ip = rho vp, Is = rho vs, rho."""
    vp, vs, rho = (np.asarray(a, np.float64) for a in bg[:3])
    r2 = rho * (1 + eps * dm[2])
    if para == 1:
        lam, mu = rho * (vp ** 2 - 2 * vs ** 2) * (1 + eps * dm[0]), rho * vs ** 2 * (1 + eps * dm[1])
        return np.sqrt((lam + 2 * mu) / r2), np.sqrt(mu / r2), r2
    if para == 3:
        return rho * vp * (1 + eps * dm[0]) / r2, rho * vs * (1 + eps * dm[1]) / r2, r2
    return vp * (1 + eps * dm[0]), vs * (1 + eps * dm[1]), r2


@pytest.mark.parametrize("L,para", [(0, 2), (1, 2), (3, 2), (0, 1), (3, 1), (0, 3), (3, 3)])
def test_born_matches_finite_difference(oracle2d_mod, L, para):
    """Born data of each parameterisation against a centred finite difference
    of modelling with the perturbation applied to that parameterisation's own
    parameters (a wrong chain rule between the parameter sets fails)."""
    p, bg = _lsrtm_case(L)
    p.inversion_para = para
    rng = np.random.default_rng(7)
    w = cases2d.taper_mask(p)
    dm = (rng.standard_normal((3, p.nx, p.nz)) * w).astype(np.float32)
    o = oracle2d_mod.Oracle2D(p)
    o.load_model(1, *bg)
    lin = o.born(dm).astype(np.float64)
    recs = {}
    h = 1e-4
    for sgn in (+1, -1):
        of = oracle2d_mod.Oracle2D(p)
        vp, vs, rho = perturbed_model(bg, dm, sgn * h, para)
        of.load_model(0, vp.astype(np.float32), vs.astype(np.float32), rho.astype(np.float32), bg[3], bg[4])
        of.step(0, 0, p.lt)
        recs[sgn] = of.record()
    fd = (recs[1] - recs[-1]) / (2 * h)
    err = np.linalg.norm(fd - lin) / np.linalg.norm(lin)
    # float32 model arrays limit the perturbation resolution to ~1e-7/h
    assert err < 5e-3, err


@pytest.mark.parametrize("L,R,para", [(0, 4, 2), (1, 4, 2), (3, 4, 2), (0, 6, 2), (3, 4, 1), (3, 4, 3)])
def test_adjoint_is_exact_transpose(oracle2d_mod, L, R, para):
    p, bg = _lsrtm_case(L, R=R, lt=120)
    p.inversion_para = para
    rng = np.random.default_rng(11)
    dm = rng.standard_normal((3, p.nx, p.nz)).astype(np.float32)
    r = rng.standard_normal((2, p.lt, p.nrec)).astype(np.float32)
    o = oracle2d_mod.Oracle2D(p)
    o.load_model(1, *bg)
    Ld = o.born(dm).astype(np.float64)
    Ltr = o.adjoint(r).astype(np.float64)
    lhs = float(np.sum(Ld * r.astype(np.float64)))
    rhs = float(np.sum(dm.astype(np.float64) * Ltr))
    # outputs are rounded to float32 on return, hence 1e-6 rather than 1e-12
    assert abs(lhs - rhs) / max(abs(lhs), abs(rhs)) < 1e-6, (lhs, rhs)


@pytest.mark.parametrize("R,para", [(4, 2), (6, 2), (4, 1)])
def test_adjoint_from_strips_equals_the_history_adjoint(oracle2d_mod, R, para):
    """adjoint_source = 2: the background stepped back from the ring strips
    (reversed velocity then stress, ring restored before each) recomputes the
    E and F the history stores, so in FP64 the gradient equals the history
    adjoint's to rounding (3e-16 / 1e-14 measured for R = 4 / 6) and the
    dot-product test against Born still holds; leaving out the velocity-side
    restore (planted in a copy of the oracle) gives 3e-5.  In FP32 the
    reversal adds rounding only: 1.2e-6 against FP64 (the FP32 history
    adjoint: 7e-7)."""
    p, bg = _lsrtm_case(0, R=R, lt=120)
    p.inversion_para = para
    rng = np.random.default_rng(13)
    dm = rng.standard_normal((3, p.nx, p.nz)).astype(np.float32)
    r = rng.standard_normal((2, p.lt, p.nrec)).astype(np.float32)
    o = oracle2d_mod.Oracle2D(p)
    o.load_model(1, *bg)
    gh = o.adjoint(r).astype(np.float64)
    gs = o.adjoint_strips(r).astype(np.float64)
    assert np.linalg.norm(gs - gh) / np.linalg.norm(gh) < 1e-12
    Ld = o.born(dm).astype(np.float64)
    lhs, rhs = float(np.sum(Ld * r.astype(np.float64))), float(np.sum(dm.astype(np.float64) * gs))
    assert abs(lhs - rhs) / max(abs(lhs), abs(rhs)) < 1e-6, (lhs, rhs)
    o32 = oracle2d_mod.Oracle2D(p, prec=32)
    o32.load_model(1, *bg)
    g32 = o32.adjoint_strips(r).astype(np.float64)
    assert np.linalg.norm(g32 - gh) / np.linalg.norm(gh) < 1e-5


def test_adjoint_strips_rejects_viscoelastic_media(oracle2d_mod):
    p, bg = _lsrtm_case(1, lt=20)
    o = oracle2d_mod.Oracle2D(p)
    o.load_model(1, *bg)
    with pytest.raises(Exception):
        o.adjoint_strips(np.zeros((2, p.lt, p.nrec), np.float32))


def test_rtm2d_reflector_depth(oracle2d_mod):
    nz, zi = 50, 32
    p = cases2d.params(nx=60, nz=nz, b=14, R=4, L=0, lt=700, freq=25.0, coe=250.0, src=(30, 3), rec_dx=1, rec_z=2)
    z = np.arange(nz)[None, :]
    vp = (np.where(z < zi, 2000.0, 3000.0) * np.ones((p.nx, 1))).astype(np.float32)
    rho = (np.where(z < zi, 2000.0, 2400.0) * np.ones((p.nx, 1))).astype(np.float32)
    vs = (vp / 1.732).astype(np.float32)
    o = oracle2d_mod.Oracle2D(p)
    o.load_model(0, vp, vs, rho)
    s = np.full_like(vp, 2000.0)
    o.load_model(1, s, (s / 1.732).astype(np.float32), np.full_like(vp, 2000.0))
    o.step(0, 0, p.lt)
    rec = o.record()
    o.zero(7)
    o.step(1, 0, p.lt)
    o.zero(1)
    o.step(2, p.lt - 1, p.lt)
    tp, pp, ps = o.images()
    col = np.abs(pp[26:35, 10:]).sum(axis=0)
    assert abs(int(np.argmax(col)) + 10 - zi) <= 2


def test_q_compensated_receiver_pass_restores_amplitude(oracle2d_mod):
    """Receiver back-propagation of data recorded in a strongly attenuating
    medium (Q = 15, one SLS mechanism): flipping the memory-term sign
    (Q-compensated propagation) leaves more energy in the back-propagated
    field than propagating with attenuation, and the flag's effect vanishes
    like 1/Q (x100 smaller from Q = 1e4 to 1e6)."""
    def energy_after_receiver_pass(q, qcomp):
        p = cases2d.params(nx=50, nz=40, b=12, R=4, L=1, lt=260, freq=25.0, src=(25, 30), rec_dx=1, rec_z=2)
        p.qcomp = qcomp
        vp = np.full((p.nx, p.nz), 2200.0, np.float32)
        vs = (vp / 1.732).astype(np.float32)
        rho = np.full_like(vp, 2100.0)
        qp = np.full_like(vp, q)
        qs = (0.6 * qp).astype(np.float32)
        o = oracle2d_mod.Oracle2D(p)
        o.load_model(0, vp, vs, rho, qp, qs)
        o.load_model(1, vp, vs, rho, qp, qs)
        o.step(0, 0, p.lt)
        o.zero(1)
        o.step(2, p.lt - 1, p.lt)
        return float(np.sum(o.get_field("vx") ** 2 + o.get_field("vz") ** 2))

    e_att, e_comp = energy_after_receiver_pass(15.0, 0), energy_after_receiver_pass(15.0, 1)
    assert e_comp > 1.5 * e_att, (e_comp, e_att)
    # the flag acts through the memory term only, which scales as 1/Q
    d4 = abs(energy_after_receiver_pass(1e4, 1) / energy_after_receiver_pass(1e4, 0) - 1)
    d6 = abs(energy_after_receiver_pass(1e6, 1) / energy_after_receiver_pass(1e6, 0) - 1)
    assert d6 < 1e-3 and 30 < d4 / d6 < 300, (d4, d6)


def _illum_case():
    p = cases2d.params(nx=40, nz=30, b=10, R=4, L=0, lt=150, freq=25.0, src=(20, 2))
    vp, vs, rho, _, _ = cases2d.model(p, seed=21, q=False)
    bg = [cases2d.smooth2(a) for a in (vp, vs, rho)]
    return p, (vp, vs, rho), bg


def test_illumination_additive_background_only_and_peaked_at_source(oracle2d_mod):
    """Oracle source illumination: the sum over shots is the sum of the
    per-shot images (the state is reset between shots), only the background
    model enters (a different true model changes nothing), it is
    non-negative and peaks next to the source."""
    p, true, bg = _illum_case()
    o = oracle2d_mod.Oracle2D(p)
    o.load_model(0, *true)
    o.load_model(1, *bg)
    s1, s2 = [(10, 2)], [(30, 3)]
    w1, w2, w12 = o.illumination(s1), o.illumination(s2), o.illumination(s1 + s2)
    np.testing.assert_allclose(w12, w1 + w2, rtol=1e-12, atol=1e-30)
    assert w1.min() >= 0 and w1.max() > 0
    i, k = np.unravel_index(np.argmax(w1), w1.shape)
    assert abs(i - 10) <= 3 and abs(k - 2) <= 3, (i, k)
    o2 = oracle2d_mod.Oracle2D(p)
    o2.load_model(0, *[(a * 1.3).astype(np.float32) for a in true])
    o2.load_model(1, *bg)
    np.testing.assert_array_equal(o2.illumination(s1), w1)
    o3 = oracle2d_mod.Oracle2D(p)
    o3.load_model(0, *true)
    o3.load_model(1, *[(a * 1.1).astype(np.float32) for a in bg])
    assert cases2d_rel(o3.illumination(s1), w1) > 1e-3


def cases2d_rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _flat_reflector_rtm(oracle2d_mod, nx=121):
    """Excitation RTM of one centred shot over a flat reflector at depth 30
    (vp 2000 -> 3000, critical angle 41.8 deg), homogeneous migration model."""
    nz, zi = 50, 30
    p = cases2d.params(nx=nx, nz=nz, b=14, R=4, L=0, lt=900, freq=25.0, coe=250.0, src=(60, 3), rec_dx=1,
                       rec_z=2, nrec=nx - 2)
    z = np.arange(nz)[None, :]
    vp = (np.where(z < zi, 2000.0, 3000.0) * np.ones((p.nx, 1))).astype(np.float32)
    rho = (np.where(z < zi, 2000.0, 2400.0) * np.ones((p.nx, 1))).astype(np.float32)
    vs = (vp / 1.732).astype(np.float32)
    s = np.full_like(vp, 2000.0)
    bg = (s, (s / 1.732).astype(np.float32), np.full_like(vp, 2000.0))
    o = oracle2d_mod.Oracle2D(p)
    o.load_model(0, vp, vs, rho)
    o.load_model(1, *bg)
    o.step(0, 0, p.lt)
    o.zero(7)
    o.step(1, 0, p.lt)
    o.zero(1)
    o.step(2, p.lt - 1, p.lt)
    return p, (vp, vs, rho), bg, o, zi


def test_angle_gathers_flat_reflector(oracle2d_mod):
    """Oracle angle gathers: at the reflector, the reflection angle and the
    gathers' peak bin follow the geometric incidence atan(offset / depth)
    below the critical angle -- within 9 deg everywhere and 4 deg on average
    (Poynting directions read just above a reflector mix the incident and
    reflected waves, which biases them 2-4 deg toward the vertical), with the
    right sign; every imaged cell lands in exactly one bin; the PS image of
    the vector condition is mirror-symmetric about the source (no polarity
    reversal to correct)."""
    p, _, _, o, zi = _flat_reflector_rtm(oracle2d_mod)
    tp, pp, ps = o.images()
    nang, dang = 30, 3.0
    gpp, gps, ppc, psc, g = o.angle_gathers(nang, dang)
    band = slice(zi - 5, zi + 6)
    err_g, err_b = [], []
    for x in list(range(40, 58, 2)) + list(range(62, 82, 2)):  # incidence 4-36.5 deg
        expect = np.degrees(np.arctan((x - 60) / (zi - 3)))
        k = 20 + int(np.argmax(np.abs(pp[x, 20:40])))
        e = (gpp[:, x, band] ** 2).sum(axis=1)
        peak = -0.5 * nang * dang + dang * (int(np.argmax(e)) + 0.5)
        err_g.append(g[x, k] - expect)
        err_b.append(peak - expect)
        if abs(expect) >= 8:
            assert np.sign(g[x, k]) == np.sign(expect) == np.sign(peak), (x, g[x, k], peak, expect)
    for err in (err_g, err_b):
        assert np.max(np.abs(err)) <= 9.0 and np.mean(np.abs(err)) <= 4.0, err
    inrange = np.abs(g) < 0.5 * nang * dang
    np.testing.assert_allclose(gpp.sum(axis=0)[inrange], pp[inrange], rtol=0, atol=0)
    np.testing.assert_allclose(gps.sum(axis=0)[inrange], ps[inrange], rtol=0, atol=0)
    num = sum(float((ps[60 + d, band] * ps[60 - d, band]).sum()) for d in range(4, 30))
    den = sum(float(np.sqrt((ps[60 + d, band] ** 2).sum() * (ps[60 - d, band] ** 2).sum())) for d in range(4, 30))
    assert num / den > 0.9, num / den
    assert np.isfinite(ppc).all() and np.isfinite(psc).all()


@pytest.mark.parametrize("R", [4, 6])
def test_discrete_energy_is_conserved_elastic(oracle2d_mod, R):
    """Elastic 2D (L = 0), no damping, no source: the staggered leapfrog
    conserves E^n = 1/2 sum rho |v^n|^2 + 1/2 sum sigma^{n+1/2} . S
    sigma^{n-1/2} to roundoff, with S the plane-strain compliance (normal
    block [[M, -lam], [-lam, M]] / (M^2 - lam^2), shear 1/mu).  It holds only
    if each velocity-side difference is minus the transpose of the matching
    stress-side one on the updated cells and the same per-cell moduli and
    density enter both updates; the same functional with a wrong lambda
    (M / 2) must drift, so the check is not satisfied trivially."""
    p = cases2d.params(nx=40, nz=36, b=0, R=R, L=0, lt=10, src=(20, 18), nrec=2, rec_z=18)
    vp, vs, rho, _, _ = cases2d.model(p, seed=2, q=False)
    o = oracle2d_mod.Oracle2D(p)
    o.load_model(1, vp, vs, rho)
    nnx, nnz = p.padded_shape
    rng = np.random.default_rng(4)
    v_prev = []
    for name in ("vx", "vz"):
        a = np.zeros((nnx, nnz))
        a[R:-R, R:-R] = rng.standard_normal((nnx - 2 * R, nnz - 2 * R))
        o.set_field(name, a)
        v_prev.append(a)
    r = rho.astype(np.float64)
    M = r * vp.astype(np.float64) ** 2
    mu = r * vs.astype(np.float64) ** 2

    def energy(v, sn, so, lam):
        normal = (M * (sn[0] * so[0] + sn[1] * so[1]) - lam * (sn[0] * so[1] + sn[1] * so[0])) / (M * M - lam * lam)
        return 0.5 * float(np.sum(r * (v[0] ** 2 + v[1] ** 2))) + 0.5 * float(np.sum(normal + sn[2] * so[2] / mu))

    s_prev = [np.zeros((nnx, nnz))] * 3
    E, Es = [], []
    for _ in range(40):
        o.step(1, 100000, 1)  # state: sigma^{n+1/2}, v^{n+1}
        s_new = [o.get_field(k) for k in ("txx", "tzz", "txz")]
        E.append(energy(v_prev, s_new, s_prev, M - 2.0 * mu))
        Es.append(energy(v_prev, s_new, s_prev, 0.5 * M))  # a wrong lambda
        v_prev = [o.get_field(k) for k in ("vx", "vz")]
        s_prev = s_new
    E, Es = np.array(E), np.array(Es)
    drift = np.abs(E - E[0]).max() / E[0]
    print(f"R={R}: energy {E[0]:.6e}, max relative drift over 40 steps {drift:.2e}")
    assert drift < 1e-10
    assert np.abs(Es - Es[0]).max() / abs(Es[0]) > 1e-3


def _source_snapshots(oracle2d_mod, p, seed=2):
    vp, vs, rho, qp, qs = cases2d.model(p, seed=seed, q=p.nsls > 0)
    o = oracle2d_mod.Oracle2D(p)
    o.load_model(1, vp, vs, rho, qp, qs)
    snap = {k: [] for k in ("vx", "vz", "vxp", "vzp", "tp")}
    for it in range(p.lt_forward):
        o.step(1, it, 1)
        for k in snap:
            snap[k].append(o.get_field(k))
    return o, {k: np.array(v) for k, v in snap.items()}


@pytest.mark.parametrize("avg", [0, 1])
def test_excitation_and_imaging_closed_forms_2d(oracle2d_mod, avg):
    """2D excitation and imaging against source-pass snapshots.
    Excitation: ex_time is the first step of maximum |vP| and ex_tp_time the
    first step of maximum |tp| (numpy argmax over the snapshots), ex_tp_amp
    the signed tp there, the stored vP components those of that step (vzp
    averaged over (i, k+-1), (i+-1, k) for the _new variant).  Imaging: a
    receiver field equal at every step to alpha x the source vector P plus an
    S part gamma x the source vector P, and tp_r = alpha x tp_s, images
    I_PP = alpha, I_PS = gamma, I_tp = alpha exactly at every excited cell; a
    wrong time pairing, a |ex| normalisation, a PS sign flip or swapped PP/PS
    images would not (shown on the same inputs)."""
    p = cases2d.params(nx=30, nz=26, b=8, R=4, L=0, lt=60, freq=25.0, src=(15, 6), ex_vzp_avg=avg)
    o, S = _source_snapshots(oracle2d_mod, p)
    R = 4
    inner = np.zeros(p.padded_shape, bool)
    inner[R:-R, R:-R] = True
    amp = np.sqrt(S["vxp"] ** 2 + S["vzp"] ** 2)
    et = o.get_field("ex_time").astype(int)
    want = np.where(amp.max(axis=0) > 0, np.argmax(amp, axis=0), -1)
    np.testing.assert_array_equal(et[inner], want[inner])
    ett = o.get_field("ex_tp_time").astype(int)
    want_t = np.where(np.abs(S["tp"]).max(axis=0) > 0, np.argmax(np.abs(S["tp"]), axis=0), -1)
    np.testing.assert_array_equal(ett[inner], want_t[inner])
    hit_t = inner & (ett >= 0)
    tp_at = np.take_along_axis(S["tp"], np.maximum(ett, 0)[None], axis=0)[0]
    np.testing.assert_array_equal(o.get_field("ex_tp_amp")[hit_t], tp_at[hit_t])

    def zmean(a):  # (i, k+-1), (i+-1, k) mean of the _new variant
        q = np.zeros_like(a)
        q[..., 1:-1, 1:-1] = (a[..., 1:-1, 2:] + a[..., 1:-1, :-2] + a[..., 2:, 1:-1] + a[..., :-2, 1:-1]) / 4
        return q
    vz_of = zmean if avg else (lambda a: a)
    exz = np.take_along_axis(vz_of(S["vzp"]), np.maximum(et, 0)[None], axis=0)[0]
    hit = inner & (et >= 0)
    np.testing.assert_allclose(o.get_field("ex_vzp")[hit], exz[hit], rtol=1e-12, atol=1e-300)

    alpha, gamma = 0.7, -1.3
    o.zero(5)
    for it in range(p.lt_forward):
        o.set_field("vxp", alpha * S["vxp"][it])
        o.set_field("vzp", alpha * S["vzp"][it])
        o.set_field("vx", (alpha + gamma) * S["vxp"][it])
        o.set_field("vz", (alpha + gamma) * S["vzp"][it])
        o.set_field("tp", alpha * S["tp"][it])
        o.image(it)
    exv = np.stack([o.get_field("ex_vxp"), o.get_field("ex_vzp")])
    den = (exv ** 2).sum(axis=0)
    good = hit & (den > 0)
    assert good.sum() > 0.5 * inner.sum()
    ipp, ips, itp = o.get_field("img_pp"), o.get_field("img_ps"), o.get_field("img_tp")
    np.testing.assert_allclose(ipp[good], alpha, rtol=1e-10)
    np.testing.assert_allclose(ips[good], gamma, rtol=1e-10)
    good_t = hit_t & (o.get_field("ex_tp_amp") != 0)
    np.testing.assert_allclose(itp[good_t], alpha, rtol=1e-12)
    # plausible mistakes on the same inputs
    rpp = alpha * den
    assert not np.allclose(rpp[good] / np.sqrt(den[good]), alpha, rtol=1e-3)
    assert not np.allclose(-gamma * np.ones(good.sum()), gamma)
    shifted = np.take_along_axis(S["tp"], np.clip(ett - 1, 0, None)[None], axis=0)[0]
    sel = good_t & (ett > 0)
    assert not np.allclose(alpha * shifted[sel] / tp_at[sel], alpha, rtol=1e-3)  # off-by-one step
    # ex_tp_shift records the tp excitation time s steps earlier, amplitude unchanged
    q = cases2d.params(nx=30, nz=26, b=8, R=4, L=0, lt=60, freq=25.0, src=(15, 6), ex_vzp_avg=avg, ex_tp_shift=5)
    o2, _ = _source_snapshots(oracle2d_mod, q)
    np.testing.assert_array_equal(o2.get_field("ex_tp_time")[hit_t], ett[hit_t] - 5)
    np.testing.assert_array_equal(o2.get_field("ex_tp_amp")[hit_t], o.get_field("ex_tp_amp")[hit_t])


def test_decom_discrete_identities(oracle2d_mod):
    """div/curl (decom, elastic_2D_kernel_2.cu:621-666) on staggered fields
    built from random potentials: v = (-D-z psi, D+x psi) from psi on the sxz
    nodes has zero divergence, v = (D-x phi, D+z phi) from phi on the stress
    nodes zero curl, to roundoff (the operators of different axes commute);
    a D+/D- swap in either breaks both.  Linear fields: vx = a x, vz = b z give
    div = a + b and curl = 0; vx = c z, vz = e x give curl = c - e (the
    coefficients sum to 1 over their odd weights).  decom_new (:671-791)
    scales div by vp and curl by the mean of vs over (i,k), (i,k+1), (i+1,k),
    (i+1,k+1): for vs linear in x and z that mean is vs at (i+1/2, k+1/2)."""
    R = 4
    p = cases2d.params(nx=24, nz=22, b=8, R=R, L=0, lt=4)
    nnx, nnz = p.padded_shape
    o = oracle2d_mod.Oracle2D(p)
    x = np.arange(nnx)[:, None] * p.dx * np.ones((1, nnz))
    z = np.arange(nnz)[None, :] * p.dz * np.ones((nnx, 1))
    vp0 = np.full((p.nx, p.nz), 2500.0, np.float32)
    vs_lin = (1500.0 + 3.0 * np.arange(p.nx)[:, None] + 5.0 * np.arange(p.nz)[None, :]).astype(np.float32)
    o.load_model(1, vp0, vs_lin, np.full_like(vp0, 2000.0))
    c = oracle2d_mod.oracle3d.coefficients(R)
    rng = np.random.default_rng(7)

    def dpx(f):  # D+ along x (axis 0), same shape, interior only
        out = np.zeros_like(f)
        for m in range(1, R + 1):
            out[R:-R, :] += c[m] * (np.roll(f, -m, 0) - np.roll(f, m - 1, 0))[R:-R, :]
        return out

    def dmx(f):
        out = np.zeros_like(f)
        for m in range(1, R + 1):
            out[R:-R, :] += c[m] * (np.roll(f, -(m - 1), 0) - np.roll(f, m, 0))[R:-R, :]
        return out

    def dpz(f):
        return dpx(f.T).T

    def dmz(f):
        return dmx(f.T).T

    inner = np.zeros((nnx, nnz), bool)
    inner[2 * R:-2 * R, 2 * R:-2 * R] = True
    psi = rng.standard_normal((nnx, nnz))
    o.set_field("vx", -dmz(psi) / p.dz)
    o.set_field("vz", dpx(psi) / p.dx)
    dv, cu = o.get_field("div"), o.get_field("curl")
    assert np.max(np.abs(dv[inner])) < 1e-12 * np.max(np.abs(cu[inner]))
    phi = rng.standard_normal((nnx, nnz))
    o.set_field("vx", dmx(phi) / p.dx)
    o.set_field("vz", dpz(phi) / p.dz)
    dv, cu = o.get_field("div"), o.get_field("curl")
    assert np.max(np.abs(cu[inner])) < 1e-12 * np.max(np.abs(dv[inner]))
    # a swapped operator (D+ for D- along z in the S potential) leaves a divergence
    o.set_field("vx", -dpz(psi) / p.dz)
    o.set_field("vz", dpx(psi) / p.dx)
    assert np.max(np.abs(o.get_field("div")[inner])) > 1e-3 * np.max(np.abs(o.get_field("curl")[inner]))
    a, b, cc, e = 0.3, -0.7, 0.45, 0.2
    rin = np.zeros((nnx, nnz), bool)
    rin[R:-R, R:-R] = True
    o.set_field("vx", a * (x - 0.5 * p.dx))   # vx lives at x - dx/2
    o.set_field("vz", b * (z + 0.5 * p.dz))   # vz at z + dz/2
    np.testing.assert_allclose(o.get_field("div")[rin], a + b, rtol=1e-9)
    np.testing.assert_allclose(o.get_field("curl")[rin], 0.0, atol=1e-9)
    o.set_field("vx", cc * z)
    o.set_field("vz", e * x)
    np.testing.assert_allclose(o.get_field("curl")[rin], cc - e, rtol=1e-9)
    np.testing.assert_allclose(o.get_field("div")[rin], 0.0, atol=1e-9)
    # decom_new: vp * div, vs(i+1/2, k+1/2) * curl
    o.set_field("vx", a * (x - 0.5 * p.dx) + cc * z)
    o.set_field("vz", b * (z + 0.5 * p.dz) + e * x)
    np.testing.assert_allclose(o.get_field("pdiv")[rin], 2500.0 * (a + b), rtol=1e-9)
    ii = np.clip(np.arange(nnx) - p.bl, 0, p.nx - 1)[:, None]
    kk = np.clip(np.arange(nnz) - p.bu, 0, p.nz - 1)[None, :]
    ii1 = np.clip(np.arange(nnx) + 1 - p.bl, 0, p.nx - 1)[:, None]
    kk1 = np.clip(np.arange(nnz) + 1 - p.bu, 0, p.nz - 1)[None, :]
    vs_mid = 1500.0 + 3.0 * (ii + ii1) / 2 + 5.0 * (kk + kk1) / 2
    np.testing.assert_allclose(o.get_field("scurl")[rin], vs_mid[rin] * (cc - e), rtol=1e-6)


def _window_case(L=0):
    p = cases2d.params(nx=70, nz=36, b=10, R=4, L=L, lt=260, freq=25.0, rec_dx=2, rec_z=2)
    vp, vs, rho, qp, qs = cases2d.model(p, seed=8, q=L > 0)
    bg = [cases2d.smooth2(a) for a in (vp, vs, rho)] + ([qp, qs] if L else [])
    tr = [vp, vs, rho] + ([qp, qs] if L else [])
    return p, tr, bg


def test_rtm_windows_composition(oracle2d_mod):
    """Windowed multi-shot RTM (cuda_get_partly_mode_boundary[_z1_z2] +
    cuda_sum_new_acqusition[_illum], viscoelastic_equation.cu:2606-2754), as
    restated: a window spanning the whole model equals the full-aperture RTM
    of the same shot; a shot's contribution lies inside its window (zero
    elsewhere) and equals the RTM of that window cut alone; stacking is the
    sum of the shots; the z2 cut equals migrating in the model with every
    row below z2 replaced by row z2."""
    p, tr, bg = _window_case()
    full = oracle2d_mod.rtm_windows(p, tr, bg, [(35, 3, 0, 70)])
    q = cases2d.params(nx=70, nz=36, b=10, R=4, L=0, lt=260, freq=25.0, rec_dx=2, rec_z=2, src=(35, 3))
    q.nrec, q.rec_x0 = 35, 0
    o = oracle2d_mod.Oracle2D(q)
    o.load_model(0, *tr)
    o.load_model(1, *bg)
    o.step(0, 0, q.lt)
    o.zero(7)
    o.step(1, 0, q.lt)
    o.zero(1)
    o.step(2, q.lt - 1, q.lt)
    for a, b in zip(full[:3], o.images()):
        np.testing.assert_array_equal(a, b)
    wins = [(20, 3, 6, 30), (50, 3, 30, 36)]
    each = [oracle2d_mod.rtm_windows(p, tr, bg, [w]) for w in wins]
    both = oracle2d_mod.rtm_windows(p, tr, bg, wins)
    for k in range(4):
        np.testing.assert_allclose(both[k], each[0][k] + each[1][k], rtol=1e-12, atol=1e-300)
    for (sx, sz, x0, wd), e in zip(wins, each):
        outside = np.ones(p.nx, bool)
        outside[x0:x0 + wd] = False
        assert all(np.all(e[k][outside] == 0) for k in range(4))
        assert np.count_nonzero(e[1][x0:x0 + wd]) > 0 and np.all(e[3][x0:x0 + wd] > 0)
    z2 = 20
    cut = oracle2d_mod.rtm_windows(p, tr, bg, [wins[0]], z2=z2)
    bgz = [np.array(a, copy=True) for a in bg]
    for a in bgz:
        a[:, z2 + 1:] = a[:, z2][:, None]
    ref = oracle2d_mod.rtm_windows(p, tr, bgz, [wins[0]])
    for k in range(4):
        np.testing.assert_array_equal(cut[k], ref[k])
    assert not np.allclose(cut[1], each[0][1])


# ------------------------------------------------------- reference-mode adjoint
@pytest.mark.parametrize("L,R", [(0, 4), (0, 6), (1, 4), (1, 6), (3, 4)])
def test_ref_adjoint_propagation_is_the_forward_system(oracle2d_mod, L, R):
    """The reference's adjoint velocities (adjoint_fwd_vx/vz_viscoelastic,
    viscoelastic_equation.cu:1923-2120) and memory (adjoint_fwd_memory_new,
    :2190-2196) with the strain-increment stress update, in a homogeneous
    medium without damping, are the forward viscoelastic system of a medium
    with density 1 and the same vp, vs, Q: starting both from the same
    velocities (zero stresses and memory), the adjoint velocities equal the
    forward ones and the moduli-weighted adjoint stresses
    Mp s_xx + (Mp-2Ms) s_zz + sum_l P_l r_xx + (P_l-2S_l) r_zz (and the zz, xz
    counterparts) equal the forward stresses at every step, to roundoff
    (derivation: S = C s + K r, Q_l = K_l (s - r_l/tau_l)/(1 - dt/tau_l) follow
    the forward recurrences exactly).  This pins the adjoint stencils' D+/D-
    directions and the placement of Mp, Ms, P_l, S_l against the forward
    oracle pinned above."""
    p = cases2d.params(nx=40, nz=36, b=10, R=R, L=L, lt=400, coe=0.0)
    vp = np.full((p.nx, p.nz), 2400.0, np.float32)
    vs = (vp / 1.732).astype(np.float32)
    rho = np.full_like(vp, 2.3)
    qp = np.full_like(vp, 15.0) if L else None
    qs = (0.6 * qp).astype(np.float32) if L else None
    o = oracle2d_mod.Oracle2D(p)
    o.load_model(1, vp, vs, rho, qp, qs)
    nnx, nnz = o.shape
    rng = np.random.default_rng(3)
    vx, vz = np.zeros((nnx, nnz)), np.zeros((nnx, nnz))
    vx[R:-R, R:-R] = rng.standard_normal((nnx - 2 * R, nnz - 2 * R))
    vz[R:-R, R:-R] = rng.standard_normal((nnx - 2 * R, nnz - 2 * R))
    f = oracle2d_mod.Oracle2D(p)
    f.load_model(1, vp, vs, np.ones_like(rho), qp, qs)
    f.set_field("vx", vx)
    f.set_field("vz", vz)
    f.step(1, 380, 25)  # past the wavelet: no source
    av, S = o.ref_adjoint_steps(vx, vz, 25)
    for a, name in ((av[0], "vx"), (av[1], "vz"), (S[0], "txx"), (S[1], "tzz"), (S[2], "txz")):
        b = f.get_field(name)
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b), name
    # the same check sees a 1% change of vp
    g = oracle2d_mod.Oracle2D(p)
    g.load_model(1, (1.01 * vp).astype(np.float32), vs, np.ones_like(rho), qp, qs)
    g.set_field("vx", vx)
    g.set_field("vz", vz)
    g.step(1, 380, 25)
    assert np.linalg.norm(av[0] - g.get_field("vx")) > 1e-3 * np.linalg.norm(av[0])


@pytest.mark.parametrize("L,q", [(0, 0.0), (1, 5.0), (1, 40.0), (3, 5.0)])
def test_ref_adjoint_gradient_against_born(oracle2d_mod, L, q):
    """The reference gradient (cal_gradient_in_viscoelastic_media_new,
    viscoelastic_equation.cu:2248-2295) of the adjoint driven by d = L dm
    against the Born operator (cuda_cal_dem_parameter_viscoelastic_media_new,
    :2398-2505, inversion_para 1), one parameter class at a time:
    <L dm, L dm> = (rho_r / dt) <dm_abs, g> with dm_abs = (lambda m1, mu m2,
    rho m3) and rho_r the density at the receivers.  The moduli classes agree
    to 1% (measured 0.1-0.3%: the continuous adjoint is not the exact
    transpose), so a dropped term, a wrong factor 2 in g_mu, a sign or a
    lambda/mu swap fails; Q = 5 makes the memory terms (np, ns) a large part
    of the gradient.  The density class comes out with the opposite sign: the
    reference accumulates +rv.v_t while its Born source is -drho v_t (both
    restated as shipped)."""
    lt = 360
    p = cases2d.params(nx=60, nz=48, b=16, R=4, L=L, lt=lt, coe=200.0, freq=12.0, src=(30, 4), rec_z=2)
    p.inversion_para = 1
    vp, vs, rho, qp, qs = cases2d.model(p, seed=2, contrast=False, q=L > 0)
    vp = cases2d.smooth2(vp, 20)
    vs = (vp / 1.732).astype(np.float32)
    rho = np.full_like(vp, 1800.0)
    if L:
        qp = np.full_like(vp, q)
        qs = (0.6 * qp).astype(np.float32)
    o = oracle2d_mod.Oracle2D(p)
    o.load_model(1, vp, vs, rho, qp, qs)
    rng = np.random.default_rng(5)
    w = cases2d.taper_mask(p, 6)
    base = cases2d.smooth2((rng.standard_normal((p.nx, p.nz)) * w).astype(np.float32), 3) * w
    v64, s64, r64 = (a.astype(np.float64) for a in (vp, vs, rho))
    scale = [r64 * (v64 ** 2 - 2 * s64 ** 2), r64 * s64 ** 2, r64]
    ratios = []
    for c in range(3):
        dm = np.zeros((3, p.nx, p.nz), np.float32)
        dm[c] = base
        d = o.born(dm)
        g = o.adjoint_ref(d).astype(np.float64)
        lhs = float(np.sum(d.astype(np.float64) ** 2))
        rhs = float(np.sum(dm[c] * scale[c] * g[c]))
        ratios.append(lhs / rhs * 1800.0 / (p.dt * 1e-3))
    assert abs(ratios[0] - 1) < 0.01 and abs(ratios[1] - 1) < 0.01, ratios
    assert abs(ratios[2] + 1) < 0.01, ratios


def test_ref_adjoint_mapped_is_an_approximate_transpose(oracle2d_mod):
    """adjoint_mode 1 (the reference gradient with the receiver scaling, the
    density sign and the parameter map; oracle2d.ref_adjoint_mapped) against
    the Born operator for each parameterisation: <L dm, L dm> = <dm, g> to 1%
    with dm in all three parameters at once."""
    p = cases2d.params(nx=60, nz=48, b=16, R=4, L=1, lt=360, coe=200.0, freq=12.0, src=(30, 4), rec_z=2)
    vp, vs, rho, qp, qs = cases2d.model(p, seed=2, contrast=False, q=True)
    vp = cases2d.smooth2(vp, 20)
    vs = (vp / 1.732).astype(np.float32)
    rho = (230.0 * vp.astype(np.float64) ** 0.25).astype(np.float32)
    rng = np.random.default_rng(8)
    w = cases2d.taper_mask(p, 6)
    dm = np.stack([cases2d.smooth2((rng.standard_normal((p.nx, p.nz)) * w).astype(np.float32), 3) * w
                   for _ in range(3)]).astype(np.float32)
    for para in (1, 2, 3):
        p.inversion_para = para
        o = oracle2d_mod.Oracle2D(p)
        o.load_model(1, vp, vs, rho, qp, qs)
        d = o.born(dm).astype(np.float64)
        g = oracle2d_mod.ref_adjoint_mapped(o, (vp, vs, rho), d)
        lhs, rhs = float(np.sum(d * d)), float(np.sum(dm.astype(np.float64) * g))
        assert abs(rhs / lhs - 1) < 0.01, (para, lhs, rhs)


# ------------------------------------------------ cross-correlation RTM strategies
def _xcorr_case(L=0, R=4, coe=200.0, boundary=0, nx=48, nz=40, lt=300):
    p = cases2d.params(nx=nx, nz=nz, b=14, R=R, L=L, lt=lt, coe=coe, freq=18.0, src=(nx // 2, 3), rec_z=2)
    p.boundary = boundary
    vp, vs, rho, qp, qs = cases2d.model(p, seed=4, contrast=True, q=L > 0)
    return p, (vp, vs, rho, qp, qs)


def _xcorr_records(oracle2d_mod, p, m):
    o = oracle2d_mod.Oracle2D(p)
    o.load_model(0, *m)
    o.step(0, 0, p.lt)
    return o.record()


def _smooth_bg(m):
    return tuple(cases2d.smooth2(a, 8) if a is not None else None for a in m[:3]) + tuple(m[3:])


@pytest.mark.parametrize("R,boundary", [(4, 0), (6, 0), (4, 1)])
def test_xcorr_strips_reconstruct_the_history(oracle2d_mod, R, boundary):
    """Boundary strips + reversed propagation rebuild the source wavefield of
    the model region exactly (elastic): all five images equal those of the
    stored history (damping and C-PML boundaries, both radii)."""
    p, m = _xcorr_case(R=R, boundary=boundary)
    rec = _xcorr_records(oracle2d_mod, p, m)
    o = oracle2d_mod.Oracle2D(p)
    o.load_model(1, *_smooth_bg(m))
    h = o.rtm_xcorr(rec, 0)
    s = o.rtm_xcorr(rec, 2)
    for k in range(5):
        assert np.linalg.norm(s[k] - h[k]) <= 1e-5 * np.linalg.norm(h[k]), k


def test_xcorr_reverse_is_exact_without_absorption(oracle2d_mod):
    """The reversed step (rfwd_txxzzxzpp_viscoelastic_and_memory_old: derivative
    and memory terms negated, the wavelet taken off where it was added) is the
    exact inverse of the forward step where nothing is damped: with no
    absorbing boundary at all the reversal-only strategy equals the history;
    with the damping boundary it does not (the boundary keeps absorbing)."""
    p, m = _xcorr_case(coe=0.0)
    rec = _xcorr_records(oracle2d_mod, p, m)
    o = oracle2d_mod.Oracle2D(p)
    o.load_model(1, *_smooth_bg(m))
    h, r = o.rtm_xcorr(rec, 0), o.rtm_xcorr(rec, 1)
    for k in range(4):
        assert np.linalg.norm(r[k] - h[k]) <= 1e-5 * np.linalg.norm(h[k]), k
    p2, m2 = _xcorr_case(coe=200.0)
    rec2 = _xcorr_records(oracle2d_mod, p2, m2)
    o2 = oracle2d_mod.Oracle2D(p2)
    o2.load_model(1, *_smooth_bg(m2))
    h2, r2 = o2.rtm_xcorr(rec2, 0), o2.rtm_xcorr(rec2, 1)
    assert np.linalg.norm(r2[0] - h2[0]) > 1e-2 * np.linalg.norm(h2[0])


def _window_counts(vote, interior):
    """9x9 window counts of each vote class over interior cells (numpy)."""
    out = []
    for cls in (1, 2, 3):
        a = np.pad(((vote == cls) & interior).astype(np.int64), 4)
        c = np.cumsum(np.cumsum(a, 0), 1)
        c = np.pad(c, ((1, 0), (1, 0)))
        nx, nz = vote.shape
        out.append(c[9:9 + nx, 9:9 + nz] - c[:nx, 9:9 + nz] - c[9:9 + nx, :nz] + c[:nx, :nz])
    return out


def test_xcorr_images_match_snapshot_correlation(oracle2d_mod):
    """The five images against a numpy restatement built from per-step
    snapshots of a source and a receiver propagation through the pinned
    stepping API (pins the time alignment, the model-region masking of the
    source fields, the vote classes, the 9x9 window and the tie rules)."""
    p, m = _xcorr_case(nx=36, nz=30, lt=160)
    rec = _xcorr_records(oracle2d_mod, p, m)
    bg = _smooth_bg(m)
    o = oracle2d_mod.Oracle2D(p)
    o.load_model(1, *bg)
    got = o.rtm_xcorr(rec, 0)
    R, sl = 4, (slice(p.bl, p.bl + p.nx), slice(p.bu, p.bu + p.nz))
    src = oracle2d_mod.Oracle2D(p)
    src.load_model(1, *bg)
    hist = []
    for it in range(p.lt):
        src.step(1, it, 1)
        hist.append({k: src.get_field(k) for k in ("vxp", "vzp", "div")})
    rcv = oracle2d_mod.Oracle2D(p)
    rcv.load_model(1, *bg)
    rcv.record(rec)
    nnx, nnz = rcv.shape
    interior = np.zeros((nnx, nnz), bool)
    interior[R:-R, R:-R] = True
    model = np.zeros((nnx, nnz), bool)
    model[sl] = True
    img = np.zeros((5, nnx, nnz))
    sign = np.zeros((nnx, nnz))
    for it in range(p.lt - 1, -1, -1):
        rcv.step(2, it, 1)
        f = {k: rcv.get_field(k) for k in ("vx", "vz", "vxp", "vzp", "div", "curl")}
        s = {k: np.where(model, v, 0.0) for k, v in hist[it].items()}
        rxs, rzs = f["vx"] - f["vxp"], f["vz"] - f["vzp"]
        dot = s["vxp"] * rxs + s["vzp"] * rzs
        flip = np.where(s["vzp"] >= 0, np.where(rxs >= 0, -1.0, 1.0), np.where(rxs > 0, 1.0, np.where(rxs < 0, -1.0, 0.0)))
        st = flip * dot
        vote = np.where(st > 0, 1, np.where(st < 0, 2, 3))
        c1, c2, c3 = _window_counts(vote, interior)
        new = np.where((c1 >= c2) & (c1 >= c3), 1.0, np.where((c2 > c1) & (c2 > c3), -1.0,
                                                              np.where((c3 > c1) & (c3 > c2), 0.0, sign)))
        sign = np.where(interior, new, sign)
        prod = s["div"] * f["curl"]
        img[0] += s["vxp"] * f["vxp"] + s["vzp"] * f["vzp"]
        img[1] += dot
        img[2] += s["div"] * f["div"]
        img[3] += prod
        img[4] += sign * prod
    for k in range(5):
        ref = img[k][sl]
        assert np.linalg.norm(got[k] - ref) <= 1e-5 * np.linalg.norm(ref), k


def test_xcorr_polarity_correction_makes_ps_mirror_symmetric(oracle2d_mod):
    """In a model symmetric about the source (layers and a centred anomaly,
    receivers symmetric), the div-curl PS image reverses polarity across the
    source (mirror correlation < -0.8) and the polarity-corrected image does
    not (> +0.8), as the PP images (> +0.8); a sign rule that ignored the
    side of the source would leave the corrected image antisymmetric."""
    nx, nz = 81, 60
    p = cases2d.params(nx=nx, nz=nz, b=16, R=4, L=0, lt=500, coe=200.0, freq=15.0, src=(40, 3), rec_z=2,
                       rec_dx=2, nrec=40)
    x, z = np.arange(nx)[:, None], np.arange(nz)[None, :]
    vp = (np.where(z < 30, 2000.0, 2700.0) + 300 * np.exp(-((x - 40) ** 2 + (z - 45) ** 2) / 50.0)).astype(np.float32)
    vs = (vp / 1.732).astype(np.float32)
    rho = (230 * vp.astype(np.float64) ** 0.25).astype(np.float32)
    rec = _xcorr_records(oracle2d_mod, p, (vp, vs, rho, None, None))
    o = oracle2d_mod.Oracle2D(p)
    o.load_model(1, *[cases2d.smooth2(a, 8) for a in (vp, vs, rho)])
    img = o.rtm_xcorr(rec, 0)

    def mirror_corr(a):
        a = a[:, 5:].astype(np.float64)
        return float(np.sum(a * a[::-1]) / np.sum(a * a))

    assert mirror_corr(img[0]) > 0.8 and mirror_corr(img[2]) > 0.8
    assert mirror_corr(img[3]) < -0.8, mirror_corr(img[3])
    assert mirror_corr(img[4]) > 0.8, mirror_corr(img[4])
