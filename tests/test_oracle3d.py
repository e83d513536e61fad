"""Pins the CPU oracle (oracle/oracle3d.cpp) before it is trusted as the checker.

The reference ships no tests, fixtures or golden vectors, and its 3D kernel
bodies are absent, so every check here comes from what the mathematics fixes:
coefficient identities, single-step closed forms on polynomial fields (which
pin each of the 21 staggered derivatives' direction, the dt/dx scaling and
the roles of M, lambda, mu and buoyancy), exact damping decay, mirror and
transposition symmetries, plane-wave speeds, exact P/S separation of a
transverse wave, excitation time against travel time, conservation of the
discrete leapfrog energy, and FP32 vs FP64.
A plausible mistake (a D+ where the reference has D-, a swapped lambda/M,
a dropped buoyancy, a transposed y stencil) fails at least one of them.
"""
import numpy as np
import pytest

import cases

PASS_MODEL, PASS_SOURCE, PASS_RECEIVER = 0, 1, 2
NO_SOURCE_IT = 100000  # time index past the wavelet: no source injection


# --------------------------------------------------------------------------- coefficients
def test_coefficients_radius6_reference_sum(oracle):
    c = oracle.coefficients(6)
    s = sum((2 * m - 1) * c[m] for m in range(1, 7))
    assert abs(s - 1.0) < 1e-9  # 0.99999999995 for the zzzzz:205-210 set


@pytest.mark.parametrize("R", [4, 6])
def test_coefficients_polynomial_exactness(oracle, R):
    """D+ of i^p equals p (i+1/2)^(p-1): the radius-4 Taylor set is exact up to
    degree 8 (its error term is the 9th derivative) and not beyond; the
    radius-6 optimized set is only approximately exact."""
    c = oracle.coefficients(R)
    i = 5.0
    for p in range(1, 2 * R + 2):
        d = sum(c[m] * ((i + m) ** p - (i - m + 1) ** p) for m in range(1, R + 1))
        exact = p * (i + 0.5) ** (p - 1)
        err = abs(d - exact) / abs(exact)
        if R == 4 and p <= 8:
            assert err < 1e-13, (p, err)
        if R == 4 and p == 9:
            assert err > 1e-9  # order is exactly 8
        if R == 6 and p <= 3:
            assert err < 1e-6


# --------------------------------------------------------------------------- ricker
def test_ricker_closed_form(oracle):
    w = oracle.ricker(25.0, 1.0)
    n = (len(w) - 1) // 2
    assert w[n] == 1.0
    t = (np.arange(len(w)) - n) * 1e-3
    a = (np.pi * 25.0 * t) ** 2
    np.testing.assert_allclose(w, (1 - 2 * a) * np.exp(-a), atol=2e-7)
    B = np.sqrt(6.0) / (np.pi * 25.0)
    assert n == int(np.ceil(1.35 * B / 1e-3))


# --------------------------------------------------------------------------- damping profile
def test_damping_profile_shape(oracle):
    p = cases.params(nx=20, ny=20, nz=20, b=7, coe=300.0)
    ax = oracle.damping_axis(p, 0)
    left = ax[:7]
    assert abs(left[0] - 300.0) < 1e-3          # 1-cos(pi/2) = 1 at the outer cell
    assert left[6] == 0.0                        # inner edge
    assert abs(left[3] - 300.0 * (1 - np.cos(np.pi / 4))) < 1e-3   # s = 1/2
    assert np.all(np.diff(left) < 0)
    np.testing.assert_allclose(ax[-7:], left[::-1], rtol=1e-12)   # right face mirrors the left
    assert np.all(ax[7:27] == 0)


def test_damping_exact_decay(oracle):
    """Constant stress, zero velocity: every cell decays by ((1-h)/(1+h))^n."""
    p = cases.params(nx=12, ny=10, nz=9, b=5, coe=400.0, R=4, dt=1.0)
    o = oracle.Oracle3D(p)
    vp, vs, rho = cases.homogeneous(p)
    o.load_model(1, vp, vs, rho)
    shape = o.shape
    c0 = 3.0
    o.set_field("txx", np.full(shape, c0))
    n = 7
    o.step(PASS_SOURCE, NO_SOURCE_IT, n)
    got = o.get_field("txx")
    d = (oracle.damping_axis(p, 2)[:, None, None] + oracle.damping_axis(p, 1)[None, :, None]
         + oracle.damping_axis(p, 0)[None, None, :])
    h = d * (p.dt / 1000.0) / 2
    want = c0 * ((1 - h) / (1 + h)) ** n
    R = 4
    sl = (slice(R, shape[0] - R), slice(R, shape[1] - R), slice(R, shape[2] - R))
    np.testing.assert_allclose(got[sl], want[sl], rtol=1e-13)
    assert np.all(o.get_field("vx") == 0)


# --------------------------------------------------------------------------- single-step closed forms
STRESS_CASES = [
    # (velocity comp, axis, {stress: (modulus, half-shift)}); shift +1/2 = D+, -1/2 = D-
    ("vx", 0, {"txx": ("M", +0.5), "tyy": ("L", +0.5), "tzz": ("L", +0.5), "tp": ("M", +0.5)}),
    ("vx", 1, {"txy": ("U", -0.5)}),
    ("vx", 2, {"txz": ("U", +0.5)}),
    ("vy", 0, {"txy": ("U", -0.5)}),
    ("vy", 1, {"tyy": ("M", +0.5), "txx": ("L", +0.5), "tzz": ("L", +0.5), "tp": ("M", +0.5)}),
    ("vy", 2, {"tyz": ("U", +0.5)}),
    ("vz", 0, {"txz": ("U", -0.5)}),
    ("vz", 1, {"tyz": ("U", -0.5)}),
    ("vz", 2, {"tzz": ("M", -0.5), "txx": ("L", -0.5), "tyy": ("L", -0.5), "tp": ("M", -0.5)}),
]
VEL_CASES = [
    # (stress comp, axis, {velocity: half-shift})
    ("txx", 0, {"vx": -0.5}),
    ("txy", 1, {"vx": +0.5}),
    ("txz", 2, {"vx": -0.5}),
    ("txy", 0, {"vy": +0.5}),
    ("tyy", 1, {"vy": -0.5}),
    ("tyz", 2, {"vy": -0.5}),
    ("txz", 0, {"vz": +0.5}),
    ("tyz", 1, {"vz": +0.5}),
    ("tzz", 2, {"vz": +0.5}),
    ("tp", 0, {"vxp": -0.5}),
    ("tp", 1, {"vyp": -0.5}),
    ("tp", 2, {"vzp": +0.5}),
]
ALL_STRESS = ["txx", "tyy", "tzz", "txy", "txz", "tyz", "tp"]
ALL_VEL = ["vx", "vy", "vz", "vxp", "vyp", "vzp"]


def _poly_setup(p):
    vp, vs, rho = cases.homogeneous(p, vp=2500.0, vs=1300.0, rho=1800.0)
    M = 1800.0 * 2500.0 ** 2
    U = 1800.0 * 1300.0 ** 2
    return (vp, vs, rho), {"M": M, "U": U, "L": M - 2 * U}, 1.0 / 1800.0


def _coord(shape, axis, d):
    idx = np.indices(shape)[2 - axis].astype(np.float64)  # array axes are (z, y, x)
    return idx, idx * d


def closed_form_stress(engine_factory, R, d=10.0):
    p = cases.params(nx=18, ny=17, nz=16, b=0, R=R, dt=1.0, d=d, src=(9, 8, 8))
    model, mod, _ = _poly_setup(p)
    delta = p.dt / 1000.0
    worst = 0.0
    for comp, axis, expect in STRESS_CASES:
        e = engine_factory(p)
        e.load_model(1, *model)
        shape = tuple(p.padded_shape)
        idx, x = _coord(shape, axis, d)
        e.set_field(comp, x ** 2)
        e.step(PASS_SOURCE, NO_SOURCE_IT, 1)
        sl = (slice(R, shape[0] - R), slice(R, shape[1] - R), slice(R, shape[2] - R))
        for s in ALL_STRESS:
            got = np.asarray(e.get_field(s), np.float64)[sl]
            if s in expect:
                key, shift = expect[s]
                want = mod[key] * delta * 2.0 * (idx + shift) * d
                want = want[sl]
                err = np.max(np.abs(got - want)) / np.max(np.abs(want))
            else:
                err = np.max(np.abs(got)) / (mod["M"] * delta * d * shape[0])
            worst = max(worst, err)
            yield comp, axis, s, err


def closed_form_velocity(engine_factory, R, d=10.0):
    p = cases.params(nx=18, ny=17, nz=16, b=0, R=R, dt=1.0, d=d, src=(9, 8, 8))
    model, _, bo = _poly_setup(p)
    delta = p.dt / 1000.0
    for comp, axis, expect in VEL_CASES:
        e = engine_factory(p)
        e.load_model(1, *model)
        shape = tuple(p.padded_shape)
        idx, x = _coord(shape, axis, d)
        e.set_field(comp, x ** 2)
        e.step(PASS_SOURCE, NO_SOURCE_IT, 1)
        sl = (slice(R, shape[0] - R), slice(R, shape[1] - R), slice(R, shape[2] - R))
        for v in ALL_VEL:
            got = np.asarray(e.get_field(v), np.float64)[sl]
            if v in expect:
                want = (bo * delta * 2.0 * (idx + expect[v]) * d)[sl]
                err = np.max(np.abs(got - want)) / np.max(np.abs(want))
            else:
                err = np.max(np.abs(got)) / (bo * delta * d * shape[0])
            yield comp, axis, v, err


def test_closed_form_stress_derivatives_radius4(oracle):
    for comp, axis, s, err in closed_form_stress(lambda p: oracle.Oracle3D(p), 4):
        assert err < 1e-12, (comp, axis, s, err)


def test_closed_form_velocity_derivatives_radius4(oracle):
    for comp, axis, v, err in closed_form_velocity(lambda p: oracle.Oracle3D(p), 4):
        assert err < 1e-12, (comp, axis, v, err)


def test_closed_form_radius6_reference_coefficients(oracle):
    """Radius 6 (reference coefficients) is exact on quadratics up to its 5e-11 sum defect."""
    for comp, axis, s, err in closed_form_stress(lambda p: oracle.Oracle3D(p), 6):
        assert err < 1e-9, (comp, axis, s, err)
    for comp, axis, v, err in closed_form_velocity(lambda p: oracle.Oracle3D(p), 6):
        assert err < 1e-9, (comp, axis, v, err)


# --------------------------------------------------------------------------- symmetries
def _centered(R, n=37, b=6, steps=14):
    nn = n  # model size; padded = n + 2b (odd)
    p = cases.params(nx=nn, ny=nn, nz=nn, b=b, R=R, dt=1.0, d=10.0, freq=30.0,
                     src=(nn // 2, nn // 2, nn // 2), lt=steps, coe=200.0)
    return p


@pytest.mark.parametrize("R", [4, 6])
def test_mirror_and_transpose_symmetry(oracle, R):
    p = _centered(R)
    o = oracle.Oracle3D(p)
    o.load_model(0, *cases.homogeneous(p, vp=2000.0))
    o.step(PASS_MODEL, 0, p.lt)
    f = {k: o.get_field(k) for k in ["vx", "vy", "vz", "txx", "tyy", "tzz", "txy", "txz", "tyz", "tp"]}
    s = p.padded_shape[2] // 2
    scale = max(np.max(np.abs(v)) for v in f.values() if v is not None)
    tol = 1e-9 * scale

    def mirror_x(a, stag):   # integer nodes: i -> 2s-i ; staggered (i-1/2): i -> 2s-i+1
        if stag:
            out = np.zeros_like(a)
            out[..., 1:] = a[..., :0:-1]
            return out
        return a[..., ::-1]

    # x-mirror: sxx/syy/szz/tp/syz even at integer x; vx odd at i-1/2; sxy/sxz odd at i-1/2; vy/vz even
    for k in ["txx", "tyy", "tzz", "tp", "tyz", "vy", "vz"]:
        assert np.max(np.abs(f[k] - mirror_x(f[k], False))) < tol, k
    for k in ["vx", "txy", "txz"]:
        m = mirror_x(f[k], True)
        core = (slice(None), slice(None), slice(2, -2))
        assert np.max(np.abs((f[k] + m)[core])) < tol, k
    # transposition x <-> y
    T = lambda a: np.swapaxes(a, 1, 2)
    pairs = [("txx", "tyy"), ("tzz", "tzz"), ("tp", "tp"), ("txy", "txy"), ("txz", "tyz"), ("vx", "vy"), ("vz", "vz")]
    for a, b in pairs:
        assert np.max(np.abs(f[a] - T(f[b]))) < tol, (a, b)
    assert s == p.bl + p.shot_start_x


# --------------------------------------------------------------------------- plane waves
def _plane_wave_peak_shift(oracle, comp, c_expected, R=4):
    nx, ny, nz = 120, 52, 52
    p = cases.params(nx=nx, ny=ny, nz=nz, b=0, R=R, dt=1.0, d=10.0)
    p.src = None
    o = oracle.Oracle3D(p)
    vp, vs, rho = cases.homogeneous(p, vp=3000.0, vs=1700.0, rho=2100.0)
    o.load_model(1, vp, vs, rho)
    shape = tuple(p.padded_shape)
    i = np.arange(shape[2], dtype=np.float64)
    x0 = 60.0
    g = np.exp(-((i - x0) / 4.0) ** 2)
    f0 = np.zeros(shape)
    f0[R:-R, R:-R, :] = g  # the R-cell margins are never updated: keep them zero
    o.set_field(comp, f0)
    n = 50
    o.step(PASS_SOURCE, NO_SOURCE_IT, n)
    line = o.get_field(comp)[shape[0] // 2, shape[1] // 2, :]
    right = line.copy()
    right[: int(x0)] = -np.inf
    k = int(np.argmax(right))
    y0, y1, y2 = line[k - 1], line[k], line[k + 1]
    frac = 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)
    shift = k + frac - x0
    # stress starts at t = -dt/2 (leapfrog), which advances the right-going
    # half of the split pulse by half a step: peak at x0 + c (n + 1/2) dt
    expected = c_expected * (n + 0.5) * (p.dt / 1000.0) / p.dx
    return shift, expected, o, line


def test_plane_p_wave_speed(oracle):
    shift, expected, _, _ = _plane_wave_peak_shift(oracle, "vx", 3000.0)
    assert abs(shift - expected) < 0.05, (shift, expected)


def test_plane_s_wave_speed(oracle):
    shift, expected, o, _ = _plane_wave_peak_shift(oracle, "vy", 1700.0)
    assert abs(shift - expected) < 0.05, (shift, expected)


def test_transverse_wave_has_no_p_component(oracle):
    """vy = g(x) w(y): divergence-free where w = 1, so tp and vector P stay at
    roundoff in the core until P waves from the taper (30 cells away, 12 cells
    of travel in 40 steps) could arrive.  The z-margins only launch S waves."""
    p = cases.params(nx=100, ny=110, nz=12, b=0, R=4, dt=1.0, d=10.0, src=(50, 55, 6))
    o = oracle.Oracle3D(p)
    o.load_model(1, *cases.homogeneous(p, vp=3000.0, vs=1700.0, rho=2100.0))
    shape = tuple(p.padded_shape)
    i = np.arange(shape[2], dtype=np.float64)
    j = np.arange(shape[1], dtype=np.float64)
    g = np.exp(-((i - 50.0) / 4.0) ** 2)
    r = np.abs(j - shape[1] // 2)
    w = np.where(r < 30, 1.0, np.where(r < 45, 0.5 * (1 + np.cos(np.pi * (r - 30) / 15)), 0.0))
    f0 = np.zeros(shape)
    f0[4:-4] = w[:, None] * g[None, :]
    o.set_field("vy", f0)
    o.step(PASS_SOURCE, NO_SOURCE_IT, 40)
    cy = shape[1] // 2
    core = (slice(None), slice(cy - 5, cy + 6), slice(None))
    ref_t = np.max(np.abs(o.get_field("txy")[core]))
    ref_v = np.max(np.abs(o.get_field("vy")[core]))
    assert ref_t > 0 and ref_v > 0
    # relative 1e-8: the smooth taper's P precursor is ~5e-10; a wrong stagger gives O(1)
    assert np.max(np.abs(o.get_field("tp")[core])) < 1e-8 * ref_t
    for k in ["vxp", "vyp", "vzp"]:
        assert np.max(np.abs(o.get_field(k)[core])) < 1e-8 * ref_v, k


# --------------------------------------------------------------------------- P/S split, excitation
def test_explosive_source_is_p_only_far_field(oracle):
    p = _centered(4, n=45, b=8, steps=60)
    p.lt_forward = p.lt
    o = oracle.Oracle3D(p)
    o.load_model(1, *cases.homogeneous(p, vp=2000.0))
    o.step(PASS_SOURCE, 0, p.lt)
    shape = np.array(p.padded_shape)
    zz, yy, xx = np.indices(tuple(shape))
    c = shape // 2
    r = np.sqrt((zz - c[0]) ** 2 + (yy - c[1]) ** 2 + (xx - c[2]) ** 2)
    far = r > 6
    v2 = sum(o.get_field(k) ** 2 for k in ["vx", "vy", "vz"])
    s2 = sum(o.get_field(k) ** 2 for k in ["vxs", "vys", "vzs"])
    ratio = s2[far].sum() / v2[far].sum()
    assert ratio < 1e-4, ratio


def test_excitation_time_follows_travel_time(oracle):
    p = _centered(4, n=45, b=8, steps=130)
    p.lt_forward = p.lt
    vp0 = 2000.0
    o = oracle.Oracle3D(p)
    o.load_model(1, *cases.homogeneous(p, vp=vp0))
    o.step(PASS_SOURCE, 0, p.lt)
    t = o.get_field("ex_time")
    shape = np.array(p.padded_shape)
    c = shape // 2
    expected = p.dx / (vp0 * p.dt / 1000.0)  # steps per cell
    # per-cell increments along x, y and z: the arrival moves dx/vp per cell;
    # |vP| has two lobes and the larger one can switch once with distance
    # (near-field term), which shows up as a single larger jump
    for line in (t[c[0], c[1], c[2] + 3: c[2] + 14], t[c[0], c[1] + 3: c[1] + 14, c[2]],
                 t[c[0] + 3: c[0] + 14, c[1], c[2]]):
        d = np.diff(line)
        assert np.all(line >= 0)
        assert np.median(d) == expected, d
        assert np.sum(np.abs(d - expected) <= 1) >= len(d) - 1, d


# --------------------------------------------------------------------------- precision, imaging
def test_fp32_instantiation_tracks_fp64(oracle):
    p = cases.params(nx=30, ny=26, nz=24, b=6, R=4, lt=150, freq=25.0)
    vp, vs, rho = cases.layered(p)
    out = {}
    for prec in (64, 32):
        o = oracle.Oracle3D(p, prec)
        o.load_model(0, vp, vs, rho)
        o.step(PASS_MODEL, 0, p.lt)
        out[prec] = {k: o.get_field(k) for k in ["vx", "vz", "txx", "txz", "tp", "vxp", "vzp"]}
    for k in out[64]:
        err = cases.rel_l2(out[32][k], out[64][k])
        # vector P integrates grad(tp) every step and tp keeps a static
        # near-source residue (the source enters sxx/syy/szz, not tp, as in
        # add_source_3D), so its FP32 rounding grows linearly in time
        assert err < (1e-4 if k in ("vxp", "vzp") else 1e-5), (k, err)


def test_rtm_images_reflector_depth(oracle):
    """PP image of a flat reflector peaks at the interface depth below the shot."""
    nz, zi = 48, 30
    p = cases.params(nx=40, ny=40, nz=nz, b=10, R=4, lt=420, dt=1.0, freq=25.0, coe=250.0,
                     src=(20, 20, 3), rec=(19, 19, 2, 2, 2, 2, 2))
    p.lt_forward = p.lt
    shape = (p.nx, p.ny, p.nz)
    z = np.arange(nz)[None, None, :]
    vp = np.where(z < zi, 2000.0, 3000.0) * np.ones(shape)
    rho = np.where(z < zi, 2000.0, 2400.0) * np.ones(shape)
    vp = vp.astype(np.float32)
    rho = rho.astype(np.float32)
    vs = (vp / 1.732).astype(np.float32)
    o = oracle.Oracle3D(p)
    o.load_model(0, vp, vs, rho)
    s = np.full(shape, 2000.0, np.float32)
    o.load_model(1, s, (s / 1.732).astype(np.float32), np.full(shape, 2000.0, np.float32))
    o.rtm_shot()
    tp, pp, ps = o.images()
    col = np.abs(pp[18:23, 18:23, 10:]).sum(axis=(0, 1))
    k = int(np.argmax(col)) + 10
    assert abs(k - zi) <= 2, (k, zi)


def test_xcorr_images_reflector_depth(oracle):
    """Cross-correlation PP image (full source history) of a flat reflector:
    below the shot its energy peaks at the interface depth.  Pins the zero-lag
    correlation restatement (a wrong time pairing of source and receiver
    fields moves or smears the peak)."""
    nz, zi = 36, 22
    p = cases.params(nx=28, ny=28, nz=nz, b=8, R=4, lt=260, dt=1.0, freq=25.0, coe=250.0,
                     src=(14, 14, 3), rec=(13, 13, 1, 1, 2, 2, 2))
    p.lt_forward = p.lt
    shape = (p.nx, p.ny, p.nz)
    z = np.arange(nz)[None, None, :]
    vp = (np.where(z < zi, 2000.0, 3000.0) * np.ones(shape)).astype(np.float32)
    rho = (np.where(z < zi, 2000.0, 2400.0) * np.ones(shape)).astype(np.float32)
    vs = (vp / 1.732).astype(np.float32)
    o = oracle.Oracle3D(p)
    o.load_model(0, vp, vs, rho)
    s = np.full(shape, 2000.0, np.float32)
    o.load_model(1, s, (s / 1.732).astype(np.float32), np.full(shape, 2000.0, np.float32))
    tp, pp, ps = o.xcorr_shot()
    col = np.abs(pp[12:17, 12:17, 12:]).sum(axis=(0, 1))
    k = int(np.argmax(col)) + 12
    assert abs(k - zi) <= 2, (k, zi)
    assert np.isfinite(ps).all() and np.abs(ps).max() > 0


def test_direct_wave_mute_window(oracle):
    """The oracle's direct-wave mute: zero-offset cut = wavelet length + margin,
    the cut grows with offset, it covers the peak of the direct arrival in a
    homogeneous medium, and samples at and after the cut are untouched."""
    p = cases.params(nx=24, ny=20, nz=18, b=8, R=4, lt=160, dt=1.0, freq=25.0, src=(12, 10, 2),
                     rec=(6, 1, 0, 10, 4, 1, 2))
    p.cut_direct_wave = 3
    vp, vs, rho = cases.homogeneous(p)
    o = oracle.Oracle3D(p, 64)
    o.load_model(0, vp, vs, rho)
    o.step(0, 0, p.lt)
    rec = o.get_record()
    m = oracle.mute_direct(p, rec, vp)
    wl = len(oracle.ricker(p.freq, p.dt))
    cut = [int(np.argmax(np.any(m[:, :, 0, r] != 0, axis=0))) for r in range(p.receiver_num_x)]
    offs = [abs(p.receiver_start_x + r * p.receiver_interval_x - 12) for r in range(p.receiver_num_x)]
    zero = offs.index(0)
    assert cut[zero] == wl + 3, (cut, wl)
    order = np.argsort(offs)
    assert all(cut[order[i]] <= cut[order[i + 1]] for i in range(len(order) - 1)), (offs, cut)
    for r in range(p.receiver_num_x):
        peak = int(np.argmax(np.abs(rec[:, :, 0, r]).max(axis=0)))
        assert peak < cut[r], (r, peak, cut[r])
        np.testing.assert_array_equal(m[:, cut[r]:, 0, r], rec[:, cut[r]:, 0, r])
        assert not np.any(m[:, :cut[r], 0, r])
    q = cases.params(nx=24, ny=20, nz=18, b=8, R=4, lt=160, dt=1.0, freq=25.0, src=(12, 10, 2))
    np.testing.assert_array_equal(oracle.mute_direct(q, rec[:, :, :1, :1], vp), rec[:, :, :1, :1])  # off by default


def _leapfrog_energy(v, s_new, s_old, rho, lam, mu):
    """E^n = 1/2 sum rho |v^n|^2 + 1/2 sum sigma^{n+1/2} . S sigma^{n-1/2},
    S = C^-1 of the isotropic stiffness acting on (exx, eyy, ezz) and on
    engineering shear strains: normal block (1/2mu)(I - lam/(2mu+3lam) 11^T),
    shear 1/mu."""
    tr_new = s_new[0] + s_new[1] + s_new[2]
    tr_old = s_old[0] + s_old[1] + s_old[2]
    normal = (s_new[0] * s_old[0] + s_new[1] * s_old[1] + s_new[2] * s_old[2]
              - lam / (2.0 * mu + 3.0 * lam) * tr_new * tr_old) / (2.0 * mu)
    shear = (s_new[3] * s_old[3] + s_new[4] * s_old[4] + s_new[5] * s_old[5]) / mu
    kinetic = rho * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2)
    return 0.5 * float(np.sum(kinetic)) + 0.5 * float(np.sum(normal + shear))


@pytest.mark.parametrize("R", [4, 6])
def test_discrete_energy_is_conserved(oracle, R):
    """Without damping or source the staggered leapfrog conserves the discrete
    energy above to roundoff: on the updated cells (the outer R cells stay
    zero) every velocity-side difference is minus the transpose of the
    matching stress-side one, and the stress update is C times the strain
    increment with the same per-cell M, lambda, mu, while the velocity update
    divides by the same per-cell density.  A D+ where a D- belongs, a swapped
    lambda/M, a buoyancy or modulus from the wrong cell, or a different dt/dx
    on the two sides of a derivative pair breaks it.  Heterogeneous model,
    random initial velocity, 30 steps.  The same functional with lambda and M
    swapped must drift, so the check is not satisfied trivially."""
    p = cases.params(nx=24, ny=22, nz=20, b=0, R=R, dt=1.0, d=10.0, lt=10, src=(12, 11, 8))
    vp, vs, rho = cases.layered(p, seed=3)
    o = oracle.Oracle3D(p)
    o.load_model(1, vp, vs, rho)
    shape = tuple(p.padded_shape)
    rng = np.random.default_rng(1)
    v0 = []
    for name in ("vx", "vy", "vz"):
        a = np.zeros(shape)
        a[R:-R, R:-R, R:-R] = rng.standard_normal((shape[0] - 2 * R, shape[1] - 2 * R, shape[2] - 2 * R))
        o.set_field(name, a)
        v0.append(a)
    # the model as the oracle holds it: float32 inputs, field layout (z, y, x)
    f64 = lambda a: a.astype(np.float64).transpose(2, 1, 0)
    r, M, mu = f64(rho), f64(rho) * f64(vp) ** 2, f64(rho) * f64(vs) ** 2
    sig_names = ("txx", "tyy", "tzz", "txy", "txz", "tyz")
    zero = [np.zeros(shape)] * 6
    v_prev, s_prev = v0, zero
    E, E_swapped = [], []
    for _ in range(30):
        o.step(PASS_SOURCE, NO_SOURCE_IT, 1)  # state: sigma^{n+1/2}, v^{n+1}
        s_new = [o.get_field(k) for k in sig_names]
        E.append(_leapfrog_energy(v_prev, s_new, s_prev, r, M - 2.0 * mu, mu))
        E_swapped.append(_leapfrog_energy(v_prev, s_new, s_prev, r, M, mu))  # lambda := M
        v_prev = [o.get_field(k) for k in ("vx", "vy", "vz")]
        s_prev = s_new
    E = np.array(E)
    drift = np.abs(E - E[0]).max() / E[0]
    print(f"R={R}: energy {E[0]:.6e}, max relative drift over 30 steps {drift:.2e}")
    assert drift < 1e-10
    Es = np.array(E_swapped)
    assert np.abs(Es - Es[0]).max() / abs(Es[0]) > 1e-3


def _excited_state(oracle, p, vp, vs, rho):
    o = oracle.Oracle3D(p)
    o.load_model(1, vp, vs, rho)
    o.step(1, 0, p.lt_forward)
    ex = {k: o.get_field(k) for k in ("ex_time", "ex_tp", "ex_vxp", "ex_vyp", "ex_vzp")}
    return o, ex


def test_imaging_condition_closed_forms(oracle):
    """The excitation-amplitude imaging condition on prescribed receiver
    fields.  With the receiver vector P equal to alpha times the source vector
    P of each cell's excitation step and the receiver S part equal to gamma
    times it, I_PP = alpha and I_PS = gamma exactly; with tp_r = alpha ex_tp,
    I_tp = alpha; an S part orthogonal to ex_vP images 0; cells never excited
    image 0.  A |ex_vP| (not |ex_vP|^2) normalisation, a PS sign flip (vP - v
    for vS) or swapped PP/PS images fail these checks (shown on the same
    inputs below)."""
    p = cases.params(nx=20, ny=18, nz=16, b=6, R=4, lt=40, dt=1.0, freq=25.0, src=(10, 9, 5))
    p.lt_forward = p.lt
    vp, vs, rho = cases.layered(p, seed=3)
    o, ex = _excited_state(oracle, p, vp, vs, rho)
    et = ex["ex_time"]
    exv = np.stack([ex["ex_vxp"], ex["ex_vyp"], ex["ex_vzp"]])
    den = (exv ** 2).sum(axis=0)
    R = 4
    inner = np.zeros(et.shape, bool)
    inner[R:-R, R:-R, R:-R] = True
    hit = inner & (et >= 0) & (den > 0) & (ex["ex_tp"] != 0)
    assert hit.sum() > 0.5 * inner.sum()
    alpha, gamma = 0.7, -1.3
    names = ("vx", "vy", "vz")

    def image_with(vp_r, vs_r, tp_r):
        o.zero(wavefield=True, images=True)
        for c in range(3):
            o.set_field(names[c], vp_r[c] + vs_r[c])
            o.set_field("v" + "xyz"[c] + "p", vp_r[c])
        o.set_field("tp", tp_r)
        for it in range(p.lt_forward):
            o.image(it)
        return o.get_field("img_tp"), o.get_field("img_pp"), o.get_field("img_ps")

    itp, ipp, ips = image_with(alpha * exv, gamma * exv, alpha * ex["ex_tp"])
    np.testing.assert_allclose(ipp[hit], alpha, rtol=1e-12)
    np.testing.assert_allclose(ips[hit], gamma, rtol=1e-12)
    np.testing.assert_allclose(itp[hit], alpha, rtol=1e-12)
    never = inner & (et < 0)
    assert np.all(ipp[never] == 0) and np.all(ips[never] == 0) and np.all(itp[never] == 0)
    # S part orthogonal to ex_vP: (ex_vP x e) for a fixed unit e
    e = np.array([0.3, -0.5, 0.8])[:, None, None, None]
    orth = np.cross(exv, np.broadcast_to(e, exv.shape), axis=0) * 1e3
    _, ipp2, ips2 = image_with(alpha * exv, orth, alpha * ex["ex_tp"])
    scale = np.sqrt((orth ** 2).sum(axis=0) / np.where(den > 0, den, 1))
    np.testing.assert_allclose(ipp2[hit], alpha, rtol=1e-12)
    assert np.max(np.abs(ips2[hit]) / (1 + scale[hit])) < 1e-12
    # plausible mistakes give other values on the same inputs
    rpp = (exv * (alpha * exv)).sum(axis=0)
    rps = (exv * (gamma * exv)).sum(axis=0)
    wrong_norm = rpp[hit] / np.sqrt(den[hit])          # |ex_vP| instead of |ex_vP|^2
    assert not np.allclose(wrong_norm, alpha, rtol=1e-3)
    assert not np.allclose(-rps[hit] / den[hit], gamma, rtol=1e-3)   # vP - v for vS
    assert not np.allclose(rps[hit] / den[hit], alpha, rtol=1e-3)    # PP and PS swapped
    o.close()


def test_poynting_vectors_closed_forms(oracle):
    """Receiver Poynting vectors at the imaging step (poynting = 1, the 3D form
    of sum_poynting) on the local state of plane waves with unit direction k,
    prescribed at every cell and imaged at each cell's excitation step:
      P wave (strain f k k^T): tp = M f, v = vP = -c_p f k, vS = 0:
        rp = M c_p f^2 k, rs = 0;
      S wave (polarisation e, e.k = 0): sigma = mu g (e k^T + k e^T), tp = 0,
        v = vS = -c_s g e:  rp = 0, rs = mu c_s g^2 k;
      both together: rp as the P wave's, rs = mu c_s g^2 k + 2 mu f vS (the
        P stress minus tp I, 2 mu f (k k^T - I), acting on vS).
    The source vector is -ex_tp ex_vP of the excitation state.  A missing
    '- tp I', a transposed stress index or a sign flip changes these values
    (the first is checked below on the same inputs)."""
    p = cases.params(nx=16, ny=14, nz=12, b=6, R=4, lt=30, dt=1.0, freq=25.0, src=(8, 7, 4))
    p.lt_forward = p.lt
    p.poynting = 1
    vp, vs, rho = cases.layered(p, seed=5)
    o, ex = _excited_state(oracle, p, vp, vs, rho)
    et = ex["ex_time"]
    hit = et >= 0
    assert hit.sum() > 100
    rng = np.random.default_rng(4)
    lam, mu = 2.0e9, 1.5e9
    M = lam + 2 * mu
    cp, cs = 3000.0, 1800.0
    names = ("vx", "vy", "vz")

    def prescribed(f, g, k, e):
        kk = np.outer(k, k)
        eps_p = f * kk
        sig = lam * np.trace(eps_p) * np.eye(3) + 2 * mu * eps_p + mu * g * (np.outer(e, k) + np.outer(k, e))
        vP = -cp * f * k
        vS = -cs * g * e
        o.zero(wavefield=True, images=True)
        shape = et.shape
        for c in range(3):
            o.set_field(names[c], np.full(shape, vP[c] + vS[c]))
            o.set_field("v" + "xyz"[c] + "p", np.full(shape, vP[c]))
        for name, (a, b) in zip(("txx", "tyy", "tzz", "txy", "txz", "tyz"),
                                ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))):
            o.set_field(name, np.full(shape, sig[a, b]))
        o.set_field("tp", np.full(shape, M * f))
        for it in range(p.lt_forward):
            o.image(it)
        rp = np.stack([o.get_field(n)[hit] for n in ("poyn_rpx", "poyn_rpy", "poyn_rpz")])
        rs = np.stack([o.get_field(n)[hit] for n in ("poyn_rsx", "poyn_rsy", "poyn_rsz")])
        return rp, rs, sig, vS

    for _ in range(3):
        k = rng.standard_normal(3)
        k /= np.linalg.norm(k)
        e = np.cross(k, rng.standard_normal(3))
        e /= np.linalg.norm(e)
        f, g = 1e-6 * rng.uniform(0.5, 2), 1e-6 * rng.uniform(0.5, 2)
        rp, rs, _, _ = prescribed(f, 0.0, k, e)
        np.testing.assert_allclose(rp, np.broadcast_to((M * cp * f * f * k)[:, None], rp.shape), rtol=1e-10)
        assert np.max(np.abs(rs)) <= 1e-12 * M * cp * f * f
        rp, rs, _, _ = prescribed(0.0, g, k, e)
        assert np.max(np.abs(rp)) == 0.0
        np.testing.assert_allclose(rs, np.broadcast_to((mu * cs * g * g * k)[:, None], rs.shape),
                                   rtol=1e-9, atol=1e-12 * mu * cs * g * g)
        rp, rs, sig, vS = prescribed(f, g, k, e)
        want = mu * cs * g * g * k + 2 * mu * f * vS
        np.testing.assert_allclose(rp, np.broadcast_to((M * cp * f * f * k)[:, None], rp.shape), rtol=1e-10)
        np.testing.assert_allclose(rs, np.broadcast_to(want[:, None], rs.shape), rtol=1e-9,
                                   atol=1e-12 * np.abs(want).max())
        wrong = -sig @ vS  # without '- tp I'
        assert not np.allclose(wrong, want, rtol=1e-3)
    # cells never imaged keep zero vectors; the source vector is -ex_tp ex_vP
    assert np.all(o.get_field("poyn_rpx")[~hit] == 0)
    for c, n in enumerate(("poyn_sx", "poyn_sy", "poyn_sz")):
        np.testing.assert_array_equal(o.get_field(n), -ex["ex_tp"] * ex["ex_v" + "xyz"[c] + "p"])
    o.close()


def test_div_curl_discrete_identities_3d(oracle):
    """3D divergence / curl decomposition (the 3D form of decom,
    elastic_2D_kernel_2.cu:621-666).  Staggered fields from random
    potentials: v = curl A with A on the edge nodes (Ax at syz, Ay at sxz, Az
    at sxy) has zero divergence; v = grad phi with phi on the stress nodes has
    zero curl, to roundoff (operators of different axes commute; a D+/D-
    mix-up breaks it, shown below).  Linear fields give div = a + b + e and
    curl = (f - g, h - l, m - r) exactly (coefficient sum 1 over the odd
    weights)."""
    R = 4
    p = cases.params(nx=14, ny=13, nz=12, b=6, R=R, lt=2)
    o = oracle.Oracle3D(p)
    nnz, nny, nnx = p.padded_shape
    c = oracle.coefficients(R)
    rng = np.random.default_rng(11)

    def D(f, axis, plus):
        out = np.zeros_like(f)
        for m in range(1, R + 1):
            if plus:
                out += c[m] * (np.roll(f, -m, axis) - np.roll(f, m - 1, axis))
            else:
                out += c[m] * (np.roll(f, -(m - 1), axis) - np.roll(f, m, axis))
        return out
    X, Y, Z = 2, 1, 0  # array axes of (z, y, x) volumes
    inner = np.zeros((nnz, nny, nnx), bool)
    inner[2 * R:-2 * R, 2 * R:-2 * R, 2 * R:-2 * R] = True
    ax, ay, az = (rng.standard_normal((nnz, nny, nnx)) for _ in range(3))
    vx = D(az, Y, True) / p.dy - D(ay, Z, False) / p.dz
    vy = D(ax, Z, False) / p.dz - D(az, X, True) / p.dx
    vz = D(ay, X, True) / p.dx - D(ax, Y, True) / p.dy
    for name, a in (("vx", vx), ("vy", vy), ("vz", vz)):
        o.set_field(name, a)
    dv = o.get_field("div")
    cmax = max(np.max(np.abs(o.get_field(k)[inner])) for k in ("curlx", "curly", "curlz"))
    assert np.max(np.abs(dv[inner])) < 1e-12 * cmax
    phi = rng.standard_normal((nnz, nny, nnx))
    o.set_field("vx", D(phi, X, False) / p.dx)
    o.set_field("vy", D(phi, Y, False) / p.dy)
    o.set_field("vz", D(phi, Z, True) / p.dz)
    dmax = np.max(np.abs(o.get_field("div")[inner]))
    for k in ("curlx", "curly", "curlz"):
        assert np.max(np.abs(o.get_field(k)[inner])) < 1e-12 * dmax, k
    o.set_field("vz", D(phi, Z, False) / p.dz)   # a D+ / D- mix-up in z
    assert np.max(np.abs(o.get_field("curlx")[inner])) > 1e-3 * dmax
    # linear fields at their staggered positions
    k3, j3, i3 = np.meshgrid(np.arange(nnz), np.arange(nny), np.arange(nnx), indexing="ij")
    x, y, z = i3 * p.dx, j3 * p.dy, k3 * p.dz
    a, b, e, f, g, h, l, m, r = 0.3, -0.2, 0.5, 0.7, -0.4, 0.25, 0.6, -0.35, 0.15
    o.set_field("vx", a * (x - p.dx / 2) + h * (z) + r * (y))           # vx at (i-1/2, j, k)
    o.set_field("vy", b * (y - p.dy / 2) + g * (z) + m * (x))           # vy at (i, j-1/2, k)
    o.set_field("vz", e * (z + p.dz / 2) + f * (y) + l * (x))           # vz at (i, j, k+1/2)
    rin = np.zeros((nnz, nny, nnx), bool)
    rin[R:-R, R:-R, R:-R] = True
    np.testing.assert_allclose(o.get_field("div")[rin], a + b + e, rtol=1e-9)
    np.testing.assert_allclose(o.get_field("curlx")[rin], f - g, rtol=1e-9)
    np.testing.assert_allclose(o.get_field("curly")[rin], h - l, rtol=1e-9)
    np.testing.assert_allclose(o.get_field("curlz")[rin], m - r, rtol=1e-9)
    assert np.all(o.get_field("div")[~rin] == 0)
    o.close()


def test_decimated_xcorr_history_converges(oracle):
    """The decimated cross-correlation (every third step, weighted by 3) is a
    coarser quadrature of the same zero-lag sum: at a wavelet sampled ~40
    steps per period it stays within 1e-3 of the full sum (a weight of 1 in
    place of 3 would miss by 2/3)."""
    import test_gpu_reconstruct as T
    p, true, sm = T._case(1)
    o = oracle.Oracle3D(p, 64)
    o.load_model(0, *true)
    o.load_model(1, *sm)
    full = o.xcorr_shot(decim=1)
    dec = o.xcorr_shot(decim=3)
    for a, b in zip(dec, full):
        assert np.linalg.norm(a - b) <= 1e-3 * np.linalg.norm(b)
