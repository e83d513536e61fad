"""GPU parity of the 3D elastic RTM hot path (liberm.so, gfx950) against the
FP64 oracle, through the C ABI.

Tolerances (FP32 kernels vs FP64 oracle):
  particle velocity, stress, tp, receiver record: rel L2 <= 1e-5;
  vector-P fields: rel L2 <= 1e-4 and within 10x of what the FP32 CPU
    instantiation of the oracle achieves (vP integrates grad tp, whose static
    near-source residue makes FP32 rounding grow linearly in time);
  images: rel L2 <= 1e-4 over cells whose excitation time agrees and whose
    excitation amplitude exceeds 1e-3 of the maximum (the excitation time is
    an argmax, so FP32 can pick a neighbouring step where two samples are
    within rounding; the fraction of such cells is bounded too).
"""
import os
import subprocess

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

PASS_MODEL, PASS_SOURCE, PASS_RECEIVER = 0, 1, 2
FIELDS_TIGHT = ["vx", "vy", "vz", "txx", "tyy", "tzz", "txy", "txz", "tyz", "tp"]
FIELDS_VP = ["vxp", "vyp", "vzp"]


def _say(msg):
    print(msg, flush=True)


# --------------------------------------------------------------------------- closed forms on the GPU
@pytest.mark.parametrize("R", [4, 6])
def test_gpu_closed_form_derivatives(erm, R):
    import test_oracle3d as T
    _say(f"closed-form stencil check on GPU, R={R}")
    for comp, axis, s, err in T.closed_form_stress(lambda p: erm.RTM3D(p), R):
        assert err < 2e-5, (comp, axis, s, err)
    for comp, axis, v, err in T.closed_form_velocity(lambda p: erm.RTM3D(p), R):
        assert err < 2e-5, (comp, axis, v, err)


# --------------------------------------------------------------------------- wavefields vs FP64
PARITY_CASES = {
    # name: (params kwargs, steps)
    "r4_ragged": (dict(nx=70, ny=37, nz=29, b=8, R=4, lt=120), 120),
    "r6_reference_coeffs": (dict(nx=50, ny=44, nz=40, b=10, R=6, lt=100), 100),
    "r4_many_chunks_ty4": (dict(nx=45, ny=30, nz=52, b=6, R=4, lt=90, zchunk=5, tile_x=64, tile_y=4), 90),
    "r6_src_near_corner": (dict(nx=33, ny=29, nz=27, b=4, R=6, lt=80, src=(3, 4, 3)), 80),
}


@pytest.mark.parametrize("name", list(PARITY_CASES))
def test_gpu_wavefields_match_fp64_oracle(erm, oracle, name):
    kw, steps = PARITY_CASES[name]
    p = cases.params(**kw)
    vp, vs, rho = cases.layered(p, seed=11)
    _say(f"wavefield parity {name}: padded {tuple(p.padded_shape)}, {steps} steps")
    g = erm.RTM3D(p)
    g.load_model(0, vp, vs, rho)
    g.zero_wavefield()
    half = steps // 2
    g.step(PASS_MODEL, 0, half)
    mid = {k: g.get_field(k) for k in ["vz", "txx", "tp"]}
    g.step(PASS_MODEL, half, steps - half)
    o64 = oracle.Oracle3D(p, 64)
    o32 = oracle.Oracle3D(p, 32)
    for o in (o64, o32):
        o.load_model(0, vp, vs, rho)
        o.step(PASS_MODEL, 0, half)
    for k, a in mid.items():
        e = cases.rel_l2(a, o64.get_field(k))
        assert e < 1e-5, ("mid", k, e)
    for o in (o64, o32):
        o.step(PASS_MODEL, half, steps - half)
    for k in FIELDS_TIGHT:
        e = cases.rel_l2(g.get_field(k), o64.get_field(k))
        assert e < 1e-5, (k, e)
    for k in FIELDS_VP:
        ref = o64.get_field(k)
        e = cases.rel_l2(g.get_field(k), ref)
        e32 = cases.rel_l2(o32.get_field(k), ref)
        assert e < 1e-4 and e <= 10 * e32 + 1e-6, (k, e, e32)
    rec = g.get_record()
    er = cases.rel_l2(rec[:, :steps], o64.get_record()[:, :steps])
    assert er < 1e-5, er
    # P/S decomposition by divergence and curl of the final velocity
    for k in ("div", "curlx", "curly", "curlz"):
        e = cases.rel_l2(g.get_field(k), o64.get_field(k))
        assert e < 1e-5, (k, e)
    g.close()


def test_gpu_full_config_grid_parity(erm, oracle):
    """The 256x256x128 model + 20-cell C-PML grid (296x296x168 padded) of the
    benchmark, with its launch shape (32x8 tiles, z sweep 6, XCD y-band order,
    region-split plain / C-PML tile lists) and random fields in every cell:
    three steps on the GPU against the FP64 oracle, so every tile, z chunk,
    memory-variable slab and edge of the production launch is compared."""
    p = cases.params(nx=256, ny=256, nz=128, b=20, R=4, lt=3, dt=0.8, freq=15.0, src=(128, 128, 1),
                     rec=(128, 128, 0, 0, 2, 2, 1), boundary=1, cpml_r=1e-3, tile_x=32, tile_y=8, zchunk=6, xcd_map=2)
    vp, vs, rho = erm.make_model3d(1, 256, 256, 128, 10.0, 10.0, 10.0, 2)
    _say(f"full-grid parity: padded {tuple(p.padded_shape)}")
    rng = np.random.default_rng(5)
    init = {k: (rng.standard_normal(tuple(p.padded_shape)) * s).astype(np.float32)
            for k, s in [("vx", 1e-3), ("vz", 1e-3), ("txx", 1e3), ("txz", 1e3), ("tp", 1e3)]}
    g = erm.RTM3D(p)
    g.load_model(0, vp, vs, rho)
    for k, a in init.items():
        g.set_field(k, a)
    g.step(PASS_MODEL, 0, 3)
    got = {k: g.get_field(k) for k in ["vx", "vy", "vz", "txx", "tyy", "tzz", "txy", "txz", "tyz", "tp",
                                       "vxp", "vyp", "vzp"]}
    g.close()
    o = oracle.Oracle3D(p, 64)
    o.load_model(0, vp, vs, rho)
    for k, a in init.items():
        o.set_field(k, a.astype(np.float64))
    o.step(PASS_MODEL, 0, 3)
    for k, a in got.items():
        e = cases.rel_l2(a, o.get_field(k))
        assert e < 1e-5, (k, e)
    o.close()


def test_gpu_discrete_energy_conserved_full_grid(erm):
    """A size-independent property at the full 296x296x168 padded grid: with
    no damping (zero boundary widths) and no source (time index past the
    wavelet), the GPU's staggered leapfrog conserves the discrete energy
    1/2 sum rho |v^n|^2 + 1/2 sum sigma^{n+1/2} . C^-1 sigma^{n-1/2} of every
    cell, whatever the launch shape.  Tolerance 1e-5 relative over 6 steps
    (FP32 state; the moduli in the functional are the FP64 values of the
    FP32 inputs; the FP32 CPU oracle drifts by 2e-9 here).  The functional
    with a wrong lambda (M / 2) must drift by more than 1e-3, so the check is
    not satisfied trivially."""
    R = 4
    p = cases.params(nx=296, ny=296, nz=168, b=0, R=R, lt=4, dt=0.8, freq=15.0, src=(148, 148, 8),
                     rec=(4, 4, 100, 100, 20, 20, 8))
    vp, vs, rho = erm.make_model3d(1, 296, 296, 168, 10.0, 10.0, 10.0, 2)
    _say(f"full-grid energy: padded {tuple(p.padded_shape)}")
    g = erm.RTM3D(p)
    g.load_model(1, vp, vs, rho)
    shape = g.shape
    rng = np.random.default_rng(8)
    v_prev = []
    for k in ("vx", "vy", "vz"):
        a = np.zeros(shape, np.float32)
        a[R:-R, R:-R, R:-R] = rng.standard_normal((shape[0] - 2 * R, shape[1] - 2 * R, shape[2] - 2 * R))
        g.set_field(k, a)
        v_prev.append(a.ravel())
    f64 = lambda a: np.ascontiguousarray(a.astype(np.float64).transpose(2, 1, 0)).ravel()
    r = f64(rho)
    M = r * f64(vp) ** 2
    mu = r * f64(vs) ** 2
    w_n, w_s = 0.5 / mu, 1.0 / mu
    w_t = {"right": (M - 2 * mu) / (2 * mu + 3 * (M - 2 * mu)) * w_n, "wrong": 0.5 * M / (2 * mu + 1.5 * M) * w_n}
    dot = lambda a, b, w: float(np.einsum("i,i,i->", a, b, w))

    def energies(v, sn, so):
        kin = sum(dot(x, x, r) for x in v)
        base = sum(dot(sn[c], so[c], w_n) for c in range(3)) + sum(dot(sn[c], so[c], w_s) for c in range(3, 6))
        tn, to = sn[0] + sn[1] + sn[2], so[0] + so[1] + so[2]
        return {k: 0.5 * kin + 0.5 * (base - dot(tn, to, wt)) for k, wt in w_t.items()}

    s_prev = [np.zeros(M.size, np.float32)] * 6
    E = {"right": [], "wrong": []}
    for n in range(6):
        g.step(PASS_SOURCE, 100000 + n, 1)  # state: sigma^{n+1/2}, v^{n+1}
        s_new = [g.get_field(k).ravel() for k in ("txx", "tyy", "tzz", "txy", "txz", "tyz")]
        for k, e in energies(v_prev, s_new, s_prev).items():
            E[k].append(e)
        v_prev = [g.get_field(k).ravel() for k in ("vx", "vy", "vz")]
        s_prev = s_new
    g.close()
    e, ew = np.array(E["right"]), np.array(E["wrong"])
    drift = np.abs(e - e[0]).max() / e[0]
    _say(f"full-grid energy: {e[0]:.6e}, max relative drift over 6 steps {drift:.2e}")
    assert drift < 1e-5
    assert np.abs(ew - ew[0]).max() / abs(ew[0]) > 1e-3


# --------------------------------------------------------------------------- full RTM shot
def _reflector_case():
    nz, zi = 48, 30
    p = cases.params(nx=40, ny=40, nz=nz, b=10, R=4, lt=420, dt=1.0, freq=25.0, coe=250.0,
                     src=(20, 20, 3), rec=(19, 19, 2, 2, 2, 2, 2))
    p.lt_forward = p.lt
    shape = (p.nx, p.ny, p.nz)
    z = np.arange(nz)[None, None, :]
    vp = (np.where(z < zi, 2000.0, 3000.0) * np.ones(shape)).astype(np.float32)
    rho = (np.where(z < zi, 2000.0, 2400.0) * np.ones(shape)).astype(np.float32)
    vs = (vp / 1.732).astype(np.float32)
    s = np.full(shape, 2000.0, np.float32)
    smodel = (s, (s / 1.732).astype(np.float32), np.full(shape, 2000.0, np.float32))
    return p, (vp, vs, rho), smodel, zi


def test_gpu_rtm_shot_images_match_fp64_oracle(erm, oracle):
    p, model, smodel, zi = _reflector_case()
    _say(f"RTM shot parity: padded {tuple(p.padded_shape)}, lt={p.lt}")
    g = erm.RTM3D(p)
    g.load_model(0, *model)
    g.load_model(1, *smodel)
    rec = g.forward_real()
    g.forward_smoothed()
    gt = g.get_field("ex_time")
    ga = g.get_field("ex_amp")
    g.backward_smoothed()
    gimg = g.get_images()
    gfull = {k: g.get_field(k) for k in ("img_tp", "img_pp", "img_ps", "ex_tp")}
    g.close()
    o = oracle.Oracle3D(p, 64)
    o.load_model(0, *model)
    o.load_model(1, *smodel)
    o.rtm_shot()
    er = cases.rel_l2(rec, o.get_record())
    assert er < 1e-5, er
    ot = o.get_field("ex_time")
    oa = o.get_field("ex_amp")
    strong = oa > 1e-3 * oa.max()
    same = (gt == ot)
    frac_diff = 1.0 - same[strong].mean()
    _say(f"  excitation time differs on {frac_diff:.2e} of excited cells")
    assert frac_diff < 0.02
    assert cases.rel_l2(ga[strong & same], oa[strong & same]) < 1e-5
    mask = strong & same
    for k in ("img_pp", "img_ps"):
        e = cases.rel_l2(gfull[k][mask], o.get_field(k)[mask])
        assert e < 1e-4, (k, e)
    etp = o.get_field("ex_tp")
    mtp = mask & (np.abs(etp) > 1e-3 * np.abs(etp).max())
    e = cases.rel_l2(gfull["img_tp"][mtp], o.get_field("img_tp")[mtp])
    assert e < 1e-4, ("img_tp", e)
    # model-region cut in disk order and the reflector depth
    oimg = o.images()
    pp = gimg[1]
    np.testing.assert_array_equal(pp.shape, (p.nx, p.ny, p.nz))
    col = np.abs(pp[18:23, 18:23, 10:]).sum(axis=(0, 1))
    assert abs(int(np.argmax(col)) + 10 - zi) <= 2
    assert cases.rel_l2(oimg[1][mask[p.bu:p.bu + p.nz, p.bf:p.bf + p.ny, p.bl:p.bl + p.nx].transpose(2, 1, 0)],
                        pp[mask[p.bu:p.bu + p.nz, p.bf:p.bf + p.ny, p.bl:p.bl + p.nx].transpose(2, 1, 0)]) < 1e-4


def test_gpu_record_roundtrip_and_join_shot(erm):
    """A record set from the host (join_shot ingest) images like the one the
    modeling pass left on the device."""
    p, model, smodel, _ = _reflector_case()
    p.lt = p.lt_forward = 200
    g = erm.RTM3D(p)
    g.load_model(0, *model)
    g.load_model(1, *smodel)
    rec = g.forward_real()
    g.forward_smoothed()
    g.backward_smoothed()
    a = g.get_images()
    g.set_record(np.zeros_like(rec))
    g.backward_smoothed()
    z = g.get_images()
    assert all(np.all(x == 0) for x in z)
    g.backward_smoothed(rec)
    b = g.get_images()
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(g.get_record(), rec)
    g.close()


# --------------------------------------------------------------------------- errors and files
def test_gpu_error_paths(erm):
    p = cases.params(nx=20, ny=20, nz=20, b=4, R=4, lt=10, shot_interval_x=6)
    g = erm.RTM3D(p)
    with pytest.raises(erm.ErmError) as ei:
        g.step(PASS_MODEL, 0, 1)            # no model loaded
    assert ei.value.code == erm.E_STATE
    g.load_model(1, *cases.homogeneous(p))
    with pytest.raises(erm.ErmError) as ei:
        g.backward_smoothed()              # no record yet
    assert ei.value.code == erm.E_STATE
    with pytest.raises(erm.ErmError) as ei:
        g.set_shot(3, 0, 0)                # 10 + 3*6 = 28: outside the 20-cell model
    assert ei.value.code == erm.E_PARAM
    with pytest.raises(erm.ErmError):
        g.set_field("ex_time", np.zeros(g.shape, np.float32))
    g.close()
    # one z sweep plus its 2R halo planes must span < 4 GiB (32-bit offsets);
    # both cases are rejected before anything is allocated
    big = cases.params(nx=4000, ny=4000, nz=80, b=0, R=4, lt=10)
    big.zchunk = 60                        # 64 MB planes: at most 4 GiB / 64 MB - 8 = 59
    with pytest.raises(erm.ErmError) as ei:
        erm.RTM3D(big)
    assert ei.value.code == erm.E_PARAM and "zchunk" in str(ei.value)
    huge = cases.params(nx=11000, ny=11000, nz=9, b=0, R=4, lt=10)  # 9 planes of 484 MB > 4 GiB
    with pytest.raises(erm.ErmError) as ei:
        erm.RTM3D(huge)
    assert ei.value.code == erm.E_SIZE


def test_gpu_write_results_format(erm, tmp_path):
    p, model, smodel, _ = _reflector_case()
    p.lt = p.lt_forward = 120
    g = erm.RTM3D(p)
    g.load_model(0, *model)
    g.load_model(1, *smodel)
    g.forward_real(want_record=False)
    g.forward_smoothed()
    g.write_excitation(str(tmp_path))
    g.backward_smoothed()
    g.write_results(str(tmp_path))
    imgs = g.get_images()
    sx, sy = p.shot_start_x, p.shot_start_y
    for name, a in zip(("tp", "pp", "ps"), imgs):
        f = tmp_path / f"result-{name}-{sy}-{sx}"
        assert f.stat().st_size == p.nx * p.ny * p.nz * 4
        np.testing.assert_array_equal(np.fromfile(f, np.float32).reshape(p.nx, p.ny, p.nz), a)
    t = np.fromfile(tmp_path / f"cut-ex-time-{sy}-{sx}", np.float32).reshape(p.nx, p.ny, p.nz)
    full = g.get_field("ex_time")
    np.testing.assert_array_equal(t, full[p.bu:p.bu + p.nz, p.bf:p.bf + p.ny, p.bl:p.bl + p.nx].transpose(2, 1, 0))
    g.close()


def test_gpu_cli_matches_api(erm, tmp_path):
    """The rtm3d CLI (par file + raw model files) writes the same images as the API."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "2d_and_3d_elastic_reverse_time_migration_amd", "build", "rtm3d")
    p, model, smodel, _ = _reflector_case()
    lt = 150
    names = {}
    for tag, arrs in (("", model), ("s_", smodel)):
        for key, a in zip(("velocity", "velocity1", "density"), arrs):
            f = str(tmp_path / f"{tag}{key}.bin")
            erm.write_raw(f, a)
            names[tag + key] = f
    keys = dict(RTM=1, GPU_switch=1, GPU_N=1, GPU_start=0, exchange_device_bool=0, lt=lt, dt=p.dt, freq=p.freq,
                join_vs=1, join_den=1, join_shot=0, join_wavefield=0, wavefield_interval=100,
                infile_shot_name="none", nx=p.nx, ny=p.ny, nz=p.nz, dx=p.dx, dy=p.dy, dz=p.dz,
                shot_num_x=1, shot_num_y=1, shot_num_z=1, shot_start_x=p.shot_start_x, shot_interval_x=0,
                shot_start_y=p.shot_start_y, shot_interval_y=0, shot_start_z=p.shot_start_z, shot_interval_z=0,
                receiver_num_x=p.receiver_num_x, receiver_num_y=p.receiver_num_y, receiver_num_z=1,
                receiver_start_x=p.receiver_start_x, receiver_interval_x=p.receiver_interval_x,
                receiver_start_y=p.receiver_start_y, receiver_interval_y=p.receiver_interval_y,
                receiver_start_z=p.receiver_start_z, receiver_interval_z=0, receiver_offset=0, receiver_mark=0,
                boundary_up=p.bu, boundary_down=p.bd, boundary_left=p.bl, boundary_right=p.br,
                boundary_front=p.bf, boundary_back=p.bb, coe_attenuation=p.coe_attenuation,
                velocity_multiple=1, smooth_time_vp=0, smooth_time_vs=0, smooth_time_density=0,
                add_receiver_bool=1, cut_direct_wave=0, laplace=0, laplace_compensate=0, vsp=0, vsp_precon=0,
                precon=0, precon_z1=0, precon_z2=0, fd_order=8, lt_forward=lt, **names)
    par = tmp_path / "rtm.par"
    par.write_text("\n".join(f"{k}={v}" for k, v in keys.items()) + "\n")
    r = subprocess.run([exe, f"par={par}"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    p.lt = p.lt_forward = lt
    g = erm.RTM3D(p)
    g.load_model(0, *model)
    g.load_model(1, *smodel)
    g.forward_real(want_record=False)
    g.forward_smoothed()
    g.backward_smoothed()
    imgs = g.get_images()
    g.close()
    for name, a in zip(("tp", "pp", "ps"), imgs):
        f = tmp_path / "result" / f"result-{name}-{p.shot_start_y}-{p.shot_start_x}"
        np.testing.assert_array_equal(np.fromfile(f, np.float32).reshape(a.shape), a)
    # a second run resumes: the shot's results exist, so it is skipped
    r2 = subprocess.run([exe, f"par={par}"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r2.returncode == 0 and "skipped" in r2.stderr


def test_gpu_wavefield_snapshots(erm, tmp_path, monkeypatch):
    """join_wavefield / wavefield_interval (output_3d_wavefiled_*,
    3D_output_file.cu:2-151, gated at 3D_forward_or_back_together.cu:50): after
    the step of time index it with (it + 1) % interval == 0 the padded-grid
    and model-region snapshots of tp, vx, vz, vzp, vzs hold the fields of that
    step in disk order; the directory becomes wavefield1 after the model
    pass."""
    monkeypatch.chdir(tmp_path)
    p = cases.params(nx=20, ny=18, nz=16, b=6, R=4, lt=30)
    p.join_wavefield, p.wavefield_interval = 1, 10
    vp, vs, rho = cases.layered(p, seed=2)
    _say("wavefield snapshots: model pass, interval 10")
    with erm.RTM3D(p) as g:
        g.load_model(0, vp, vs, rho)
        g.zero_wavefield()
        want = {}
        for it in range(p.lt):
            g.step(PASS_MODEL, it, 1)
            if (it + 1) % 10 == 0:
                want[it + 1] = {k: g.get_field(k) for k in ("tp", "vx", "vz", "vzp", "vzs")}
    d = tmp_path / "wavefield1"
    assert d.is_dir() and not (tmp_path / "wavefield").exists()
    nnz, nny, nnx = p.padded_shape
    for label, fields in want.items():
        for name, a in fields.items():
            full = np.fromfile(d / f"{name}-{label}", np.float32).reshape(nnx, nny, nnz)
            np.testing.assert_array_equal(full, a.transpose(2, 1, 0))
            np.testing.assert_array_equal(np.fromfile(d / f"{name}-0-{label}", np.float32), full.ravel())
            cut = np.fromfile(d / f"cut-{name}-{label}", np.float32).reshape(p.nx, p.ny, p.nz)
            np.testing.assert_array_equal(cut, full[p.bl:p.bl + p.nx, p.bf:p.bf + p.ny, p.bu:p.bu + p.nz])
