"""Host-side checks of the C ABI (no GPU needed): the library loads and
exports every entry point include/erm.h declares, the getpar-compatible
parser reads the reference's keys (3D_elastic_modeling_parameter_sh:6-162),
errors come back as codes + messages, and the host utilities (raw I/O,
model generator) behave."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

REF_KEYS = dict(
    RTM=1, GPU_switch=1, GPU_N=1, GPU_start=0, exchange_device_bool=0, lt=100, dt=0.8, freq=15,
    join_vs=1, join_den=1, join_shot=0, join_wavefield=0, wavefield_interval=100, infile_shot_name="shot.bin",
    nx=20, ny=22, nz=24, dx=10, dy=10, dz=10, shot_num_x=1, shot_num_y=1, shot_num_z=1,
    shot_start_x=10, shot_interval_x=5, shot_start_y=11, shot_interval_y=5, shot_start_z=3, shot_interval_z=0,
    receiver_num_x=4, receiver_num_y=5, receiver_num_z=1, receiver_start_x=1, receiver_interval_x=3,
    receiver_start_y=2, receiver_interval_y=2, receiver_start_z=1, receiver_interval_z=0,
    receiver_offset=0, receiver_mark=0, boundary_up=10, boundary_down=10, boundary_left=10,
    boundary_right=10, boundary_front=10, boundary_back=10, coe_attenuation=200.0,
    velocity_multiple=1.0, smooth_time_vp=0, smooth_time_vs=0, smooth_time_density=0,
    velocity="vp.bin", velocity1="vs.bin", density="rho.bin", s_velocity="svp.bin",
    s_velocity1="svs.bin", s_density="srho.bin", add_receiver_bool=1, cut_direct_wave=0, laplace=0,
    laplace_compensate=0, vsp=0, vsp_precon=0, precon=0, precon_z1=0, precon_z2=0,
)


def header_functions():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(erm2?d?_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol(erm):
    names = header_functions()
    assert len(names) >= 55 and "erm2d_adjoint" in names and "erm_unet_train" in names
    for n in names:
        assert hasattr(erm.lib, n), n
        ctypes.cast(getattr(erm.lib, n), ctypes.c_void_p)


def test_params_struct_matches_c(erm):
    assert erm.lib.erm_params_size() == ctypes.sizeof(erm.Params)
    p = erm.default_params()
    assert p.fd_order == 12 and p.boundary == 0 and p.lt_forward == -1


def test_parse_par_file_and_overrides(erm, tmp_path):
    par = tmp_path / "model.par"
    lines = [f"{k}={v}" for k, v in REF_KEYS.items()]
    par.write_text("# reference keys\n" + "\n".join(lines) + "\n")
    p = erm.parse_par([f"par={par}", "nx=33", "fd_order=8"])
    assert p.nx == 33 and p.ny == 22 and p.nz == 24
    assert abs(p.dt - 0.8) < 1e-7 and p.lt == 100
    assert p.bu == 10 and p.bl == 10 and p.bb == 10
    assert p.velocity == b"vp.bin" and p.s_density == b"srho.bin"
    assert p.fd_order == 8
    assert p.receiver_interval_x == 3


@pytest.mark.parametrize("missing, msg", [("nx", "nx is not set!"), ("velocity", "can not read vp model!"),
                                          ("boundary_up", "boundary_up is not set!")])
def test_parse_par_missing_key_is_an_error(erm, missing, msg):
    args = [f"{k}={v}" for k, v in REF_KEYS.items() if k != missing]
    with pytest.raises(erm.ErmError) as ei:
        erm.parse_par(args)
    assert msg in str(ei.value)
    assert ei.value.code == erm.E_PARAM


def test_smoothed_names_required_only_without_smoothing(erm):
    args = [f"{k}={v}" for k, v in REF_KEYS.items() if k not in ("s_velocity",)]
    with pytest.raises(erm.ErmError):
        erm.parse_par(args)
    args2 = [a for a in args if not a.startswith("smooth_time_vp=")] + ["smooth_time_vp=3"]
    p = erm.parse_par(args2)
    assert p.smooth_time_vp == 3


def test_create_rejects_bad_grids(erm):
    p = erm.default_params(nx=0, ny=10, nz=10, dx=10, dy=10, dz=10, dt=1, lt=10, freq=20)
    with pytest.raises(erm.ErmError) as ei:
        erm.RTM3D(p)
    assert ei.value.code == erm.E_PARAM
    p = erm.default_params(nx=20, ny=20, nz=20, dx=10, dy=10, dz=10, dt=1, lt=10, freq=20, bu=1)
    with pytest.raises(erm.ErmError) as ei:
        erm.RTM3D(p)
    assert "0 or >= 2" in str(ei.value)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="host without a GPU only")
def test_create_without_gpu_reports_device_error(erm):
    p = erm.default_params(nx=20, ny=20, nz=20, dx=10, dy=10, dz=10, dt=1, lt=10, freq=20,
                           shot_start_x=10, shot_start_y=10, shot_start_z=10)
    with pytest.raises(erm.ErmError) as ei:
        erm.RTM3D(p)
    assert ei.value.code == erm.E_DEVICE


def test_raw_io_roundtrip_and_missing_file(erm, tmp_path):
    a = np.random.default_rng(0).standard_normal((5, 6, 7)).astype(np.float32)
    f = str(tmp_path / "a.bin")
    erm.write_raw(f, a)
    assert os.path.getsize(f) == a.size * 4
    np.testing.assert_array_equal(erm.read_raw(f, a.shape), a)
    np.testing.assert_array_equal(np.fromfile(f, np.float32).reshape(a.shape), a)
    with pytest.raises(erm.ErmError) as ei:
        erm.read_raw(str(tmp_path / "nope.bin"), (3,))
    assert ei.value.code == erm.E_IO


def test_model_generator_is_seeded_and_consistent(erm):
    vp, vs, rho = erm.make_model3d(1, 16, 14, 12, 10, 10, 10, 2)
    vp2, _, _ = erm.make_model3d(1, 16, 14, 12, 10, 10, 10, 2)
    vp3, _, _ = erm.make_model3d(1, 16, 14, 12, 10, 10, 10, 3)
    np.testing.assert_array_equal(vp, vp2)
    assert not np.array_equal(vp, vp3)
    np.testing.assert_allclose(vs, vp / 1.732, rtol=1e-6)
    np.testing.assert_allclose(rho, 230.0 * vp.astype(np.float64) ** 0.25, rtol=1e-6)
    vp0, _, _ = erm.make_model3d(0, 200, 4, 60, 10, 10, 10, 1)
    assert abs(vp0[0, 0, 2] - 2000.0) < 1.0   # first layer, 1 km from the anomaly
    assert abs(vp0[0, 0, 58] - 3500.0) < 1.0
    # +300 m/s Gaussian anomaly at half width, 0.55 depth, on the 2800 m/s layer
    assert abs(vp0[100, 2, 33] - 3100.0) < 1e-3      # the anomaly centre (1000, 20, 330) m


def test_headers_are_c99_and_link(erm, tmp_path):
    """The public headers are plain C: a C99 translation unit that includes all
    of them and takes the address of every declared function compiles with
    -pedantic and links against liberm.so (the binding a C, cgo or FFI caller
    makes)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    names = header_functions()
    src = tmp_path / "abi.c"
    src.write_text(
        "#include \"erm.h\"\n#include \"erm2d.h\"\n#include \"erm_unet.h\"\n#include \"erm_batch.h\"\n"
        "#include <stdio.h>\n"
        "typedef void (*fn)(void);\n"
        "int main(void) {\n    fn f[] = {\n"
        + "".join(f"        (fn)&{n},\n" for n in names)
        + "    };\n    printf(\"%d\\n\", (int)(sizeof f / sizeof f[0]));\n    return 0;\n}\n")
    libdir = os.path.join(ROOT, "2d_and_3d_elastic_reverse_time_migration_amd")
    exe = tmp_path / "abi"
    r = subprocess.run([cc, "-std=c99", "-Wall", "-pedantic", "-Werror", "-I", os.path.join(ROOT, "include"),
                        str(src), "-o", str(exe), "-L", libdir, "-lerm", "-Wl,-rpath," + libdir],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


@pytest.mark.parametrize("key,val", [("laplace", 1), ("laplace_compensate", 1), ("precon", 1), ("vsp_precon", 1),
                                     ("velocity_multiple", 1.5)])
def test_create_refuses_unsupported_reference_keys(erm, key, val):
    """Reference keys whose code is not in the library are refused when set to
    a non-neutral value (checked before any device work), not ignored."""
    p = erm.default_params(nx=20, ny=20, nz=20, dx=10.0, dy=10.0, dz=10.0, dt=1.0, lt=10, freq=20.0,
                           bu=4, bd=4, bl=4, br=4, bf=4, bb=4, receiver_num_x=1, receiver_num_y=1)
    setattr(p, key, val)
    with pytest.raises(erm.ErmError) as e:
        erm.RTM3D(p)
    assert e.value.code == -1 and key in str(e.value), str(e.value)
