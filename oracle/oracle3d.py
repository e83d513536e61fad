"""ctypes wrapper of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.  The
product library never calls it.

The oracle restates the reference's 3D elastic RTM equations in C++ and is
instantiated in FP64 (the correctness oracle) and FP32 (the CPU baseline);
its header comment lists the reference lines it follows.
"""
from __future__ import annotations

import ctypes
import importlib.util
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "build", "liboracle.so")

_spec = importlib.util.spec_from_file_location(
    "_erm_params", os.path.join(ROOT, "2d_and_3d_elastic_reverse_time_migration_amd", "params.py"))
_params = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(_params)
Params = _params.Params
FIELD_ID = _params.FIELD_ID
PASS_MODEL, PASS_SOURCE, PASS_RECEIVER = 0, 1, 2


def build():
    subprocess.run(["make", "-C", HERE], check=True, capture_output=True)


def _load():
    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    P = ctypes.POINTER
    dp = P(ctypes.c_double)
    fp = P(ctypes.c_float)
    vp = ctypes.c_void_p
    sig = {
        "oracle_set_threads": (ctypes.c_int, [ctypes.c_int]),
        "oracle_coefficients": (ctypes.c_int, [ctypes.c_int, dp]),
        "oracle_ricker": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, dp, ctypes.c_int]),
        "oracle_damping_axis": (ctypes.c_int, [vp, ctypes.c_int, dp]),
        "oracle3d_create": (vp, [vp, ctypes.c_int]),
        "oracle3d_destroy": (None, [vp]),
        "oracle3d_load_model": (ctypes.c_int, [vp, ctypes.c_int, fp, fp, fp]),
        "oracle3d_set_shot": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "oracle3d_zero": (ctypes.c_int, [vp, ctypes.c_int]),
        "oracle3d_step": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "oracle3d_get_field": (ctypes.c_int, [vp, ctypes.c_int, dp]),
        "oracle3d_set_field": (ctypes.c_int, [vp, ctypes.c_int, dp]),
        "oracle3d_get_record": (ctypes.c_int, [vp, dp]),
        "oracle3d_set_record": (ctypes.c_int, [vp, dp]),
        "oracle3d_rtm_shot": (ctypes.c_int, [vp, ctypes.c_int]),
        "oracle3d_image": (ctypes.c_int, [vp, ctypes.c_int]),
        "oracle3d_xcorr_shot": (ctypes.c_int, [vp, ctypes.c_int, dp, dp, dp]),
        "oracle3d_xcorr_shot_decim": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, dp, dp, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _f32(a):
    if a is None:
        return None
    a = np.ascontiguousarray(a, np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _pp(p):
    """Address of an erm_params-layout struct (any Params class instance)."""
    assert ctypes.sizeof(p) == ctypes.sizeof(Params)
    return ctypes.addressof(p)


def set_threads(n: int) -> int:
    return lib.oracle_set_threads(n)


def coefficients(radius: int) -> np.ndarray:
    c = np.zeros(radius + 1)
    assert lib.oracle_coefficients(radius, _dp(c)) == 0
    return c


def ricker(freq: float, dt_ms: float) -> np.ndarray:
    n = lib.oracle_ricker(freq, dt_ms, None, 0)
    w = np.zeros(n)
    lib.oracle_ricker(freq, dt_ms, _dp(w), n)
    return w


def damping_axis(p: Params, axis: int) -> np.ndarray:
    n = [p.nx + p.bl + p.br, p.ny + p.bf + p.bb, p.nz + p.bu + p.bd][axis]
    out = np.zeros(n)
    rc = lib.oracle_damping_axis(_pp(p), axis, _dp(out))
    assert rc == n, rc
    return out


def mute_direct(p: Params, rec, vp, shot=(0, 0, 0)):
    """Direct-wave mute of a record [3][lt][ny_rec][nx_rec] (the library applies it
    in erm_forward_real when cut_direct_wave > 0 and vsp == 0), FP64: at each
    receiver the samples it < ceil(offset / (vp_s dt) - 1e-6) + wl + cut_direct_wave
    become 0, with offset the source-receiver distance in metres, vp_s the true
    model's vp at the source cell (vp: model region [nx][ny][nz]) and wl the
    Ricker length in samples.  The reference calls
    cut_direct_shot_3D_surface_three_new (3D_forward_or_back_together.cu:133-136)
    without its body, so the window is the library's definition, restated;
    pinned by tests/test_oracle3d.py (zero-offset cut = wl + margin, monotone in
    offset, only the direct wave's samples removed in a homogeneous medium)."""
    out = np.array(rec, np.float64, copy=True)
    if p.cut_direct_wave <= 0 or p.vsp != 0:
        return out
    sx = p.shot_start_x + shot[0] * p.shot_interval_x
    sy = p.shot_start_y + shot[1] * p.shot_interval_y
    sz = p.shot_start_z + shot[2] * p.shot_interval_z
    vps = float(np.asarray(vp)[sx, sy, sz])
    wl = len(ricker(p.freq, p.dt))
    dts = p.dt / 1000.0
    for iry in range(p.receiver_num_y):
        for irx in range(p.receiver_num_x):
            ddx = (p.receiver_start_x + irx * p.receiver_interval_x - sx) * p.dx
            ddy = (p.receiver_start_y + iry * p.receiver_interval_y - sy) * p.dy
            ddz = (p.receiver_start_z - sz) * p.dz
            t = np.sqrt(ddx * ddx + ddy * ddy + ddz * ddz) / (vps * dts)
            n = min(int(np.ceil(t - 1e-6)) + wl + p.cut_direct_wave, p.lt)
            out[:, :n, iry, irx] = 0.0
    return out


class Oracle3D:
    """FP64 (prec=64) or FP32 (prec=32) CPU restatement of the 3D path."""

    def __init__(self, p: Params, prec: int = 64):
        self.p = p
        if p.lt_forward < 0:
            p.lt_forward = 2 * p.lt // 3
        self._h = lib.oracle3d_create(_pp(p), prec)
        if not self._h:
            raise ValueError("oracle3d_create rejected the parameters")
        self.shape = p.padded_shape

    def close(self):
        if getattr(self, "_h", None):
            lib.oracle3d_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _c(self, rc):
        if rc != 0:
            raise RuntimeError(f"oracle error {rc}")

    def load_model(self, which, vp, vs=None, rho=None):
        keep = [_f32(vp), _f32(vs), _f32(rho)]
        self._c(lib.oracle3d_load_model(self._h, which, *[k[1] if k else None for k in keep]))

    def set_shot(self, isx=0, isy=0, isz=0):
        self._c(lib.oracle3d_set_shot(self._h, isx, isy, isz))

    def zero(self, wavefield=True, excitation=False, images=False):
        self._c(lib.oracle3d_zero(self._h, (1 if wavefield else 0) | (2 if excitation else 0) | (4 if images else 0)))

    def step(self, pass_, it_begin, nsteps):
        self._c(lib.oracle3d_step(self._h, pass_, it_begin, nsteps))

    def get_field(self, name):
        out = np.empty(self.shape, np.float64)
        self._c(lib.oracle3d_get_field(self._h, FIELD_ID[name] if isinstance(name, str) else name, _dp(out)))
        return out

    def set_field(self, name, a):
        a = np.ascontiguousarray(a, np.float64)
        self._c(lib.oracle3d_set_field(self._h, FIELD_ID[name], _dp(a)))

    def record_shape(self):
        p = self.p
        return (3, p.lt, p.receiver_num_y, p.receiver_num_x)

    def get_record(self):
        out = np.empty(self.record_shape(), np.float64)
        self._c(lib.oracle3d_get_record(self._h, _dp(out)))
        return out

    def set_record(self, rec):
        rec = np.ascontiguousarray(rec, np.float64)
        self._c(lib.oracle3d_set_record(self._h, _dp(rec)))

    def image(self, it):
        """Apply the excitation imaging condition of step `it` to the current
        fields (v, vP, tp) without propagating: cells whose excitation time is
        `it` get I_tp, I_PP, I_PS."""
        self._c(lib.oracle3d_image(self._h, it))

    def rtm_shot(self, use_record=False):
        self._c(lib.oracle3d_rtm_shot(self._h, 1 if use_record else 0))

    def xcorr_shot(self, use_record=False, decim=1):
        """Cross-correlation images (tp, PP, PS) with the full source history,
        [nx][ny][nz]; decim > 1 keeps the receiver steps with (it+1) % decim == 0,
        each weighted by decim."""
        p = self.p
        out = [np.empty((p.nx, p.ny, p.nz), np.float64) for _ in range(3)]
        self._c(lib.oracle3d_xcorr_shot_decim(self._h, 1 if use_record else 0, decim, *[_dp(o) for o in out]))
        return out

    def images(self):
        """PP/PS/tp images over the model region in disk order [nx][ny][nz]."""
        p = self.p
        out = []
        for name in ("img_tp", "img_pp", "img_ps"):
            f = self.get_field(name)
            cut = f[p.bu:p.bu + p.nz, p.bf:p.bf + p.ny, p.bl:p.bl + p.nx]
            out.append(np.ascontiguousarray(cut.transpose(2, 1, 0)))
        return out
