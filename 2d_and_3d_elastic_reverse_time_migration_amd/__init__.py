"""MI355X elastic RTM: Python mirror of the C ABI in include/erm.h.

The library (liberm.so, built from csrc/ by `make` in this directory) holds
the hot path: fused gfx950 HIP kernels for the 3D staggered-grid elastic
velocity-stress step with vector P/S decomposition, excitation tracking and
excitation-amplitude imaging, behind the entry points a user of the
reference would call in place of its host drivers
(3D_elastic_RTM/3D_forward_or_back_together.cu:9-453).

This module is a thin ctypes binding: the parameter struct mirrors
`erm_params`, `RTM3D` mirrors an `erm_ctx`, and every call raises
`ErmError` with the library's message when the C call fails.  There is no
CPU fallback: if liberm.so is missing, importing raises ImportError.

Because the package directory name starts with a digit, import it with
``importlib.import_module("2d_and_3d_elastic_reverse_time_migration_amd")``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liberm.so")

from .params import *  # noqa: F401,F403
from .params import PATH_MAX, Params, FIELD_ID, FIELDS  # noqa: F401
from .params2d import Params2D, FIELD2D_ID, FIELDS2D, SRC_HISTORY, SRC_REVERSE, SRC_STRIPS  # noqa: F401


class ErmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"erm error {code}: {msg}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C {_HERE}` (or __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    fp = P(ctypes.c_float)
    vp = ctypes.c_void_p
    sig = {
        "erm_last_error": (ctypes.c_char_p, [vp]),
        "erm_version": (ctypes.c_int, []),
        "erm_params_size": (ctypes.c_size_t, []),
        "erm_device_count": (ctypes.c_int, [P(ctypes.c_int)]),
        "erm_params_default": (ctypes.c_int, [vp]),
        "erm_parse_par": (ctypes.c_int, [ctypes.c_int, P(ctypes.c_char_p), vp]),
        "erm_create": (ctypes.c_int, [vp, ctypes.c_int, P(vp)]),
        "erm_destroy": (ctypes.c_int, [vp]),
        "erm_get_dims": (ctypes.c_int, [vp, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int)]),
        "erm_load_model": (ctypes.c_int, [vp, ctypes.c_int, fp, fp, fp]),
        "erm_load_model_files": (ctypes.c_int, [vp]),
        "erm_set_shot": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "erm_zero_wavefield": (ctypes.c_int, [vp]),
        "erm_zero_excitation": (ctypes.c_int, [vp]),
        "erm_forward_real": (ctypes.c_int, [vp, fp]),
        "erm_forward_smoothed": (ctypes.c_int, [vp]),
        "erm_backward_smoothed": (ctypes.c_int, [vp, fp]),
        "erm_set_record": (ctypes.c_int, [vp, fp]),
        "erm_get_record": (ctypes.c_int, [vp, fp]),
        "erm_rtm_step": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "erm_get_field": (ctypes.c_int, [vp, ctypes.c_int, fp]),
        "erm_set_field": (ctypes.c_int, [vp, ctypes.c_int, fp]),
        "erm_get_images": (ctypes.c_int, [vp, fp, fp, fp]),
        "erm_write_results": (ctypes.c_int, [vp, ctypes.c_char_p]),
        "erm_write_excitation": (ctypes.c_int, [vp, ctypes.c_char_p]),
        "erm_set_profiling": (ctypes.c_int, [vp, ctypes.c_int]),
        "erm_get_timing": (ctypes.c_int, [vp, P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_long),
                                          P(ctypes.c_double), P(ctypes.c_long)]),
        "erm_algorithmic_bytes": (ctypes.c_int, [vp, ctypes.c_int, P(ctypes.c_double), P(ctypes.c_double)]),
        "erm_synchronize": (ctypes.c_int, [vp]),
        "erm_autotune": (ctypes.c_int, [vp, ctypes.c_int, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_double)]),
        "erm_autotune_shape": (ctypes.c_int, [vp, ctypes.c_int, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int),
                                              P(ctypes.c_double)]),
        "erm_strip_bytes": (ctypes.c_int, [vp, P(ctypes.c_double)]),
        "erm_forward_smoothed_strips": (ctypes.c_int, [vp]),
        "erm_backward_smoothed_xcorr": (ctypes.c_int, [vp, fp]),
        "erm_history_bytes": (ctypes.c_int, [vp, ctypes.c_int, P(ctypes.c_double)]),
        "erm_forward_smoothed_history": (ctypes.c_int, [vp, ctypes.c_int]),
        "erm_backward_smoothed_xcorr_history": (ctypes.c_int, [vp, fp]),
        "erm_get_xcorr_images": (ctypes.c_int, [vp, fp, fp, fp]),
        "erm_reconstruct": (ctypes.c_int, [vp, ctypes.c_int]),
        "erm_get_reconstructed_field": (ctypes.c_int, [vp, ctypes.c_int, fp]),
        "erm_copy_bandwidth": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.c_int, P(ctypes.c_double)]),
        "erm_make_model3d": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_uint64,
                                            fp, fp, fp]),
        "erm_write_raw": (ctypes.c_int, [ctypes.c_char_p, fp, ctypes.c_size_t]),
        "erm_read_raw": (ctypes.c_int, [ctypes.c_char_p, fp, ctypes.c_size_t]),
    }
    ip = P(ctypes.c_int)
    sig.update({
        "erm2d_last_error": (ctypes.c_char_p, [vp]),
        "erm2d_params_size": (ctypes.c_size_t, []),
        "erm2d_params_default": (ctypes.c_int, [vp]),
        "erm2d_create": (ctypes.c_int, [vp, ctypes.c_int, P(vp)]),
        "erm2d_destroy": (ctypes.c_int, [vp]),
        "erm2d_load_model": (ctypes.c_int, [vp, ctypes.c_int, fp, fp, fp, fp, fp]),
        "erm2d_set_shot": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
        "erm2d_zero": (ctypes.c_int, [vp]),
        "erm2d_step": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "erm2d_get_field": (ctypes.c_int, [vp, ctypes.c_int, fp]),
        "erm2d_set_field": (ctypes.c_int, [vp, ctypes.c_int, fp]),
        "erm2d_get_dims": (ctypes.c_int, [vp, ip, ip]),
        "erm2d_forward": (ctypes.c_int, [vp, fp]),
        "erm2d_rtm": (ctypes.c_int, [vp, fp, fp, fp, fp]),
        "erm2d_angle_gathers_noalloc": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_float, fp, fp, fp, fp, fp]),
        "erm2d_rtm_windows": (ctypes.c_int, [vp, ctypes.c_int] + [fp] * 10 + [ctypes.c_int, vp, fp, ctypes.c_int,
                                                                          fp, fp, fp, fp]),
        "erm2d_born": (ctypes.c_int, [vp, fp, fp]),
        "erm2d_adjoint": (ctypes.c_int, [vp, fp, fp]),
        "erm2d_adjoint_ref": (ctypes.c_int, [vp, fp, fp]),
        "erm2d_rtm_xcorr": (ctypes.c_int, [vp, fp, ctypes.c_int, fp]),
        "erm2d_last_ms": (ctypes.c_int, [vp, P(ctypes.c_double)]),
        "erm2d_lsrtm_cg": (ctypes.c_int, [vp, ctypes.c_int, ip, fp, ctypes.c_int, ctypes.c_uint, fp, fp,
                                          P(ctypes.c_double), vp, vp, ctypes.c_float]),
    })
    ull = ctypes.c_ulonglong
    sig.update({
        "erm_unet_last_error": (ctypes.c_char_p, []),
        "erm_unet_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ull, P(vp)]),
        "erm_unet_destroy": (ctypes.c_int, [vp]),
        "erm_unet_num_params": (ctypes.c_size_t, [vp]),
        "erm_unet_get_params": (ctypes.c_int, [vp, fp]),
        "erm_unet_set_params": (ctypes.c_int, [vp, fp]),
        "erm_unet_forward": (ctypes.c_int, [vp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, fp]),
        "erm_unet_loss_grad": (ctypes.c_int, [vp, fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, fp, fp]),
        "erm_unet_train_step": (ctypes.c_int, [vp, fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                               fp]),
        "erm_unet_train": (ctypes.c_int, [vp, fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_float, ull, fp]),
        "erm_unet_last_stats": (ctypes.c_int, [vp, P(ctypes.c_double), P(ctypes.c_double)]),
        "erm_unet_bind_model": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
        "erm_unet_precond": (ctypes.c_int, [fp, fp, vp]),
        "erm_unet_bind_weights": (ctypes.c_int, [vp, fp]),
        "erm2d_illumination": (ctypes.c_int, [vp, ctypes.c_int, ip, fp]),
        "erm2d_angle_gathers": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_float, fp, fp, fp, fp, fp]),
        "erm2d_precond_diag": (ctypes.c_int, [fp, fp, vp]),
        "erm2d_set_workers": (ctypes.c_int, [vp, ctypes.c_int]),
        "erm2d_set_batch": (ctypes.c_int, [vp, ctypes.c_int]),
        "erm2d_born_shots": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), fp, fp]),
        "erm2d_adjoint_shots": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), fp, fp]),
        "erm_batch_last_error": (ctypes.c_char_p, [vp]),
        "erm_batch_create": (ctypes.c_int, [vp, ctypes.c_int, P(vp)]),
        "erm_batch_destroy": (ctypes.c_int, [vp]),
        "erm_batch_load_model": (ctypes.c_int, [vp, ctypes.c_int, fp, fp, fp]),
        "erm_batch_set_shots": (ctypes.c_int, [vp, vp, ctypes.c_int]),
        "erm_batch_set_launch": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "erm_batch_set_launch_shape": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "erm_batch_autotune_shape": (ctypes.c_int, [vp, ctypes.c_int, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int),
                                                    P(ctypes.c_int), P(ctypes.c_double)]),
        "erm_batch_autotune": (ctypes.c_int, [vp, ctypes.c_int, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int),
                                              P(ctypes.c_double)]),
        "erm_batch_use_graph": (ctypes.c_int, [vp, ctypes.c_int]),
        "erm_batch_set_record": (ctypes.c_int, [vp, ctypes.c_int, fp]),
        "erm_batch_run": (ctypes.c_int, [vp, fp, fp, fp]),
        "erm_batch_set_progress": (ctypes.c_int, [vp, ctypes.c_void_p, ctypes.c_void_p]),
        "erm_batch_shot_context": (ctypes.c_int, [vp, ctypes.c_int, P(vp)]),
        "erm_batch_timing": (ctypes.c_int, [vp, P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_double),
                                            P(ctypes.c_double), P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int)]),
        "erm2d_set_cg_monitor": (ctypes.c_int, [vp, vp, vp]),
    })
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.erm2d_params_size() != ctypes.sizeof(Params2D):
        raise ImportError("erm2d_params size mismatch")
    if lib.erm_params_size() != ctypes.sizeof(Params):
        raise ImportError(f"erm_params size mismatch: C {lib.erm_params_size()} vs Python {ctypes.sizeof(Params)}")
    return lib


lib = _load()


def _fptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"], "float32 C-contiguous array required"
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _pp(p):
    """Address of an erm_params-layout struct (any Params class instance)."""
    if ctypes.sizeof(p) != ctypes.sizeof(Params):
        raise TypeError("not an erm_params struct")
    return ctypes.addressof(p)


def _check(rc: int, ctx=None):
    if rc != 0:
        raise ErmError(rc, lib.erm_last_error(ctx).decode(errors="replace"))


def default_params(**kw) -> Params:
    p = Params()
    _check(lib.erm_params_default(_pp(p)))
    return p.update(**kw)


def parse_par(args) -> Params:
    """getpar-compatible parsing of `key=value` / `par=file` arguments."""
    p = Params()
    arr = (ctypes.c_char_p * len(args))(*[a.encode() for a in args])
    _check(lib.erm_parse_par(len(args), arr, _pp(p)))
    return p


def device_count() -> int:
    n = ctypes.c_int(0)
    lib.erm_device_count(ctypes.byref(n))
    return n.value


def copy_bandwidth(nbytes, reps=20, device=0):
    """GB/s of a plain streaming copy (2*nbytes of HBM traffic per copy)."""
    ms = ctypes.c_double()
    _check(lib.erm_copy_bandwidth(device, nbytes, reps, ctypes.byref(ms)))
    return 2.0 * nbytes / (ms.value * 1e-3) / 1e9, ms.value


def make_model3d(kind, nx, ny, nz, dx, dy, dz, seed):
    """Synthetic model, model region in disk order [nx][ny][nz]."""
    vp = np.empty((nx, ny, nz), np.float32)
    vs = np.empty_like(vp)
    rho = np.empty_like(vp)
    _check(lib.erm_make_model3d(kind, nx, ny, nz, dx, dy, dz, seed, _fptr(vp), _fptr(vs), _fptr(rho)))
    return vp, vs, rho


def write_raw(path, a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    _check(lib.erm_write_raw(path.encode(), _fptr(a), a.size))


def read_raw(path, shape):
    a = np.empty(shape, np.float32)
    _check(lib.erm_read_raw(path.encode(), _fptr(a), a.size))
    return a


class RTM3D:
    """One device-resident 3D elastic RTM context (≙ GPUdevice)."""

    def __init__(self, params: Params, device: int = 0, _borrow=None):
        self.params = params
        self._owned = _borrow is None
        if _borrow is None:
            h = ctypes.c_void_p()
            _check(lib.erm_create(_pp(params), device, ctypes.byref(h)))
        else:
            h = _borrow
        self._h = h
        nnx, nny, nnz, wl = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib.erm_get_dims(h, ctypes.byref(nnx), ctypes.byref(nny), ctypes.byref(nnz), ctypes.byref(wl))
        self.shape = (nnz.value, nny.value, nnx.value)
        self.wavelet_length = wl.value
        self.nrec = params.receiver_num_x * params.receiver_num_y

    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owned", True):
                lib.erm_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, rc):
        _check(rc, self._h)

    def load_model(self, which, vp, vs=None, rho=None):
        self._c(lib.erm_load_model(self._h, which, _fptr(np.ascontiguousarray(vp, np.float32)),
                                   _fptr(None if vs is None else np.ascontiguousarray(vs, np.float32)),
                                   _fptr(None if rho is None else np.ascontiguousarray(rho, np.float32))))

    def load_model_files(self):
        self._c(lib.erm_load_model_files(self._h))

    def set_shot(self, isx=0, isy=0, isz=0):
        self._c(lib.erm_set_shot(self._h, isx, isy, isz))

    def zero_wavefield(self):
        self._c(lib.erm_zero_wavefield(self._h))

    def zero_excitation(self):
        self._c(lib.erm_zero_excitation(self._h))

    def record_shape(self):
        p = self.params
        return (3, p.lt, p.receiver_num_y, p.receiver_num_x)

    def forward_real(self, want_record=True):
        rec = np.empty(self.record_shape(), np.float32) if want_record else None
        self._c(lib.erm_forward_real(self._h, _fptr(rec)))
        return rec

    def forward_smoothed(self):
        self._c(lib.erm_forward_smoothed(self._h))

    def backward_smoothed(self, rec=None):
        self._c(lib.erm_backward_smoothed(self._h, _fptr(None if rec is None else np.ascontiguousarray(rec, np.float32))))

    def set_record(self, rec):
        self._c(lib.erm_set_record(self._h, _fptr(np.ascontiguousarray(rec, np.float32))))

    def get_record(self):
        rec = np.empty(self.record_shape(), np.float32)
        self._c(lib.erm_get_record(self._h, _fptr(rec)))
        return rec

    def step(self, pass_, it_begin, nsteps):
        self._c(lib.erm_rtm_step(self._h, pass_, it_begin, nsteps))

    def get_field(self, name):
        out = np.empty(self.shape, np.float32)
        self._c(lib.erm_get_field(self._h, FIELD_ID[name] if isinstance(name, str) else name, _fptr(out)))
        return out

    def set_field(self, name, a):
        a = np.ascontiguousarray(a, np.float32)
        assert a.shape == self.shape
        self._c(lib.erm_set_field(self._h, FIELD_ID[name] if isinstance(name, str) else name, _fptr(a)))

    def get_images(self):
        p = self.params
        out = [np.empty((p.nx, p.ny, p.nz), np.float32) for _ in range(3)]
        self._c(lib.erm_get_images(self._h, *[_fptr(o) for o in out]))
        return out

    def write_results(self, d):
        self._c(lib.erm_write_results(self._h, d.encode()))

    def write_excitation(self, d):
        self._c(lib.erm_write_excitation(self._h, d.encode()))

    def set_profiling(self, on=True):
        self._c(lib.erm_set_profiling(self._h, 1 if on else 0))

    def timing(self):
        a, b, d = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        n1, n2 = ctypes.c_long(), ctypes.c_long()
        self._c(lib.erm_get_timing(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n1),
                                   ctypes.byref(d), ctypes.byref(n2)))
        return {"pass_ms": a.value, "stress_ms": b.value, "stress_launches": n1.value,
                "velocity_ms": d.value, "velocity_launches": n2.value}

    def algorithmic_bytes(self, pass_):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._c(lib.erm_algorithmic_bytes(self._h, pass_, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def synchronize(self):
        self._c(lib.erm_synchronize(self._h))

    def strip_bytes(self):
        v = ctypes.c_double()
        self._c(lib.erm_strip_bytes(self._h, ctypes.byref(v)))
        return v.value

    def forward_smoothed_strips(self):
        """Source pass that also saves the boundary strips for reconstruction."""
        self._c(lib.erm_forward_smoothed_strips(self._h))

    def backward_smoothed_xcorr(self, rec=None):
        """Receiver pass with source reconstruction and cross-correlation imaging."""
        self._c(lib.erm_backward_smoothed_xcorr(self._h, _fptr(None if rec is None else np.ascontiguousarray(rec, np.float32))))

    def history_bytes(self, decim):
        b = ctypes.c_double()
        self._c(lib.erm_history_bytes(self._h, decim, ctypes.byref(b)))
        return b.value

    def forward_smoothed_history(self, decim=1):
        """Source pass storing tp and vP of the model region every decim steps."""
        self._c(lib.erm_forward_smoothed_history(self._h, decim))

    def backward_smoothed_xcorr_history(self, rec=None):
        """Receiver pass with cross-correlation imaging against the stored source steps."""
        self._c(lib.erm_backward_smoothed_xcorr_history(
            self._h, _fptr(None if rec is None else np.ascontiguousarray(rec, np.float32))))

    def get_xcorr_images(self):
        p = self.params
        out = [np.empty((p.nx, p.ny, p.nz), np.float32) for _ in range(3)]
        self._c(lib.erm_get_xcorr_images(self._h, *[_fptr(o) for o in out]))
        return out

    def reconstruct(self, it_end):
        self._c(lib.erm_reconstruct(self._h, it_end))

    def get_reconstructed_field(self, name):
        out = np.empty(self.shape, np.float32)
        self._c(lib.erm_get_reconstructed_field(self._h, FIELD_ID[name] if isinstance(name, str) else name, _fptr(out)))
        return out

    def autotune(self, steps=4):
        """Pick the fastest tile shape / z sweep depth (clobbers the wavefield)."""
        tx, ty, kc, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._c(lib.erm_autotune_shape(self._h, steps, ctypes.byref(tx), ctypes.byref(ty), ctypes.byref(kc),
                                       ctypes.byref(ms)))
        return {"tile_x": tx.value, "tile_y": ty.value, "zchunk": kc.value, "ms_per_step": ms.value}


def _check2(rc, ctx=None):
    if rc != 0:
        raise ErmError(rc, lib.erm2d_last_error(ctx).decode(errors="replace"))


class _DiagPrecondS(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("nz", ctypes.c_int), ("w", ctypes.POINTER(ctypes.c_float))]


class DiagPrecond:
    """Diagonal LSRTM preconditioner z = w * g (erm2d_precond_diag), w [nx][nz]."""

    def __init__(self, w):
        self.w = np.ascontiguousarray(w, np.float32)
        self.shape = self.w.shape
        self._s = _DiagPrecondS(self.shape[0], self.shape[1], self.w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))


class Elastic2D:
    """2D elastic / viscoelastic context (include/erm2d.h): modeling, RTM,
    and the LSRTM Born / adjoint operators."""

    def __init__(self, p: Params2D, device: int = 0):
        if ctypes.sizeof(p) != ctypes.sizeof(Params2D):
            raise TypeError("not an erm2d_params struct")
        self.p = p
        h = ctypes.c_void_p()
        _check2(lib.erm2d_create(ctypes.addressof(p), device, ctypes.byref(h)))
        self._h = h
        nx, nz = ctypes.c_int(), ctypes.c_int()
        lib.erm2d_get_dims(h, ctypes.byref(nx), ctypes.byref(nz))
        self.shape = (nx.value, nz.value)

    def close(self):
        if getattr(self, "_h", None):
            lib.erm2d_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, rc):
        _check2(rc, self._h)

    def load_model(self, which, vp, vs, rho, qp=None, qs=None):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (vp, vs, rho, qp, qs)]
        self._c(lib.erm2d_load_model(self._h, which, *[_fptr(a) for a in arrs]))

    def set_shot(self, sx, sz):
        self._c(lib.erm2d_set_shot(self._h, sx, sz))

    def zero(self):
        self._c(lib.erm2d_zero(self._h))

    def step(self, pass_, it_begin, nsteps):
        self._c(lib.erm2d_step(self._h, pass_, it_begin, nsteps))

    def get_field(self, name):
        out = np.empty(self.shape, np.float32)
        self._c(lib.erm2d_get_field(self._h, FIELD2D_ID[name] if isinstance(name, str) else name, _fptr(out)))
        return out

    def set_field(self, name, a):
        a = np.ascontiguousarray(a, np.float32)
        assert a.shape == self.shape
        self._c(lib.erm2d_set_field(self._h, FIELD2D_ID[name], _fptr(a)))

    def record_shape(self):
        return (2, self.p.lt, self.p.nrec)

    def forward(self):
        rec = np.empty(self.record_shape(), np.float32)
        self._c(lib.erm2d_forward(self._h, _fptr(rec)))
        return rec

    def rtm(self, rec=None):
        p = self.p
        out = [np.empty((p.nx, p.nz), np.float32) for _ in range(3)]
        r = None if rec is None else np.ascontiguousarray(rec, np.float32)
        self._c(lib.erm2d_rtm(self._h, _fptr(r), *[_fptr(o) for o in out]))
        return out

    def born(self, dm):
        d = np.empty(self.record_shape(), np.float32)
        self._c(lib.erm2d_born(self._h, _fptr(np.ascontiguousarray(dm, np.float32)), _fptr(d)))
        return d

    def adjoint(self, r):
        p = self.p
        g = np.empty((3, p.nx, p.nz), np.float32)
        self._c(lib.erm2d_adjoint(self._h, _fptr(np.ascontiguousarray(r, np.float32)), _fptr(g)))
        return g

    def rtm_xcorr(self, rec, strategy=0):
        """Cross-correlation RTM (erm2d_rtm_xcorr) with source strategy 0
        (history), 1 (reversed propagation) or 2 (boundary strips); images
        [5][nx][nz]: PP, PS, PPd, PSc, PSc polarity-corrected."""
        p = self.p
        out = np.empty((5, p.nx, p.nz), np.float32)
        r = None if rec is None else np.ascontiguousarray(rec, np.float32)
        self._c(lib.erm2d_rtm_xcorr(self._h, _fptr(r), strategy, _fptr(out)))
        return out

    def adjoint_ref(self, r):
        """The reference's continuous-adjoint equations and gradient
        (erm2d_adjoint_ref): (g_lambda, g_mu, g_rho) [3][nx][nz]."""
        p = self.p
        g = np.empty((3, p.nx, p.nz), np.float32)
        self._c(lib.erm2d_adjoint_ref(self._h, _fptr(np.ascontiguousarray(r, np.float32)), _fptr(g)))
        return g

    def set_workers(self, n):
        """Run the shots of lsrtm_cg on n concurrent contexts/streams."""
        self._c(lib.erm2d_set_workers(self._h, n))

    def set_batch(self, nb):
        """Batched shots (erm2d_set_batch): born_shots / adjoint_shots and
        lsrtm_cg run nb shots per launch (1: one shot per launch)."""
        self._c(lib.erm2d_set_batch(self._h, nb))

    @staticmethod
    def _shots(shots):
        a = np.ascontiguousarray(np.asarray(shots, np.int32).reshape(-1, 2))
        return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))

    def born_shots(self, shots, dm):
        """Born records of every shot [(x, z), ...]: [nshots][2][lt][nrec]."""
        a, ptr = self._shots(shots)
        d = np.empty((len(a),) + tuple(self.record_shape()), np.float32)
        self._c(lib.erm2d_born_shots(self._h, len(a), ptr, _fptr(np.ascontiguousarray(dm, np.float32)), _fptr(d)))
        return d

    def adjoint_shots(self, shots, r):
        """Sum over shots of the adjoint of residual records r [nshots][2][lt][nrec]."""
        a, ptr = self._shots(shots)
        p = self.p
        g = np.empty((3, p.nx, p.nz), np.float32)
        self._c(lib.erm2d_adjoint_shots(self._h, len(a), ptr, _fptr(np.ascontiguousarray(r, np.float32)), _fptr(g)))
        return g

    def angle_gathers(self, nang=30, dang=3.0):
        """Angle-domain gathers of the last rtm() (erm2d_angle_gathers):
        (pp [nang][nx][nz], ps [nang][nx][nz], angle-compensated pp and ps
        [nx][nz], reflection angle [nx][nz] in degrees); bins of dang degrees
        centred on the vertical."""
        p = self.p
        pp = np.empty((nang, p.nx, p.nz), np.float32)
        ps = np.empty_like(pp)
        ppc = np.empty((p.nx, p.nz), np.float32)
        psc = np.empty_like(ppc)
        gam = np.empty_like(ppc)
        self._c(lib.erm2d_angle_gathers(self._h, nang, dang, _fptr(pp), _fptr(ps), _fptr(ppc), _fptr(psc),
                                        _fptr(gam)))
        return pp, ps, ppc, psc, gam

    def illumination(self, shots):
        """Source illumination of the background model summed over shots
        [(x, z), ...] (erm2d_illumination): [nx][nz] float32."""
        p = self.p
        shots_a = np.ascontiguousarray(np.asarray(shots, np.int32).reshape(-1, 2))
        out = np.empty((p.nx, p.nz), np.float32)
        self._c(lib.erm2d_illumination(self._h, len(shots_a), shots_a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                       _fptr(out)))
        return out

    def lsrtm_cg(self, shots, dobs, iters, pmask=3, m0=None, tol=0.0, precond=None, monitor=None, weights=None):
        """CGLS over shots [(x, z), ...] with observed data [nshots][2][lt][nrec].
        precond: a UNet (applied through erm_unet_precond, to weights * g when
        weights [nx][nz] are given), a DiagPrecond (erm2d_precond_diag) or a
        python callable g -> z on [3][nx][nz] float32 arrays.  monitor:
        optional callable (it, m, g, misfit) called with copies of the model
        and gradient.  Returns (m, misfit history, seconds history, iterations)."""
        p = self.p
        shots_a = np.ascontiguousarray(np.asarray(shots, np.int32).reshape(-1, 2))
        d = np.ascontiguousarray(dobs, np.float32)
        m = np.zeros((3, p.nx, p.nz), np.float32) if m0 is None else np.ascontiguousarray(m0, np.float32).copy()
        mis = np.zeros(iters + 1, np.float32)
        sec = np.zeros(iters + 1, np.float64)
        shape = (3, p.nx, p.nz)
        n = 3 * p.nx * p.nz

        def view(ptr):
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_float)), shape=(n,)).reshape(shape)

        cb, user = None, None
        if isinstance(precond, UNet):
            _checku(lib.erm_unet_bind_model(precond._h, p.nx, p.nz))
            if weights is not None:
                w = np.ascontiguousarray(weights, np.float32)
                assert w.shape == (p.nx, p.nz)
                _checku(lib.erm_unet_bind_weights(precond._h, _fptr(w)))
            cb, user = ctypes.cast(lib.erm_unet_precond, ctypes.c_void_p), precond._h
        elif isinstance(precond, DiagPrecond):
            if precond.shape != (p.nx, p.nz):
                raise ValueError("DiagPrecond shape does not match the model")
            cb, user = ctypes.cast(lib.erm2d_precond_diag, ctypes.c_void_p), ctypes.addressof(precond._s)
        elif precond is not None:
            def _cb(gp, zp, _user):
                view(zp)[...] = precond(view(gp).copy())
                return 0
            cb = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)(_cb)
        mon = None
        if monitor is not None:
            def _mon(it, mp, gp, mis_, _user):
                monitor(it, view(mp).copy(), view(gp).copy(), mis_)
                return 0
            mon = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                   ctypes.c_void_p)(_mon)
        self._c(lib.erm2d_set_cg_monitor(self._h, ctypes.cast(mon, ctypes.c_void_p) if mon else None, None))
        try:
            it = lib.erm2d_lsrtm_cg(self._h, len(shots_a), shots_a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                    _fptr(d), iters, pmask, _fptr(m), _fptr(mis),
                                    sec.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                    ctypes.cast(cb, ctypes.c_void_p) if isinstance(cb, ctypes._CFuncPtr) else cb,
                                    user, tol)
        finally:
            lib.erm2d_set_cg_monitor(self._h, None, None)
        if it < 0:
            msg = lib.erm2d_last_error(self._h).decode(errors="replace")
            if isinstance(precond, UNet):
                msg += " / " + lib.erm_unet_last_error().decode(errors="replace")
            raise ErmError(it, msg)
        return m, mis[:it + 1], sec[:it + 1], it

    def last_ms(self):
        v = ctypes.c_double()
        self._c(lib.erm2d_last_ms(self._h, ctypes.byref(v)))
        return v.value


class Window2D(ctypes.Structure):
    """Mirror of `erm2d_window`."""
    _fields_ = [("shot_x", ctypes.c_int), ("shot_z", ctypes.c_int), ("x0", ctypes.c_int), ("width", ctypes.c_int)]


def rtm_windows(p, true, background, windows, records=None, z2=-1, device=0):
    """Windowed multi-shot 2D RTM over a full model (erm2d_rtm_windows):
    true / background = (vp, vs, rho[, qp, qs]) [nx][nz]; windows =
    [(shot_x, shot_z, x0, width), ...]; records: None (modelled) or the shots'
    records concatenated.  Returns stacked (tp, pp, ps, illumination) [nx][nz]."""
    if ctypes.sizeof(p) != ctypes.sizeof(Params2D):
        raise TypeError("not an erm2d_params struct")
    arrs = []
    for m in (true, background):
        m = list(m) + [None] * (5 - len(m))
        arrs += [None if a is None else np.ascontiguousarray(a, np.float32) for a in m]
    w = (Window2D * len(windows))(*[Window2D(*map(int, t)) for t in windows])
    rec = None if records is None else np.ascontiguousarray(np.concatenate([np.ravel(r) for r in records]), np.float32)
    outs = [np.zeros((p.nx, p.nz), np.float32) for _ in range(4)]
    _check2(lib.erm2d_rtm_windows(ctypes.addressof(p), device, *[_fptr(a) for a in arrs], len(windows),
                                  ctypes.cast(w, ctypes.c_void_p), _fptr(rec), int(z2), *[_fptr(o) for o in outs]))
    return outs


def _checku(rc):
    if rc != 0:
        raise ErmError(rc, lib.erm_unet_last_error().decode(errors="replace"))


class UNet:
    """U-Net preconditioner for LSRTM (include/erm_unet.h): BF16-MFMA
    convolutions, FP32 master weights, Adam training on the device.
    Images are [n][2][h][w] float32 with h, w multiples of 8."""

    def __init__(self, channels=(16, 32, 64, 64), seed=0, device=0):
        h = ctypes.c_void_p()
        _checku(lib.erm_unet_create(device, *[int(c) for c in channels], seed, ctypes.byref(h)))
        self._h = h
        self.channels = tuple(channels)
        self.num_params = lib.erm_unet_num_params(h)

    def close(self):
        if getattr(self, "_h", None):
            lib.erm_unet_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def get_params(self):
        out = np.empty(self.num_params, np.float32)
        _checku(lib.erm_unet_get_params(self._h, _fptr(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params
        _checku(lib.erm_unet_set_params(self._h, _fptr(p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        n, c, hh, ww = x.shape
        assert c == 2
        y = np.empty_like(x)
        _checku(lib.erm_unet_forward(self._h, _fptr(x), n, hh, ww, _fptr(y)))
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        assert x.shape == t.shape and x.shape[1] == 2
        n, _, hh, ww = x.shape
        loss = np.zeros(1, np.float32)
        g = np.empty(self.num_params, np.float32)
        _checku(lib.erm_unet_loss_grad(self._h, _fptr(x), _fptr(t), n, hh, ww, _fptr(loss), _fptr(g)))
        return float(loss[0]), g

    def train_step(self, x, t, lr):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        n, _, hh, ww = x.shape
        loss = np.zeros(1, np.float32)
        _checku(lib.erm_unet_train_step(self._h, _fptr(x), _fptr(t), n, hh, ww, lr, _fptr(loss)))
        return float(loss[0])

    def train(self, x, t, batch, epochs, lr, seed=0):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        count, _, hh, ww = x.shape
        losses = np.zeros(epochs, np.float32)
        _checku(lib.erm_unet_train(self._h, _fptr(x), _fptr(t), count, hh, ww, batch, epochs, lr, seed,
                                   _fptr(losses)))
        return losses

    def last_stats(self):
        ms, fl = ctypes.c_double(), ctypes.c_double()
        _checku(lib.erm_unet_last_stats(self._h, ctypes.byref(ms), ctypes.byref(fl)))
        return ms.value, fl.value


class ShotWindow(ctypes.Structure):
    """erm_shot_window: source and lateral window in full-model cells."""
    _fields_ = [(n, ctypes.c_int) for n in ("sx", "sy", "sz", "x0", "y0", "wx", "wy")]


def _checkb(rc, h=None):
    if rc != 0:
        raise ErmError(rc, lib.erm_batch_last_error(h).decode(errors="replace"))


class Batch3D:
    """Batched multi-shot 3D RTM (include/erm_batch.h): all shots' windows
    step together, one launch per kernel per step, captured in a hipGraph."""

    def __init__(self, params: Params, device: int = 0):
        self.params = params
        h = ctypes.c_void_p()
        _checkb(lib.erm_batch_create(_pp(params), device, ctypes.byref(h)))
        self._h = h
        self.windows = []

    def close(self):
        if getattr(self, "_h", None):
            lib.erm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, rc):
        _checkb(rc, self._h)

    _PROGRESS = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_double, ctypes.c_void_p)

    def set_progress(self, fn):
        """fn(pass, ms) after every completed pass of run() (None removes it)."""
        self._progress_cb = None if fn is None else self._PROGRESS(lambda ps, ms, user: fn(ps, ms))
        self._c(lib.erm_batch_set_progress(self._h, ctypes.cast(self._progress_cb, ctypes.c_void_p)
                                           if self._progress_cb else None, None))

    def load_model(self, which, vp, vs, rho):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (vp, vs, rho)]
        self._c(lib.erm_batch_load_model(self._h, which, *[_fptr(a) for a in arrs]))

    def set_shots(self, windows):
        """windows: iterable of (sx, sy, sz, x0, y0, wx, wy)."""
        arr = (ShotWindow * len(windows))(*[ShotWindow(*map(int, w)) for w in windows])
        self._c(lib.erm_batch_set_shots(self._h, ctypes.cast(arr, ctypes.c_void_p), len(windows)))
        self.windows = [tuple(map(int, w)) for w in windows]

    def set_launch(self, tile_y, zchunk, grid_blocks=0, tile_x=64):
        self._c(lib.erm_batch_set_launch_shape(self._h, tile_x, tile_y, zchunk, grid_blocks))

    def autotune(self, steps=4):
        tx, ty, kc, gr, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._c(lib.erm_batch_autotune_shape(self._h, steps, ctypes.byref(tx), ctypes.byref(ty), ctypes.byref(kc),
                                             ctypes.byref(gr), ctypes.byref(ms)))
        return {"tile_x": tx.value, "tile_y": ty.value, "zchunk": kc.value, "grid_blocks": gr.value,
                "ms_per_step": ms.value}

    def use_graph(self, on=True):
        self._c(lib.erm_batch_use_graph(self._h, 1 if on else 0))

    def set_record(self, shot, rec):
        self._c(lib.erm_batch_set_record(self._h, shot, _fptr(np.ascontiguousarray(rec, np.float32))))

    def run(self, want_images=True):
        p = self.params
        out = [np.empty((p.nx, p.ny, p.nz), np.float32) for _ in range(3)] if want_images else [None] * 3
        self._c(lib.erm_batch_run(self._h, *[_fptr(o) for o in out]))
        return out

    def shot(self, s):
        """RTM3D view (not owned) of shot s's window context."""
        h = ctypes.c_void_p()
        self._c(lib.erm_batch_shot_context(self._h, s, ctypes.byref(h)))
        sx, sy, sz, x0, y0, wx, wy = self.windows[s]
        q = Params()
        ctypes.memmove(ctypes.addressof(q), ctypes.addressof(self.params), ctypes.sizeof(Params))
        q.nx, q.ny = wx, wy
        rix, riy = max(1, q.receiver_interval_x), max(1, q.receiver_interval_y)
        q.receiver_num_x, q.receiver_num_y = (wx + rix - 1) // rix, (wy + riy - 1) // riy
        return RTM3D(q, _borrow=h)

    def timing(self):
        d = [ctypes.c_double() for _ in range(4)]
        n = [ctypes.c_int() for _ in range(3)]
        self._c(lib.erm_batch_timing(self._h, *[ctypes.byref(v) for v in d], *[ctypes.byref(v) for v in n]))
        return {"ms_model": d[0].value, "ms_source": d[1].value, "ms_receiver": d[2].value,
                "cells_per_step": d[3].value, "steps_model": n[0].value, "steps_source": n[1].value,
                "steps_receiver": n[2].value}
