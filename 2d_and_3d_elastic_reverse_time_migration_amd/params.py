"""Parameter struct and constants of the C ABI (include/erm.h), without
loading the library, so that host-only tools can build parameter sets."""
import ctypes

PATH_MAX = 256

E_PARAM, E_IO, E_DEVICE, E_STATE, E_SIZE, E_NOMEM = -1, -2, -3, -4, -5, -6
BOUNDARY_DAMP, BOUNDARY_CPML = 0, 1
PASS_MODEL, PASS_SOURCE, PASS_RECEIVER = 0, 1, 2

FIELDS = [
    "vx", "vy", "vz", "txx", "tyy", "tzz", "txy", "txz", "tyz", "tp", "vxp", "vyp", "vzp",
    "vxs", "vys", "vzs", "ex_time", "ex_amp", "ex_tp", "ex_vxp", "ex_vyp", "ex_vzp",
    "img_tp", "img_pp", "img_ps", "div", "curlx", "curly", "curlz",
    "poyn_sx", "poyn_sy", "poyn_sz", "poyn_rpx", "poyn_rpy", "poyn_rpz", "poyn_rsx", "poyn_rsy", "poyn_rsz",
]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}

_INT_KEYS_1 = ["RTM", "GPU_switch", "GPU_N", "GPU_start", "exchange_device_bool", "lt"]
_PARAM_FIELDS = (
    [(k, ctypes.c_int) for k in _INT_KEYS_1]
    + [("dt", ctypes.c_float), ("freq", ctypes.c_float)]
    + [(k, ctypes.c_int) for k in ["join_vs", "join_den", "join_shot", "join_wavefield", "wavefield_interval"]]
    + [("infile_shot_name", ctypes.c_char * PATH_MAX)]
    + [(k, ctypes.c_int) for k in ["nx", "ny", "nz"]]
    + [(k, ctypes.c_float) for k in ["dx", "dy", "dz"]]
    + [(k, ctypes.c_int) for k in [
        "shot_num_x", "shot_num_y", "shot_num_z",
        "shot_start_x", "shot_interval_x", "shot_start_y", "shot_interval_y", "shot_start_z", "shot_interval_z",
        "receiver_num_x", "receiver_num_y", "receiver_num_z",
        "receiver_start_x", "receiver_interval_x", "receiver_start_y", "receiver_interval_y",
        "receiver_start_z", "receiver_interval_z", "receiver_offset", "receiver_mark",
        "bu", "bd", "bl", "br", "bf", "bb"]]
    + [(k, ctypes.c_float) for k in ["coe_attenuation", "velocity_multiple", "smooth_time_vp",
                                      "smooth_time_vs", "smooth_time_density"]]
    + [(k, ctypes.c_char * PATH_MAX) for k in ["velocity", "velocity1", "density",
                                                "s_velocity", "s_velocity1", "s_density"]]
    + [(k, ctypes.c_int) for k in ["add_receiver_bool", "cut_direct_wave", "laplace", "laplace_compensate",
                                    "vsp", "vsp_precon", "precon", "precon_z1", "precon_z2",
                                    "fd_order", "boundary", "lt_forward"]]
    + [("cpml_r", ctypes.c_float), ("cpml_vmax", ctypes.c_float)]
    + [(k, ctypes.c_int) for k in ["tile_x", "tile_y", "zchunk", "pitch_align", "xcd_map", "cpml_split",
                                   "region_streams", "poynting"]]
    + [("reserved", ctypes.c_int * 3)]
)


class Params(ctypes.Structure):
    """Mirror of `erm_params` (reference getpar keys + extensions)."""

    _fields_ = _PARAM_FIELDS

    def update(self, **kw) -> "Params":
        for k, v in kw.items():
            if not hasattr(self, k):
                raise KeyError(k)
            if isinstance(v, str):
                v = v.encode()
            setattr(self, k, v)
        return self

    @property
    def padded_shape(self):
        """(nnz, nny, nnx): the padded grid, x fastest in memory."""
        return (self.nz + self.bu + self.bd, self.ny + self.bf + self.bb, self.nx + self.bl + self.br)
