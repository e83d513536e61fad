"""Parameter struct of the 2D C ABI (include/erm2d.h), without loading the library."""
import ctypes

FIELDS2D = ["vx", "vz", "txx", "tzz", "txz", "tp", "vxp", "vzp", "div", "curl",
            "ex_time", "ex_amp", "img_tp", "img_pp", "img_ps", "ex_tp", "ex_vxp", "ex_vzp", "r_vxp", "r_vzp",
            "ex_tp_time", "ex_tp_amp", "pdiv", "scurl"]
FIELD2D_ID = {n: i for i, n in enumerate(FIELDS2D)}
# source-wavefield strategies (ERM2D_SRC_*): erm2d_rtm_xcorr's strategy and
# erm2d_params.adjoint_source
SRC_HISTORY, SRC_REVERSE, SRC_STRIPS = 0, 1, 2


class Params2D(ctypes.Structure):
    """Mirror of `erm2d_params`."""

    _fields_ = ([("nx", ctypes.c_int), ("nz", ctypes.c_int), ("dx", ctypes.c_float), ("dz", ctypes.c_float),
                 ("dt", ctypes.c_float), ("freq", ctypes.c_float), ("lt", ctypes.c_int), ("lt_forward", ctypes.c_int)]
                + [(k, ctypes.c_int) for k in ["bu", "bd", "bl", "br"]]
                + [("coe_attenuation", ctypes.c_float)]
                + [(k, ctypes.c_int) for k in ["fd_order", "nsls", "shot_x", "shot_z", "nrec", "rec_x0", "rec_dx",
                                                "rec_z"]]
                + [("qcomp", ctypes.c_int), ("ex_tp_shift", ctypes.c_int), ("ex_vzp_avg", ctypes.c_int),
                   ("boundary", ctypes.c_int), ("cpml_r", ctypes.c_float), ("cpml_vmax", ctypes.c_float),
                   ("inversion_para", ctypes.c_int), ("adjoint_mode", ctypes.c_int),
                   ("adjoint_source", ctypes.c_int), ("tile_rows", ctypes.c_int)])

    def update(self, **kw):
        for k, v in kw.items():
            if not hasattr(self, k):
                raise KeyError(k)
            setattr(self, k, v)
        return self

    @property
    def padded_shape(self):
        """(nnx, nnz): z fastest in memory."""
        return (self.nx + self.bl + self.br, self.nz + self.bu + self.bd)
