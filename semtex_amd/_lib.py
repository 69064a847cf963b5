"""ctypes binding of libsemtex_amd.so (the C ABI in include/semtex_amd.h).

The product path has no CPU fallback: if the HIP library is missing or
fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsemtex_amd.so")

SX_OK = 0
SX_ERR_ARG, SX_ERR_HIP, SX_ERR_GEOM, SX_ERR_FACTOR, SX_ERR_LIMIT, SX_ERR_SYNC = -1, -2, -3, -4, -5, -6
SX_FP64, SX_MIXED, SX_JACPCG = 0, 1, 2
SX_SKEW, SX_ALTSKEW = 0, 1
SX_FORWARD, SX_INVERSE = 1, -1
SX_NOALLOC_DOUBLES = 10 * 1024 * 1024   # fixed workspace of sx_field_solve_mixed_noalloc

# Every symbol include/semtex_amd.h declares, with (restype, argtypes).
_vp, _i, _d, _ip, _dp = C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_void_p
SIGNATURES = {
    "sx_last_error": (C.c_char_p, []),
    "sx_device_info": (_i, [C.c_char_p, _i]),
    "sx_gll": (_i, [_i, _dp, _dp, _dp]),
    "sx_mesh_create": (_i, [_i, _i, _i, _dp, _dp, _ip, _ip, _ip, _i, _i, _i,
                            _i, _ip, _ip, _ip, C.POINTER(_vp)]),
    "sx_mesh_destroy": (_i, [_vp]),
    "sx_mesh_nclass": (_i, [_vp]),
    "sx_msys_create": (_i, [_vp, _d, _d, _i, _i, _i, _vp, C.POINTER(_vp)]),
    "sx_msys_create_named": (_i, [_vp, _d, _d, _i, _i, C.c_char, _i, _vp, C.POINTER(_vp)]),
    "sx_msys_destroy": (_i, [_vp]),
    "sx_couple": (_i, [_i, _i, _i, _dp, _dp, _i, _i, _vp]),
    "sx_msys_nsolve": (_i, [_vp, _i]),
    "sx_msys_layout": (_i, [_vp, C.POINTER(C.c_int), _i]),
    "sx_msys_status": (_i, [_vp, _vp]),
    "sx_field_solve": (_i, [_vp, _i, _i, _dp, _dp, _dp, _vp]),
    "sx_field_solve_mixed": (_i, [_vp, _i, _i, _dp, _dp, _dp, _d, _i,
                                  C.POINTER(_i), _vp]),
    "sx_field_solve_mixed_noalloc": (_i, [_vp, _i, _i, _dp, _dp, _dp, _d, _i,
                                          C.POINTER(_i), _vp]),
    "sx_field_solve_host": (_i, [_vp, _i, _i, _dp, _dp, _dp]),
    "sx_helmholtz_op": (_i, [_vp, _i, _i, _dp, _dp, _vp]),
    "sx_fourier_transform": (_i, [_i, _i, _dp, _dp, _i, _vp]),
    "sx_dns_set_startup": (_i, [_vp, _i, _d, _i, _vp]),
    "sx_dns_set_velocity_bc": (_i, [_vp, _dp, _vp]),
    "sx_dns_set_dealias": (_i, [_vp, _i, _vp]),
    "sx_field_solve_weak": (_i, [_vp, _i, _i, _dp, _dp, _vp]),
    "sx_batch_create": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, C.POINTER(_vp)]),
    "sx_batch_run": (_i, [_vp, _vp]),
    "sx_batch_destroy": (_i, [_vp]),
    "sx_batch_tune": (_i, [_vp, _i, _vp]),
    "sx_batch_info": (_i, [_vp, C.c_char_p, _i]),
    "sx_mass_inner": (_i, [_vp, _i, _i, _dp, _dp, _dp, _vp]),
    "sx_mass_weight": (_i, [_vp, _i, _i, _dp, _dp, _d, _vp]),
    "sx_grad_planes": (_i, [_vp, _i, _i, _dp, _dp, _i, _i, _vp]),
    "sx_smooth_planes": (_i, [_vp, _i, _i, _dp, _dp, _i, _vp]),
    "sx_dz_planes": (_i, [_i, _i, _i, _d, _dp, _dp, _i, _vp]),
    "sx_field_solve_pcg": (_i, [_vp, _i, _i, _dp, _dp, _dp, _d, _i, C.POINTER(_i), _vp]),
    "sx_dns_create": (_i, [_vp, _vp, _i, _d, _d, _d, _dp, _i, _i, _vp, C.POINTER(_vp)]),
    "sx_dns_destroy": (_i, [_vp]),
    "sx_dns_set_state": (_i, [_vp, _dp, _vp]),
    "sx_dns_step": (_i, [_vp, _i, _vp]),
    "sx_dns_get": (_i, [_vp, _i, _dp, C.POINTER(_i), _vp]),
    "sx_dns_plane_stride": (_i, [_vp]),
    "sx_field_write": (_i, [C.c_char_p, _vp, _vp, _i, _i]),
    "sx_field_read_header": (_i, [C.c_char_p, _i, _vp]),
    "sx_field_read": (_i, [C.c_char_p, _i, C.c_char, _dp, _i]),
    "sx_enumerate": (_i, [_i, _i, _vp, _i, _vp, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "sx_num_write": (_i, [C.c_char_p, C.c_char_p, _i, _vp, _i, _i, _vp, _vp, _vp, _vp, _vp,
                          _vp]),
    "sx_num_read": (_i, [C.c_char_p, _i, _i, _i, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sx_timing_enable": (_i, [_i]),
    "sx_timing_read": (_i, [_dp, _ip, _i]),
}
SX_NPHASE = 5
PHASES = ("essential", "condense", "assemble", "band_solve", "recover")

_lib = None


class SxError(RuntimeError):
    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


def load():
    """Load the HIP library; raises if it is absent (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SxError(f"{LIB_PATH} is missing: run `python -m semtex_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != SX_OK:
        msg = load().sx_last_error().decode()
        raise SxError(f"{what} failed ({rc}): {msg}", rc)
    return rc
