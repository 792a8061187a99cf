"""ctypes binding of libmpdata_hip.so (C ABI in include/mpdata_hip.h).

The shared library is built in-tree by ``__graft_entry__.build()``
(``make -C advectionhpctester_amd/csrc``).  Importing this module without it
raises: there is no CPU fallback for the product path.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# MPDATA_HIP_LIB names a library built elsewhere (same C ABI)
LIB_PATH = os.environ.get("MPDATA_HIP_LIB") or os.path.join(HERE, "libmpdata_hip.so")

MPDATA_OK, MPDATA_EINVAL, MPDATA_EHIP, MPDATA_ENOMEM, MPDATA_EUNSUPPORTED = 0, 1, 2, 3, 4
MPDATA_ESTOP = 5
UPWIND, GAUGE, STANDARD = 0, 1, 2
PATH_AUTO, PATH_STAGES, PATH_FUSED, PATH_FAST = 0, 1, 2, 3

_P, _I, _D, _F, _LL = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_longlong


class MpdataError(RuntimeError):
    """Non-zero status from libmpdata_hip; ``code`` is the MPDATA_E* value."""

    def __init__(self, code, msg):
        super().__init__(f"mpdata error {code}: {msg}")
        self.code = code


def _proto(real):
    """Argument lists per entry point (without the _f64/_f32 suffix)."""
    R = real
    IP, RP = ctypes.POINTER(_I), ctypes.POINTER(real)
    return {
        "mpdata_upwind3d_gpubc": [_P] * 12 + [_I] * 8 + [R, _P],
        "mpdata_antidiff3d_gauge_gpubc": [_P, _I] + [_P] * 7 + [_I] * 9 + [_P],
        "mpdata_antidiff3d_standard_gpubc": [_P, _I] + [_P] * 7 + [_I] * 9 + [_P],
        "mpdata_advec_dwarf": [_P] + [_I] * 12 + [_P] * 9,
        "mpdata_step": [_P, _I, _I, _I] + [_P] * 9,
        "mpdata_step_oop": [_P, _I, _I, _I] + [_P] * 10,
        "mpdata_c_timeloop": [_P, _I, _I, _I, _I] + [_P] * 9,
        "mpdata_velprd_traj0": [_P] * 8 + [R] * 5 + [_P],
        "mpdata_velprdA_to_C": [_P] * 8,
        "mpdata_velprd_c": [_P] * 8 + [R] * 5 + [_P] * 4,
        "mpdata_velprd_rotate": [_P] * 8 + [R, _P],
        "mpdata_compute_courants": [_P] * 5 + [_I] + [R] * 4 + [_P, _P],
        "mpdata_compute_courants_noalloc": [_P] * 4 + [_I] + [R] * 4 + [_I] * 4 + [_P, _P],
        "mpdata_norm_sums": [_P] * 5,
        "mpdata_ce_advection": [_P, _I, _I, _I] + [_P] * 4 + [_P] * 5 + [_P],
        "mpdata_ce_advection_oop": [_P, _I, _I, _I] + [_P] * 5 + [_P] * 5 + [_P],
        "mpdata_velprd_traj0_adapt_dt": [_P] * 5 + [R] * 5 + [_P],
        "mpdata_velprd_driver": [_P, IP, IP, _I, RP, RP, _I, _I, _I, R, R] + [_P] * 10 +
                                [_I] * 8 + [R] * 4 + [_P],
        "mpdata_rho_gauge_core": [_P, _I] + [_P] * 13 + [_P],
        "mpdata3d_rho_gauge_gpubc": [_P, _I, IP, IP, _I, RP, RP, _I, _I, _I, R, R] + [_P] * 19 +
                                    [_I] * 8 + [R] * 4 + [_I, _P],
    }


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(LIB_PATH)
    lib.mpdata_version.restype = ctypes.c_char_p
    lib.mpdata_last_error.restype = ctypes.c_char_p
    lib.mpdata_ctx_create.argtypes = [_I] * 8 + [ctypes.POINTER(_P)]
    lib.mpdata_ctx_destroy.argtypes = [_P]
    lib.mpdata_ctx_set_path.argtypes = [_P, _I]
    lib.mpdata_ctx_scratch_bytes.argtypes = [_P]
    lib.mpdata_ctx_scratch_bytes.restype = _LL
    lib.upwind_c_timeloop.argtypes = [_P] * 6 + [_I] * 5
    IP = ctypes.POINTER(_I)
    lib.mpdata_ens_create.argtypes = [_I, IP, IP, IP, _I, _I, ctypes.POINTER(_P)]
    lib.mpdata_ens_destroy.argtypes = [_P]
    lib.mpdata_ens_set_fields_f64.argtypes = [_P] + [ctypes.POINTER(_P)] * 9
    lib.mpdata_ens_set_fields_f32.argtypes = [_P] + [ctypes.POINTER(_P)] * 9
    lib.mpdata_ens_nconfigs.argtypes = []
    lib.mpdata_ens_config_name.argtypes = [_I]
    lib.mpdata_ens_config_name.restype = ctypes.c_char_p
    lib.mpdata_ens_set_tiling.argtypes = [_P, _I, _I]
    lib.mpdata_ens_info.argtypes = [_P, IP, IP, IP]
    lib.mpdata_ens_set_path.argtypes = [_P, _I]
    lib.mpdata_ens_autotune.argtypes = [_P, _I, _P, IP, IP, ctypes.POINTER(_D)]
    lib.mpdata_ens_step.argtypes = [_P, _I, _P]
    lib.mpdata_ens_run.argtypes = [_P, _I, _I, _P]
    lib.mpdata_tl_step_f64.argtypes = [_P, _I, _I, _I] + [_P] * 12 + [_P]
    lib.mpdata_adj_reserve.argtypes = [_P, _I]
    lib.mpdata_adj_record_f64.argtypes = [_P, _I] + [_P] * 9 + [_P]
    lib.mpdata_adj_record_variant_f64.argtypes = [_P, _I, _I] + [_P] * 9 + [_P]
    lib.mpdata_adj_step_f64.argtypes = [_P, _I] + [_P] * 9 + [_D, _P]
    lib.mpdata_dev_alloc.argtypes = [_LL, ctypes.POINTER(_P)]
    lib.mpdata_dev_free.argtypes = [_P]
    lib.mpdata_copy.argtypes = [_P, _P, _LL, _I]
    lib.mpdata_sync.argtypes = [_P]
    lib.mpdata_ttini.argtypes = []
    lib.mpdata_ttbeg.argtypes = [_I, _P]
    lib.mpdata_ttend.argtypes = [_I, _P]
    lib.mpdata_ttget.argtypes = [_I, ctypes.POINTER(_I), ctypes.POINTER(_D)]
    lib.mpdata_ttprt.argtypes = []
    lib.mpdata_serializer_open.argtypes = [ctypes.c_char_p]
    lib.mpdata_serializer_close.argtypes = []
    for sfx in ("_f64", "_f32"):
        fn = getattr(lib, "mpdata_serialize_field" + sfx)
        fn.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _P, _I, _I, _I, _P]
        fn.restype = _I
    for sfx, real in (("_f64", _D), ("_f32", _F)):
        for name, args in _proto(real).items():
            fn = getattr(lib, name + sfx)
            fn.argtypes = args
            fn.restype = _I
    return lib


LIB = _load()


def check(rc):
    """Raise MpdataError for a non-zero status."""
    if rc != MPDATA_OK:
        raise MpdataError(rc, LIB.mpdata_last_error().decode())
    return rc


def exported_symbols():
    """Names of the functions include/mpdata_hip.h declares."""
    names = ["mpdata_version", "mpdata_last_error", "mpdata_ctx_create", "mpdata_ctx_destroy",
             "mpdata_ctx_set_path", "mpdata_ctx_scratch_bytes", "upwind_c_timeloop",
             "mpdata_serializer_open", "mpdata_serializer_close", "mpdata_serialize_field_f64",
             "mpdata_serialize_field_f32", "mpdata_ttini", "mpdata_ttbeg", "mpdata_ttend",
             "mpdata_ttget", "mpdata_ttprt", "mpdata_addr", "mpdata_is_device_ptr"]
    for base in _proto(_D):
        names += [base + "_f64", base + "_f32"]
    return names
