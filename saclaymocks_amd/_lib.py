"""ctypes binding of libsaclay_hip.so (the C ABI declared in include/saclaymocks_hip.h).

The library is loaded lazily, once.  A missing library is an error: there is
no CPU fallback for any compute entry point.
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsaclay_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "saclaymocks_hip.h")

# layout of the structs passed by pointer (smk_skewers, smk_box_view): must
# equal SMK_ABI_VERSION of include/saclaymocks_hip.h and the library's
ABI_VERSION = 9

SMK_OK = 0
SMK_ERR_ARG = -1
SMK_ERR_OOM = -2
SMK_ERR_NULLBOX = -3
SMK_ERR_HIP = -4
SMK_ERR_UNSUPPORTED = -5
SMK_ERR_RANGE = -6

AUX_KINDS = {
    None: 0, "none": 0,
    "eta_xx": 1, "eta_yy": 2, "eta_zz": 3, "eta_xy": 4, "eta_xz": 5, "eta_yz": 6,
    "vx": 7, "vy": 8, "vz": 9,
}


class SmkError(RuntimeError):
    """A libsaclay_hip call failed (HIP runtime, allocation)."""


_lib = None

_c = ctypes
_vp = _c.c_void_p
_i = _c.c_int
_u64 = _c.c_uint64
_d = _c.c_double
_sz = _c.c_size_t

# name -> (restype, argtypes)
_SIGNATURES = {
    "smk_abi_version": (_i, []),
    "smk_last_error": (_c.c_char_p, []),
    "smk_fft_length_supported": (_i, [_i]),
    "smk_box_plan_create": (_i, [_c.POINTER(_vp), _i, _i, _i, _d]),
    "smk_box_plan_destroy": (_i, [_vp]),
    "smk_box_workspace_bytes": (_sz, [_vp]),
    "smk_grf_box": (_i, [_vp, _u64, _vp, _c.POINTER(_d), _i, _d, _vp, _vp, _vp, _vp, _vp, _vp]),
    "smk_draw_grf_boxk": (_i, [_vp, _u64, _vp, _vp, _vp]),
    "smk_box_enable_pass_timing": (_i, [_vp, _i]),
    "smk_box_pass_times": (_i, [_vp, _c.POINTER(_d), _i, _c.POINTER(_i), _i]),
    "smk_box_set_schedule": (_i, [_vp, _i]),
    "smk_box_get_schedule": (_i, [_vp]),
    "smk_box_variant_count": (_i, []),
    "smk_box_variant_valid": (_i, [_vp, _i, _i]),
    "smk_box_set_variant": (_i, [_vp, _i, _i]),
    "smk_box_get_variant": (_i, [_vp, _i]),
    "smk_box_autotune": (_i, [_vp, _u64, _vp, _c.POINTER(_d), _i, _d, _vp, _vp, _vp, _vp, _vp, _vp, _i,
                              _c.POINTER(_d)]),
    "smk_fft_and_store": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "smk_fft_and_store_workspace_bytes": (_sz, [_vp]),
    "smk_forest_plan_create": (_i, [_c.POINTER(_vp)]),
    "smk_forest_plan_destroy": (_i, [_vp]),
    "smk_forest_workspace_bytes": (_sz, [_i]),
    "smk_forest_gather": (_i, [_vp, _vp, _i, _vp, _c.c_int64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "smk_forest_missing_power": (_i, [_vp, _vp, _vp, _c.POINTER(_c.c_int64), _u64, _vp, _vp, _vp, _vp, _vp]),
    "smk_forest_plan_set_graph_mode": (_i, [_vp, _vp]),
    "smk_forest_missing_power_sets": (_i, [_vp, _vp, _vp, _c.POINTER(_c.c_int64), _c.c_int64, _c.c_int32, _vp,
                                           _vp, _u64, _vp, _vp, _vp, _vp, _vp]),
    "smk_fgpa_flux": (_i, [_vp, _vp, _vp, _vp, _c.c_int64, _vp, _vp, _vp]),
    "smk_forest_step_workspace_bytes": (_sz, []),
    "smk_forest_plan_set_step_variant": (_i, [_vp, _i]),
    "smk_forest_step_noalloc_scratch_bytes": (_sz, []),
    "smk_forest_step_noalloc": (_i, [_vp, _vp, _vp, _vp, _c.POINTER(_c.c_int64), _c.c_int64, _c.c_int32, _vp, _vp,
                                     _u64, _vp, _c.c_int64, _c.c_int32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "smk_forest_step": (_i, [_vp, _vp, _vp, _vp, _c.POINTER(_c.c_int64), _c.c_int64, _c.c_int32, _vp, _vp, _u64,
                             _vp, _c.c_int64, _c.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "smk_qso_draw": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _u64, _vp, _c.c_int64, _vp, _vp]),
    "smk_pk_table": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _c.c_float, _vp, _vp]),
    "smk_p1d_rsd_rows": (_i, [_vp, _vp, _i, _vp, _i, _d, _i, _vp, _d, _d, _vp, _vp]),
    "smk_reduce_work_len": (_c.c_int64, [_i, _i, _i]),
    "smk_reduce_by_set": (_i, [_vp, _vp, _i, _i, _vp, _vp, _vp, _c.c_int64, _vp]),
    "smk_p1d_binned": (_i, [_vp, _vp, _vp, _d, _i, _vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                            _c.POINTER(_c.c_int64), _vp]),
    "smk_p1d_table_len": (_c.c_int64, [_i]),
    "smk_p1d_tables": (_i, [_vp, _vp, _i, _vp, _vp]),
    "smk_p1d_adjoint": (_i, [_vp, _vp, _vp, _vp, _c.c_int64, _vp, _vp, _d, _i, _vp, _i, _i, _vp, _vp, _vp, _vp,
                             _vp, _vp, _vp, _c.POINTER(_c.c_int64), _vp]),
    "smk_fgpa_flux_jac": (_i, [_vp, _vp, _vp, _vp, _c.c_int64, _vp, _vp, _vp]),
    "smk_p1d_jacobian": (_i, [_vp, _vp, _vp, _vp, _c.c_int64, _vp, _vp, _vp, _d, _i, _vp, _i, _i, _vp, _vp, _vp,
                              _vp, _vp, _c.POINTER(_c.c_int64), _vp]),
}


def declared_symbols(header=HEADER):
    """Names of the functions the C header declares (used by the ABI test)."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(smk_[a-z0-9_]+)\s*\(", text)))


def lib():
    """Load libsaclay_hip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m saclaymocks_amd.build` "
            "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if L.smk_abi_version() != ABI_VERSION:
        raise ImportError(f"{LIB_PATH} has ABI {L.smk_abi_version()}, this binding expects {ABI_VERSION}: "
                          "rebuild it with `python -m saclaymocks_amd.build`")
    _lib = L
    return L


def check(rc, what=""):
    """Turn a negative SMK_ERR_* return code into a Python exception."""
    if rc == SMK_OK:
        return
    msg = lib().smk_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc in (SMK_ERR_ARG, SMK_ERR_RANGE, SMK_ERR_NULLBOX, SMK_ERR_UNSUPPORTED):
        raise ValueError(text)
    raise SmkError(text)


def ptr(t):
    """Device pointer of a torch tensor (or None -> NULL)."""
    if t is None:
        return None
    return _vp(t.data_ptr())


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return _vp(s.cuda_stream)
