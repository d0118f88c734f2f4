"""ctypes binding of libdmrg_hip.so (the C ABI declared in include/chemtensor_hip.h).

The shared library is built in-tree by ``chemtensor_mojo_amd.build`` (hipcc,
gfx950) and loaded from the package directory.  There is no fallback: when the
library or a GPU is missing every entry point raises.
"""
from __future__ import annotations

import atexit
import ctypes
import itertools
import os
import weakref

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdmrg_hip.so")

PREC_FP64, PREC_FP32, PREC_BF16X3, PREC_BF16X6 = 0, 1, 2, 3
F64, F32 = 0, 1
SVD_GESVD, SVD_GRAM = 0, 1

_i64 = ctypes.c_int64
_i64p = ctypes.POINTER(ctypes.c_int64)
_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_int = ctypes.c_int
_dbl = ctypes.c_double


class DMRGParamsC(ctypes.Structure):
    _fields_ = [("num_sweeps", _int), ("chi_max", _i64), ("eps_trunc", _dbl), ("max_krylov_iter", _int),
                ("krylov_tol", _dbl), ("energy_tol", _dbl), ("jacobi_max_rot", _int), ("svd_mode", _int),
                ("prec", _int), ("refine", _int),
                ("env_prec", _int), ("verbose", _int), ("graphs", _int)]


class DMRGStatsC(ctypes.Structure):
    _fields_ = [("energy", _dbl), ("sweeps", _int), ("local_updates", _int), ("trunc_weight_sum", _dbl),
                ("trunc_weight_max", _dbl), ("max_bond", _i64), ("t_total", _dbl), ("t_lanczos", _dbl),
                ("t_svd", _dbl), ("t_env", _dbl), ("t_theta", _dbl), ("matvecs", _i64),
                ("matvec_flops", _dbl), ("env_flops", _dbl), ("graph_replays", _i64), ("kept_start", _i64)]


# name -> argtypes (all return int status unless listed in _RESTYPES)
SIGNATURES = {
    "ct_handle_create": [_int, _vp, ctypes.POINTER(_vp)],
    "ct_handle_destroy": [_vp],
    "ct_handle_set_stream": [_vp, _vp],
    "ct_synchronize": [_vp],
    "ct_last_error": [],
    "ct_version": [],
    "ct_tensordot": [_vp, _int, _int, _vp, _int, _i64p, _vp, _int, _i64p, _int, _int, _int, _vp, _int, _i64p],
    "ct_heff_create": [_vp, _int, _int, _vp, _vp, _vp, _vp] + [_i64] * 8 + [_vp],
    "ct_heff_apply_prepared": [_vp, _vp, _vp],
    "ct_heff_destroy": [_vp],
    "ct_split_bf16": [_vp, _int, _vp, _i64, _i64, _i64, _i64, _vp, _i64],
    "ct_gemm_bf16x3_nt": [_vp, _vp, _i64, _vp, _i64, _int, _vp, _i64, _i64, _i64, _i64],
    "ct_split_bf16x6": [_vp, _int, _vp, _i64, _i64, _i64, _i64, _vp, _i64, _vp, _i64],
    "ct_gemm_bf16x6_nt": [_vp, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _int, _vp, _i64, _i64, _i64, _i64],
    "ct_permute": [_vp, _int, _vp, _int, _i64p, _ip, _vp],
    "ct_norm_sq": [_vp, _int, _vp, _i64, _dp],
    "ct_dot": [_vp, _int, _vp, _vp, _i64, _dp],
    "ct_axpy": [_vp, _int, _dbl, _vp, _vp, _i64],
    "ct_scale": [_vp, _int, _dbl, _vp, _i64],
    "ct_convert": [_vp, _int, _vp, _int, _vp, _i64],
    "ct_heff_apply": [_vp, _int, _int, _vp, _vp, _vp, _vp, _vp, _vp] + [_i64] * 8,
    "ct_heff_apply_noalloc": [_vp, _int, _int, _vp, _vp, _vp, _vp, _vp, _vp] + [_i64] * 8,
    "ct_env_update_left": [_vp, _int, _int, _vp, _vp, _vp, _vp] + [_i64] * 5,
    "ct_env_update_right": [_vp, _int, _int, _vp, _vp, _vp, _vp] + [_i64] * 5,
    "ct_two_site_theta": [_vp, _int, _int, _vp, _vp, _vp] + [_i64] * 4,
    "ct_lanczos_two_site": [_vp, _int, _int, _vp, _vp, _vp, _vp, _vp] + [_i64] * 6
    + [_int, _dbl, _int, _int, _dp, _ip, _dp, _dp],
    "ct_svd_trunc": [_vp, _int, _vp, _i64, _i64, _i64, _dbl, _vp, _vp, _vp, _i64p, _dp],
    "ct_theta_split": [_vp, _int, _int, _vp, _i64, _i64, _i64, _i64, _dbl, _int, _vp, _vp, _i64p, _dp],
    "ct_sytrd": [_vp, _i64, _vp, _vp, _vp, _vp],
    "ct_syevd": [_vp, _i64, _vp, _vp, _vp, _int],
    "ct_dmrg_create": [_vp, ctypes.POINTER(DMRGParamsC), _int, _i64, _int, _i64p, ctypes.POINTER(_vp), _i64p,
                       ctypes.POINTER(_vp), ctypes.POINTER(_vp)],
    "ct_dmrg_destroy": [_vp],
    "ct_dmrg_run": [_vp, _int, _dp, _dp, _ip],
    "ct_dmrg_get_stats": [_vp, ctypes.POINTER(DMRGStatsC)],
    "ct_dmrg_get_dims": [_vp, _i64p],
    "ct_dmrg_get_site": [_vp, _int, _int, _vp],
    "ct_heff_adjoint": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp] + [_i64] * 6,
    "ct_env_adjoint": [_vp, _vp, _vp, _vp, _vp] + [_i64] * 5,
    "ct_random_mps_c_compatible": [_int, _i64, _i64, ctypes.c_uint64, _i64p, _dp, _i64p],
}
_RESTYPES = {"ct_last_error": ctypes.c_char_p, "ct_version": ctypes.c_char_p}

_lib = None


class CTError(RuntimeError):
    """A non-zero status from libdmrg_hip.so (message from ct_last_error)."""

    def __init__(self, fn: str, status: int, msg: str):
        super().__init__(f"{fn} failed with status {status}: {msg}")
        self.status = status


def load():
    """Load libdmrg_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with chemtensor_mojo_amd.build.build()")
        lib = ctypes.CDLL(LIB_PATH)
        for name, argt in SIGNATURES.items():
            f = getattr(lib, name)
            f.argtypes = argt
            f.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = lib
    return _lib


def call(name: str, *args) -> None:
    """Call entry point `name` and raise CTError on a non-zero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.ct_last_error().decode(errors="replace")
        raise CTError(name, rc, msg)


# Library objects still open (handles, prepared operators, DMRG states), in
# creation order.  At interpreter exit they are closed newest first - every
# operator and sweep state before the handle it was created on - while the HIP
# runtime and torch's allocator are still fully alive, instead of in whatever
# order garbage collection reaches them during shutdown.
_live: dict[int, weakref.ref] = {}
_seq = itertools.count()


def track(obj) -> None:
    """Register an object with a close() method for the ordered teardown at exit."""
    k = next(_seq)
    obj._ct_track_key = k
    _live[k] = weakref.ref(obj)


def untrack(obj) -> None:
    _live.pop(getattr(obj, "_ct_track_key", None), None)


@atexit.register
def _close_all() -> None:
    for k in sorted(_live, reverse=True):
        o = _live.get(k)
        o = o() if o is not None else None
        if o is not None:
            try:
                o.close()
            except Exception:
                pass
    _live.clear()


def last_error() -> str:
    return load().ct_last_error().decode(errors="replace")


def i64_array(values):
    arr = (ctypes.c_int64 * max(len(values), 1))(*[int(v) for v in values])
    return arr
