"""ctypes binding of libcgm_align.so (declarations in include/cgm_align.h).

The library must have been built (``build.py``); importing this module
without it raises, so no caller can silently fall back to a CPU path.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcgm_align.so")

CGM_OK = 0
CGM_ERR_INVALID = 1
CGM_ERR_HIP = 2
CGM_ERR_NOMEM = 3
CGM_ERR_BUSY = 4
CGM_ERR_NODEV = 5
CGM_ERR_NOTFOUND = 6
CGM_BAD_PAIR = -(2 ** 31)

# cgm_tuning fields (CGM_TUNE_* order)
TUNE_FIELDS = ("blocks_per_cu", "lds_bytes", "tile_steps", "wave_max_width", "reg_max_band",
               "lane_max_len", "sw_dual", "sw_split")

LAYOUT_AUTO = 0
LAYOUT_PLAIN = 1
LAYOUT_LANE = 2
LAYOUT_WAVE = 3

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_pi32 = ctypes.POINTER(ctypes.c_int32)
_pi64 = ctypes.POINTER(ctypes.c_int64)
_pf = ctypes.POINTER(ctypes.c_float)
_pvp = ctypes.POINTER(ctypes.c_void_p)

# name -> argtypes (all return int status unless listed in _RESTYPE)
SIGNATURES = {
    "cgm_last_error": [],
    "cgm_version": [],
    "cgm_device_count": [_pi32],
    "cgm_set_device": [ctypes.c_int],
    "cgm_device_info": [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, _pi32, _pi32, _pi64, _pi32],
    "cgm_malloc": [_pvp, _i64],
    "cgm_free": [_vp],
    "cgm_host_malloc": [_pvp, _i64],
    "cgm_host_free": [_vp],
    "cgm_memcpy_htod": [_vp, _vp, _i64, _vp],
    "cgm_memcpy_dtoh": [_vp, _vp, _i64, _vp],
    "cgm_memcpy_dtod": [_vp, _vp, _i64, _vp],
    "cgm_stream_create": [_pvp],
    "cgm_stream_destroy": [_vp],
    "cgm_stream_synchronize": [_vp],
    "cgm_device_synchronize": [],
    "cgm_event_create": [_pvp],
    "cgm_event_destroy": [_vp],
    "cgm_event_record": [_vp, _vp],
    "cgm_event_synchronize": [_vp],
    "cgm_event_elapsed_ms": [_vp, _vp, _pf],
    "cgm_sw_banded_launch": [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _i32,
                             _i32, _i32, _i32, _i32, _vp, _vp, _vp, _i32, _vp],
    "cgm_sw_banded_launch_noalloc": [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _i32,
                                     _i32, _i32, _i32, _i32, _vp, _vp, _vp, _vp],
    "cgm_sw_banded_start_launch": [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _i32,
                                   _i32, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _i32, _vp],
    "cgm_sw_banded_start_host": [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _i32,
                                 _i32, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _i32],
    "cgm_wfa_affine_launch": [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _i32,
                              _i32, _i32, _i32, _vp, _i32, _vp],
    "cgm_plan_create": [_i64, _i32, _i32, _i32, _i32, _i32, _i32, _pvp],
    "cgm_plan_destroy": [_vp],
    "cgm_align_mixed": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _i32, _vp, _i32,
                        _vp, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _vp, _i32,
                        _vp],
    "cgm_plan_capture": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _i32, _vp, _i32,
                         _vp, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _vp, _vp],
    "cgm_plan_replay": [_vp, _vp],
    "cgm_plan_kernel_ms": [_vp, _pf],
    "cgm_plan_wfa_kernel_ms": [_vp, _pf],
    "cgm_plan_get_tuning": [_vp, _vp],
    "cgm_plan_set_tuning": [_vp, _vp],
    "cgm_tuning_get_default": [_vp],
    "cgm_tuning_set_default": [_vp],
    "cgm_plan_autotune": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _i32, _vp,
                          _i32, _vp, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp,
                          _vp, _i32, _vp, _pf, _pf, _vp],
    "cgm_plan_autotune_cached": [_vp, ctypes.c_char_p, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64,
                                 _i64, _vp, _i32, _vp, _i32, _vp, _i32, _i32, _i32, _i32, _i32,
                                 _i32, _i32, _i32, _vp, _vp, _vp, _i32, _vp, _pf, _pf, _vp, _vp],
    "cgm_plan_profile_key": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _i32, _vp,
                             _i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32, ctypes.c_char_p,
                             _i32, _vp],
    "cgm_engine_create": [_i64, _i32, _i64, _i64, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32,
                          _pvp],
    "cgm_engine_destroy": [_vp],
    "cgm_engine_submit": [_vp, _vp, _i64, _vp, _i64, _vp, _vp, _vp, _vp, _i32, _vp, _vp, _vp,
                          ctypes.c_double, _pi64],
    "cgm_engine_submit_device": [_vp, _vp, _i64, _vp, _i64, _vp, _vp, _vp, _vp, _i32, _vp, _vp,
                                 _vp, ctypes.c_double, _pi64],
    "cgm_engine_poll": [_vp, _i64, _pi32],
    "cgm_engine_wait": [_vp, _i64, _i32, _vp, _vp, _vp, ctypes.POINTER(ctypes.c_double)],
    "cgm_engine_stats": [_vp, _pi64],
    "cgm_sw_banded_batch": [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _i64, _i64, _vp, _i32,
                            _i32, _i32, _i32, _i32, _vp, _vp, _vp, _vp],
    "cgm_wfa_affine_batch": [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _i64, _i64, _vp, _i32,
                             _i32, _i32, _i32, _vp, _vp],
    "cgm_engine_get_tuning": [_vp, _i32, _vp],
    "cgm_engine_set_tuning": [_vp, _i32, _vp],
    "cgm_engine_now": [],
    "cgm_serve_run": [_vp, _vp, ctypes.c_double, ctypes.c_double, _i32, _i32, _i32, _i32, _u64,
                      _vp],
    "cgm_sw_banded_host": [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _i32,
                           _i32, _i32, _i32, _i32, _vp, _vp, _vp, _i32],
    "cgm_wfa_affine_host": [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _i32,
                            _i32, _i32, _i32, _vp, _i32],
    "cgm_nw_banded_launch": [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _i32,
                             _i32, _i32, _i32, _i32, _vp, _vp, _vp, _i32, _vp],
    "cgm_nw_banded_host": [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _i32,
                           _i32, _i32, _i32, _i32, _vp, _vp, _vp, _i32],
    "cgm_wfa_cigar_launch": [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _i32, _i32,
                             _i32, _i32, _vp, _vp, _vp, _vp, _vp],
    "cgm_wfa_cigar_host": [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _i32, _i32, _i32,
                           _i32, _vp, _vp, _vp, _vp],
    "cgm_encode_ascii_launch": [_vp, _i64, _vp, _vp, _i64, _vp, _vp, _i64, _vp, _vp, _vp, _vp],
    "cgm_encode_ascii_host": [_vp, _i64, _vp, _vp, _i64, _vp, _vp, _i64, _vp, _vp, _pi64],
    "cgm_gen_create": [_i32, _i64, _u64, _pvp],
    "cgm_gen_sizes": [_vp, _pi64, _pi64, _pi64],
    "cgm_gen_copy": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "cgm_gen_destroy": [_vp],
}
_RESTYPE = {"cgm_last_error": ctypes.c_char_p, "cgm_engine_now": ctypes.c_double}

_lib = None


class CgmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cgm error {code}: {msg}")
        self.code = code


def load():
    """Loads libcgm_align.so (raises FileNotFoundError if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it with build.py (hipcc, gfx950) first")
    L = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = _RESTYPE.get(name, ctypes.c_int)
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != CGM_OK:
        msg = load().cgm_last_error()
        raise CgmError(rc, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))


def ptr(a):
    """Data pointer of a numpy array / DeviceArray / int / None for ctypes."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if hasattr(a, "dptr"):
        return a.dptr
    return a.ctypes.data_as(ctypes.c_void_p)
