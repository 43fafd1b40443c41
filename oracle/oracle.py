"""ctypes loader for the CPU oracle (oracle/cgm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker or as the timed CPU
baseline -- never as a code path of the shipped library.

Functions follow the reference Mojo sources:
  sw_banded    genomics/cpu/align_cpu.mojo:229-309 (smith_waterman_banded)
  wfa_affine   genomics/cpu/align_cpu.mojo:75-226  (wfa_affine_cpu)
  pack_ascii   genomics/core/sequence.mojo:38-66   (SequenceBatch.add_sequence)
plus sw_banded_start (SW with the library's start coordinates) and the
independent cross-checks sw_antidiag and gotoh_global.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libcgm_oracle.so")
_FAST_PATH = os.path.join(_HERE, "libcgm_oracle_fast.so")
_lib = None


def host_has_v3() -> bool:
    """Whether this host's CPU runs x86-64-v3 code (AVX2, BMI2, FMA)."""
    try:
        flags = next(l for l in open("/proc/cpuinfo") if l.startswith("flags")).split()
    except (OSError, StopIteration):
        return False
    return all(f in flags for f in ("avx2", "bmi2", "fma", "movbe"))


def select_fast_build() -> str:
    """Load the -O3 -march=x86-64-v3 build instead of the portable one when
    the host supports it (call before the first use); returns the flags of
    the build that will be used.  For timing the CPU baseline only."""
    global _LIB_PATH
    if _lib is None and host_has_v3() and os.path.exists(_FAST_PATH):
        _LIB_PATH = _FAST_PATH
    # the flags of the build in use (a second call sees the first call's choice)
    return "-O3 -march=x86-64-v3" if _LIB_PATH == _FAST_PATH else "-O3 -march=x86-64-v2"

_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_vp = ctypes.c_void_p


def build(force: bool = False) -> str:
    """Compile oracle/cgm_oracle.c with its Makefile (gcc)."""
    if force and os.path.exists(_LIB_PATH):
        os.remove(_LIB_PATH)
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_pack_ascii.restype = _i64
        L.or_pack_ascii.argtypes = [ctypes.c_char_p, _i64, _vp, _vp]
        L.or_get_base_public.restype = ctypes.c_int
        L.or_get_base_public.argtypes = [_vp, _i64]
        for name in ("or_sw_banded", "or_sw_antidiag", "or_nw_banded"):
            f = getattr(L, name)
            f.restype = None
            f.argtypes = [_vp, _i64, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _vp]
        for name in ("or_wfa_affine", "or_gotoh_global"):
            f = getattr(L, name)
            f.restype = _i64
            f.argtypes = [_vp, _i64, _vp, _i64, _i64, _i64, _i64, _i64]
        L.or_wfa_affine_cigar.restype = _i64
        L.or_wfa_affine_cigar.argtypes = [_vp, _i64, _vp, _i64, _i64, _i64, _i64, _i64, _vp, _i64,
                                          ctypes.POINTER(_i64)]
        L.or_wfa_cigar_batch.restype = ctypes.c_int
        L.or_wfa_cigar_batch.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _i32, _i32, _i32,
                                         _i32, _vp, _vp, _vp, _vp, ctypes.c_int]
        L.or_sw_banded_start.restype = None
        L.or_sw_banded_start.argtypes = [_vp, _i64, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _vp]
        L.or_sw_start_batch.restype = ctypes.c_int
        L.or_sw_start_batch.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _i32, _i32, _i32,
                                        _i32, _i32, _vp, _vp, _vp, _vp, _vp, ctypes.c_int]
        L.or_batch.restype = ctypes.c_int
        L.or_batch.argtypes = [ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp,
                               _vp, _i32, _i32, _i32, _i32, _i32, _vp, _i32, _i32, _i32, _i32,
                               _vp, _vp, _vp, ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def pack_ascii(seq: str | bytes):
    """Returns (words uint64[ceil(len/32)], nmask uint64[same])."""
    b = seq.encode() if isinstance(seq, str) else bytes(seq)
    nw = (len(b) + 31) // 32
    words = np.zeros(max(nw, 1), dtype=np.uint64)
    nmask = np.zeros(max(nw, 1), dtype=np.uint64)
    lib().or_pack_ascii(b, len(b), _ptr(words), _ptr(nmask))
    return words[:nw].copy(), nmask[:nw].copy()


def _words(seq):
    if isinstance(seq, (str, bytes)):
        w, _ = pack_ascii(seq)
        return w, len(seq)
    w, n = seq
    return np.ascontiguousarray(w, dtype=np.uint64), int(n)


def _nonempty(w):
    return w if w.size else np.zeros(1, dtype=np.uint64)


def sw_banded(q, r, match=2, mismatch=-3, gap_open=-5, gap_extend=-2, band=32):
    """(score, query_end, ref_end) of align_cpu.mojo:229-309.  q/r are ASCII
    strings or (packed words, length) tuples.  Defaults = default_config()."""
    qw, ql = _words(q)
    rw, rl = _words(r)
    out = np.zeros(3, dtype=np.int64)
    lib().or_sw_banded(_ptr(_nonempty(qw)), ql, _ptr(_nonempty(rw)), rl,
                       match, mismatch, gap_open, gap_extend, band, _ptr(out))
    return int(out[0]), int(out[1]), int(out[2])


def sw_banded_start(q, r, match=2, mismatch=-3, gap_open=-5, gap_extend=-2, band=32):
    """(score, query_start, ref_start, query_end, ref_end): sw_banded plus the
    start of the latest-starting best alignment ending at the end cell (the
    reference reports starts 0, align_cpu.mojo:302-305)."""
    qw, ql = _words(q)
    rw, rl = _words(r)
    out = np.zeros(5, dtype=np.int64)
    lib().or_sw_banded_start(_ptr(_nonempty(qw)), ql, _ptr(_nonempty(rw)), rl,
                             match, mismatch, gap_open, gap_extend, band, _ptr(out))
    return tuple(int(v) for v in out)


def sw_start_batch(q_packed, r_packed, q_off, r_off, q_len, r_len, *, band=None, band_scalar=-1,
                   match=2, mismatch=-4, gap_open=-4, gap_extend=-2, threads=1):
    """Whole-batch sw_banded_start: (score, q_start, r_start, q_end, r_end)
    int32 arrays."""
    n = int(len(q_len))
    outs = [np.zeros(n, dtype=np.int32) for _ in range(5)]
    a = [np.ascontiguousarray(x) for x in (q_packed, r_packed)]
    i32 = [None if x is None else np.ascontiguousarray(x, dtype=np.int32)
           for x in (q_off, r_off, q_len, r_len, band)]
    rc = lib().or_sw_start_batch(_ptr(_nonempty(a[0])), _ptr(_nonempty(a[1])), _ptr(i32[0]),
                                 _ptr(i32[1]), _ptr(i32[2]), _ptr(i32[3]), n, _ptr(i32[4]),
                                 band_scalar, match, mismatch, gap_open, gap_extend,
                                 *[_ptr(o) for o in outs], int(threads))
    if rc != 0:
        raise RuntimeError("oracle batch failed")
    return tuple(outs)


def sw_antidiag(q, r, match=2, mismatch=-3, gap_open=-5, gap_extend=-2, band=32):
    qw, ql = _words(q)
    rw, rl = _words(r)
    out = np.zeros(3, dtype=np.int64)
    lib().or_sw_antidiag(_ptr(_nonempty(qw)), ql, _ptr(_nonempty(rw)), rl,
                         match, mismatch, gap_open, gap_extend, band, _ptr(out))
    return int(out[0]), int(out[1]), int(out[2])


def nw_banded(q, r, match=2, mismatch=-3, gap_open=-5, gap_extend=-2, band=32):
    """(score, query_end, ref_end) of needleman_wunsch_banded,
    align_cpu.mojo:312-377 (defaults = default_config())."""
    qw, ql = _words(q)
    rw, rl = _words(r)
    out = np.zeros(3, dtype=np.int64)
    lib().or_nw_banded(_ptr(_nonempty(qw)), ql, _ptr(_nonempty(rw)), rl, match, mismatch,
                       gap_open, gap_extend, band, _ptr(out))
    return int(out[0]), int(out[1]), int(out[2])


def wfa_affine(q, r, x=4, o=6, e=2, max_error=200):
    """Cost or -1 of align_cpu.mojo:75-226.  Defaults = default_wfa_config()."""
    qw, ql = _words(q)
    rw, rl = _words(r)
    return int(lib().or_wfa_affine(_ptr(_nonempty(qw)), ql, _ptr(_nonempty(rw)), rl,
                                   x, o, e, max_error))


def gotoh_global(q, r, x=4, o=6, e=2, max_error=-1):
    qw, ql = _words(q)
    rw, rl = _words(r)
    return int(lib().or_gotoh_global(_ptr(_nonempty(qw)), ql, _ptr(_nonempty(rw)), rl,
                                     x, o, e, max_error))


def batch(mode, q_packed, r_packed, q_off, r_off, q_len, r_len, *, kind=None,
          band=None, band_scalar=-1, match=2, mismatch=-4, gap_open=-4, gap_extend=-2,
          max_error=None, max_error_scalar=-1, x=4, o=6, e=2, threads=1):
    """Whole-batch driver (mode 0 = SW, 1 = WFA, kind[] overrides per pair).
    Returns (score, q_end, r_end) int32 arrays."""
    n = int(len(q_len))
    score = np.zeros(n, dtype=np.int32)
    qe = np.zeros(n, dtype=np.int32)
    re_ = np.zeros(n, dtype=np.int32)
    args = [np.ascontiguousarray(a) for a in (q_packed, r_packed)]
    i32 = [None if a is None else np.ascontiguousarray(a, dtype=np.int32)
           for a in (q_off, r_off, q_len, r_len, kind, band, max_error)]
    rc = lib().or_batch(mode, _ptr(_nonempty(args[0])), _ptr(_nonempty(args[1])),
                        _ptr(i32[0]), _ptr(i32[1]), _ptr(i32[2]), _ptr(i32[3]), n,
                        _ptr(i32[4]), _ptr(i32[5]), band_scalar, match, mismatch, gap_open,
                        gap_extend, _ptr(i32[6]), max_error_scalar, x, o, e,
                        _ptr(score), _ptr(qe), _ptr(re_), int(threads))
    if rc != 0:
        raise RuntimeError("oracle batch failed")
    return score, qe, re_


CIGAR_OPS = "=XID"


def runs_to_cigar(runs) -> str:
    """Run-length operations (count << 2 | op) as the reference's CIGAR
    string (align_cpu.mojo:602-621): '=' match, 'X' mismatch, 'I', 'D'."""
    return "".join(f"{int(v) >> 2}{CIGAR_OPS[int(v) & 3]}" for v in runs)


def wfa_affine_cigar(q, r, x=4, o=6, e=2, max_error=200):
    """wfa_affine_cigar_cpu (align_cpu.mojo:380-623): (cost or -1, CIGAR)."""
    qw, ql = _words(q)
    rw, rl = _words(r)
    cap = ql + rl + 8
    runs = np.zeros(cap, np.uint32)
    n = ctypes.c_int64(0)
    c = lib().or_wfa_affine_cigar(_ptr(_nonempty(qw)), ql, _ptr(_nonempty(rw)), rl, x, o, e,
                                  max_error, _ptr(runs), cap, ctypes.byref(n))
    assert n.value >= 0
    return int(c), runs_to_cigar(runs[:n.value])


def wfa_cigar_batch(q_packed, r_packed, q_off, r_off, q_len, r_len, *, max_error, x=4, o=6, e=2,
                    threads=8):
    """or_wfa_cigar_batch: (costs int32[n], list of CIGAR strings)."""
    n = int(len(q_len))
    q_len = np.ascontiguousarray(q_len, np.int32)
    r_len = np.ascontiguousarray(r_len, np.int32)
    cap = q_len.astype(np.int64) + r_len + 8
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum(cap)
    runs = np.zeros(max(int(off[-1]), 1), np.uint32)
    nr = np.zeros(n, np.int32)
    score = np.zeros(n, np.int32)
    me = np.ascontiguousarray(max_error, np.int32) if np.ndim(max_error) else None
    lib().or_wfa_cigar_batch(_ptr(_nonempty(np.ascontiguousarray(q_packed, np.uint64))),
                             _ptr(_nonempty(np.ascontiguousarray(r_packed, np.uint64))),
                             _ptr(np.ascontiguousarray(q_off, np.int32)),
                             _ptr(np.ascontiguousarray(r_off, np.int32)), _ptr(q_len),
                             _ptr(r_len), n, _ptr(me) if me is not None else None,
                             int(max_error) if me is None else 0, x, o, e, _ptr(score),
                             _ptr(runs), _ptr(off), _ptr(nr), threads)
    assert np.all(nr >= 0)
    return score, [runs_to_cigar(runs[off[i]:off[i] + nr[i]]) for i in range(n)]
