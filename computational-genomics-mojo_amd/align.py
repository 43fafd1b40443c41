"""Alignment API: the reference's CPU signatures and GPU launchers, run on
the gfx950 kernels of libcgm_align.so.

Reference interfaces mirrored (Computational-Genomics-Mojo):
  AlignConfig / AlignResult / default_config   genomics/cpu/align_cpu.mojo:18-53
  WFAConfig / default_wfa_config               genomics/cpu/align_cpu.mojo:62-72
  smith_waterman_banded(query, reference, cfg) genomics/cpu/align_cpu.mojo:229-309
  wfa_affine_cpu(query, ref_seq, cfg)          genomics/cpu/align_cpu.mojo:75-226
  launch_banded_sw(device, ...)                genomics/gpu/align_gpu.mojo:460-493
  launch_wfa_affine(device, ...)               genomics/gpu/align_gpu.mojo:496-535
Results are bit-identical to align_cpu.mojo (score, query_end, ref_end for
SW; cost or -1 for WFA); query_start/ref_start are 0 and the CIGAR is empty,
as in the reference.  Every function runs on the GPU and raises when the
library or the device is missing -- there is no CPU fallback.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import os

import numpy as np

from . import _lib
from .device import DeviceArray, GenomicsDevice
from .sequence import SequenceBatch, SequenceView


@dataclass
class AlignConfig:
    """align_cpu.mojo:18-25 -- penalties are negative values."""
    match_score: int
    mismatch: int
    gap_open: int
    gap_extend: int
    band_width: int


@dataclass
class AlignResult:
    """align_cpu.mojo:28-36"""
    score: int
    query_start: int
    ref_start: int
    query_end: int
    ref_end: int
    cigar: list = field(default_factory=list)


@dataclass
class WFAConfig:
    """align_cpu.mojo:62-68 -- positive costs."""
    x: int
    o: int
    e: int
    max_error: int


def default_config() -> AlignConfig:
    """align_cpu.mojo:46-53"""
    return AlignConfig(match_score=2, mismatch=-3, gap_open=-5, gap_extend=-2, band_width=32)


def default_wfa_config() -> WFAConfig:
    """align_cpu.mojo:71-72"""
    return WFAConfig(x=4, o=6, e=2, max_error=200)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def _nonempty(a):
    return a if a.size else np.zeros(1, dtype=a.dtype)


def sw_banded_batch(q_packed, q_off, q_len, r_packed, r_off, r_len, cfg: AlignConfig,
                    band=None, layout=_lib.LAYOUT_AUTO):
    """Aligns pair i = (query i, reference i) for every i; host arrays in,
    host arrays (score, query_end, ref_end) out.  band: per-pair array or
    None for cfg.band_width.  A pair the library refuses (lengths/offsets
    outside the buffers, effective band above 1600, match * min(qlen, rlen)
    > 2^30) gets score CGM_BAD_PAIR (-2^31) and ends -1."""
    q_packed = np.ascontiguousarray(q_packed, np.uint64)
    r_packed = np.ascontiguousarray(r_packed, np.uint64)
    ql, rl, qo, ro, bd = _i32(q_len), _i32(r_len), _i32(q_off), _i32(r_off), _i32(band)
    n = int(ql.size)
    score = np.zeros(n, np.int32)
    qe = np.zeros(n, np.int32)
    re_ = np.zeros(n, np.int32)
    if n:
        _lib.call("cgm_sw_banded_host", _lib.ptr(_nonempty(q_packed)), _lib.ptr(_nonempty(r_packed)),
                  _lib.ptr(ql), _lib.ptr(rl), _lib.ptr(qo), _lib.ptr(ro), n, q_packed.size,
                  r_packed.size, _lib.ptr(bd), cfg.band_width, cfg.match_score, cfg.mismatch,
                  cfg.gap_open, cfg.gap_extend, _lib.ptr(score), _lib.ptr(qe), _lib.ptr(re_),
                  layout)
    return score, qe, re_


def sw_banded_start_batch(q_packed, q_off, q_len, r_packed, r_off, r_len, cfg: AlignConfig,
                          band=None, layout=_lib.LAYOUT_AUTO):
    """sw_banded_batch plus start coordinates: host arrays (score,
    query_start, ref_start, query_end, ref_end).  The start is that of the
    latest-starting best alignment ending at the reported end (the
    reference reports 0, align_cpu.mojo:302-305); refused pairs carry
    CGM_BAD_PAIR and -1 as in sw_banded_batch."""
    q_packed = np.ascontiguousarray(q_packed, np.uint64)
    r_packed = np.ascontiguousarray(r_packed, np.uint64)
    ql, rl, qo, ro, bd = _i32(q_len), _i32(r_len), _i32(q_off), _i32(r_off), _i32(band)
    n = int(ql.size)
    outs = [np.zeros(n, np.int32) for _ in range(5)]
    if n:
        _lib.call("cgm_sw_banded_start_host", _lib.ptr(_nonempty(q_packed)),
                  _lib.ptr(_nonempty(r_packed)), _lib.ptr(ql), _lib.ptr(rl), _lib.ptr(qo),
                  _lib.ptr(ro), n, q_packed.size, r_packed.size, _lib.ptr(bd), cfg.band_width,
                  cfg.match_score, cfg.mismatch, cfg.gap_open, cfg.gap_extend,
                  *[_lib.ptr(o) for o in outs], layout)
    return tuple(outs)


def nw_banded_batch(q_packed, q_off, q_len, r_packed, r_off, r_len, cfg: AlignConfig,
                    band=None, layout=_lib.LAYOUT_AUTO):
    """Banded NW (needleman_wunsch_banded, align_cpu.mojo:312-377) of every
    pair on the GPU; host arrays in, (score, query_end, ref_end) out.
    Refused pairs get score CGM_BAD_PAIR (-2^31) and ends -1."""
    q_packed = np.ascontiguousarray(q_packed, np.uint64)
    r_packed = np.ascontiguousarray(r_packed, np.uint64)
    ql, rl, qo, ro, bd = _i32(q_len), _i32(r_len), _i32(q_off), _i32(r_off), _i32(band)
    n = int(ql.size)
    score, qe, re_ = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.int32)
    if n:
        _lib.call("cgm_nw_banded_host", _lib.ptr(_nonempty(q_packed)), _lib.ptr(_nonempty(r_packed)),
                  _lib.ptr(ql), _lib.ptr(rl), _lib.ptr(qo), _lib.ptr(ro), n, q_packed.size,
                  r_packed.size, _lib.ptr(bd), cfg.band_width, cfg.match_score, cfg.mismatch,
                  cfg.gap_open, cfg.gap_extend, _lib.ptr(score), _lib.ptr(qe), _lib.ptr(re_),
                  layout)
    return score, qe, re_


def wfa_affine_batch(q_packed, q_off, q_len, r_packed, r_off, r_len, cfg: WFAConfig,
                     max_error=None, layout=_lib.LAYOUT_AUTO):
    """Gap-affine WFA cost (or -1) of every pair; host arrays in and out.
    Refused pairs (lengths/offsets outside the buffers, max_error < 0 or
    above 2^20) get CGM_BAD_PAIR (-2^31)."""
    q_packed = np.ascontiguousarray(q_packed, np.uint64)
    r_packed = np.ascontiguousarray(r_packed, np.uint64)
    ql, rl, qo, ro, me = _i32(q_len), _i32(r_len), _i32(q_off), _i32(r_off), _i32(max_error)
    n = int(ql.size)
    cost = np.zeros(n, np.int32)
    if n:
        _lib.call("cgm_wfa_affine_host", _lib.ptr(_nonempty(q_packed)), _lib.ptr(_nonempty(r_packed)),
                  _lib.ptr(ql), _lib.ptr(rl), _lib.ptr(qo), _lib.ptr(ro), n, q_packed.size,
                  r_packed.size, _lib.ptr(me), cfg.max_error, cfg.x, cfg.o, cfg.e,
                  _lib.ptr(cost), layout)
    return cost


CIGAR_OPS = "=XID"


@dataclass
class WFAResult:
    """align_cpu.mojo:40-43: the cost (or -1 above max_error) and the
    run-length extended CIGAR ('=', 'X', 'I', 'D')."""
    score: int
    cigar: str


def runs_to_cigar(runs) -> str:
    """Run-length operations (count << 2 | op; op 0 '=', 1 'X', 2 'I', 3 'D')
    as a CIGAR string, e.g. [12, 5, 4] -> '3=1X1='."""
    return "".join(f"{int(v) >> 2}{CIGAR_OPS[int(v) & 3]}" for v in runs)


def cigar_capacity(max_error, cfg: WFAConfig) -> np.ndarray:
    """Runs a CIGAR of cost <= max_error can have: every run other than '='
    costs at least min(x, o + e), and '=' runs alternate with them."""
    me = np.maximum(np.asarray(max_error, np.int64), 0)
    return 2 * (me // max(1, min(cfg.x, cfg.o + cfg.e))) + 1


def wfa_affine_cigar_batch(q_packed, q_off, q_len, r_packed, r_off, r_len, cfg: WFAConfig,
                           max_error=None):
    """Cost and CIGAR of every pair (wfa_affine_cigar_cpu, align_cpu.mojo:
    380-623, on the GPU); host arrays in, (costs, list of CIGAR strings) out."""
    q_packed = np.ascontiguousarray(q_packed, np.uint64)
    r_packed = np.ascontiguousarray(r_packed, np.uint64)
    ql, rl, qo, ro, me = _i32(q_len), _i32(r_len), _i32(q_off), _i32(r_off), _i32(max_error)
    n = int(ql.size)
    cost = np.zeros(n, np.int32)
    if not n:
        return cost, []
    cap = cigar_capacity(me if me is not None else np.full(n, cfg.max_error), cfg)
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum(cap)
    runs = np.zeros(max(int(off[-1]), 1), np.uint32)
    nr = np.zeros(n, np.int32)
    _lib.call("cgm_wfa_cigar_host", _lib.ptr(_nonempty(q_packed)), _lib.ptr(_nonempty(r_packed)),
              _lib.ptr(ql), _lib.ptr(rl), _lib.ptr(qo), _lib.ptr(ro), n, q_packed.size,
              r_packed.size, _lib.ptr(me), cfg.max_error, cfg.x, cfg.o, cfg.e, _lib.ptr(cost),
              _lib.ptr(runs), _lib.ptr(off), _lib.ptr(nr))
    if np.any(nr < 0):
        raise _lib.CgmError(_lib.CGM_ERR_INVALID, "CIGAR capacity exceeded")
    return cost, [runs_to_cigar(runs[off[i]:off[i] + nr[i]]) for i in range(n)]


def wfa_affine_cigar_cpu(query: SequenceView, ref_seq: SequenceView, cfg: WFAConfig) -> WFAResult:
    """align_cpu.mojo:380-623 (same name and signature), computed on the GPU."""
    qp, qo, ql = _view_arrays(query)
    rp, ro, rl = _view_arrays(ref_seq)
    c, cig = wfa_affine_cigar_batch(qp, qo, ql, rp, ro, rl, cfg)
    return WFAResult(int(c[0]), cig[0])


def _view_arrays(v: SequenceView):
    return np.ascontiguousarray(v.packed, np.uint64), np.zeros(1, np.int32), \
        np.array([v.length], np.int32)


def _refused(score, what):
    """The single-pair signatures return a result or raise, as the
    reference returns a result: a pair the library refuses (effective band
    above 1600, match * min(qlen, rlen) > 2^30, max_error > 2^20) raises."""
    if int(score) == _lib.CGM_BAD_PAIR:
        raise _lib.CgmError(_lib.CGM_ERR_INVALID, f"{what}: pair outside the supported limits")


def smith_waterman_banded(query: SequenceView, reference: SequenceView,
                          cfg: AlignConfig, with_start: bool = False) -> AlignResult:
    """align_cpu.mojo:229-309, computed by the GPU kernel.  query_start and
    ref_start are 0 as in the reference unless with_start is set; then they
    hold the start of the latest-starting best alignment that ends at
    (query_end, ref_end)."""
    qp, qo, ql = _view_arrays(query)
    rp, ro, rl = _view_arrays(reference)
    if with_start:
        s, qs, rs, qe, re_ = sw_banded_start_batch(qp, qo, ql, rp, ro, rl, cfg)
        _refused(s[0], "smith_waterman_banded")
        return AlignResult(int(s[0]), int(qs[0]), int(rs[0]), int(qe[0]), int(re_[0]), [])
    s, qe, re_ = sw_banded_batch(qp, qo, ql, rp, ro, rl, cfg)
    _refused(s[0], "smith_waterman_banded")
    return AlignResult(int(s[0]), 0, 0, int(qe[0]), int(re_[0]), [])


def needleman_wunsch_banded(query: SequenceView, reference: SequenceView,
                            cfg: AlignConfig) -> AlignResult:
    """align_cpu.mojo:312-377 (same name and signature), computed on the GPU."""
    qp, qo, ql = _view_arrays(query)
    rp, ro, rl = _view_arrays(reference)
    s, qe, re_ = nw_banded_batch(qp, qo, ql, rp, ro, rl, cfg)
    _refused(s[0], "needleman_wunsch_banded")
    return AlignResult(int(s[0]), 0, 0, int(qe[0]), int(re_[0]), [])


def wfa_affine_cpu(query: SequenceView, ref_seq: SequenceView, cfg: WFAConfig) -> int:
    """align_cpu.mojo:75-226 (same name and signature), computed by the GPU
    kernel."""
    qp, qo, ql = _view_arrays(query)
    rp, ro, rl = _view_arrays(ref_seq)
    c = int(wfa_affine_batch(qp, qo, ql, rp, ro, rl, cfg)[0])
    _refused(c, "wfa_affine_cpu")
    return c


wfa_affine = wfa_affine_cpu


class SwResult:
    """Device outputs of launch_banded_sw: scores (the reference's return
    buffer) plus the end coordinates the CPU result carries."""

    def __init__(self, n: int):
        self.scores = DeviceArray(max(n, 1), np.int32)
        self.query_end = DeviceArray(max(n, 1), np.int32)
        self.ref_end = DeviceArray(max(n, 1), np.int32)


def launch_banded_sw(device: GenomicsDevice, queries_buf: DeviceArray, refs_buf: DeviceArray,
                     q_lengths_buf: DeviceArray, r_lengths_buf: DeviceArray,
                     q_offsets_buf: DeviceArray, r_offsets_buf: DeviceArray, n_pairs: int,
                     total_words: int, band: int = 64, cfg: AlignConfig | None = None,
                     band_buf: DeviceArray | None = None, r_total_words: int | None = None,
                     layout: int = _lib.LAYOUT_AUTO) -> SwResult:
    """align_gpu.mojo:460-493.  Asynchronous on device.stream; synchronise
    (device.synchronize()) before reading the result.  total_words bounds the
    query buffer and, unless r_total_words is given, the reference buffer."""
    cfg = cfg or default_config()
    out = SwResult(n_pairs)
    rw = total_words if r_total_words is None else r_total_words
    _lib.call("cgm_sw_banded_launch", queries_buf.dptr, refs_buf.dptr, q_lengths_buf.dptr,
              r_lengths_buf.dptr, q_offsets_buf.dptr, r_offsets_buf.dptr, n_pairs, total_words,
              rw, band_buf.dptr if band_buf is not None else None, band, cfg.match_score,
              cfg.mismatch, cfg.gap_open, cfg.gap_extend, out.scores.dptr, out.query_end.dptr,
              out.ref_end.dptr, layout, device.stream)
    return out


class SwStartResult(SwResult):
    """Device outputs of launch_banded_sw_start: SwResult plus starts."""

    def __init__(self, n: int):
        super().__init__(n)
        self.query_start = DeviceArray(max(n, 1), np.int32)
        self.ref_start = DeviceArray(max(n, 1), np.int32)


def launch_banded_sw_start(device: GenomicsDevice, queries_buf: DeviceArray,
                           refs_buf: DeviceArray, q_lengths_buf: DeviceArray,
                           r_lengths_buf: DeviceArray, q_offsets_buf: DeviceArray,
                           r_offsets_buf: DeviceArray, n_pairs: int, total_words: int,
                           band: int = 64, cfg: AlignConfig | None = None,
                           band_buf: DeviceArray | None = None,
                           r_total_words: int | None = None,
                           layout: int = _lib.LAYOUT_AUTO) -> SwStartResult:
    """launch_banded_sw (align_gpu.mojo:460-493) plus start coordinates
    (cgm_sw_banded_start_launch).  Asynchronous on device.stream."""
    cfg = cfg or default_config()
    out = SwStartResult(n_pairs)
    rw = total_words if r_total_words is None else r_total_words
    _lib.call("cgm_sw_banded_start_launch", queries_buf.dptr, refs_buf.dptr, q_lengths_buf.dptr,
              r_lengths_buf.dptr, q_offsets_buf.dptr, r_offsets_buf.dptr, n_pairs, total_words,
              rw, band_buf.dptr if band_buf is not None else None, band, cfg.match_score,
              cfg.mismatch, cfg.gap_open, cfg.gap_extend, out.scores.dptr, out.query_start.dptr,
              out.ref_start.dptr, out.query_end.dptr, out.ref_end.dptr, layout, device.stream)
    return out


def launch_wfa_affine(device: GenomicsDevice, queries_buf: DeviceArray, refs_buf: DeviceArray,
                      q_lengths_buf: DeviceArray, r_lengths_buf: DeviceArray,
                      q_offsets_buf: DeviceArray, r_offsets_buf: DeviceArray, n_pairs: int,
                      total_words: int, max_error: int = 200, x: int = 4, o: int = 6, e: int = 2,
                      max_error_buf: DeviceArray | None = None, r_total_words: int | None = None,
                      layout: int = _lib.LAYOUT_AUTO) -> DeviceArray:
    """align_gpu.mojo:496-535 -- returns the device buffer of costs (-1 when
    the cost exceeds max_error).  Asynchronous on device.stream."""
    scores = DeviceArray(max(n_pairs, 1), np.int32)
    rw = total_words if r_total_words is None else r_total_words
    _lib.call("cgm_wfa_affine_launch", queries_buf.dptr, refs_buf.dptr, q_lengths_buf.dptr,
              r_lengths_buf.dptr, q_offsets_buf.dptr, r_offsets_buf.dptr, n_pairs, total_words,
              rw, max_error_buf.dptr if max_error_buf is not None else None, max_error, x, o, e,
              scores.dptr, layout, device.stream)
    return scores


def batch_arrays(queries: SequenceBatch, refs: SequenceBatch):
    """(q_packed, q_off, q_len, r_packed, r_off, r_len) of two batches."""
    return (queries.packed, queries.offsets, queries.lengths, refs.packed, refs.offsets,
            refs.lengths)


def _mixed_args(bufs, n, kind, band, max_error, cfg: AlignConfig, out: SwResult, pen):
    """kind: a device array of per-pair kinds, or an int (0: every pair SW,
    1: every pair WFA; the WFA-only form launches no SW kernel)."""
    q, r, ql, rl, qo, ro = bufs
    x, o, e = pen
    ptr = lambda a: a.dptr if a is not None else None  # noqa: E731
    kind_arr, kind_scalar = (None, int(kind)) if isinstance(kind, int) else (kind, 0)
    return (q.dptr, r.dptr, ql.dptr, rl.dptr, qo.dptr, ro.dptr, n, q.count, r.count,
            ptr(kind_arr), kind_scalar, ptr(band), cfg.band_width, ptr(max_error), 0,
            cfg.match_score, cfg.mismatch, cfg.gap_open, cfg.gap_extend, x, o, e,
            out.scores.dptr, out.query_end.dptr, out.ref_end.dptr)


class MixedPlan:
    """Preallocated workspace of the persistent mixed SW + WFA layout
    (cgm_plan_create): sort keys, task queue and HBM wavefront slabs for WFA
    pairs up to max_error_cap.  align() enqueues one call without any
    allocation or host synchronisation; capture()/replay() record it into a
    hipGraph and launch it again.  One device, one call at a time."""

    def __init__(self, max_pairs: int, max_error_cap: int, x: int = 4, o: int = 6, e: int = 2,
                 waves_per_cu: int = 0, lds_bytes: int = 0):
        import ctypes
        h = ctypes.c_void_p()
        _lib.call("cgm_plan_create", int(max_pairs), int(max_error_cap), x, o, e, waves_per_cu,
                  lds_bytes, ctypes.byref(h))
        self.handle = h.value
        self.pen = (x, o, e)

    def align(self, device: GenomicsDevice, bufs, n: int, kind, band, max_error, cfg: AlignConfig,
              out: SwResult, layout: int = _lib.LAYOUT_AUTO) -> None:
        _lib.call("cgm_align_mixed", self.handle,
                  *_mixed_args(bufs, n, kind, band, max_error, cfg, out, self.pen), layout,
                  device.stream)

    def capture(self, device: GenomicsDevice, bufs, n: int, kind, band, max_error,
                cfg: AlignConfig, out: SwResult) -> None:
        _lib.call("cgm_plan_capture", self.handle,
                  *_mixed_args(bufs, n, kind, band, max_error, cfg, out, self.pen), device.stream)

    def replay(self, device: GenomicsDevice) -> None:
        _lib.call("cgm_plan_replay", self.handle, device.stream)

    def kernel_ms(self) -> float:
        """Duration of the SW persistent kernel of the last call or replay."""
        import ctypes
        ms = ctypes.c_float()
        _lib.call("cgm_plan_kernel_ms", self.handle, ctypes.byref(ms))
        return float(ms.value)

    def wfa_kernel_ms(self) -> float:
        """Duration of the WFA kernels of the last call or replay (0 when it
        had no WFA pair)."""
        import ctypes
        ms = ctypes.c_float()
        _lib.call("cgm_plan_wfa_kernel_ms", self.handle, ctypes.byref(ms))
        return float(ms.value)

    def tuning(self) -> dict:
        """The plan's layout choices (cgm_plan_get_tuning) by field name."""
        a = np.zeros(len(_lib.TUNE_FIELDS), np.int32)
        _lib.call("cgm_plan_get_tuning", self.handle, _lib.ptr(a))
        return dict(zip(_lib.TUNE_FIELDS, (int(v) for v in a)))

    def set_tuning(self, **fields) -> None:
        """Changes some layout choices (others keep their value)."""
        t = self.tuning()
        for k, v in fields.items():
            if k not in t:
                raise KeyError(k)
            t[k] = int(v)
        a = np.array([t[k] for k in _lib.TUNE_FIELDS], np.int32)
        _lib.call("cgm_plan_set_tuning", self.handle, _lib.ptr(a))

    def autotune(self, device: GenomicsDevice, bufs, n: int, kind, band, max_error,
                 cfg: AlignConfig, out: SwResult, reps: int = 2):
        """cgm_plan_autotune on one batch: returns (best tuning dict, best
        ms, starting ms); the plan keeps the best tuning."""
        import ctypes
        a = np.zeros(len(_lib.TUNE_FIELDS), np.int32)
        best, start = ctypes.c_float(), ctypes.c_float()
        _lib.call("cgm_plan_autotune", self.handle,
                  *_mixed_args(bufs, n, kind, band, max_error, cfg, out, self.pen), int(reps),
                  _lib.ptr(a), ctypes.byref(best), ctypes.byref(start), device.stream)
        return dict(zip(_lib.TUNE_FIELDS, (int(v) for v in a))), float(best.value), \
            float(start.value)

    def autotune_cached(self, device: GenomicsDevice, cache_path: str, bufs, n: int, kind, band,
                        max_error, cfg: AlignConfig, out: SwResult, reps: int = 2):
        """cgm_plan_autotune_cached: the tuning kept in `cache_path` for this
        batch's length-bin profile, or a measured one that is appended there.
        Returns (tuning dict, best ms, starting ms or -1, from_cache)."""
        import ctypes
        a = np.zeros(len(_lib.TUNE_FIELDS), np.int32)
        best, start = ctypes.c_float(), ctypes.c_float()
        hit = ctypes.c_int32()
        args = _mixed_args(bufs, n, kind, band, max_error, cfg, out, self.pen)
        _lib.call("cgm_plan_autotune_cached", self.handle, os.fsencode(cache_path), *args,
                  int(reps), _lib.ptr(a), ctypes.byref(best), ctypes.byref(start),
                  ctypes.byref(hit), device.stream)
        return (dict(zip(_lib.TUNE_FIELDS, (int(v) for v in a))), float(best.value),
                float(start.value), bool(hit.value))

    def profile_key(self, device: GenomicsDevice, bufs, n: int, kind, band,
                    cfg: AlignConfig) -> str:
        """The batch's length-bin profile key (cgm_plan_profile_key)."""
        import ctypes
        q, r, ql, rl, qo, ro = bufs
        ptr = lambda v: v.dptr if v is not None else None  # noqa: E731
        kind_arr, kind_scalar = (None, int(kind)) if isinstance(kind, int) else (kind, 0)
        buf = ctypes.create_string_buffer(4096)
        x, o, e = self.pen
        _lib.call("cgm_plan_profile_key", self.handle, q.dptr, r.dptr, ql.dptr, rl.dptr, qo.dptr,
                  ro.dptr, n, q.count, r.count, ptr(kind_arr), kind_scalar, ptr(band),
                  cfg.band_width,
                  cfg.match_score, cfg.mismatch, cfg.gap_open, cfg.gap_extend, x, o, e, buf,
                  len(buf), device.stream)
        return buf.value.decode()

    def close(self) -> None:
        if self.handle:
            _lib.load().cgm_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_tuning() -> dict:
    """The tuning new plans start from (cgm_tuning_get_default)."""
    a = np.zeros(len(_lib.TUNE_FIELDS), np.int32)
    _lib.call("cgm_tuning_get_default", _lib.ptr(a))
    return dict(zip(_lib.TUNE_FIELDS, (int(v) for v in a)))


def set_default_tuning(**fields) -> None:
    """Changes the tuning of plans created afterwards (including the
    library's own plans and serving engines)."""
    t = default_tuning()
    for k, v in fields.items():
        if k not in t:
            raise KeyError(k)
        t[k] = int(v)
    _lib.call("cgm_tuning_set_default", _lib.ptr(np.array([t[k] for k in _lib.TUNE_FIELDS],
                                                            np.int32)))


def align_mixed_plain(device: GenomicsDevice, bufs, n: int, kind, band, max_error,
                      cfg: AlignConfig, out: SwResult, pen=(4, 6, 2)) -> None:
    """The one-workgroup-per-pair baseline for a mixed batch (no plan)."""
    _lib.call("cgm_align_mixed", None, *_mixed_args(bufs, n, kind, band, max_error, cfg, out, pen),
              _lib.LAYOUT_PLAIN, device.stream)


def align_mixed_batch(device: GenomicsDevice, w, cfg: AlignConfig, x: int = 4, o: int = 6,
                      e: int = 2, layout: int = _lib.LAYOUT_AUTO, plan: "MixedPlan | None" = None):
    """Host convenience: uploads a workload (packed buffers, kind, band,
    max_error arrays), aligns every pair (SW or WFA by kind) and returns host
    (score, query_end, ref_end); WFA pairs carry their cost and ends -1."""
    nz = lambda a: a if a.size else np.zeros(1, a.dtype)  # noqa: E731
    n = int(w.q_len.size)
    bufs = tuple(device.upload(nz(a)) for a in (w.q_packed, w.r_packed, w.q_len, w.r_len,
                                                 w.q_off, w.r_off))
    kind, band, me = (device.upload(nz(a)) for a in (w.kind, w.band, w.max_error))
    out = SwResult(n)
    if layout == _lib.LAYOUT_PLAIN:
        align_mixed_plain(device, bufs, n, kind, band, me, cfg, out, (x, o, e))
    else:
        own = plan is None
        if own:
            plan = MixedPlan(max(n, 1), int(max(int(w.max_error.max(initial=0)), 1)), x, o, e)
        plan.align(device, bufs, n, kind, band, me, cfg, out, layout)
    device.synchronize()
    return (device.download(out.scores, n), device.download(out.query_end, n),
            device.download(out.ref_end, n))
