"""Serving engine: continuous batching of alignment requests on one GPU
(cgm_engine_* in libcgm_align.so; see include/cgm_align.h)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

STAT_NAMES = ["submitted", "completed", "refused", "pairs_completed", "elapsed_s", "pairs_per_s",
              "requests_per_s", "p50_ms", "p90_ms", "p99_ms", "max_ms", "mean_ms", "batches",
              "mean_batch_pairs", "bad_results", "submit_phase_s", "sw_only_requests",
              "sw_only_p50_ms", "sw_only_p99_ms", "with_wfa_requests", "with_wfa_p50_ms",
              "with_wfa_p99_ms", "mean_sw_batch_pairs", "mean_wfa_batch_pairs",
              "submit_busy_frac", "submit_throttled_frac", "mean_in_flight"]


class ServingEngine:
    """Two bins (SW pairs, WFA pairs), each with n_slots preallocated batch
    slots within hbm_budget_bytes and its own dispatcher thread, which
    launches the bin's next batch whenever one of its slots is free."""

    def __init__(self, hbm_budget_bytes: int = 8 << 30, n_slots: int = 4,
                 max_batch_pairs: int = 16384, max_batch_words: int = 1 << 19,
                 max_error_cap: int = 8192, match: int = 2, mismatch: int = -4,
                 gap_open: int = -4, gap_extend: int = -2, x: int = 4, o: int = 6, e: int = 2):
        h = ctypes.c_void_p()
        _lib.call("cgm_engine_create", int(hbm_budget_bytes), n_slots, int(max_batch_pairs),
                  int(max_batch_words), max_error_cap, match, mismatch, gap_open, gap_extend,
                  x, o, e, ctypes.byref(h))
        self.handle = h.value

    def submit(self, q_packed, r_packed, q_len, r_len, q_off, r_off, kind, band, max_error,
               arrival_s: float = 0.0) -> int:
        """Queues one request; raises CgmError (code CGM_ERR_BUSY) when the
        engine refuses it."""
        arrs = [np.ascontiguousarray(a, np.int32) for a in (q_len, r_len, q_off, r_off, kind,
                                                             band, max_error)]
        qp = np.ascontiguousarray(q_packed, np.uint64)
        rp = np.ascontiguousarray(r_packed, np.uint64)
        rid = ctypes.c_int64()
        _lib.call("cgm_engine_submit", self.handle, _lib.ptr(qp), int(qp.size), _lib.ptr(rp),
                  int(rp.size), *[_lib.ptr(a) for a in arrs[:4]], int(arrs[0].size),
                  *[_lib.ptr(a) for a in arrs[4:]], float(arrival_s), ctypes.byref(rid))
        return rid.value

    def submit_device(self, q_dev, r_dev, q_len, r_len, q_off, r_off, kind, band, max_error,
                      arrival_s: float = 0.0) -> int:
        """Queues one request whose packed words are already on the device
        (q_dev / r_dev: DeviceArray of uint64, kept unchanged until the
        request is done); the metadata arrays are host arrays."""
        arrs = [np.ascontiguousarray(a, np.int32) for a in (q_len, r_len, q_off, r_off, kind,
                                                             band, max_error)]
        rid = ctypes.c_int64()
        _lib.call("cgm_engine_submit_device", self.handle, q_dev.dptr, int(q_dev.count),
                  r_dev.dptr, int(r_dev.count), *[_lib.ptr(a) for a in arrs[:4]],
                  int(arrs[0].size), *[_lib.ptr(a) for a in arrs[4:]], float(arrival_s),
                  ctypes.byref(rid))
        return rid.value

    def submit_pairs(self, w, idx, arrival_s: float = 0.0) -> int:
        """Queues pairs idx of a gen.Workload as one request."""
        idx = np.asarray(idx, np.int64)
        return self.submit(w.q_packed, w.r_packed, w.q_len[idx], w.r_len[idx], w.q_off[idx],
                           w.r_off[idx], w.kind[idx], w.band[idx], w.max_error[idx], arrival_s)

    def poll(self, request_id: int) -> bool:
        """Whether a submitted, not yet collected request has finished."""
        done = ctypes.c_int32()
        _lib.call("cgm_engine_poll", self.handle, request_id, ctypes.byref(done))
        return bool(done.value)

    def wait(self, request_id: int, n_pairs: int, timeout_ms: int = -1):
        """(score, query_end, ref_end, latency_ms) of a finished request."""
        s = np.zeros(n_pairs, np.int32)
        q = np.zeros(n_pairs, np.int32)
        r = np.zeros(n_pairs, np.int32)
        lat = ctypes.c_double()
        _lib.call("cgm_engine_wait", self.handle, request_id, timeout_ms, _lib.ptr(s),
                  _lib.ptr(q), _lib.ptr(r), ctypes.byref(lat))
        return s, q, r, lat.value

    def stats(self) -> dict:
        out = np.zeros(8, np.int64)
        _lib.call("cgm_engine_stats", self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        keys = ["batches", "batched_pairs", "done_requests", "done_pairs", "refused",
                "queued_pairs", "device_bytes", "worker_error"]
        return dict(zip(keys, out.tolist()))

    def get_tuning(self, bin: int) -> dict:
        """Tuning of bin 0 (SW pairs) or 1 (WFA pairs), CGM_TUNE_* fields."""
        a = np.zeros(len(_lib.TUNE_FIELDS), np.int32)
        _lib.call("cgm_engine_get_tuning", self.handle, bin, _lib.ptr(a))
        return dict(zip(_lib.TUNE_FIELDS, (int(v) for v in a)))

    def set_tuning(self, bin: int, **fields) -> dict:
        """Changes the given fields of a bin's tuning (from its next launch)
        and returns the new tuning."""
        t = self.get_tuning(bin)
        for k, v in fields.items():
            if k not in t:
                raise KeyError(k)
            t[k] = int(v)
        a = np.array([t[k] for k in _lib.TUNE_FIELDS], np.int32)
        _lib.call("cgm_engine_set_tuning", self.handle, bin, _lib.ptr(a))
        return t

    def run_load(self, gen_handle, rate_rps: float, duration_s: float,
                 max_pairs_per_request: int = 64, max_outstanding: int = 512,
                 seed: int = 1, clients: int = 1, device_inputs: bool = False) -> dict:
        """Native load generator (cgm_serve_run): `clients` client threads
        with Poisson arrivals at rate_rps in total (<= 0: closed loop with
        max_outstanding requests over all clients); device_inputs: the
        sequences are uploaded once and submitted as device pointers."""
        st = np.zeros(len(STAT_NAMES), np.float64)
        _lib.call("cgm_serve_run", self.handle, gen_handle, float(rate_rps), float(duration_s),
                  max_pairs_per_request, max_outstanding, clients, int(bool(device_inputs)),
                  seed, _lib.ptr(st))
        return dict(zip(STAT_NAMES, st.tolist()))

    def close(self) -> None:
        if self.handle:
            _lib.load().cgm_engine_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GeneratorHandle:
    """A batch generated inside the library (cgm_gen_create), kept as a
    handle for the native load generator."""

    def __init__(self, config: int, n_pairs: int, seed: int):
        h = ctypes.c_void_p()
        _lib.call("cgm_gen_create", config, int(n_pairs), seed, ctypes.byref(h))
        self.handle = h.value

    def close(self):
        if self.handle:
            _lib.load().cgm_gen_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
