"""Seeded synthetic pair batches (cgm_gen_* in libcgm_align.so).

Workloads (the five benchmark workloads, by id):
  SW_150BP      plain SW: 150 bp targets, 2% subs, 0.5% indels of 1-3 bp, band 32
  SW_100_300BP  batched SW: targets 100-300 bp, divergence 1-5% (sub:indel 3:1,
                indel length geometric p=0.5), band 16-64 per pair
  WFA_1_10KB    WFA: targets log-uniform 1-10 kb, divergence 0.5-5% (geometric p=0.3)
  MIXED         mixed: 50% SW 100-300 bp band 32, 40% SW 1-2 kb band 128,
                10% WFA 5-20 kb at 1-10% divergence
  SERVING       serving pair mix: length log-normal (median 250, sigma 1) clipped to
                100-20000 bp, SW band 64 below 2 kb, WFA above
All workloads put 0.1% N positions into the query (packed as A + NMask bit).
WFA pairs carry max_error = planted edit cost + 4 * #N + 8.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib

SW_SCORING = dict(match=2, mismatch=-4, gap_open=-4, gap_extend=-2)
WFA_PENALTIES = dict(x=4, o=6, e=2)
# Workload ids of cgm_gen_create (CGM_WL_* in cgm_align.h).
SW_150BP, SW_100_300BP, WFA_1_10KB, MIXED, SERVING = 1, 2, 3, 4, 5
DEFAULT_PAIRS = {SW_150BP: 1000, SW_100_300BP: 200_000, WFA_1_10KB: 20_000, MIXED: 100_000,
                 SERVING: 200_000}
DEFAULT_SEED = 20260402


@dataclass
class Workload:
    workload: int
    seed: int
    q_packed: np.ndarray
    q_nmask: np.ndarray
    q_off: np.ndarray
    q_len: np.ndarray
    r_packed: np.ndarray
    r_nmask: np.ndarray
    r_off: np.ndarray
    r_len: np.ndarray
    kind: np.ndarray
    band: np.ndarray
    max_error: np.ndarray

    @property
    def n_pairs(self) -> int:
        return int(self.q_len.size)

    def subset(self, idx) -> "Workload":
        """A new workload holding pairs idx (re-packed contiguously)."""
        idx = np.asarray(idx, dtype=np.int64)

        def take(packed, nmask, off, ln):
            nw = (ln[idx].astype(np.int64) + 31) // 32
            new_off = np.zeros(idx.size, np.int64)
            if idx.size:
                new_off[1:] = np.cumsum(nw)[:-1]
            total = int(nw.sum())
            p = np.zeros(total, np.uint64)
            m = np.zeros(total, np.uint64)
            for k, i in enumerate(idx):
                a, c = int(off[i]), int(nw[k])
                p[new_off[k]:new_off[k] + c] = packed[a:a + c]
                m[new_off[k]:new_off[k] + c] = nmask[a:a + c]
            return p, m, new_off.astype(np.int32), ln[idx].astype(np.int32)

        qp, qn, qo, ql = take(self.q_packed, self.q_nmask, self.q_off, self.q_len)
        rp, rn, ro, rl = take(self.r_packed, self.r_nmask, self.r_off, self.r_len)
        return Workload(self.workload, self.seed, qp, qn, qo, ql, rp, rn, ro, rl,
                        self.kind[idx].copy(), self.band[idx].copy(), self.max_error[idx].copy())

    def sw_cells(self) -> np.ndarray:
        """Valid band cells per pair: sum over rows of the in-matrix band
        width (the cells align_cpu.mojo:265-268 does not skip)."""
        return valid_band_cells(self.q_len, self.r_len, self.band)


def concat(parts) -> "Workload":
    """One workload holding the pairs of several, in order."""
    def cat(name):
        return np.concatenate([getattr(p, name) for p in parts])

    def offs(name_off, name_len, packed):
        out, base = [], 0
        for p in parts:
            out.append(getattr(p, name_off).astype(np.int64) + base)
            base += getattr(p, packed).size
        return np.concatenate(out).astype(np.int32)

    return Workload(0, 0, cat("q_packed"), cat("q_nmask"), offs("q_off", "q_len", "q_packed"),
                    cat("q_len"), cat("r_packed"), cat("r_nmask"),
                    offs("r_off", "r_len", "r_packed"), cat("r_len"), cat("kind"), cat("band"),
                    cat("max_error"))


def valid_band_cells(q_len, r_len, band) -> np.ndarray:
    """sum_{qi<qlen} max(0, min(rlen-1, qi+band) - max(0, qi-band) + 1) per
    pair, in closed form (vectorised)."""
    n = np.asarray(q_len, np.int64)
    r = np.asarray(r_len, np.int64)
    b = np.broadcast_to(np.asarray(band, np.int64), n.shape).astype(np.int64)
    c = r - 1
    # sum_{qi<n} min(c, qi + b) + 1
    k = np.clip(c - b + 1, 0, n)
    s_min = k * (k - 1) // 2 + k * b + (n - k) * c + n
    # sum_{qi<n} max(0, qi - b)
    m = np.maximum(0, n - b - 1)
    s_max = m * (m + 1) // 2
    # rows with qi >= rlen + band contribute negative widths above: add back
    neg = np.maximum(0, n - (r + b))
    total = s_min - s_max + neg * (neg - 1) // 2
    return np.where((n > 0) & (r > 0), total, 0)


def wfa_wavefront_cells(q_len, r_len, cost, max_error, o=6, e=2) -> np.ndarray:
    """Algorithmic WFA work per pair: the reachable wavefront cells
    sum_{s=1}^{c} n(s), n(s) = 1 for s < o + e, else
    min(qlen, m_s, me) + min(rlen, m_s, me) + 1 with m_s = (s - o) // e, over
    the scores s = 1 .. c that wfa_affine_cpu (align_cpu.mojo:148-224)
    evaluates: c = the cost, or max_error when the result is -1."""
    q = np.asarray(q_len, np.int64)
    r = np.asarray(r_len, np.int64)
    me = np.asarray(max_error, np.int64)
    c = np.asarray(cost, np.int64)
    c = np.where(c < 0, me, c)
    total = np.zeros(q.shape, np.int64)
    for s in range(1, int(c.max(initial=0)) + 1):
        if s < o + e:
            n = np.ones_like(q)
        else:
            ms = (s - o) // e
            n = np.minimum(np.minimum(q, ms), me) + np.minimum(np.minimum(r, ms), me) + 1
        total += np.where(s <= c, n, 0)
    return total


def generate(workload: int, n_pairs: int | None = None, seed: int = DEFAULT_SEED) -> Workload:
    n = DEFAULT_PAIRS[workload] if n_pairs is None else int(n_pairs)
    h = ctypes.c_void_p()
    _lib.call("cgm_gen_create", workload, n, seed, ctypes.byref(h))
    try:
        npairs, qw, rw = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.call("cgm_gen_sizes", h, ctypes.byref(npairs), ctypes.byref(qw), ctypes.byref(rw))
        qp = np.zeros(qw.value, np.uint64)
        qn = np.zeros(qw.value, np.uint64)
        rp = np.zeros(rw.value, np.uint64)
        rn = np.zeros(rw.value, np.uint64)
        arrs = [np.zeros(n, np.int32) for _ in range(7)]
        qo, ql, ro, rl, kind, band, me = arrs
        P = _lib.ptr
        _lib.call("cgm_gen_copy", h, P(qp) if qp.size else None, P(qn) if qn.size else None,
                  P(qo), P(ql), P(rp) if rp.size else None, P(rn) if rn.size else None, P(ro),
                  P(rl), P(kind), P(band), P(me))
    finally:
        _lib.load().cgm_gen_destroy(h)
    return Workload(workload, seed, qp, qn, qo, ql, rp, rn, ro, rl, kind, band, me)
