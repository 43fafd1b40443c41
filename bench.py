"""Benchmark: banded SW GCUPS on the batched workload (200,000 pairs,
100-300 bp, band 16-64 per pair, scores 2/-4/-4/-2) on one MI355X, plus the
other workloads of the path as extras (plain SW on 1,000 x 150 bp, WFA on
20,000 x 1-10 kb, the mixed batch, serving, ingest, CIGAR, NW), each with
its roofline fraction and a CPU baseline where it has one.

A step is one pass of the alignment path over the whole batch, with the
packed sequences already resident in HBM.  Prints ONE JSON line:
  value       GCUPS = valid band cells (the cells align_cpu.mojo:265-268
              computes) / second, over the K timed steps
  roofline    integer-VALU roofline of the dominant kernel: achieved int ops
              (11 per cell) / that kernel's mean launch time (HIP events on
              the stream it runs on), against 256 CU x 64 lanes x 2.4 GHz
              int32 ops, or twice that when the register lanes compute two
              cells per packed 16-bit instruction (tuning sw_dual; `dtype`
              says which)
  cpu_baseline  the CPU oracle (C restatement of align_cpu.mojo) on this
              host, all pairs, `cores` threads
Extra objects report the plain one-workgroup-per-pair kernel on the same
input, band-area GCUPS, pairs/s and WFA pairs/s.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "computational-genomics-mojo_amd")

VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12   # int32 lane-ops/s, MI355X spec
HBM_PEAK_GBS = 8000.0
SW_OPS_PER_CELL = 11
WFA_OPS_PER_CELL = 8
# HBM bytes per launch of the persistent kernel on this workload, measured
# by rocprofv3 FETCH_SIZE / WRITE_SIZE passes (each scaled by what the
# counter reads on a 256 MiB device copy) and read from the summary that
# holds those passes and their commands (its "traffic_bytes:" line).
TRAFFIC_SOURCE = "profiles/sw_persistent_200k_100to300bp_rocprof.md"
WFA_TRAFFIC_SOURCE = "profiles/wfa_reg_20k_1to10kb_rocprof.md"


def traffic_from_profile(path=TRAFFIC_SOURCE):
    try:
        with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), path)) as f:
            m = re.search(r"traffic_bytes:\s*(\d+)", f.read())
    except OSError:
        return None
    return int(m.group(1)) if m else None


SW = dict(match=2, mismatch=-4, gap_open=-4, gap_extend=-2)
WFA_PEN = dict(x=4, o=6, e=2)
# the host's CPU share for one GPU: at most this many worker threads
CPU_THREADS_CAP = 16


def host_info():
    """nproc, the threads this process may use, the CPU model."""
    model = ""
    try:
        model = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo")
                     if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "allowed_cpus": allowed, "cpu_model": model,
            "threads_used": max(1, min(CPU_THREADS_CAP, allowed))}


def load_oracle():
    """The CPU oracle (the checker; here the timed CPU baseline), its tuned
    build where the host supports it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    flags = oracle.select_fast_build()
    return oracle, flags


def load_package():
    spec = importlib.util.spec_from_file_location(
        "cgm_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["cgm_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class SwRunner:
    """Device-resident batch + one launch per step on the device stream."""

    def __init__(self, m, dev, w):
        self.m, self.dev, self.w = m, dev, w
        self.n = w.n_pairs
        self.q = dev.upload(w.q_packed)
        self.r = dev.upload(w.r_packed)
        self.ql = dev.upload(w.q_len)
        self.rl = dev.upload(w.r_len)
        self.qo = dev.upload(w.q_off)
        self.ro = dev.upload(w.r_off)
        self.band = dev.upload(w.band)
        self.cfg = m.AlignConfig(SW["match"], SW["mismatch"], SW["gap_open"], SW["gap_extend"], 0)
        self.out = m.SwResult(self.n)

    def launch(self, layout):
        m = self.m
        m._lib.call("cgm_sw_banded_launch", self.q.dptr, self.r.dptr, self.ql.dptr, self.rl.dptr,
                    self.qo.dptr, self.ro.dptr, self.n, self.q.count, self.r.count, self.band.dptr,
                    0, SW["match"], SW["mismatch"], SW["gap_open"], SW["gap_extend"],
                    self.out.scores.dptr, self.out.query_end.dptr, self.out.ref_end.dptr, layout,
                    self.dev.stream)

    def results(self):
        d = self.dev
        return (d.download(self.out.scores, self.n), d.download(self.out.query_end, self.n),
                d.download(self.out.ref_end, self.n))


def wfa_extra(m, dev, gen, n_pairs=20_000, cpu=True):
    """Gap-affine WFA on 20,000 pairs of 1-10 kb (x=4, o=6, e=2): pairs/s of
    the mixed persistent layout on an autotuned plan, of the automatic
    launcher (the WFA persistent workgroup-per-pair kernel) and of the plain
    one-workgroup-per-pair baseline."""
    w = gen.generate(gen.WFA_1_10KB, n_pairs)
    up = {k: dev.upload(getattr(w, k)) for k in ("q_packed", "r_packed", "q_len", "r_len",
                                                  "q_off", "r_off", "max_error", "kind", "band")}
    out = m.DeviceArray(w.n_pairs, np.int32)
    res = {}
    costs = {}
    n = w.n_pairs
    bufs = tuple(up[k] for k in ("q_packed", "r_packed", "q_len", "r_len", "q_off", "r_off"))
    cfg = m.AlignConfig(SW["match"], SW["mismatch"], SW["gap_open"], SW["gap_extend"], 0)
    plan = m.MixedPlan(n, int(w.max_error.max()))
    pout = m.SwResult(n)
    # every pair is a WFA pair: the plan takes the kind as one value
    tuning, best_ms, start_ms = plan.autotune(dev, bufs, n, 1, up["band"], up["max_error"], cfg,
                                              pout, reps=2)
    wall, kms = time_steps(m, dev, lambda: plan.align(dev, bufs, n, 1, up["band"],
                                                      up["max_error"], cfg, pout), 3, 1)
    costs["persistent"] = dev.download(pout.scores, n)
    res["persistent"] = {"pairs_per_s": n * 3 / wall, "ms_per_step": wall / 3 * 1e3,
                         "tuning": tuning, "autotune_start_ms": start_ms,
                         "autotune_best_ms": best_ms}
    # the WFA kernels' own span, from the plan's events
    wfa_ms = []
    for _ in range(3):
        plan.align(dev, bufs, n, 1, up["band"], up["max_error"], cfg, pout)
        wfa_ms.append(plan.wfa_kernel_ms())
    kms_wfa = float(np.median(wfa_ms))
    log(f"[bench] wfa persistent (autotuned): {res['persistent']}, WFA kernels {kms_wfa:.3f} ms")
    plan.close()

    def launch(layout):
        m._lib.call("cgm_wfa_affine_launch", up["q_packed"].dptr, up["r_packed"].dptr,
                    up["q_len"].dptr, up["r_len"].dptr, up["q_off"].dptr, up["r_off"].dptr,
                    w.n_pairs, up["q_packed"].count, up["r_packed"].count, up["max_error"].dptr,
                    0, 4, 6, 2, out.dptr, layout, dev.stream)

    for name, layout, steps in (("auto", 0, 3), ("plain", 1, 1)):
        wall, kms = time_steps(m, dev, lambda: launch(layout), steps, 1)
        costs[name] = dev.download(out, w.n_pairs)
        res[name] = {"pairs_per_s": w.n_pairs * steps / wall, "ms_per_step": wall / steps * 1e3}
        log(f"[bench] wfa {name}: {res[name]}")
    res["identical"] = all(bool(np.array_equal(costs[k], costs["plain"]))
                           for k in ("persistent", "auto"))
    res["speedup_auto_vs_plain"] = res["auto"]["pairs_per_s"] / res["plain"]["pairs_per_s"]
    res["mean_cost"] = float(costs["plain"].mean())
    res["pairs"] = w.n_pairs
    # roofline of the WFA kernels: 8 int ops per algorithmic wavefront cell
    cells = gen.wfa_wavefront_cells(w.q_len, w.r_len, costs["plain"], w.max_error)
    total = int(cells.sum())
    achieved = WFA_OPS_PER_CELL * total / (kms_wfa * 1e-3) / 1e12
    in_bytes = int(w.q_packed.nbytes + w.r_packed.nbytes + 5 * 4 * n)
    res["roofline"] = {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TOPS,
                       "unit": "TOP/s (int32)", "frac": achieved / VALU_PEAK_TOPS,
                       "traffic": traffic_from_profile(WFA_TRAFFIC_SOURCE),
                       "traffic_source": WFA_TRAFFIC_SOURCE,
                       "kernel": "WFA class kernels (plan events)",
                       "kernel_ms": kms_wfa, "cells": total, "ops_per_cell": WFA_OPS_PER_CELL,
                       "hbm_algorithmic_bytes": in_bytes + 4 * n,
                       "hbm_frac": (in_bytes + 4 * n) / (kms_wfa * 1e-3) / 1e9 / HBM_PEAK_GBS}
    log(f"[bench] wfa roofline: {res['roofline']}")
    if cpu:
        res["cpu_baseline"] = wfa_cpu_baseline(w, costs["plain"])
    return res


def wfa_cpu_baseline(w, gpu_costs, n_multi=6000, n_single=400):
    """or_wfa_affine (the C restatement of align_cpu.mojo:75-226) on this
    host: a random sample of the workload on the host's threads, and a smaller
    one on 1 thread; pairs/s of each, checked against the GPU costs."""
    oracle, flags = load_oracle()
    info = host_info()
    rng = np.random.default_rng(5)
    res = {"unit": "pairs/s", "kind": "port", "oracle_build": flags}
    for name, k, thr in (("multi", n_multi, info["threads_used"]), ("single", n_single, 1)):
        idx = np.sort(rng.choice(w.n_pairs, min(k, w.n_pairs), replace=False))
        sub = w.subset(idx)
        t = time.perf_counter()
        c = oracle.batch(1, sub.q_packed, sub.r_packed, sub.q_off, sub.r_off, sub.q_len,
                         sub.r_len, max_error=sub.max_error, threads=thr, **WFA_PEN)[0]
        sec = time.perf_counter() - t
        res[name] = {"pairs_per_s": len(idx) / sec, "threads": thr, "seconds": sec,
                     "sample": f"{len(idx)} random pairs of the workload",
                     "bit_exact_vs_gpu": bool(np.array_equal(c, gpu_costs[idx]))}
    res["value"] = res["multi"]["pairs_per_s"]
    res["cores"] = res["multi"]["threads"]
    res["sample"] = res["multi"]["sample"]
    log(f"[bench] wfa cpu: {res}")
    return res


def sw_150bp_extra(m, dev, gen, cpu, steps=20):
    """Plain banded SW (the CPU reference's own case): 1,000 pairs of 150 bp,
    band 32, 2/-4/-4/-2: GCUPS of the one-workgroup-per-pair kernel and of
    the automatic layout, and the oracle on 1 thread over all pairs."""
    w = gen.generate(gen.SW_150BP, 1000)
    run = SwRunner(m, dev, w)
    cells = int(gen.valid_band_cells(w.q_len, w.r_len, w.band).sum())
    res = {"pairs": w.n_pairs, "valid_cells": cells}
    got = {}
    for name, layout in (("plain", 1), ("auto", 0)):
        wall, kms = time_steps(m, dev, lambda: run.launch(layout), steps, 3)
        got[name] = run.results()
        res[name] = {"gcups": cells * steps / wall / 1e9, "ms_per_step": wall / steps * 1e3,
                     "pairs_per_s": w.n_pairs * steps / wall}
    res["identical"] = all(bool(np.array_equal(a, b)) for a, b in zip(got["plain"], got["auto"]))
    if cpu:
        oracle, flags = load_oracle()
        t = time.perf_counter()
        c = oracle.batch(0, w.q_packed, w.r_packed, w.q_off, w.r_off, w.q_len, w.r_len,
                         band=w.band, threads=1, **SW)
        sec = time.perf_counter() - t
        res["cpu_baseline"] = {"value": cells / sec / 1e9, "unit": "GCUPS", "cores": 1,
                               "kind": "port", "oracle_build": flags, "seconds": sec,
                               "sample": "all 1000 pairs, 1 thread",
                               "bit_exact_vs_gpu": all(bool(np.array_equal(a, b))
                                                       for a, b in zip(c, got["plain"]))}
    log(f"[bench] sw 1k x 150 bp: {res}")
    return res


def mixed_extra(m, dev, gen, n_pairs=100_000, cpu=True):
    """Mixed batch (50% SW 100-300 bp band 32, 40% SW 1-2 kb band 128, 10%
    WFA 5-20 kb): pairs/s of the persistent task-queue kernel on an
    autotuned plan replayed from a hipGraph, of the same enqueued eagerly,
    and of the one-workgroup-per-pair kernels."""
    w = gen.generate(gen.MIXED, n_pairs)
    n = w.n_pairs
    bufs = tuple(dev.upload(a) for a in (w.q_packed, w.r_packed, w.q_len, w.r_len, w.q_off,
                                          w.r_off))
    kind, band, me = (dev.upload(a) for a in (w.kind, w.band, w.max_error))
    cfg = m.AlignConfig(SW["match"], SW["mismatch"], SW["gap_open"], SW["gap_extend"], 0)
    out = m.SwResult(n)
    plan = m.MixedPlan(n, int(w.max_error.max()))
    res = {"pairs": n}
    t0 = time.perf_counter()
    tuning, best_ms, start_ms = plan.autotune(dev, bufs, n, kind, band, me, cfg, out, reps=1)
    res["autotune"] = {"tuning": tuning, "start_ms": start_ms, "best_ms": best_ms,
                       "seconds": time.perf_counter() - t0}
    log(f"[bench] mixed autotune: {res['autotune']}")
    plan.capture(dev, bufs, n, kind, band, me, cfg, out)
    wall, kms = time_steps(m, dev, lambda: plan.replay(dev), 3, 1)
    graph_res = [dev.download(a, n) for a in (out.scores, out.query_end, out.ref_end)]
    res["persistent_graph"] = {"pairs_per_s": n * 3 / wall, "ms_per_step": wall / 3 * 1e3}
    wall, kms = time_steps(m, dev, lambda: plan.align(dev, bufs, n, kind, band, me, cfg, out), 2, 0)
    res["persistent_eager"] = {"pairs_per_s": n * 2 / wall, "ms_per_step": wall / 2 * 1e3}
    sw_ms, wfa_ms = [], []
    for _ in range(2):
        plan.align(dev, bufs, n, kind, band, me, cfg, out)
        sw_ms.append(plan.kernel_ms())
        wfa_ms.append(plan.wfa_kernel_ms())
    log(f"[bench] mixed persistent: {res['persistent_graph']} eager {res['persistent_eager']}")
    # algorithmic work: SW valid band cells x 11 ops + WFA wavefront cells x 8
    is_wfa = w.kind == 1
    sw_cells = int(gen.valid_band_cells(w.q_len[~is_wfa], w.r_len[~is_wfa],
                                        w.band[~is_wfa]).sum())
    wfa_cells = int(gen.wfa_wavefront_cells(w.q_len[is_wfa], w.r_len[is_wfa],
                                            graph_res[0][is_wfa], w.max_error[is_wfa]).sum())
    step_ms = res["persistent_graph"]["ms_per_step"]
    ops = SW_OPS_PER_CELL * sw_cells + WFA_OPS_PER_CELL * wfa_cells
    res["roofline"] = {"bound": "valu", "achieved": ops / (step_ms * 1e-3) / 1e12,
                       "peak": VALU_PEAK_TOPS, "unit": "TOP/s (int32)",
                       "frac": ops / (step_ms * 1e-3) / 1e12 / VALU_PEAK_TOPS,
                       "over": "the whole captured step (graph replay)", "traffic": None,
                       "sw_cells": sw_cells, "wfa_cells": wfa_cells,
                       "sw_kernel_ms": float(np.median(sw_ms)),
                       "wfa_kernels_ms": float(np.median(wfa_ms)),
                       "sw_kernel_frac": SW_OPS_PER_CELL * sw_cells /
                       (float(np.median(sw_ms)) * 1e-3) / 1e12 / VALU_PEAK_TOPS,
                       "wfa_kernels_frac": WFA_OPS_PER_CELL * wfa_cells /
                       (float(np.median(wfa_ms)) * 1e-3) / 1e12 / VALU_PEAK_TOPS}
    log(f"[bench] mixed roofline: {res['roofline']}")
    wall, kms = time_steps(m, dev, lambda: m.align_mixed_plain(dev, bufs, n, kind, band, me, cfg,
                                                               out), 1, 0)
    plain_res = [dev.download(a, n) for a in (out.scores, out.query_end, out.ref_end)]
    res["plain_wg_per_pair"] = {"pairs_per_s": n / wall, "ms_per_step": wall * 1e3}
    res["identical"] = all(bool(np.array_equal(a, b)) for a, b in zip(graph_res, plain_res))
    res["speedup_vs_plain"] = res["persistent_graph"]["pairs_per_s"] / \
        res["plain_wg_per_pair"]["pairs_per_s"]
    log(f"[bench] mixed plain: {res['plain_wg_per_pair']} speedup {res['speedup_vs_plain']:.2f}")
    plan.close()
    # the same batch split by kind into the two drop-in launchers' automatic
    # layouts (cgm_sw_banded_launch, cgm_wfa_affine_launch), one after the
    # other on the stream: what the mixed persistent form adds beyond them
    sw_idx, wfa_idx = np.nonzero(~is_wfa)[0], np.nonzero(is_wfa)[0]
    sw_run = SwRunner(m, dev, w.subset(sw_idx))
    ww = w.subset(wfa_idx)
    wup = {k: dev.upload(getattr(ww, k)) for k in ("q_packed", "r_packed", "q_len", "r_len",
                                                    "q_off", "r_off", "max_error")}
    wout = m.DeviceArray(ww.n_pairs, np.int32)

    def per_kind():
        sw_run.launch(0)
        m._lib.call("cgm_wfa_affine_launch", wup["q_packed"].dptr, wup["r_packed"].dptr,
                    wup["q_len"].dptr, wup["r_len"].dptr, wup["q_off"].dptr, wup["r_off"].dptr,
                    ww.n_pairs, wup["q_packed"].count, wup["r_packed"].count,
                    wup["max_error"].dptr, 0, WFA_PEN["x"], WFA_PEN["o"], WFA_PEN["e"],
                    wout.dptr, 0, dev.stream)

    wall, kms = time_steps(m, dev, per_kind, 2, 1)
    sw_got = sw_run.results()
    wfa_got = dev.download(wout, ww.n_pairs)
    res["per_kind_auto"] = {
        "pairs_per_s": n * 2 / wall, "ms_per_step": wall / 2 * 1e3,
        "identical": bool(all(np.array_equal(a, b[sw_idx]) for a, b in zip(sw_got, graph_res))
                          and np.array_equal(wfa_got, graph_res[0][wfa_idx]))}
    res["speedup_vs_per_kind_auto"] = res["persistent_graph"]["pairs_per_s"] / \
        res["per_kind_auto"]["pairs_per_s"]
    log(f"[bench] mixed per-kind automatic launchers: {res['per_kind_auto']}")
    if cpu:
        oracle, flags = load_oracle()
        thr = host_info()["threads_used"]
        idx = np.sort(np.random.default_rng(9).choice(n, 1500, replace=False))
        sub = w.subset(idx)
        t = time.perf_counter()
        c = oracle.batch(0, sub.q_packed, sub.r_packed, sub.q_off, sub.r_off, sub.q_len,
                         sub.r_len, kind=sub.kind, band=sub.band, max_error=sub.max_error,
                         threads=thr, **SW, **WFA_PEN)
        sec = time.perf_counter() - t
        res["cpu_baseline"] = {
            "value": len(idx) / sec, "unit": "pairs/s", "cores": thr, "kind": "port",
            "oracle_build": flags, "seconds": sec,
            "sample": f"{len(idx)} random pairs of the mixed workload ({int(sub.kind.sum())} WFA)",
            "bit_exact_vs_gpu": all(bool(np.array_equal(a, b[idx]))
                                    for a, b in zip(c, graph_res))}
        log(f"[bench] mixed cpu: {res['cpu_baseline']}")
    return res


def serving_extra(m, duration=3.0):
    """Serving loop on requests of 1-64 pairs (log-normal lengths,
    SW band 64 below 2 kb, WFA above): peak sustainable pairs/s in a closed
    loop, then Poisson arrivals at 70% and at 10% of that peak (p50/p99 over
    all requests, over SW-only requests and over requests holding a WFA
    pair), then the 50,000 requests/s arrival rate of the workload
    definition (an overload: admission and back-pressure)."""
    eng = m.ServingEngine(hbm_budget_bytes=16 << 30, n_slots=2, max_batch_pairs=65536,
                          max_batch_words=1 << 19, max_error_cap=8192)
    g = m.GeneratorHandle(5, 400_000, 11)
    res = {"hbm_budget_bytes": 16 << 30, "device_bytes": eng.stats()["device_bytes"],
           "bins": "SW pairs and WFA pairs batched apart, 2 slots (streams) and a dispatcher "
                   "thread each: one hardware queue per slot",
           "sw_bin_tuning": eng.get_tuning(0),
           "client": "one native client thread; closed loop with up to 16,384 requests "
                     "in flight for the peak, Poisson arrivals otherwise"}
    eng.run_load(g.handle, 0.0, 0.5, max_outstanding=16384)  # warm-up
    peak = eng.run_load(g.handle, 0.0, duration, max_outstanding=16384)
    res["peak_closed_loop"] = peak
    for name, frac in (("at_70pct_peak", 0.7), ("at_10pct_peak", 0.1)):
        rate = frac * peak["requests_per_s"]
        res[name] = eng.run_load(g.handle, rate, duration)
        res[name]["offered_rps"] = rate
    res["at_50k_rps"] = eng.run_load(g.handle, 50_000.0, duration)
    # the same load with the sequences resident in HBM (uploaded once,
    # submitted as device pointers: only metadata crosses PCIe per request)
    res["device_resident_peak"] = eng.run_load(g.handle, 0.0, duration, max_outstanding=16384,
                                               device_inputs=True)
    rate = 0.7 * res["device_resident_peak"]["requests_per_s"]
    res["device_resident_at_70pct_peak"] = eng.run_load(g.handle, rate, duration,
                                                        device_inputs=True)
    res["device_resident_at_70pct_peak"]["offered_rps"] = rate
    for k in ("peak_closed_loop", "at_70pct_peak", "at_10pct_peak", "at_50k_rps",
              "device_resident_peak", "device_resident_at_70pct_peak"):
        r = res[k]
        log(f"[bench] serving {k}: {r['requests_per_s']:.0f} req/s {r['pairs_per_s']:.0f} pairs/s "
            f"p50 {r['p50_ms']:.2f} ms p99 {r['p99_ms']:.2f} ms | SW-only p50 "
            f"{r['sw_only_p50_ms']:.2f} p99 {r['sw_only_p99_ms']:.2f} | with WFA p50 "
            f"{r['with_wfa_p50_ms']:.2f} p99 {r['with_wfa_p99_ms']:.2f} | batches "
            f"{r['mean_sw_batch_pairs']:.0f} / {r['mean_wfa_batch_pairs']:.0f} pairs | "
            f"refused {r['refused']:.0f}")
    g.close()
    eng.close()
    return res


def nw_extra(m, dev, w, total_cells, steps=5):
    """Banded NW on the same 200,000-pair batch (band 16-64, 2/-4/-4/-2): GCUPS
    over the same valid band cells, automatic (register lanes) and plain
    (one workgroup per pair) layouts, equal on every pair."""
    up = {k: dev.upload(getattr(w, k)) for k in ("q_packed", "r_packed", "q_len", "r_len",
                                                  "q_off", "r_off", "band")}
    n = w.n_pairs
    out = [m.DeviceArray(n, np.int32) for _ in range(3)]
    res = {}
    got = {}
    for name, layout, st in (("auto", 0, steps), ("plain", 1, 1)):
        def step():
            m._lib.call("cgm_nw_banded_launch", up["q_packed"].dptr, up["r_packed"].dptr,
                        up["q_len"].dptr, up["r_len"].dptr, up["q_off"].dptr, up["r_off"].dptr, n,
                        w.q_packed.size, w.r_packed.size, up["band"].dptr, 0, SW["match"],
                        SW["mismatch"], SW["gap_open"], SW["gap_extend"], out[0].dptr,
                        out[1].dptr, out[2].dptr, layout, dev.stream)
        wall, kms = time_steps(m, dev, step, st, 1)
        got[name] = [dev.download(a, n) for a in out]
        res[name] = {"gcups": total_cells * st / wall / 1e9, "ms_per_step": wall / st * 1e3}
    res["identical"] = all(bool(np.array_equal(a, b)) for a, b in zip(got["auto"], got["plain"]))
    log(f"[bench] nw: {res}")
    return res


def cigar_extra(m, dev, gen, n_pairs=20_000, steps=2):
    """WFA with traceback on 20,000 pairs of 1-10 kb (x=4, o=6, e=2): pairs/s
    of cgm_wfa_cigar_launch (costs, history sizing, history + backtrace),
    beside the cost-only automatic launch on the same batch."""
    w = gen.generate(gen.WFA_1_10KB, n_pairs)
    n = w.n_pairs
    up = {k: dev.upload(getattr(w, k)) for k in ("q_packed", "r_packed", "q_len", "r_len",
                                                  "q_off", "r_off", "max_error")}
    cfg = m.WFAConfig(4, 6, 2, int(w.max_error.max()))
    cap = m.cigar_capacity(w.max_error, cfg)
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum(cap)
    d_off = dev.upload(off)
    d_runs = m.DeviceArray(int(off[-1]), np.uint32)
    d_cost, d_len = m.DeviceArray(n, np.int32), m.DeviceArray(n, np.int32)

    def step():
        m._lib.call("cgm_wfa_cigar_launch", up["q_packed"].dptr, up["r_packed"].dptr,
                    up["q_len"].dptr, up["r_len"].dptr, up["q_off"].dptr, up["r_off"].dptr, n,
                    w.q_packed.size, w.r_packed.size, up["max_error"].dptr, 0, 4, 6, 2,
                    d_cost.dptr, d_runs.dptr, d_off.dptr, d_len.dptr, dev.stream)
    wall, kms = time_steps(m, dev, step, steps, 1)
    cost = dev.download(d_cost)
    lens = dev.download(d_len)
    # history traffic of the traceback kernel: three int16/int32 planes of
    # (cost / 2 + 1) rows x (2 R(cost) + 1) diagonals per pair, written once
    K = np.minimum(np.minimum(w.max_error, np.maximum(w.q_len, w.r_len)),
                   np.where(cost >= 8, (cost - 6) // 2, 0)).astype(np.int64)
    cells = ((cost.astype(np.int64) // 2 + 1) * (2 * K + 1)).sum()
    res = {"pairs": n, "pairs_per_s": n * steps / wall, "ms_per_call": wall / steps * 1e3,
           "mean_runs": float(lens.mean()), "history_cells": int(cells),
           "all_cigars_present": bool(np.all(lens > 0))}
    log(f"[bench] cigar: {res}")
    return res


def ingest_extra(m, dev, gen, n_pairs=200_000, steps=5):
    """Device ingest: the queries and references of the batched SW workload
    (200,000 + 200,000 sequences of ~100-300 bp) as ASCII in HBM, encoded
    to packed words + NMask + word offsets by cgm_encode_ascii_launch;
    GB/s of algorithmic bytes (ASCII in, 12 B per sequence in, 16 B per
    32-base word and 8 B per sequence out) against the 8 TB/s HBM peak."""
    w = gen.generate(gen.SW_100_300BP, n_pairs)
    seqs = m.decode_ascii(w.q_packed, w.q_nmask, w.q_off, w.q_len) + \
        m.decode_ascii(w.r_packed, w.r_nmask, w.r_off, w.r_len)
    buf, starts, lens = m.concat_ascii(seqs)
    n = len(lens)
    cap = n + int(buf.size) // 32 + 1
    d_buf, d_st, d_ln = dev.upload(buf), dev.upload(starts), dev.upload(lens)
    d_packed, d_mask = m.DeviceArray(cap, np.uint64), m.DeviceArray(cap, np.uint64)
    d_off, d_len = m.DeviceArray(n, np.int32), m.DeviceArray(n, np.int32)
    d_tot = m.DeviceArray(1, np.int64)
    step = lambda: m.encode_ascii_launch(dev, d_buf, d_st, d_ln, n, cap, d_packed,  # noqa: E731
                                         d_mask, d_off, d_len, d_tot)
    wall, kms = time_steps(m, dev, step, steps, 2)
    total = int(dev.download(d_tot)[0])
    exp = np.concatenate([w.q_packed, w.r_packed])
    same = total == exp.size and bool(np.array_equal(dev.download(d_packed, total), exp))
    alg = int(buf.size) + 12 * n + 16 * total + 8 * n
    ms = float(np.median(kms))
    res = {"sequences": n, "bases": int(buf.size), "words": total, "ms_per_call": ms,
           "gbases_per_s": buf.size / (ms * 1e-3) / 1e9,
           "algorithmic_bytes": alg, "achieved_gb_s": alg / (ms * 1e-3) / 1e9,
           "hbm_frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "identical_to_generator": same}
    log(f"[bench] ingest: {res}")
    return res


def time_steps(m, dev, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    dev.synchronize()
    timers = [m.Timer(dev.stream) for _ in range(steps)]
    t0 = time.perf_counter()
    for t in timers:
        t.start()
        fn()
        t.stop()
    dev.synchronize()
    wall = time.perf_counter() - t0
    kms = [t.elapsed_ms() for t in timers]
    return wall, kms


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--layout", type=int, default=0, help="0 auto, 1 plain, 2 lane")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extras", action="store_true", help="skip plain / WFA extras")
    ap.add_argument("--pairs", type=int, default=200_000)
    ap.add_argument("--sw-path", choices=["persistent", "buckets"], default="persistent",
                    help="persistent: the task-queue kernel (the automatic layout); buckets: "
                         "one register kernel per band bucket")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run 3 device-to-device copies of this many bytes (counter scale)")
    ap.add_argument("--only-extra", choices=["wfa_20k_1to10kb", "ingest_200k_pairs", "wfa_cigar_20k_1to10kb",
                                             "nw_200k_100to300bp", "serving_1to64_pairs",
                                             "mixed_100k", "sw_1k_150bp"],
                    help="run only this extra workload and print its result (for profiling "
                         "its kernels alone)")
    args = ap.parse_args()
    if args.only_extra:
        m = load_package()
        from cgm_amd import gen
        dev = m.GenomicsDevice(0)
        if args.only_extra == "serving_1to64_pairs":
            res = serving_extra(m)
        elif args.only_extra == "nw_200k_100to300bp":
            w = gen.generate(gen.SW_100_300BP, args.pairs)
            cells = int(gen.valid_band_cells(w.q_len, w.r_len, w.band).sum())
            res = nw_extra(m, dev, w, cells)
        elif args.only_extra == "mixed_100k":
            res = mixed_extra(m, dev, gen, cpu=not args.no_cpu)
        elif args.only_extra == "sw_1k_150bp":
            res = sw_150bp_extra(m, dev, gen, not args.no_cpu)
        else:
            fn = {"wfa_20k_1to10kb": wfa_extra, "ingest_200k_pairs": ingest_extra,
                  "wfa_cigar_20k_1to10kb": cigar_extra}[args.only_extra]
            res = fn(m, dev, gen)
        print(json.dumps({"extra": args.only_extra, **res}), flush=True)
        return
    if args.copy_probe:
        m = load_package()
        dev = m.GenomicsDevice(0)
        a = m.DeviceArray(args.copy_probe, np.uint8)
        b = m.DeviceArray(args.copy_probe, np.uint8)
        for _ in range(3):
            m._lib.call("cgm_memcpy_dtod", b.dptr, a.dptr, args.copy_probe, dev.stream)
        dev.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "copies": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU only; --gpus must be 1 "
                          "(pairs are independent: shard a batch across GPUs by calling the "
                          "library once per GPU)"}))
        sys.exit(2)

    m = load_package()
    from cgm_amd import gen
    dev = m.GenomicsDevice(0)
    info = dev.info()
    host = host_info()
    log(f"[bench] device {info} host {host}")
    t = time.perf_counter()
    w = gen.generate(gen.SW_100_300BP, args.pairs)
    cells = gen.valid_band_cells(w.q_len, w.r_len, w.band)
    total_cells = int(cells.sum())
    band_area = int((w.q_len.astype(np.int64) * (2 * w.band.astype(np.int64) + 1)).sum())
    log(f"[bench] generated {w.n_pairs} pairs, {total_cells / 1e9:.3f} G valid cells "
        f"in {time.perf_counter() - t:.1f}s")
    run = SwRunner(m, dev, w)
    plan = None
    autotune = None
    if args.sw_path == "buckets" or args.layout != 0:
        step = lambda: run.launch(args.layout if args.sw_path == "persistent" else 2)  # noqa: E731
    else:
        # the automatic layout's path with an explicit plan, so the persistent
        # kernel's own duration can be read from the plan's events
        plan = m.MixedPlan(w.n_pairs, 0)
        bufs = (run.q, run.r, run.ql, run.rl, run.qo, run.ro)
        step = lambda: plan.align(dev, bufs, w.n_pairs, None, run.band, None,  # noqa: E731
                                  run.cfg, run.out)
        # autotuned layout (untimed): workgroup shape, band split, switch-over
        t = time.perf_counter()
        tuning, best_ms, start_ms = plan.autotune(dev, bufs, w.n_pairs, None, run.band, None,
                                                  run.cfg, run.out, reps=5)
        autotune = {"tuning": tuning, "start_ms": start_ms, "best_ms": best_ms,
                    "seconds": time.perf_counter() - t}
        log(f"[bench] autotune: {autotune}")

    wall, kms = time_steps(m, dev, step, args.steps, args.warmup)
    kernel_ms = []
    if plan is not None:
        # the same steps again, each followed by a read of the kernel's events
        for _ in range(args.steps):
            step()
            kernel_ms.append(plan.kernel_ms())
    gpu_res = run.results()
    ms_step = wall / args.steps * 1e3
    gcups = total_cells * args.steps / wall / 1e9
    step_ms = float(np.mean(kms))
    kmean = float(np.mean(kernel_ms)) if kernel_ms else step_ms
    achieved_tops = SW_OPS_PER_CELL * total_cells / (kmean * 1e-3) / 1e12
    in_bytes = int(w.q_packed.nbytes + w.r_packed.nbytes + 5 * 4 * w.n_pairs)
    out_bytes = 12 * w.n_pairs
    log(f"[bench] layout {args.layout} {args.sw_path}: {ms_step:.3f} ms/step, {gcups:.1f} GCUPS, "
        f"step events {step_ms:.3f} ms, persistent kernel {kmean:.3f} ms")

    extras = {}
    if not args.no_extras:
        extras["sw_1k_150bp"] = sw_150bp_extra(m, dev, gen, not args.no_cpu)
        extras["wfa_20k_1to10kb"] = wfa_extra(m, dev, gen, cpu=not args.no_cpu)
        extras["mixed_100k"] = mixed_extra(m, dev, gen, cpu=not args.no_cpu)
        extras["serving_1to64_pairs"] = serving_extra(m)
        extras["ingest_200k_pairs"] = ingest_extra(m, dev, gen)
        extras["wfa_cigar_20k_1to10kb"] = cigar_extra(m, dev, gen)
        extras["nw_200k_100to300bp"] = nw_extra(m, dev, w, total_cells)
        wall_p, kms_p = time_steps(m, dev, lambda: run.launch(1), max(2, args.steps // 5), 1)
        plain_res = run.results()
        same = all(np.array_equal(a, b) for a, b in zip(plain_res, gpu_res))
        steps_p = max(2, args.steps // 5)
        extras["plain_wg_per_pair"] = {
            "gcups": total_cells * steps_p / wall_p / 1e9,
            "ms_per_step": wall_p / steps_p * 1e3,
            "pairs_per_s": w.n_pairs * steps_p / wall_p,
            "identical_to_batched": bool(same)}
        log(f"[bench] plain: {extras['plain_wg_per_pair']}")

    cpu = None
    if not args.no_cpu:
        oracle, oflags = load_oracle()  # the checker, here the timed CPU baseline
        threads = host["threads_used"]
        t = time.perf_counter()
        cres = oracle.batch(0, w.q_packed, w.r_packed, w.q_off, w.r_off, w.q_len, w.r_len,
                            band=w.band, threads=threads, **SW)
        cpu_s = time.perf_counter() - t
        match = all(np.array_equal(a, b) for a, b in zip(cres, gpu_res))
        n1 = min(w.n_pairs, 20_000)
        sub = w.subset(np.arange(n1))
        t = time.perf_counter()
        oracle.batch(0, sub.q_packed, sub.r_packed, sub.q_off, sub.r_off, sub.q_len, sub.r_len,
                     band=sub.band, threads=1, **SW)
        cpu1_s = time.perf_counter() - t
        cpu = {"value": total_cells / cpu_s / 1e9, "unit": "GCUPS", "cores": threads,
               "kind": "port", "oracle_build": oflags, "host": host,
               "sample": f"all {w.n_pairs} pairs of the workload, {threads} threads",
               "single_thread_gcups": float(cells[:n1].sum() / cpu1_s / 1e9),
               "single_thread_sample": f"first {n1} pairs, 1 thread",
               "bit_exact_vs_gpu": bool(match)}
        log(f"[bench] cpu: {cpu}")

    # the register lanes of the persistent kernel compute two cells per
    # packed 16-bit instruction when the chosen tuning enables them, in IEEE
    # half precision for every pair of this workload (all values are even
    # integers within +-4096: match * 300 + match <= 2048 * 2), which the
    # packed units run at the int16 rate
    dual = bool(autotune and autotune["tuning"].get("sw_dual", 0))
    peak = VALU_PEAK_TOPS * (2 if dual else 1)
    line = {
        "metric": "SW GCUPS (batched banded SW, valid band cells/s)",
        "value": gcups,
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16 (exact integers)" if dual else "int32",
        "data": "synthetic (seeded splitmix64 generator, inputs resident in HBM)",
        "config": {"workload": "batched banded SW: 200,000 pairs, targets 100-300 bp, "
                               "divergence 1-5%, 0.1% N, band 16-64 per pair, 2/-4/-4/-2",
                   "pairs": w.n_pairs, "valid_cells": total_cells, "band_area_cells": band_area,
                   "layout": args.layout, "sw_path": args.sw_path, "autotune": autotune},
        "roofline": {"bound": "valu", "achieved": achieved_tops, "peak": peak,
                     "unit": "TOP/s (packed 16-bit)" if dual else "TOP/s (int32)",
                     "frac": achieved_tops / peak,
                     "frac_of_int32_peak": achieved_tops / VALU_PEAK_TOPS,
                     "traffic": traffic_from_profile() if plan is not None else None,
                     "traffic_source": TRAFFIC_SOURCE if plan is not None else None,
                     "kernel": "mix_persistent_kernel" if plan is not None else "step",
                     "kernel_ms": kmean, "step_ms": step_ms,
                     "ops_per_cell": SW_OPS_PER_CELL,
                     "hbm_algorithmic_bytes": in_bytes + out_bytes,
                     "hbm_frac": (in_bytes + out_bytes) / (kmean * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
        "band_area_gcups": band_area * args.steps / wall / 1e9,
        "pairs_per_s": w.n_pairs * args.steps / wall,
        "device": info["name"],
        "host": host,
        "extras": extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
