"""GPU parity of the mixed SW + WFA batch path (one persistent kernel over a
work-sorted task queue, plan workspace, hipGraph capture/replay) against the
CPU oracle and against the one-workgroup-per-pair kernels."""
import numpy as np
import pytest

from gpu_helpers import CONFIG_SW, Uploaded, assert_same

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gen_mod(cgm):
    from cgm_amd import gen
    return gen


def mixed_sample(gen, seed=5):
    """SW pairs of 100-300 bp (band 16-64), SW pairs of 1-2 kb (band 128) and
    WFA pairs of 1-10 kb with small costs, interleaved."""
    sw = gen.generate(gen.SW_100_300BP, 3000, seed=seed)
    long_ = gen.generate(gen.MIXED, 3000, seed=seed)
    long_ = long_.subset(np.nonzero((long_.kind == 0) & (long_.band == 128))[0][:300])
    wfa = gen.generate(gen.WFA_1_10KB, 600, seed=seed)
    wfa = wfa.subset(np.argsort(wfa.max_error)[:250])
    w = gen.concat([sw, long_, wfa])
    perm = np.random.default_rng(seed).permutation(w.n_pairs)
    return w.subset(perm)


def oracle_mixed(orc, w):
    return orc.batch(0, w.q_packed, w.r_packed, w.q_off, w.r_off, w.q_len, w.r_len,
                     kind=w.kind, band=w.band, max_error=w.max_error, threads=8, **CONFIG_SW)


def sw_cfg(cgm):
    return cgm.AlignConfig(CONFIG_SW["match"], CONFIG_SW["mismatch"], CONFIG_SW["gap_open"],
                           CONFIG_SW["gap_extend"], 0)


@pytest.mark.parametrize("layout", ["plain", "auto"])
def test_mixed_sample_parity(cgm, dev, orc, gen_mod, layout):
    w = mixed_sample(gen_mod)
    print(f"[case] mixed sample {w.n_pairs} pairs layout={layout}", flush=True)
    got = cgm.align_mixed_batch(dev, w, sw_cfg(cgm), layout={"plain": 1, "auto": 0}[layout])
    assert_same(got, oracle_mixed(orc, w), f"mixed {layout}")


def test_mixed_100k_full_size(cgm, dev, orc, gen_mod):
    """Mixed batch (100,000 pairs incl. 5-20 kb WFA pairs): the persistent
    layout equals the per-pair kernels on every pair; a sample of SW pairs
    and of the cheaper WFA pairs equals the oracle."""
    w = gen_mod.generate(gen_mod.MIXED, 100_000)
    print("[case] mixed full 100000 pairs", flush=True)
    auto = cgm.align_mixed_batch(dev, w, sw_cfg(cgm), layout=0)
    plain = cgm.align_mixed_batch(dev, w, sw_cfg(cgm), layout=1)
    assert_same(auto, plain, "mixed 100k persistent vs plain")
    rng = np.random.default_rng(1)
    sw_idx = rng.choice(np.nonzero(w.kind == 0)[0], 1500, replace=False)
    wfa_idx = np.nonzero(w.kind == 1)[0]
    wfa_idx = wfa_idx[np.argsort(w.max_error[wfa_idx])[:40]]
    idx = np.concatenate([sw_idx, wfa_idx])
    assert_same(tuple(a[idx] for a in auto), oracle_mixed(orc, w.subset(idx)), "mixed 100k sample")


def test_mixed_graph_replay_matches_eager(cgm, dev, gen_mod):
    """Capture one mixed call into a hipGraph and replay it several times,
    with eager calls and synchronisations in between; every replay must
    reproduce the eager result (the queue counter and class counts are
    cleared by kernels inside the graph, not by memset nodes)."""
    w = mixed_sample(gen_mod, seed=9)
    n = w.n_pairs
    nz = lambda a: a if a.size else np.zeros(1, a.dtype)  # noqa: E731
    bufs = tuple(dev.upload(nz(a)) for a in (w.q_packed, w.r_packed, w.q_len, w.r_len, w.q_off,
                                             w.r_off))
    kind, band, me = (dev.upload(a) for a in (w.kind, w.band, w.max_error))
    plan = cgm.MixedPlan(n, int(w.max_error.max()))
    eager_out = cgm.SwResult(n)
    plan.align(dev, bufs, n, kind, band, me, sw_cfg(cgm), eager_out)
    dev.synchronize()
    eager = tuple(dev.download(a, n) for a in (eager_out.scores, eager_out.query_end,
                                               eager_out.ref_end))
    graph_out = cgm.SwResult(n)
    plan.capture(dev, bufs, n, kind, band, me, sw_cfg(cgm), graph_out)
    for rep in range(4):
        fill = np.full(n, 12345, np.int32)
        for a in (graph_out.scores, graph_out.query_end, graph_out.ref_end):
            cgm._lib.call("cgm_memcpy_htod", a.dptr, fill.ctypes.data, fill.nbytes, dev.stream)
        dev.synchronize()
        plan.replay(dev)
        dev.synchronize()
        got = tuple(dev.download(a, n) for a in (graph_out.scores, graph_out.query_end,
                                                 graph_out.ref_end))
        assert_same(got, eager, f"graph replay {rep}")
        # eager work between replays on the same plan
        plan.align(dev, bufs, n, kind, band, me, sw_cfg(cgm), eager_out)
        dev.synchronize()
    plan.close()


def test_mixed_is_deterministic(cgm, dev, gen_mod):
    w = mixed_sample(gen_mod, seed=3)
    a = cgm.align_mixed_batch(dev, w, sw_cfg(cgm))
    b = cgm.align_mixed_batch(dev, w, sw_cfg(cgm))
    assert_same(a, b, "run twice")


TUNINGS = [
    {},
    {"wave_max_width": 0},
    {"wave_max_width": 1 << 30, "blocks_per_cu": 2, "lds_bytes": 79 * 1024},
    {"blocks_per_cu": 1, "lds_bytes": 159 * 1024, "tile_steps": 64},
    {"reg_max_band": 32, "wave_max_width": 512},
    {"lane_max_len": 1024, "tile_steps": 16, "lds_bytes": 24 * 1024},
    {"sw_dual": 0},
    {"sw_dual": 1, "blocks_per_cu": 1, "lds_bytes": 96 * 1024},
    {"lane_max_len": 1},     # every SW pair of band <= 64 on one wavefront per pair
    {"lane_max_len": 200},
]


def upload_mixed(dev, w):
    nz = lambda a: a if a.size else np.zeros(1, a.dtype)  # noqa: E731
    bufs = tuple(dev.upload(nz(a)) for a in (w.q_packed, w.r_packed, w.q_len, w.r_len, w.q_off,
                                             w.r_off))
    return bufs, tuple(dev.upload(nz(a)) for a in (w.kind, w.band, w.max_error))


def run_plan(cgm, dev, plan, bufs, arrs, n):
    out = cgm.SwResult(n)
    plan.align(dev, bufs, n, *arrs, sw_cfg(cgm), out)
    dev.synchronize()
    return tuple(dev.download(a, n) for a in (out.scores, out.query_end, out.ref_end))


def test_mixed_every_tuning_equals_oracle(cgm, dev, orc, gen_mod):
    """Every layout choice (per-wavefront vs per-workgroup WFA, LDS budget,
    tile height, band split, per-lane vs per-workgroup SW) gives the
    oracle's results on a batch that exercises all of them."""
    big = gen_mod.generate(gen_mod.MIXED, 4000, seed=21)
    wf = np.nonzero(big.kind == 1)[0]
    big = big.subset(wf[np.argsort(big.max_error[wf])][[0, 5, 20, 60, 150, 250]])
    w = gen_mod.concat([mixed_sample(gen_mod, seed=13), big])
    exp = oracle_mixed(orc, w)
    bufs, arrs = upload_mixed(dev, w)
    plan = cgm.MixedPlan(w.n_pairs, int(w.max_error.max()))
    base = plan.tuning()
    for t in TUNINGS:
        print(f"[case] tuning {t}", flush=True)
        plan.set_tuning(**{**base, **t})
        assert plan.tuning() == {**base, **t}
        assert_same(run_plan(cgm, dev, plan, bufs, arrs, w.n_pairs), exp, f"tuning {t}")
    plan.close()


def test_mixed_autotune_keeps_results_and_tuning(cgm, dev, orc, gen_mod):
    w = mixed_sample(gen_mod, seed=17)
    exp = oracle_mixed(orc, w)
    bufs, arrs = upload_mixed(dev, w)
    plan = cgm.MixedPlan(w.n_pairs, int(w.max_error.max()))
    out = cgm.SwResult(w.n_pairs)
    print("[case] autotune mixed sample", flush=True)
    best, best_ms, start_ms = plan.autotune(dev, bufs, w.n_pairs, *arrs, sw_cfg(cgm), out, reps=1)
    dev.synchronize()
    got = tuple(dev.download(a, w.n_pairs) for a in (out.scores, out.query_end, out.ref_end))
    assert_same(got, exp, "autotune outputs")
    assert plan.tuning() == best
    assert 0 < best_ms <= start_ms * 1.5
    assert_same(run_plan(cgm, dev, plan, bufs, arrs, w.n_pairs), exp, "after autotune")
    with pytest.raises(cgm._lib.CgmError):
        plan.set_tuning(blocks_per_cu=2, lds_bytes=100 * 1024)
    with pytest.raises(cgm._lib.CgmError):
        plan.set_tuning(reg_max_band=65)
    assert plan.tuning() == best
    plan.close()


def test_mixed_autotune_cached_by_profile(cgm, dev, orc, gen_mod, tmp_path):
    """cgm_plan_autotune_cached: the first call measures and appends its
    choice for the batch's length-bin profile; a new plan on the same
    profile takes it from the file without measuring; a batch of another
    profile misses.  Results equal the oracle in every case."""
    cache = str(tmp_path / "tuning.tsv")
    w = mixed_sample(gen_mod, seed=23)
    exp = oracle_mixed(orc, w)
    bufs, arrs = upload_mixed(dev, w)
    plan = cgm.MixedPlan(w.n_pairs, int(w.max_error.max()))
    out = cgm.SwResult(w.n_pairs)
    key = plan.profile_key(dev, bufs, w.n_pairs, arrs[0], arrs[1], sw_cfg(cgm))
    print(f"[case] cached autotune, key {key}", flush=True)
    import re
    assert key.startswith("dev=") and re.search(r";sw\d+\.\d=", key) and re.search(r";wfa\d+\.0=", key)
    best, best_ms, start_ms, hit = plan.autotune_cached(dev, cache, bufs, w.n_pairs, *arrs,
                                                        sw_cfg(cgm), out, reps=1)
    assert not hit and best_ms > 0 and start_ms > 0
    lines = open(cache).read().splitlines()
    assert len(lines) == 1 and lines[0].split("\t")[0] == key
    dev.synchronize()
    assert_same(tuple(dev.download(a, w.n_pairs) for a in (out.scores, out.query_end, out.ref_end)),
                exp, "cached autotune, measured")
    plan.close()
    # a new plan on the same batch: the choice comes from the file
    plan2 = cgm.MixedPlan(w.n_pairs, int(w.max_error.max()))
    out2 = cgm.SwResult(w.n_pairs)
    best2, ms2, start2, hit2 = plan2.autotune_cached(dev, cache, bufs, w.n_pairs, *arrs,
                                                     sw_cfg(cgm), out2, reps=1)
    assert hit2 and best2 == best and start2 == -1 and abs(ms2 - best_ms) < 1e-3
    assert plan2.tuning() == best
    dev.synchronize()
    assert_same(tuple(dev.download(a, w.n_pairs) for a in (out2.scores, out2.query_end,
                                                          out2.ref_end)), exp, "from the cache")
    assert len(open(cache).read().splitlines()) == 1
    plan2.close()
    # SW pairs only: another profile, measured and appended
    sw = gen_mod.generate(gen_mod.SW_100_300BP, 3000, seed=24)
    up = Uploaded(dev, sw)
    plan3 = cgm.MixedPlan(sw.n_pairs, 0)
    out3 = cgm.SwResult(sw.n_pairs)
    sbufs = (up.q, up.r, up.ql, up.rl, up.qo, up.ro)
    key3 = plan3.profile_key(dev, sbufs, sw.n_pairs, None, up.band, sw_cfg(cgm))
    assert key3 != key and re.search(r";sw\d+\.\d=", key3) and not re.search(r";wfa\d+\.", key3)
    _, _, _, hit3 = plan3.autotune_cached(dev, cache, sbufs, sw.n_pairs, None, up.band, None,
                                          sw_cfg(cgm), out3, reps=1)
    assert not hit3 and len(open(cache).read().splitlines()) == 2
    dev.synchronize()
    from gpu_helpers import sw_oracle
    assert_same(tuple(dev.download(a, sw.n_pairs) for a in (out3.scores, out3.query_end,
                                                           out3.ref_end)),
                sw_oracle(orc, sw, CONFIG_SW), "SW-only profile")
    plan3.close()


def test_mixed_kind_as_one_value(cgm, dev, orc, gen_mod):
    """kind given as one value for the whole batch (0: every pair SW, 1:
    every pair WFA, the form that launches no SW kernel) equals the per-pair
    kind array on the SW and on the WFA pairs of a mixed sample."""
    w = mixed_sample(gen_mod, seed=29)
    exp = oracle_mixed(orc, w)
    for k in (0, 1):
        sub = w.subset(np.nonzero(w.kind == k)[0])
        bufs, arrs = upload_mixed(dev, sub)
        plan = cgm.MixedPlan(sub.n_pairs, int(sub.max_error.max()))
        print(f"[case] kind {k} as one value, {sub.n_pairs} pairs", flush=True)
        by_array = run_plan(cgm, dev, plan, bufs, arrs, sub.n_pairs)
        by_value = run_plan(cgm, dev, plan, bufs, (k, arrs[1], arrs[2]), sub.n_pairs)
        assert_same(by_value, by_array, f"kind {k} value vs array")
        idx = np.nonzero(w.kind == k)[0]
        assert_same(by_value, tuple(a[idx] for a in exp), f"kind {k} vs oracle")
        plan.close()
