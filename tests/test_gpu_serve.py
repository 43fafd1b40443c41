"""Serving engine on the GPU: results equal the oracle, a request's result
is the same alone and in a mix, admission back-pressure, the HBM budget and
the native load generator."""
import numpy as np
import pytest

from gpu_helpers import CONFIG_SW, assert_same

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gen_mod(cgm):
    from cgm_amd import gen
    return gen


@pytest.fixture(scope="module")
def serving_pairs(gen_mod):
    """Serving-workload pairs (SW band 64 below 2 kb, WFA above), WFA pairs limited
    to max_error <= 2500 so the oracle check stays short."""
    w = gen_mod.generate(gen_mod.SERVING, 6000, seed=17)
    keep = np.nonzero((w.kind == 0) | (w.max_error <= 2500))[0]
    return w.subset(keep)


def oracle(orc, w):
    return orc.batch(0, w.q_packed, w.r_packed, w.q_off, w.r_off, w.q_len, w.r_len, kind=w.kind,
                     band=w.band, max_error=w.max_error, threads=8, **CONFIG_SW)


def requests_of(n_pairs, rng, max_size=64):
    out, i = [], 0
    while i < n_pairs:
        k = int(rng.integers(1, max_size + 1))
        out.append(np.arange(i, min(i + k, n_pairs)))
        i += k
    return out


def test_engine_results_equal_oracle(cgm, dev, orc, serving_pairs):
    w = serving_pairs
    assert (w.kind == 1).any() and (w.kind == 0).any()
    eng = cgm.ServingEngine(hbm_budget_bytes=8 << 30, n_slots=3, max_batch_pairs=2048,
                            max_batch_words=1 << 17, max_error_cap=int(w.max_error.max()))
    reqs = requests_of(w.n_pairs, np.random.default_rng(0))
    ids = [eng.submit_pairs(w, idx) for idx in reqs]
    got = [np.zeros(w.n_pairs, np.int32) for _ in range(3)]
    for idx, rid in zip(reqs, ids):
        s, q, r, lat = eng.wait(rid, idx.size)
        assert lat >= 0
        got[0][idx], got[1][idx], got[2][idx] = s, q, r
    st = eng.stats()
    print(f"[case] engine {len(reqs)} requests, stats {st}", flush=True)
    assert st["done_requests"] == len(reqs) and st["worker_error"] == 0
    assert_same(tuple(got), oracle(orc, w), "engine vs oracle")
    eng.close()


@pytest.mark.parametrize("lane_max_len", [1, 150, 1 << 30])
def test_engine_bin_tuning_keeps_results(cgm, dev, orc, serving_pairs, lane_max_len):
    """Bin 0's lane/wavefront switch-over moves SW reads between the
    one-lane and the one-wavefront layouts; results stay the oracle's."""
    w = serving_pairs
    eng = cgm.ServingEngine(hbm_budget_bytes=8 << 30, n_slots=2, max_batch_pairs=2048,
                            max_batch_words=1 << 17, max_error_cap=int(w.max_error.max()))
    assert eng.get_tuning(0)["lane_max_len"] == 200
    t = eng.set_tuning(0, lane_max_len=lane_max_len)
    assert eng.get_tuning(0) == t and t["lane_max_len"] == lane_max_len
    with pytest.raises(cgm._lib.CgmError):
        eng.set_tuning(0, sw_dual=5)
    reqs = requests_of(w.n_pairs, np.random.default_rng(1))
    ids = [eng.submit_pairs(w, idx) for idx in reqs]
    got = [np.zeros(w.n_pairs, np.int32) for _ in range(3)]
    for idx, rid in zip(reqs, ids):
        s, q, r, _ = eng.wait(rid, idx.size)
        got[0][idx], got[1][idx], got[2][idx] = s, q, r
    print(f"[case] engine lane_max_len {lane_max_len}: {len(reqs)} requests", flush=True)
    assert_same(tuple(got), oracle(orc, w), f"engine lane_max_len {lane_max_len} vs oracle")
    eng.close()


def test_request_alone_equals_in_mix(cgm, dev, serving_pairs):
    w = serving_pairs
    eng = cgm.ServingEngine(hbm_budget_bytes=8 << 30, n_slots=4, max_batch_pairs=4096,
                            max_batch_words=1 << 17, max_error_cap=int(w.max_error.max()))
    probe = np.arange(100, 164)
    rid = eng.submit_pairs(w, probe)
    alone = eng.wait(rid, probe.size)[:3]
    # the same request in the middle of a burst of others
    rng = np.random.default_rng(5)
    others = requests_of(w.n_pairs, rng)
    ids = [(idx, eng.submit_pairs(w, idx)) for idx in others[: len(others) // 2]]
    rid2 = eng.submit_pairs(w, probe)
    ids += [(idx, eng.submit_pairs(w, idx)) for idx in others[len(others) // 2:]]
    mixed = eng.wait(rid2, probe.size)[:3]
    for idx, i in ids:
        eng.wait(i, idx.size)
    assert eng.stats()["batches"] >= 2
    assert_same(mixed, alone, "alone vs mix")
    eng.close()


def test_admission_back_pressure(cgm, dev, serving_pairs):
    w = serving_pairs
    eng = cgm.ServingEngine(hbm_budget_bytes=4 << 30, n_slots=1, max_batch_pairs=64,
                            max_batch_words=1 << 15, max_error_cap=int(w.max_error.max()))
    accepted, refused = [], 0
    for k in range(400):
        idx = np.arange((k * 32) % (w.n_pairs - 32), (k * 32) % (w.n_pairs - 32) + 32)
        try:
            accepted.append((idx, eng.submit_pairs(w, idx)))
        except cgm._lib.CgmError as exc:
            assert exc.code == cgm._lib.CGM_ERR_BUSY
            refused += 1
    for idx, rid in accepted:
        s = eng.wait(rid, idx.size)[0]
        assert (s != cgm._lib.CGM_BAD_PAIR).all()
    print(f"[case] back-pressure: accepted {len(accepted)} refused {refused}", flush=True)
    assert refused > 0 and len(accepted) > 0
    assert eng.stats()["refused"] == refused
    eng.close()


def test_submit_checks_buffers(cgm, dev, serving_pairs):
    """A pair whose words lie outside the caller's buffers is refused before
    anything is copied; a valid request still runs afterwards."""
    w = serving_pairs
    eng = cgm.ServingEngine(hbm_budget_bytes=4 << 30, n_slots=1, max_batch_pairs=256,
                            max_batch_words=1 << 15, max_error_cap=int(w.max_error.max()))
    idx = np.arange(0, 8)
    qp = w.q_packed[: int(w.q_off[idx].max())]   # too short for the last pair
    with pytest.raises(cgm._lib.CgmError) as ei:
        eng.submit(qp, w.r_packed, w.q_len[idx], w.r_len[idx], w.q_off[idx], w.r_off[idx],
                   w.kind[idx], w.band[idx], w.max_error[idx])
    assert ei.value.code == cgm._lib.CGM_ERR_INVALID
    rid = eng.submit_pairs(w, idx)
    s = eng.wait(rid, idx.size)[0]
    assert (s != cgm._lib.CGM_BAD_PAIR).all()
    assert eng.stats()["done_requests"] == 1
    eng.close()


def test_engine_respects_hbm_budget(cgm, dev):
    with pytest.raises(cgm._lib.CgmError) as ei:
        cgm.ServingEngine(hbm_budget_bytes=1 << 20, n_slots=2, max_batch_pairs=4096,
                          max_batch_words=1 << 16, max_error_cap=4096)
    assert ei.value.code == cgm._lib.CGM_ERR_NOMEM


def test_native_load_generator(cgm, dev):
    eng = cgm.ServingEngine(hbm_budget_bytes=8 << 30, n_slots=4, max_batch_pairs=8192,
                            max_batch_words=1 << 18, max_error_cap=8192)
    g = cgm.GeneratorHandle(5, 20000, 3)
    closed = eng.run_load(g.handle, 0.0, 0.5, max_outstanding=256)
    print(f"[case] closed loop {closed}", flush=True)
    assert closed["completed"] > 0 and closed["bad_results"] == 0
    opened = eng.run_load(g.handle, 2000.0, 0.5)
    print(f"[case] open loop 2000 rps {opened}", flush=True)
    assert opened["completed"] > 0 and opened["bad_results"] == 0
    assert 0 < opened["p50_ms"] <= opened["p99_ms"]
    # per-bin latency: SW-only requests and requests holding a WFA pair
    assert opened["sw_only_requests"] + opened["with_wfa_requests"] == opened["completed"]
    assert opened["sw_only_requests"] > 0 and 0 < opened["sw_only_p50_ms"] <= opened["sw_only_p99_ms"]
    assert opened["mean_sw_batch_pairs"] > 0
    g.close()
    eng.close()


def test_engine_concurrent_submitters(cgm, dev, orc, serving_pairs):
    """Four host threads submit at once (each request reserves its part of a
    staging buffer under the bins' locks and copies into it after releasing
    them); every result equals the oracle, and the native load generator
    with four client threads reports no bad result."""
    import threading
    import time
    w = serving_pairs
    # max_error_cap 8192: the generated serving batch below holds WFA pairs
    # beyond the fixture's (those above the cap would be refused as bad pairs)
    eng = cgm.ServingEngine(hbm_budget_bytes=8 << 30, n_slots=2, max_batch_pairs=1024,
                            max_batch_words=1 << 16, max_error_cap=8192)
    reqs = requests_of(w.n_pairs, np.random.default_rng(5))
    got = [np.zeros(w.n_pairs, np.int32) for _ in range(3)]
    errors = []

    def worker(part):
        try:
            ids = []
            for idx in part:
                while True:
                    try:
                        ids.append((idx, eng.submit_pairs(w, idx)))
                        break
                    except cgm._lib.CgmError as exc:
                        if exc.code != cgm._lib.CGM_ERR_BUSY:
                            raise
                        time.sleep(0.0005)
            for idx, rid in ids:
                s, q, r, _ = eng.wait(rid, idx.size)
                got[0][idx], got[1][idx], got[2][idx] = s, q, r
        except Exception as exc:  # reported below
            errors.append(exc)

    threads = [threading.Thread(target=worker, args=(reqs[k::4],)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    print(f"[case] concurrent submitters: {len(reqs)} requests, stats {eng.stats()}", flush=True)
    assert not errors, errors
    assert_same(tuple(got), oracle(orc, w), "concurrent submitters vs oracle")
    g = cgm.GeneratorHandle(5, 20000, 4)
    st = eng.run_load(g.handle, 0.0, 0.3, max_outstanding=256, clients=4)
    print(f"[case] 4-client closed loop {st}", flush=True)
    assert st["completed"] > 0 and st["bad_results"] == 0
    g.close()
    eng.close()


def test_engine_device_resident_requests(cgm, dev, orc, serving_pairs):
    """Requests whose packed words are already in HBM (cgm_engine_submit_device:
    gathered on the device when their batch launches) interleaved with
    host-word requests in the same batches; every result equals the oracle,
    and the load generator with device-resident inputs reports no bad
    result."""
    w = serving_pairs
    eng = cgm.ServingEngine(hbm_budget_bytes=8 << 30, n_slots=2, max_batch_pairs=2048,
                            max_batch_words=1 << 17, max_error_cap=8192)
    qd, rd = dev.upload(w.q_packed), dev.upload(w.r_packed)
    reqs = requests_of(w.n_pairs, np.random.default_rng(9))
    ids = []
    for k, idx in enumerate(reqs):
        if k % 2 == 0:
            rid = eng.submit_device(qd, rd, w.q_len[idx], w.r_len[idx], w.q_off[idx],
                                    w.r_off[idx], w.kind[idx], w.band[idx], w.max_error[idx])
        else:
            rid = eng.submit_pairs(w, idx)
        ids.append((idx, rid))
    got = [np.zeros(w.n_pairs, np.int32) for _ in range(3)]
    for idx, rid in ids:
        s, q, r, _ = eng.wait(rid, idx.size)
        got[0][idx], got[1][idx], got[2][idx] = s, q, r
    print(f"[case] device-resident and host requests: {len(reqs)}, stats {eng.stats()}", flush=True)
    assert_same(tuple(got), oracle(orc, w), "device-resident requests vs oracle")
    # a device pair outside the declared buffer length is refused before any copy
    with pytest.raises(cgm._lib.CgmError) as ei:
        idx = reqs[-1]
        eng.submit_device(dev.upload(w.q_packed[:1]), rd, w.q_len[idx], w.r_len[idx],
                          w.q_off[idx], w.r_off[idx], w.kind[idx], w.band[idx],
                          w.max_error[idx])
    assert ei.value.code == cgm._lib.CGM_ERR_INVALID
    g = cgm.GeneratorHandle(5, 20000, 6)
    st = eng.run_load(g.handle, 0.0, 0.3, max_outstanding=256, device_inputs=True)
    print(f"[case] device-resident closed loop {st}", flush=True)
    assert st["completed"] > 0 and st["bad_results"] == 0
    g.close()
    eng.close()


def test_engine_request_ids_after_collection(cgm, serving_pairs):
    """A collected request's id is unknown to poll and wait; its slot is
    reused under a new id, and ids of live requests never repeat."""
    w = serving_pairs
    eng = cgm.ServingEngine(hbm_budget_bytes=8 << 30, n_slots=2, max_batch_pairs=2048,
                            max_batch_words=1 << 17, max_error_cap=int(w.max_error.max()))
    sw = np.nonzero(w.kind == 0)[0]
    first = eng.submit_pairs(w, sw[:4])
    s1, _, _, _ = eng.wait(first, 4)
    print("[case] engine ids: collected id is unknown", flush=True)
    for call in (lambda: eng.poll(first), lambda: eng.wait(first, 4)):
        with pytest.raises(cgm._lib.CgmError) as ei:
            call()
        assert ei.value.code == cgm._lib.CGM_ERR_NOTFOUND
    ids = [eng.submit_pairs(w, sw[4 * k:4 * k + 4]) for k in range(1, 40)]
    assert first not in ids and len(set(ids)) == len(ids) and min(ids) > 0
    got = [eng.wait(i, 4)[0] for i in ids]
    s2, _, _, _ = eng.wait(eng.submit_pairs(w, sw[:4]), 4)
    assert np.array_equal(s1, s2)
    assert all(g.shape == (4,) for g in got)
    eng.close()


@pytest.mark.parametrize("reg_max_band", [64, 32])
def test_engine_sw_bands_beyond_the_register_lanes(cgm, dev, orc, gen_mod, reg_max_band):
    """SW requests whose bands reach every SW class (register lanes, LDS
    lanes above the tuning's reg_max_band, packed lanes at 128, one
    workgroup per pair beyond 150, one wavefront per pair for reads over
    200 bp): batches holding them launch those kernels, batches of register
    bands only leave them out; every result equals the oracle."""
    w = gen_mod.generate(gen_mod.SW_100_300BP, 3000, seed=41)
    rng = np.random.default_rng(reg_max_band)
    choices = np.array([0, 3, 16, 33, 64, 65, 100, 128, 150, 300], np.int32)
    w.band[:] = choices[rng.integers(0, choices.size, w.n_pairs)]
    w.band[:1500] = np.minimum(w.band[:1500], 64)  # the first requests: register bands only
    eng = cgm.ServingEngine(hbm_budget_bytes=8 << 30, n_slots=2, max_batch_pairs=1024,
                            max_batch_words=1 << 16, max_error_cap=0)
    eng.set_tuning(0, reg_max_band=reg_max_band)
    reqs = requests_of(w.n_pairs, rng)
    got = [np.zeros(w.n_pairs, np.int32) for _ in range(3)]
    for part in (reqs[: len(reqs) // 2], reqs[len(reqs) // 2:]):
        ids = [eng.submit_pairs(w, idx) for idx in part]
        for idx, rid in zip(part, ids):
            s, q, r, _ = eng.wait(rid, idx.size)
            got[0][idx], got[1][idx], got[2][idx] = s, q, r
    print(f"[case] engine SW bands 0-300, reg_max_band {reg_max_band}: {len(reqs)} requests",
          flush=True)
    exp = orc.batch(0, w.q_packed, w.r_packed, w.q_off, w.r_off, w.q_len, w.r_len, band=w.band,
                    threads=8, **CONFIG_SW)
    assert_same(tuple(got), exp, f"engine SW bands reg_max_band {reg_max_band}")
    assert eng.stats()["worker_error"] == 0
    eng.close()
