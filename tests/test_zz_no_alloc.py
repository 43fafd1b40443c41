"""cgm_sw_banded_launch_noalloc (the allocation-free SW entry point) against
cgm_sw_banded_launch on the same inputs, through the C ABI: bit-exact on
score and ends, on two calls in a row, the second with a smaller batch, so
that the shared static workspace is reused with leftovers of the first."""
import numpy as np
import pytest

from gpu_helpers import CONFIG_SW, Uploaded, assert_same, sw_gpu

pytestmark = pytest.mark.gpu


def noalloc_gpu(cgm, dev, up, scoring=CONFIG_SW):
    out = cgm.SwResult(up.n)
    cgm._lib.call("cgm_sw_banded_launch_noalloc", up.q.dptr, up.r.dptr, up.ql.dptr, up.rl.dptr,
                  up.qo.dptr, up.ro.dptr, up.n, up.q_words, up.r_words, up.band.dptr, 0,
                  scoring["match"], scoring["mismatch"], scoring["gap_open"],
                  scoring["gap_extend"], out.scores.dptr, out.query_end.dptr, out.ref_end.dptr,
                  dev.stream)
    dev.synchronize()
    return tuple(dev.download(a, up.n) for a in (out.scores, out.query_end, out.ref_end))


def test_noalloc_equals_allocating_launcher_twice(cgm, dev):
    from cgm_amd import gen
    # a batch with every class of the layout: register buckets, LDS lanes
    # (bands 65-150) and one workgroup per pair (wider bands), then a
    # smaller one
    w1 = gen.generate(gen.SW_100_300BP, 200_000, seed=61)
    rng = np.random.default_rng(3)
    wide = rng.choice(w1.n_pairs, 3000, replace=False)
    w1.band[wide[:1500]] = rng.integers(65, 151, 1500)
    w1.band[wide[1500:]] = rng.integers(151, 400, 1500)
    w2 = gen.generate(gen.SW_150BP, 777, seed=62)
    for w, name in ((w1, "200000 pairs"), (w2, "777 pairs")):
        print(f"[case] noalloc vs allocating launcher, {name}", flush=True)
        up = Uploaded(dev, w)
        exp = sw_gpu(cgm, dev, up, 2)      # CGM_LAYOUT_LANE, allocating
        got = noalloc_gpu(cgm, dev, up)
        assert_same(got, exp, f"noalloc {name}")
        assert_same(got, sw_gpu(cgm, dev, up, 0), f"noalloc {name} vs automatic layout")


def test_noalloc_rejects_batches_above_its_limit(cgm, dev):
    lib = cgm._lib.load()
    rc = lib.cgm_sw_banded_launch_noalloc(None, None, None, None, None, None, (1 << 18) + 1, 0, 0,
                                          None, 8, 2, -4, -4, -2, None, None, None, None)
    assert rc == cgm._lib.CGM_ERR_INVALID
    assert b"262144" in lib.cgm_last_error()
