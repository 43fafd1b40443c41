"""Helpers for the GPU parity tests: run a Workload through the C ABI
(device pointers, library stream) and through the CPU oracle."""
import numpy as np

CONFIG_SW = dict(match=2, mismatch=-4, gap_open=-4, gap_extend=-2)
DEFAULT_SW = dict(match=2, mismatch=-3, gap_open=-5, gap_extend=-2)
WFA_PEN = dict(x=4, o=6, e=2)


class Uploaded:
    """Device copies of one Workload's arrays."""

    def __init__(self, dev, w):
        self.n = w.n_pairs
        self.q_words = int(w.q_packed.size)
        self.r_words = int(w.r_packed.size)
        nz = lambda a: a if a.size else np.zeros(1, a.dtype)  # noqa: E731
        self.q = dev.upload(nz(w.q_packed))
        self.r = dev.upload(nz(w.r_packed))
        self.ql = dev.upload(nz(w.q_len))
        self.rl = dev.upload(nz(w.r_len))
        self.qo = dev.upload(nz(w.q_off))
        self.ro = dev.upload(nz(w.r_off))
        self.band = dev.upload(nz(w.band))
        self.me = dev.upload(nz(w.max_error))


def sw_gpu(cgm, dev, up, layout, scoring=CONFIG_SW, band_scalar=None):
    cfg = cgm.AlignConfig(scoring["match"], scoring["mismatch"], scoring["gap_open"],
                          scoring["gap_extend"], band_scalar if band_scalar is not None else 0)
    out = cgm.launch_banded_sw(dev, up.q, up.r, up.ql, up.rl, up.qo, up.ro, up.n, up.q_words,
                               band=cfg.band_width, cfg=cfg,
                               band_buf=None if band_scalar is not None else up.band,
                               r_total_words=up.r_words, layout=layout)
    dev.synchronize()
    return (dev.download(out.scores, up.n), dev.download(out.query_end, up.n),
            dev.download(out.ref_end, up.n))


def wfa_gpu(cgm, dev, up, layout, pen=WFA_PEN, me_scalar=None):
    out = cgm.launch_wfa_affine(dev, up.q, up.r, up.ql, up.rl, up.qo, up.ro, up.n, up.q_words,
                                max_error=me_scalar if me_scalar is not None else 0,
                                max_error_buf=None if me_scalar is not None else up.me,
                                r_total_words=up.r_words, layout=layout, **pen)
    dev.synchronize()
    return dev.download(out, up.n)


def sw_oracle(orc, w, scoring=CONFIG_SW, threads=8):
    return orc.batch(0, w.q_packed, w.r_packed, w.q_off, w.r_off, w.q_len, w.r_len,
                     band=w.band, threads=threads, **scoring)


def wfa_oracle(orc, w, pen=WFA_PEN, threads=8):
    return orc.batch(1, w.q_packed, w.r_packed, w.q_off, w.r_off, w.q_len, w.r_len,
                     max_error=w.max_error, threads=threads, **pen)[0]


def workload_from_strings(pairs, band=None, max_error=None):
    """A gen.Workload built from (query, reference) ASCII pairs."""
    from cgm_amd import gen
    from cgm_amd.sequence import encode_ascii, pack_codes

    def cat(seqs):
        ws, ms, offs, lens, o = [], [], [], [], 0
        for s in seqs:
            w, m = pack_codes(encode_ascii(s))
            ws.append(w)
            ms.append(m)
            offs.append(o)
            lens.append(len(s))
            o += w.size
        cat_ = lambda x: np.concatenate(x) if x else np.zeros(0, np.uint64)  # noqa: E731
        return (cat_(ws).astype(np.uint64), cat_(ms).astype(np.uint64),
                np.array(offs, np.int32), np.array(lens, np.int32))

    qp, qn, qo, ql = cat([p[0] for p in pairs])
    rp, rn, ro, rl = cat([p[1] for p in pairs])
    n = len(pairs)
    bd = np.zeros(n, np.int32) if band is None else np.asarray(band, np.int32)
    me = np.zeros(n, np.int32) if max_error is None else np.asarray(max_error, np.int32)
    return gen.Workload(0, 0, qp, qn, qo, ql, rp, rn, ro, rl, np.zeros(n, np.int32), bd, me)


def assert_same(got, exp, what):
    got = [np.asarray(g) for g in got] if isinstance(got, tuple) else [np.asarray(got)]
    exp = [np.asarray(e) for e in exp] if isinstance(exp, tuple) else [np.asarray(exp)]
    for k, (g, e) in enumerate(zip(got, exp)):
        bad = np.nonzero(g != e)[0]
        assert bad.size == 0, (f"{what}: output {k} differs at {bad.size} pairs, first "
                               f"{bad[:5].tolist()}: got {g[bad[:5]].tolist()} "
                               f"expected {e[bad[:5]].tolist()}")
