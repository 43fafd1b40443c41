// cgm_gen.cpp -- seeded synthetic pair batches for the five CGM_WL_* workloads.
//
// Every pair is generated from its own splitmix64 stream keyed by
// (seed, workload, pair index), so a batch is identical whatever the thread
// count.  Bases are iid uniform A/C/G/T.  The query is the reference mutated
// position by position: a substitution (one of the 3 other bases), an
// insertion of `len` random bases before the current base, or a deletion of
// `len` reference bases; afterwards each query position becomes N with
// probability 0.1% (packed as A, flagged in the NMask, as
// genomics/core/sequence.mojo:54-57 does).  The packed layout is that of
// SequenceBatch (sequence.mojo:38-66): each sequence starts on a new word.

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cgm_align.h"

namespace cgm {
int set_error(int code, const std::string &msg);
}

namespace {

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s;
    uint64_t next() { return mix64(s += 0x9E3779B97F4A7C15ull); }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    int64_t range(int64_t lo, int64_t hi) {  // inclusive
        return lo + (int64_t)(((unsigned __int128)next() * (uint64_t)(hi - lo + 1)) >> 64);
    }
    double normal() {  // Box-Muller
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
    int geometric(double p) {  // support >= 1
        int k = 1;
        while (uniform() >= p && k < 10000) ++k;
        return k;
    }
};

// WFA penalties used for the planted cost (mismatch 4, open 6, extend 2).
constexpr int kX = 4, kO = 6, kE = 2;

struct PairData {
    std::vector<uint8_t> q, r;  // codes 0..3, 0xFF = N
    int32_t kind = 0, band = 0, max_error = 0;
};

struct MutModel {
    double p_sub, p_indel;
    int len_kind;  // 0: uniform 1..3, 1: geometric
    double geo_p;
};

void random_bases(Rng &g, int64_t n, std::vector<uint8_t> &out) {
    out.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) out[(size_t)i] = (uint8_t)(g.next() >> 62);
}

// Returns the planted edit cost under (kX, kO, kE).
int64_t mutate(Rng &g, const std::vector<uint8_t> &t, const MutModel &m, std::vector<uint8_t> &q) {
    q.clear();
    q.reserve(t.size() + t.size() / 16 + 8);
    int64_t cost = 0;
    const int64_t L = (int64_t)t.size();
    int64_t i = 0;
    while (i < L) {
        double u = g.uniform();
        if (u < m.p_sub) {
            q.push_back((uint8_t)((t[(size_t)i] + 1 + g.range(0, 2)) & 3));
            cost += kX;
            ++i;
        } else if (u < m.p_sub + m.p_indel) {
            int len = m.len_kind == 0 ? (int)g.range(1, 3) : g.geometric(m.geo_p);
            if (g.uniform() < 0.5) {
                for (int k = 0; k < len; ++k) q.push_back((uint8_t)(g.next() >> 62));
                q.push_back(t[(size_t)i]);
                ++i;
            } else {
                if (len > L - i) len = (int)(L - i);
                i += len;
            }
            cost += kO + (int64_t)kE * len;
        } else {
            q.push_back(t[(size_t)i]);
            ++i;
        }
    }
    return cost;
}

int64_t add_ns(Rng &g, std::vector<uint8_t> &q, double p) {
    int64_t n = 0;
    for (auto &b : q)
        if (g.uniform() < p) {
            b = 0xFF;
            ++n;
        }
    return n;
}

void make_pair(int workload, uint64_t seed, int64_t idx, PairData &pd) {
    Rng g{mix64(mix64(seed ^ ((uint64_t)workload * 0xD1B54A32D192ED03ull)) + (uint64_t)idx)};
    MutModel sw_model{0, 0, 1, 0.5};
    int64_t L = 0;
    bool wfa = false;
    double d = 0;
    MutModel m{};
    switch (workload) {
        case CGM_WL_SW_150BP:
            L = 150;
            m = {0.02, 0.005, 0, 0};
            pd.band = 32;
            break;
        case CGM_WL_SW_100_300BP:
            L = g.range(100, 300);
            d = 0.01 + 0.04 * g.uniform();
            m = {0.75 * d, 0.25 * d, 1, 0.5};
            pd.band = (int32_t)g.range(16, 64);
            break;
        case CGM_WL_WFA_1_10KB:
            L = llround(exp(log(1000.0) + (log(10000.0) - log(1000.0)) * g.uniform()));
            d = 0.005 + 0.045 * g.uniform();
            m = {0.75 * d, 0.25 * d, 1, 0.3};
            wfa = true;
            break;
        case CGM_WL_MIXED: {
            double u = g.uniform();
            if (u < 0.5) {
                L = g.range(100, 300);
                pd.band = 32;
            } else if (u < 0.9) {
                L = g.range(1000, 2000);
                pd.band = 128;
            } else {
                L = g.range(5000, 20000);
                wfa = true;
            }
            d = wfa ? 0.01 + 0.09 * g.uniform() : 0.01 + 0.04 * g.uniform();
            m = {0.75 * d, 0.25 * d, 1, wfa ? 0.3 : 0.5};
            break;
        }
        case CGM_WL_SERVING: {
            double ln = log(250.0) + g.normal();
            L = llround(exp(ln));
            L = std::min<int64_t>(20000, std::max<int64_t>(100, L));
            d = 0.01 + 0.04 * g.uniform();
            m = {0.75 * d, 0.25 * d, 1, 0.5};
            wfa = L >= 2000;
            if (!wfa) pd.band = 64;
            break;
        }
        default:
            m = sw_model;
            break;
    }
    random_bases(g, L, pd.r);
    int64_t cost = mutate(g, pd.r, m, pd.q);
    int64_t nn = add_ns(g, pd.q, 0.001);
    pd.kind = wfa ? 1 : 0;
    pd.max_error = wfa ? (int32_t)(cost + kX * nn + 8) : 0;
}

struct Workload {
    int64_t n = 0, q_words = 0, r_words = 0;
    std::vector<uint64_t> qp, qn, rp, rn;
    std::vector<int32_t> qo, ql, ro, rl, kind, band, me;
};

void pack(const std::vector<uint8_t> &s, uint64_t *w, uint64_t *nm) {
    const size_t nw = (s.size() + 31) / 32;
    for (size_t k = 0; k < nw; ++k) {
        uint64_t word = 0, mask = 0;
        for (size_t i = 0; i < 32 && k * 32 + i < s.size(); ++i) {
            uint8_t b = s[k * 32 + i];
            if (b == 0xFF) {
                mask |= 1ull << i;
                b = 0;
            }
            word |= (uint64_t)b << (62 - 2 * i);
        }
        w[k] = word;
        nm[k] = mask;
    }
}

}  // namespace

extern "C" {

int cgm_gen_create(int32_t workload, int64_t n_pairs, uint64_t seed, void **handle) {
    if (!handle || n_pairs < 0 || workload < CGM_WL_SW_150BP || workload > CGM_WL_SERVING)
        return cgm::set_error(CGM_ERR_INVALID, "cgm_gen_create: bad arguments");
    Workload *w = new (std::nothrow) Workload();
    if (!w) return cgm::set_error(CGM_ERR_NOMEM, "cgm_gen_create: out of memory");
    w->n = n_pairs;
    std::vector<PairData> pairs((size_t)n_pairs);
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (int64_t i = t; i < n_pairs; i += nt) make_pair(workload, seed, i, pairs[(size_t)i]);
        });
    for (auto &x : th) x.join();
    th.clear();
    w->qo.resize((size_t)n_pairs);
    w->ql.resize((size_t)n_pairs);
    w->ro.resize((size_t)n_pairs);
    w->rl.resize((size_t)n_pairs);
    w->kind.resize((size_t)n_pairs);
    w->band.resize((size_t)n_pairs);
    w->me.resize((size_t)n_pairs);
    int64_t qw = 0, rw = 0;
    for (int64_t i = 0; i < n_pairs; ++i) {
        const PairData &p = pairs[(size_t)i];
        w->qo[(size_t)i] = (int32_t)qw;
        w->ro[(size_t)i] = (int32_t)rw;
        w->ql[(size_t)i] = (int32_t)p.q.size();
        w->rl[(size_t)i] = (int32_t)p.r.size();
        w->kind[(size_t)i] = p.kind;
        w->band[(size_t)i] = p.band;
        w->me[(size_t)i] = p.max_error;
        qw += (int64_t)(p.q.size() + 31) / 32;
        rw += (int64_t)(p.r.size() + 31) / 32;
    }
    if (qw >= (int64_t(1) << 31) || rw >= (int64_t(1) << 31)) {
        delete w;
        return cgm::set_error(CGM_ERR_INVALID, "cgm_gen_create: batch exceeds int32 word offsets");
    }
    w->q_words = qw;
    w->r_words = rw;
    w->qp.assign((size_t)qw, 0);
    w->qn.assign((size_t)qw, 0);
    w->rp.assign((size_t)rw, 0);
    w->rn.assign((size_t)rw, 0);
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (int64_t i = t; i < n_pairs; i += nt) {
                const PairData &p = pairs[(size_t)i];
                pack(p.q, w->qp.data() + w->qo[(size_t)i], w->qn.data() + w->qo[(size_t)i]);
                pack(p.r, w->rp.data() + w->ro[(size_t)i], w->rn.data() + w->ro[(size_t)i]);
            }
        });
    for (auto &x : th) x.join();
    *handle = w;
    return CGM_OK;
}

int cgm_gen_sizes(void *handle, int64_t *n_pairs, int64_t *q_words, int64_t *r_words) {
    if (!handle) return cgm::set_error(CGM_ERR_INVALID, "null generator handle");
    Workload *w = (Workload *)handle;
    if (n_pairs) *n_pairs = w->n;
    if (q_words) *q_words = w->q_words;
    if (r_words) *r_words = w->r_words;
    return CGM_OK;
}

int cgm_gen_copy(void *handle, uint64_t *q_packed, uint64_t *q_nmask, int32_t *q_off,
                 int32_t *q_len, uint64_t *r_packed, uint64_t *r_nmask, int32_t *r_off,
                 int32_t *r_len, int32_t *kind, int32_t *band, int32_t *max_error) {
    if (!handle) return cgm::set_error(CGM_ERR_INVALID, "null generator handle");
    Workload *w = (Workload *)handle;
    auto cp = [](void *dst, const void *src, size_t bytes) {
        if (dst && bytes) memcpy(dst, src, bytes);
    };
    cp(q_packed, w->qp.data(), w->qp.size() * 8);
    cp(q_nmask, w->qn.data(), w->qn.size() * 8);
    cp(r_packed, w->rp.data(), w->rp.size() * 8);
    cp(r_nmask, w->rn.data(), w->rn.size() * 8);
    cp(q_off, w->qo.data(), w->qo.size() * 4);
    cp(q_len, w->ql.data(), w->ql.size() * 4);
    cp(r_off, w->ro.data(), w->ro.size() * 4);
    cp(r_len, w->rl.data(), w->rl.size() * 4);
    cp(kind, w->kind.data(), w->kind.size() * 4);
    cp(band, w->band.data(), w->band.size() * 4);
    cp(max_error, w->me.data(), w->me.size() * 4);
    return CGM_OK;
}

int cgm_gen_destroy(void *handle) {
    delete (Workload *)handle;
    return CGM_OK;
}

}  // extern "C"
