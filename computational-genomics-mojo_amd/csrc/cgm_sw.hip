// cgm_sw.hip -- banded affine Smith-Waterman kernels for gfx950.
//
// Semantics are those of smith_waterman_banded (reference
// genomics/cpu/align_cpu.mojo:229-309).  With k = ri - qi and the band
// |k| <= band, every in-matrix band cell (qi, ri) computes
//     sub = match if q[qi] == r[ri] else mismatch
//     e   = max(E(qi, ri-1) + gap_extend, H(qi-1, ri) + gap_open + gap_extend)
//     f   = H(qi, ri-1) + gap_open + gap_extend          (F never extends)
//     h   = max(0, H(qi-1, ri-1) + sub, e, f)
// with H = E = 0 on cells outside the matrix, H = 0 outside the band, and
// E = -1000000 left of the band (align_cpu.mojo:282-283).  The result is the
// maximum h and the first cell reaching it in row-major order (strict `>`,
// align_cpu.mojo:293-296), i.e. the key (h desc, qi asc, ri asc); with no
// positive cell the result is (0, 0, 0).

#include <mutex>

#include "cgm_internal.h"

namespace cgm {

namespace {

constexpr int kNegSentinel = -1000000;  // align_cpu.mojo:282

struct BestCell {
    int h, qi, ri;
};

// (h desc, qi asc, ri asc): true if a beats b.
__device__ __forceinline__ bool better(const BestCell &a, const BestCell &b) {
    if (a.h != b.h) return a.h > b.h;
    if (a.qi != b.qi) return a.qi < b.qi;
    return a.ri < b.ri;
}

__device__ __forceinline__ void write_bad(const SwOut &o, int64_t i) {
    o.score[i] = CGM_BAD_PAIR;
    o.q_end[i] = -1;
    o.r_end[i] = -1;
}

// Pair-level validity shared by every SW layout.
__device__ __forceinline__ bool sw_pair_ok(int qlen, int rlen, int band, int band_cap,
                                           const SwScoring &sc) {
    if (band < 0 || band > band_cap) return false;
    int64_t mn = qlen < rlen ? qlen : rlen;
    int64_t m = sc.match > 0 ? sc.match : 0;
    return m * mn <= kMaxScoreProduct;
}

// ---------------------------------------------------------------------------
// Plain layout: one 128-thread workgroup per pair, anti-diagonal sweep over
// the band with a workgroup barrier per anti-diagonal.  Threads own band
// diagonals j = k + band (strided by 128); three anti-diagonals of H and two
// of E live in LDS.  Bases are read from global memory per cell.  This is
// the straightforward one-workgroup-per-pair baseline.
// ---------------------------------------------------------------------------
constexpr int kPlainThreads = 128;

__device__ void sw_plain_pair(const Batch &b, int64_t pair, const int32_t *band_arr,
                              int32_t band_scalar, int32_t band_cap, const SwScoring &sc,
                              const SwOut &out, int32_t *lds, int *red_h, int *red_q, int *red_r) {
    const int tid = threadIdx.x;
    const uint64_t *q, *r;
    int qlen, rlen;
    const bool ok = load_pair(b, pair, q, r, qlen, rlen);
    int band = band_arr ? band_arr[pair] : band_scalar;
    if (ok && band >= 0) band = effective_band(band, qlen, rlen);
    if (!ok || !sw_pair_ok(qlen, rlen, band, band_cap, sc)) {
        if (tid == 0) write_bad(out, pair);
        return;
    }
    const int W = 2 * band + 1;
    int32_t *H[3] = {lds, lds + W, lds + 2 * W};
    int32_t *E[2] = {lds + 3 * W, lds + 4 * W};
    for (int j = tid; j < 5 * W; j += kPlainThreads) lds[j] = 0;
    __syncthreads();

    const int oe = sc.gap_open + sc.gap_extend;
    BestCell best = {0, 0, 0};
    const int steps = (qlen > 0 && rlen > 0) ? qlen + rlen - 1 : 0;
    for (int t = 0; t < steps; ++t) {
        int32_t *Hc = H[t % 3], *Hp = H[(t + 2) % 3], *Hpp = H[(t + 1) % 3];
        int32_t *Ec = E[t & 1], *Ep = E[(t + 1) & 1];
        for (int j = tid; j < W; j += kPlainThreads) {
            const int k = j - band;
            if ((t + k) & 1) continue;  // no cell of this diagonal on anti-diagonal t
            const int qi = (t - k) >> 1, ri = (t + k) >> 1;
            if (qi < 0 || qi >= qlen || ri < 0 || ri >= rlen) {
                Hc[j] = 0;
                Ec[j] = 0;
                continue;
            }
            const int sub = base_at(q, qi) == base_at(r, ri) ? sc.match : sc.mismatch;
            const int hd = Hpp[j];
            const int hu = (j + 1 < W) ? Hp[j + 1] : 0;
            const int hl = (j > 0) ? Hp[j - 1] : 0;
            const int el = (j > 0) ? Ep[j - 1] : kNegSentinel;
            const int e = max(el + sc.gap_extend, hu + oe);
            const int f = hl + oe;
            const int h = max(max(0, hd + sub), max(e, f));
            Hc[j] = h;
            Ec[j] = e;
            BestCell c = {h, qi, ri};
            if (h > 0 && better(c, best)) best = c;
        }
        __syncthreads();
    }
    red_h[tid] = best.h;
    red_q[tid] = best.qi;
    red_r[tid] = best.ri;
    __syncthreads();
    for (int s = kPlainThreads / 2; s > 0; s >>= 1) {
        if (tid < s) {
            BestCell a = {red_h[tid], red_q[tid], red_r[tid]};
            BestCell c = {red_h[tid + s], red_q[tid + s], red_r[tid + s]};
            if (better(c, a)) {
                red_h[tid] = c.h;
                red_q[tid] = c.qi;
                red_r[tid] = c.ri;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        const bool pos = red_h[0] > 0;
        out.score[pair] = red_h[0];
        out.q_end[pair] = pos ? red_q[0] : 0;
        out.r_end[pair] = pos ? red_r[0] : 0;
    }
    __syncthreads();
}

// Pairs order[range[0] + i] (or pair i when order is NULL, i < n_pairs),
// one workgroup per pair, grid-stride.
__global__ __launch_bounds__(kPlainThreads) void sw_plain_kernel(Batch b, const int32_t *band_arr,
                                                                 int32_t band_scalar, int32_t band_cap,
                                                                 SwScoring sc, SwOut out,
                                                                 const int32_t *order,
                                                                 const int32_t *range,
                                                                 const int32_t *kind) {
    extern __shared__ int32_t lds[];
    __shared__ int red_h[kPlainThreads], red_q[kPlainThreads], red_r[kPlainThreads];
    const int64_t start = range ? range[0] : 0;
    const int64_t end = range ? range[1] : b.n_pairs;
    for (int64_t item = start + blockIdx.x; item < end; item += gridDim.x) {
        const int64_t pair = order ? order[item] : item;
        if (kind && kind[pair] != 0) continue;
        sw_plain_pair(b, pair, band_arr, band_scalar, band_cap, sc, out, lds, red_h, red_q, red_r);
    }
}

// ---------------------------------------------------------------------------
// Lane layout: one pair per lane, rows swept in the reference's own
// row-major order, so the strict-`>` maximum needs no cross-lane reduction.
// Each lane keeps its previous row of H in LDS (layout [j][lane], conflict
// free) and updates it in place; E, the left H and the diagonal H ride in
// registers along the row.  Reference bases come 32 at a time from a 64-bit
// window built by a funnel shift of two packed words and XORed against the
// row's query base replicated 32 times.  Only in-matrix band cells
// j in [jlo, jhi] of each row are visited: cells left of jlo are zero in
// LDS from the start and cells right of jhi are never read by the next row.
// ---------------------------------------------------------------------------
constexpr int kLaneThreads = 64;

__device__ void sw_lds_lane_pair(const Batch &b, int64_t pair, const int32_t *band_arr,
                                 int32_t band_scalar, int32_t band_cap, const SwScoring &sc,
                                 const SwOut &out, int32_t *lds) {
    const int tid = threadIdx.x;
    const uint64_t *q, *r;
    int qlen, rlen;
    const bool ok = load_pair(b, pair, q, r, qlen, rlen);
    int band = band_arr ? band_arr[pair] : band_scalar;
    if (ok && band >= 0) band = effective_band(band, qlen, rlen);
    if (!ok || !sw_pair_ok(qlen, rlen, band, band_cap, sc)) {
        write_bad(out, pair);
        return;
    }
    const int W = 2 * band + 1;
    int32_t *Hrow = lds + tid;  // H of diagonal slot j at Hrow[j * kLaneThreads]
    for (int j = 0; j <= W; ++j) Hrow[j * kLaneThreads] = 0;  // slot W stays 0 (out of band)
    const int oe = sc.gap_open + sc.gap_extend;
    const int ext = sc.gap_extend;
    const int rwords = (rlen + 31) >> 5;
    int best = 0, bq = 0, br = 0;
    for (int qi = 0; qi < qlen; ++qi) {
        const int jlo = max(0, band - qi);
        const int jhi = min(W - 1, rlen - 1 - qi + band);
        if (jlo > jhi) continue;
        const uint64_t qrep = (uint64_t)base_at(q, qi) * 0x5555555555555555ull;
        int hd = Hrow[jlo * kLaneThreads];
        int hl = 0;
        int el = jlo == 0 ? kNegSentinel : 0;
        for (int c0 = jlo; c0 <= jhi; c0 += 32) {
            const int rs = qi - band + c0;  // first reference position of the chunk, >= 0
            const int wi = rs >> 5, sh = 2 * (rs & 31);
            uint64_t win = r[wi];
            if (sh) {
                const uint64_t w1 = (wi + 1 < rwords) ? r[wi + 1] : 0ull;
                win = (win << sh) | (w1 >> (64 - sh));
            }
            const uint64_t x = win ^ qrep;  // 2-bit field u is 0 iff bases match
            const uint32_t xh = (uint32_t)(x >> 32), xl = (uint32_t)x;
            const int n = min(32, jhi - c0 + 1);
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                if (u < n) {
                    const uint32_t half = u < 16 ? xh : xl;
                    const int eq = ((half >> (30 - 2 * (u & 15))) & 3u) == 0;
                    const int j = c0 + u;
                    const int hu = Hrow[(j + 1) * kLaneThreads];
                    const int sub = eq ? sc.match : sc.mismatch;
                    const int e = max(el + ext, hu + oe);
                    const int f = hl + oe;
                    const int h = max(max(0, hd + sub), max(e, f));
                    Hrow[j * kLaneThreads] = h;
                    if (h > best) {
                        best = h;
                        bq = qi;
                        br = qi - band + j;
                    }
                    hd = hu;
                    hl = h;
                    el = e;
                }
            }
        }
    }
    out.score[pair] = best;
    out.q_end[pair] = bq;
    out.r_end[pair] = br;
}

__global__ __launch_bounds__(kLaneThreads) void sw_lane_kernel(Batch b, const int32_t *band_arr,
                                                               int32_t band_scalar, int32_t band_cap,
                                                               SwScoring sc, SwOut out,
                                                               const int32_t *order,
                                                               const int32_t *range) {
    extern __shared__ int32_t lds[];
    const int64_t start = range[0], end = range[1];
    for (int64_t base = start + (int64_t)blockIdx.x * kLaneThreads; base < end;
         base += (int64_t)gridDim.x * kLaneThreads) {
        const int64_t item = base + threadIdx.x;
        if (item < end)
            sw_lds_lane_pair(b, order[item], band_arr, band_scalar, band_cap, sc, out, lds);
    }
}

}  // namespace

int launch_sw_plain(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                    SwOut out, hipStream_t stream) {
    return launch_sw_plain_kind(b, band, band_scalar, sc, out, nullptr, stream);
}

int launch_sw_plain_kind(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                         SwOut out, const int32_t *kind, hipStream_t stream) {
    if (int rc = check_batch(b)) return rc;
    if (int rc = check_sw_scoring(sc)) return rc;
    if (b.n_pairs == 0) return CGM_OK;
    int32_t max_band = 0;
    if (int rc = device_max_band_eff(b, band, band_scalar, stream, &max_band)) return rc;
    if (max_band < 0) max_band = 0;
    int32_t cap = max_band < kPlainMaxBand ? max_band : kPlainMaxBand;
    const size_t lds_bytes = (size_t)5 * (2 * cap + 1) * sizeof(int32_t);
    hipLaunchKernelGGL(sw_plain_kernel, dim3((unsigned)b.n_pairs), dim3(kPlainThreads), lds_bytes,
                       stream, b, band, band_scalar, cap, sc, out, nullptr, nullptr, kind);
    CGM_HIP(hipGetLastError());
    return CGM_OK;
}

int launch_sw_plain_range(const Batch &b, const int32_t *band, int32_t band_scalar,
                          SwScoring sc, SwOut out, const int32_t *order, const int32_t *range,
                          int grid, hipStream_t stream) {
    const size_t lds_bytes = (size_t)5 * (2 * kPlainMaxBand + 1) * sizeof(int32_t);
    hipLaunchKernelGGL(sw_plain_kernel, dim3(grid), dim3(kPlainThreads), lds_bytes, stream, b, band,
                       band_scalar, kPlainMaxBand, sc, out, order, range, nullptr);
    CGM_HIP(hipGetLastError());
    return CGM_OK;
}

int launch_sw_lds_lane(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                       SwOut out, const int32_t *order, const int32_t *range, int grid,
                       hipStream_t stream) {
    const size_t lds_bytes = (size_t)(2 * kLdsLaneMaxBand + 2) * kLaneThreads * sizeof(int32_t);
    if (int rc = set_kernel_lds((const void *)sw_lane_kernel, (int)lds_bytes)) return rc;
    hipLaunchKernelGGL(sw_lane_kernel, dim3(grid), dim3(kLaneThreads), lds_bytes, stream, b, band,
                       band_scalar, kLdsLaneMaxBand, sc, out, order, range);
    CGM_HIP(hipGetLastError());
    return CGM_OK;
}

// Bytes of freed stream-ordered allocations the device's default pool keeps
// for reuse: the launchers allocate their temporaries (sort space, slabs,
// orders) per call with hipMallocAsync, and a pool that hands everything back
// at each synchronisation makes every call fetch (recycled, to-be-cleared)
// memory from the driver again.
constexpr uint64_t kAsyncPoolKeepBytes = uint64_t(8) << 30;

namespace {
std::mutex g_cu_mu;
int g_cu_cached[64] = {0};
}  // namespace

int device_cu_count(int *cus) {
    int dev = 0;
    CGM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_cu_mu);
    if (dev < 64 && g_cu_cached[dev]) {
        *cus = g_cu_cached[dev];
        return CGM_OK;
    }
    int n = 0;
    CGM_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    // first use on this device: keep freed async allocations for reuse
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t keep = kAsyncPoolKeepBytes;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    if (dev < 64) g_cu_cached[dev] = n;
    *cus = n;
    return CGM_OK;
}

// Band-bucket layout: classify + sort on the device, then one launch per
// class (register buckets on forked streams when `fork`, LDS lanes, one
// workgroup per pair).
int sw_bucket_kernels(const Batch &b, const int32_t *band, int32_t band_scalar,
                             SwScoring sc, SwOut out, const int32_t *order, const int32_t *ranges,
                             bool fork, hipStream_t stream) {
    int cus = 0;
    if (int rc = device_cu_count(&cus)) return rc;
    const int64_t waves = (b.n_pairs + 63) / 64;
    const int reg_grid = (int)(waves < (int64_t)cus * 16 ? waves : (int64_t)cus * 16);
    int rc = launch_sw_reg_buckets(b, band, band_scalar, sc, out, order, ranges, reg_grid, stream,
                                   fork);
    if (!rc)
        rc = launch_sw_lds_lane(b, band, band_scalar, sc, out, order, ranges + 2 * kSwClassLds,
                                reg_grid < cus * 2 ? reg_grid : cus * 2, stream);
    if (!rc)
        rc = launch_sw_plain_range(b, band, band_scalar, sc, out, order,
                                   ranges + 2 * kSwClassPlain, cus * 2, stream);
    return rc;
}

int launch_sw_buckets(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                      SwOut out, hipStream_t stream) {
    if (int rc = check_batch(b)) return rc;
    if (int rc = check_sw_scoring(sc)) return rc;
    if (b.n_pairs == 0) return CGM_OK;
    int32_t *order = nullptr, *ranges = nullptr;
    CGM_HIP(hipMallocAsync((void **)&order, (size_t)b.n_pairs * 4, stream));
    CGM_HIP(hipMallocAsync((void **)&ranges, 2 * kSwClasses * 4, stream));
    int rc = sw_classify(b, band, band_scalar, sc, out, order, ranges, stream);
    if (!rc) rc = sw_bucket_kernels(b, band, band_scalar, sc, out, order, ranges, true, stream);
    CGM_HIP(hipFreeAsync(order, stream));
    CGM_HIP(hipFreeAsync(ranges, stream));
    return rc;
}

// Automatic layout: the persistent task-queue kernel of the mixed path with
// every pair an SW pair, on a plan cached per (device, stream) and grown
// when a larger batch arrives.  A batch of at most kWaveBatchPerCu pairs per
// CU runs one wavefront per pair (bands <= 64), pair i on wavefront i: one
// pair per lane would leave most SIMDs idle and every pair's whole band area
// on one lane's serial chain, and with about as many pairs as wavefronts a
// longest-first order has nothing to balance.
constexpr int64_t kWaveBatchPerCu = 16;
int launch_sw_auto(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                   SwOut out, hipStream_t stream) {
    if (int rc = check_batch(b)) return rc;
    if (int rc = check_sw_scoring(sc)) return rc;
    if (b.n_pairs == 0) return CGM_OK;
    int cus = 0;
    if (int rc = device_cu_count(&cus)) return rc;
    if (b.n_pairs <= kWaveBatchPerCu * cus) return launch_sw_wave(b, band, band_scalar, sc, out, stream);
    const WfaPen pen = {4, 6, 2};  // unused: the batch holds no WFA pair
    return with_cached_plan(stream, b.n_pairs, 0, pen, [&](MixPlan *P) {
        return mix_enqueue(P, b, nullptr, 0, band, band_scalar, nullptr, 0, sc, pen, out, stream);
    });
}

int launch_sw_wave(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                   SwOut out, hipStream_t stream) {
    if (int rc = check_batch(b)) return rc;
    if (int rc = check_sw_scoring(sc)) return rc;
    if (b.n_pairs == 0) return CGM_OK;
    const WfaPen pen = {4, 6, 2};  // unused: the batch holds no WFA pair
    return with_cached_plan(stream, b.n_pairs, 0, pen, [&](MixPlan *P) {
        return mix_enqueue_wave_direct(P, b, band, band_scalar, sc, out, stream);
    });
}

}  // namespace cgm
