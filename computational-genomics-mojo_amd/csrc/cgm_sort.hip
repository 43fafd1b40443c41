// cgm_sort.hip -- work ordering of a pair batch on the device.
//
// Layouts that put many pairs into one wavefront (one pair per lane) waste
// lanes when the pairs of a wavefront differ in shape, and persistent
// layouts balance better when the largest work goes first.  The helpers
// here compute a per-pair sort key on the GPU and radix-sort pair indices
// by it (hipCUB), so the drop-in launchers stay device-pointer only.

#include <hipcub/hipcub.hpp>

#include "cgm_internal.h"

namespace cgm {

namespace {

__global__ void sw_class_kernel(Batch b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                                SwOut out, uint32_t *keys, int32_t *idx, int32_t *counts) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool in = i < b.n_pairs;
    const uint64_t *q, *r;
    int ql = 0, rl = 0;
    int cls = in ? kSwClassBad : -1, be = 0;
    const bool reg_scoring = sc.mismatch <= 0 && sc.gap_extend <= 0 &&
                             sc.gap_open + sc.gap_extend <= 0;
    if (in && load_pair(b, i, q, r, ql, rl)) {
        const int bd = band ? band[i] : band_scalar;
        if (bd >= 0) {
            be = effective_band(bd, ql, rl);
            const int64_t hmax = (int64_t)max(sc.match, 0) * min(ql, rl);
            if (hmax > kMaxScoreProduct) {
                cls = kSwClassBad;
            } else if (reg_scoring && be <= 64 && hmax < (1 << 23)) {
                cls = be <= 16 ? 0 : be <= 32 ? 1 : be <= 48 ? 2 : 3;
            } else if (be <= kLdsLaneMaxBand) {
                cls = kSwClassLds;
            } else if (be <= kPlainMaxBand) {
                cls = kSwClassPlain;
            }
        }
    }
    // one atomic per class per wavefront
    const int lane = threadIdx.x & 63;
    for (int c = 0; c < kSwClasses; ++c) {
        const uint64_t m = __ballot(cls == c);
        if (m && lane == __ffsll((unsigned long long)m) - 1) atomicAdd(&counts[c], __popcll(m));
    }
    if (!in) return;
    if (cls == kSwClassBad) {
        out.score[i] = CGM_BAD_PAIR;
        out.q_end[i] = -1;
        out.r_end[i] = -1;
    }
    const uint32_t len = (uint32_t)min(max(ql, rl), (1 << 17) - 1);
    keys[i] = ((uint32_t)(7 - cls) << 28) | ((uint32_t)min(be, 2047) << 17) | len;
    idx[i] = (int32_t)i;
}

__global__ void sw_ranges_kernel(const int32_t *counts, int32_t *ranges) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int32_t s = 0;
    for (int c = 0; c < kSwClasses; ++c) {
        ranges[2 * c] = s;
        s += counts[c];
        ranges[2 * c + 1] = s;
    }
}

__global__ void zero_words_kernel(int32_t *p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

}  // namespace

int sort_tmp_bytes(int64_t n, size_t *bytes) {
    *bytes = 0;
    CGM_HIP(hipcub::DeviceRadixSort::SortPairsDescending(
        nullptr, *bytes, (uint32_t *)nullptr, (uint32_t *)nullptr, (int32_t *)nullptr,
        (int32_t *)nullptr, (int)n, 0, 32, (hipStream_t)0));
    return CGM_OK;
}

int sort_desc_by_key_ws(const uint32_t *keys, uint32_t *keys_out, const int32_t *idx,
                        int32_t *order_out, int64_t n, void *tmp, size_t tmp_bytes,
                        hipStream_t stream) {
    size_t need = 0;
    if (int rc = sort_tmp_bytes(n, &need)) return rc;
    if (need > tmp_bytes) return set_error(CGM_ERR_INVALID, "sort workspace too small");
    CGM_HIP(hipcub::DeviceRadixSort::SortPairsDescending(tmp, tmp_bytes, keys, keys_out, idx,
                                                         order_out, (int)n, 0, 32, stream));
    return CGM_OK;
}

int sort_desc_by_key(uint32_t *keys, int32_t *idx, int32_t *order_out, int64_t n,
                     hipStream_t stream) {
    uint32_t *keys_out = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    if (int rc = sort_tmp_bytes(n, &tmp_bytes)) return rc;
    CGM_HIP(hipMallocAsync((void **)&keys_out, (size_t)n * 4, stream));
    CGM_HIP(hipMallocAsync(&tmp, tmp_bytes > 0 ? tmp_bytes : 4, stream));
    int rc = sort_desc_by_key_ws(keys, keys_out, idx, order_out, n, tmp,
                                 tmp_bytes > 0 ? tmp_bytes : 4, stream);
    CGM_HIP(hipFreeAsync(tmp, stream));
    CGM_HIP(hipFreeAsync(keys_out, stream));
    return rc;
}

int sw_classify_ws(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                   SwOut out, const SwClassWs &ws, hipStream_t stream) {
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, stream, ws.counts, 16);
    const unsigned blocks = (unsigned)((b.n_pairs + 255) / 256);
    hipLaunchKernelGGL(sw_class_kernel, dim3(blocks), dim3(256), 0, stream, b, band, band_scalar,
                       sc, out, ws.keys, ws.idx, ws.counts);
    hipLaunchKernelGGL(sw_ranges_kernel, dim3(1), dim3(64), 0, stream, ws.counts, ws.ranges);
    CGM_HIP(hipGetLastError());
    return sort_desc_by_key_ws(ws.keys, ws.keys_out, ws.idx, ws.order, b.n_pairs, ws.sort_tmp,
                               ws.sort_tmp_bytes, stream);
}

int sw_classify(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                SwOut out, int32_t *order, int32_t *ranges, hipStream_t stream) {
    SwClassWs ws;
    ws.order = order;
    ws.ranges = ranges;
    if (int rc = sort_tmp_bytes(b.n_pairs, &ws.sort_tmp_bytes)) return rc;
    if (ws.sort_tmp_bytes == 0) ws.sort_tmp_bytes = 4;
    CGM_HIP(hipMallocAsync((void **)&ws.keys, (size_t)b.n_pairs * 4, stream));
    CGM_HIP(hipMallocAsync((void **)&ws.keys_out, (size_t)b.n_pairs * 4, stream));
    CGM_HIP(hipMallocAsync((void **)&ws.idx, (size_t)b.n_pairs * 4, stream));
    CGM_HIP(hipMallocAsync((void **)&ws.counts, 64, stream));
    CGM_HIP(hipMallocAsync(&ws.sort_tmp, ws.sort_tmp_bytes, stream));
    int rc = sw_classify_ws(b, band, band_scalar, sc, out, ws, stream);
    CGM_HIP(hipFreeAsync(ws.keys, stream));
    CGM_HIP(hipFreeAsync(ws.keys_out, stream));
    CGM_HIP(hipFreeAsync(ws.idx, stream));
    CGM_HIP(hipFreeAsync(ws.counts, stream));
    CGM_HIP(hipFreeAsync(ws.sort_tmp, stream));
    return rc;
}

}  // namespace cgm
