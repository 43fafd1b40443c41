// cgm_noalloc.hip -- an allocation-free form of the banded SW launcher.
//
// Some callers cannot allocate device memory during a call: code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc.  For
// them cgm_sw_banded_launch_noalloc runs the band-bucket SW layout
// (CGM_LAYOUT_LANE of cgm_sw_banded_launch: classify, radix-sort, one
// register-lane kernel per band bucket, LDS lanes, one workgroup per pair)
// with every temporary in a file-scope device array, `scratch`, instead of
// stream-ordered allocations, and without forking streams.  Results are
// those of cgm_sw_banded_launch.

#include <mutex>

#include "cgm_internal.h"

namespace cgm {

// ---------------------------------------------------------------------------
// Allocation-free form of the band-bucket layout.  Its workspace (sort keys,
// pair order, class counts and ranges, the radix sort's temporary storage)
// lives in `scratch`, a file-scope device array each GPU holds once for the
// life of the process; host code reaches it only through the address
// hipGetSymbolAddress gives on each device.
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocMaxPairs = int64_t(1) << 18;
constexpr size_t kNoAllocSlot = (size_t)kNoAllocMaxPairs * 4;  // one int32 per pair
constexpr size_t kNoAllocSmall = 4096;                          // counts + ranges
constexpr size_t kNoAllocScratchBytes = size_t(12) << 20;

namespace {
__device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kScratchDevices] = {};

// The current device's copy of `scratch` (looked up once per device).
int scratch_address(void **out) {
    int dev = 0;
    CGM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices) return set_error(CGM_ERR_INVALID, "device index too large");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_error(e, "hipGetSymbolAddress(scratch)");
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return CGM_OK;
}
}  // namespace

static int launch_sw_buckets_noalloc(const Batch &b, const int32_t *band, int32_t band_scalar,
                              SwScoring sc, SwOut out, hipStream_t stream) {
    if (b.n_pairs > kNoAllocMaxPairs)
        return set_error(CGM_ERR_INVALID,
                         "n_pairs above the allocation-free limit of 262144 pairs per call");
    if (int rc = check_batch(b)) return rc;
    if (int rc = check_sw_scoring(sc)) return rc;
    if (b.n_pairs == 0) return CGM_OK;
    void *base = nullptr;
    if (int rc = scratch_address(&base)) return rc;
    uint8_t *p = (uint8_t *)base;
    SwClassWs ws;
    ws.order = (int32_t *)(p + 0 * kNoAllocSlot);
    ws.keys = (uint32_t *)(p + 1 * kNoAllocSlot);
    ws.keys_out = (uint32_t *)(p + 2 * kNoAllocSlot);
    ws.idx = (int32_t *)(p + 3 * kNoAllocSlot);
    ws.counts = (int32_t *)(p + 4 * kNoAllocSlot);
    ws.ranges = ws.counts + 64;
    ws.sort_tmp = p + 4 * kNoAllocSlot + kNoAllocSmall;
    ws.sort_tmp_bytes = kNoAllocScratchBytes - 4 * kNoAllocSlot - kNoAllocSmall;
    if (int rc = sw_classify_ws(b, band, band_scalar, sc, out, ws, stream)) return rc;
    return sw_bucket_kernels(b, band, band_scalar, sc, out, ws.order, ws.ranges, false, stream);
}

}  // namespace cgm

using namespace cgm;

extern "C" int cgm_sw_banded_launch_noalloc(
    const uint64_t *q_packed, const uint64_t *r_packed, const int32_t *q_len,
    const int32_t *r_len, const int32_t *q_off, const int32_t *r_off, int64_t n_pairs,
    int64_t q_words, int64_t r_words, const int32_t *band, int32_t band_scalar, int32_t match,
    int32_t mismatch, int32_t gap_open, int32_t gap_extend, int32_t *out_score,
    int32_t *out_q_end, int32_t *out_r_end, void *stream) {
    Batch b = {q_packed, r_packed, q_len, r_len, q_off, r_off, n_pairs, q_words, r_words};
    SwScoring sc = {match, mismatch, gap_open, gap_extend};
    SwOut out = {out_score, out_q_end, out_r_end};
    if (n_pairs > kNoAllocMaxPairs)
        return set_error(CGM_ERR_INVALID,
                         "n_pairs above the allocation-free limit of 262144 pairs per call");
    if (n_pairs > 0 && (!out_score || !out_q_end || !out_r_end))
        return set_error(CGM_ERR_INVALID, "null output array");
    if (!band && band_scalar < 0) return set_error(CGM_ERR_INVALID, "band_scalar < 0");
    return launch_sw_buckets_noalloc(b, band, band_scalar, sc, out, (hipStream_t)stream);
}
