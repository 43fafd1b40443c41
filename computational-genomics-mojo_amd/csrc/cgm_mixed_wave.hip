// cgm_mixed_wave.hip -- the SW persistent kernel's instance with the
// wavefront-per-pair phase, in its own translation unit (its device code
// is as large as the plain instance's and compiles in parallel with it).

#include <algorithm>

#include "cgm_mixed_dev.h"

namespace cgm {

namespace {

// One wavefront per SW pair, pair index = wavefront index (grid-stride): the
// small-batch and CGM_LAYOUT_WAVE form, with no classification or sort.
// Pairs the wavefront layout does not take (effective band above 64, both
// sequences empty, score bounds, bad pairs) are appended to `rest` (their
// count in rest_range[1], rest_range[0] = 0) for the one-workgroup-per-pair
// kernel.  Every test below is wavefront-uniform.
__global__ __launch_bounds__(256) void sw_wave_direct_kernel(Batch b, const int32_t *band_arr,
                                                             int32_t band_scalar, SwScoring sc,
                                                             SwOut out, int32_t *rest,
                                                             int32_t *rest_range) {
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t pair = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; pair < b.n_pairs;
         pair += nwaves) {
        const uint64_t *q = nullptr, *r = nullptr;
        int ql = 0, rl = 0, be = 0;
        bool wave = false;
        if (load_pair(b, pair, q, r, ql, rl)) {
            const int bd = band_arr ? band_arr[pair] : band_scalar;
            if (bd >= 0) {
                be = effective_band(bd, ql, rl);
                const int64_t hmax = (int64_t)max(sc.match, 0) * min(ql, rl);
                wave = max(ql, rl) > 0 && be <= kWaveMaxBand && hmax < (1 << 23) &&
                       max(ql, rl) < kWaveMaxLen;
            }
        }
        if (!wave) {
            if (lane == 0) rest[atomicAdd(rest_range + 1, 1)] = (int32_t)pair;
            continue;
        }
        int score, qe, re;
        swwave::sw_wave_pair(q, ql, r, rl, be, sc, score, qe, re);
        if (lane == 0) {
            out.score[pair] = score;
            if (out.q_end) out.q_end[pair] = qe;
            if (out.r_end) out.r_end[pair] = re;
        }
    }
}

}  // namespace

int launch_sw_wave_direct(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                          SwOut out, int32_t *rest, int32_t *rest_range, hipStream_t stream) {
    const int64_t blocks = std::min<int64_t>((b.n_pairs + 3) / 4, 65536);
    hipLaunchKernelGGL(sw_wave_direct_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, b, band,
                       band_scalar, sc, out, rest, rest_range);
    CGM_HIP(hipGetLastError());
    return CGM_OK;
}

int launch_mix_persistent_wave(unsigned grid, hipStream_t stream, const Batch &b,
                               const int32_t *band, int32_t band_scalar, SwScoring sc, SwOut out,
                               const int32_t *order, const void *tasks, const int32_t *task_order,
                               const int32_t *n_tasks, int32_t *small, int32_t sw_dual,
                               int32_t sw_split) {
    hipLaunchKernelGGL(mix_persistent_kernel<true>, dim3(grid), dim3(kMixThreads), 0, stream, b,
                       band, band_scalar, sc, out, order, (const Task *)tasks, task_order,
                       n_tasks, small, sw_dual, sw_split);
    CGM_HIP(hipGetLastError());
    return CGM_OK;
}

}  // namespace cgm
