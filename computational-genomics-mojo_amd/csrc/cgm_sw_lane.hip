// cgm_sw_lane.hip -- per-band-bucket kernels of the register-resident
// one-pair-per-lane SW layout (device code in cgm_sw_dev.h).

#include "cgm_sw_dev.h"

namespace cgm {


// Register lane kernel for bands LO..HI over order[range[0] .. range[1]),
// one wavefront per 64 consecutive entries, persistent over the range.
// Every pair it receives was classified register-eligible (valid, effective
// band in LO..HI, h < 2^23, non-positive penalties).
template <int LO, int HI, int WPE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) void sw_reg_kernel(Batch b, const int32_t *band_arr,
                                                    int32_t band_scalar, SwScoring sc, SwOut out,
                                                    const int32_t *order, const int32_t *range) {
    const int64_t start = range[0], end = range[1];
    for (int64_t base = start + (int64_t)blockIdx.x * 64; base < end;
         base += (int64_t)gridDim.x * 64) {
        const int count = (int)min((int64_t)64, end - base);
        swdev::sw_reg_group<LO, HI, false>(b, band_arr, band_scalar, sc, out, order, base, count);
    }
}

// Band buckets of the register layout, one kernel each so small bands are
// not held to the register budget of the widest: [0,16], [17,32], [33,48],
// [49,64] (the classifier in cgm_sort.hip uses the same bounds).
int launch_sw_reg_buckets(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                          SwOut out, const int32_t *order, const int32_t *ranges, int grid,
                          hipStream_t stream, bool fork) {
    // The buckets are independent: run them side by side on forked streams
    // so that together they fill the GPU.
    hipStream_t aux[4] = {stream, stream, stream, stream};
    if (fork)
        if (int rc = fork_streams(stream, 4, aux)) return rc;
    // waves per SIMD each bucket fits without spills: 5 / 4 / 3 / 2
#define CGM_REG(K, LO, HI, WPE)                                                              \
    hipLaunchKernelGGL((sw_reg_kernel<LO, HI, WPE>), dim3(grid), dim3(64), 0, aux[3 - (K)], b, \
                       band, band_scalar, sc, out, order, ranges + 2 * (K));
    CGM_REG(3, 49, 64, 2)
    CGM_REG(2, 33, 48, 3)
    CGM_REG(1, 17, 32, 4)
    CGM_REG(0, 0, 16, 5)
#undef CGM_REG
    CGM_HIP(hipGetLastError());
    return fork ? join_streams(stream, 4, aux) : CGM_OK;
}


}  // namespace cgm
