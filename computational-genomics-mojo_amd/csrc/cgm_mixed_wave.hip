// cgm_mixed_wave.hip -- the SW persistent kernel's instance with the
// wavefront-per-pair phase, in its own translation unit (its device code
// is as large as the plain instance's and compiles in parallel with it).

#include "cgm_mixed_dev.h"

namespace cgm {

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
