// cgm_mixed_dev.h -- device side of the mixed-batch SW persistent kernel,
// shared by its two instances: mix_persistent_kernel<false> (cgm_mixed.hip)
// and mix_persistent_kernel<true> with the wavefront-per-pair phase
// (cgm_mixed_wave.hip), compiled in separate translation units so that the
// two builds run in parallel.  Also the plan's constants and task record.
#pragma once

#include "cgm_sw_dev.h"
#include "cgm_sw_wave_dev.h"
#include "cgm_wfa_reg.h"

namespace cgm {

namespace {

constexpr int kMixClasses = 7;
constexpr int kMixSwReg = 0, kMixSwPacked = 1, kMixWfa = 2, kMixSwLds = 3, kMixSwPlain = 4,
              kMixBad = 5, kMixSwWave = 6;
constexpr int kMixWaves = 4;              // wavefronts per workgroup
constexpr int kMixMaxBlocksPerCu = 2;     // 256 VGPRs: two workgroups per CU at most
constexpr int kMixMaxLds = 160 * 1024;    // LDS of a CU
constexpr int kPackedBand = 128;
// One workgroup = four wavefronts: an SW task holds up to 4 x 64 pairs (one
// lane-group per wavefront).
constexpr int kMixThreads = 256;
constexpr int kMixGroup = kMixThreads;

// Layout of the plan's small device buffer (int32 words).  Register-lane SW
// tasks never straddle two effective bands: a wavefront that holds two bands
// runs once per band, and a task key taken from its first pair would then
// understate its work, which puts long tasks at the end of the queue.
constexpr int kRegBands = 65;                // effective bands 0..64
constexpr int kSmallCounts = 0;              // pairs per class [kMixClasses]
constexpr int kSmallRanges = 8;              // [start, end) per class in order[]
constexpr int kSmallNTasks = 24;             // tasks in the queue
constexpr int kSmallRegTasks = 26;           // register-lane SW tasks
constexpr int kSmallBandCount = 32;          // register-lane pairs per band [kRegBands]
constexpr int kSmallBandTask = 32 + 72;      // first task of each band, widest first [kRegBands+1]
constexpr int kSmallBandPair = 32 + 2 * 72;  // first pair (in the class) of each band [kRegBands]
constexpr int kSmallUnits = 28;              // SW kernel: long/short unit counters (64 bit)
constexpr int kSmallWaveQueue = 30;          // SW wavefront-per-pair kernel: pair counter
constexpr int kSmallWfaCounts = 256;         // WFA pairs per layout class [kWfaClasses]
constexpr int kSmallWfaRanges = kSmallWfaCounts + wfareg::kWfaClasses;  // [start, end) per class
constexpr int kSmallWfaQueues = kSmallWfaRanges + 2 * wfareg::kWfaClasses;  // class kernels' counters
constexpr int kSmallSimdSlots = 512;         // wavefronts started per SIMD [kSimdKeys]
constexpr int kSimdKeys = 8 << 10;           // XCC (3 bits), CU bits of HW_ID (8), SIMD (2)
constexpr int kSmallWords = kSmallSimdSlots + kSimdKeys;

enum TaskType : int32_t { kTaskSwReg = 0, kTaskSwPacked = 1 };

struct Task {
    int32_t type, start, count, pad;
};

// The SW persistent kernel.  Every wavefront takes 64-pair units of the
// longest-first task list by itself (unit u = part u % 4 of task u / 4), so
// no wavefront waits for the others of its workgroup.  The list is cut in
// two queues, both taken longest first: the first sw_split percent of the
// units (the long queue) and the rest (the short queue).  Of two wavefronts
// on one SIMD the older one gets most issue slots, so the first wavefront
// started on a SIMD takes the long queue and, once it is empty, the short
// one; the second takes only the short queue.  The longest units never go
// to a starved wavefront, and the last units of the launch are the shortest
// (one 64-bit counter: long-queue claims in the low half, short-queue
// claims in the high half).
// kWave: the batch may hold class-6 pairs.  The wavefront-per-pair phase is
// compiled into a separate instance because its code changes the register
// allocation of the lane loops (measured ~9 % slower on SW-only batches).
template <bool kWave>
__global__ __launch_bounds__(kMixThreads) __attribute__((amdgpu_waves_per_eu(2, 2))) void
mix_persistent_kernel(Batch b, const int32_t *band_arr, int32_t band_scalar, SwScoring sc,
                      SwOut out, const int32_t *order, const Task *tasks,
                      const int32_t *task_order, const int32_t *n_tasks, int32_t *small,
                      int32_t sw_dual, int32_t sw_split) {
    const int total = *n_tasks;
    const int lane = threadIdx.x & 63;
    const uint32_t units = (uint32_t)total * kMixWaves;
    unsigned long long *ends = (unsigned long long *)(small + kSmallUnits);
    // SW pairs in the wavefront-per-pair layout (reads longer than the
    // tuning's lane_max_len, sorted longest first): each is one serial chain
    // of 2 (qlen + rlen) steps, so every wavefront takes them before any
    // 64-pair unit, the longest first
    if constexpr (kWave) {
        const int32_t w0 = small[kSmallRanges + 2 * kMixSwWave];
        const int32_t w1 = small[kSmallRanges + 2 * kMixSwWave + 1];
        if (w1 > w0) {
            for (;;) {
                int item = 0;
                if (lane == 0) item = atomicAdd(small + kSmallWaveQueue, 1);
                item = __builtin_amdgcn_readfirstlane(item);
                if (item >= w1 - w0) break;
                const int64_t pair = order[w0 + item];
                const uint64_t *q, *r;
                int ql, rl;
                load_pair(b, pair, q, r, ql, rl);  // checked by the classifier
                const int bd = band_arr ? band_arr[pair] : band_scalar;
                int score, qe, re;
                swwave::sw_wave_pair(q, ql, r, rl, effective_band(bd, ql, rl), sc, score, qe, re);
                if (lane == 0) {
                    out.score[pair] = score;
                    if (out.q_end) out.q_end[pair] = qe;
                    if (out.r_end) out.r_end[pair] = re;
                }
            }
        }
    }
    int slot = 0;
    if (lane == 0) {
        const uint32_t hw = __builtin_amdgcn_s_getreg((31 << 11) | 4);    // HW_ID
        const uint32_t xcc = __builtin_amdgcn_s_getreg((15 << 11) | 20);  // XCC_ID
        const uint32_t key = ((xcc & 7u) << 10) | (((hw >> 8) & 0xffu) << 2) | ((hw >> 4) & 3u);
        slot = atomicAdd(small + kSmallSimdSlots + key, 1);
    }
    const bool second = (__shfl(slot, 0, 64) & 1) != 0;
    const uint32_t split = (uint32_t)(((uint64_t)units * (uint32_t)sw_split) / 100u);
    for (;;) {
        int u = -1;
        if (lane == 0) {
            if (!second) {
                const uint32_t a = (uint32_t)atomicAdd(ends, 1ull);
                if (a < split) u = (int)a;
            }
            if (u < 0) {
                const uint32_t c = (uint32_t)(atomicAdd(ends, 1ull << 32) >> 32);
                if ((uint64_t)split + c < units) u = (int)(split + c);
            }
        }
        u = __shfl(u, 0, 64);
        if (u < 0) break;
        const Task task = tasks[task_order[u / kMixWaves]];
        const int part = u % kMixWaves;
        const int64_t start = task.start + 64 * part;
        const int count = min(64, task.count - 64 * part);
        if (count <= 0) continue;
        if (task.type == kTaskSwReg)
            swdev::sw_reg_group<0, 64, false, true>(b, band_arr, band_scalar, sc, out, order,
                                                    start, count, sw_dual != 0);
        else
            swdev::sw_reg_group<kPackedBand, kPackedBand, true>(b, band_arr, band_scalar, sc, out,
                                                                order, start, count);
    }
}

}  // namespace

}  // namespace cgm
