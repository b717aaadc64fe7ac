// Time-split work queues of the resident kernels (compiled 70 x 70 grid):
// the task fetch, the wait for a unit's previous part and the hand-over,
// shared by the forward queue (wave_kernels_fixed_fwd.hip) and the adjoint
// queue (wave_kernels_fixed.hip).
#pragma once

#include "fixed_kernels.h"

namespace wg {
namespace fx {

// ===========================================================================
// Time-split work queues: each unit's forward (with snapshots) and adjoint
// are cut into `parts` time segments, and one launch runs all forward
// segments, a second all adjoint segments.  Task k = p * units + u is
// segment p of unit u (forward in time order; adjoint segments from the last
// in time to the first).  Workgroups (one per compute unit) take tasks in
// that order from a counter; a task waits for its unit's previous segment,
// whose state it restores from HBM.  Every unit still runs the single-launch
// arithmetic step for step, so results are bit-identical, and the cut units
// pack the compute units evenly where whole units (320 on 256 compute units,
// say) leave a quarter of them idle.  A task only waits for a task taken
// before it by a running workgroup, so a queue cannot deadlock; the wait is
// bounded all the same (err is set and the results are flagged invalid).
// The two bodies are separate kernels: in one kernel their registers would
// be allocated together and the forward's time loop would spill.
// ===========================================================================
// Polls of a wait before it gives up (QueueParams::spin_limit; the engine's
// default is 1 << 24, seconds of s_sleep polling where tasks take
// milliseconds).  A negative limit makes every wait give up at once, which
// lets a test drive the timeout path with every kernel still ending normally.
static __device__ __forceinline__ void wait_done(const int* flag, int* err, int limit) {
    if (limit < 0) {
        __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    for (int n = 0; __hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0; ++n) {
        if (n > limit) {
            __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        __builtin_amdgcn_s_sleep(8);
    }
}

// Takes the next task of a queue and waits for its predecessor; returns
// false when the queue is empty.  All threads of the workgroup call it.
static __device__ __forceinline__ bool next_task(int* next, const int* done, int ntask, int units, int* err,
                                                 int limit, int* ts, int& p, int& u) {
    __syncthreads();  // the previous task is done with LDS and *ts
    if (threadIdx.x == 0) *ts = __hip_atomic_fetch_add(next, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int task = *ts;
    if (task >= ntask) return false;
    p = task / units;
    u = task - p * units;
    if (p > 0) {
        // the acquiring load invalidates this compute unit's L1 and this
        // XCD's L2 before the barrier lets the other waves read
        if (threadIdx.x == 0) wait_done(done + (size_t)(p - 1) * units + u, err, limit);
        __syncthreads();
    }
    return true;
}

static __device__ __forceinline__ void task_done(int* done, int units, int p, int u) {
    // every wave waits until its stores are acknowledged by the L2, then one
    // agent-scope release writes the L2 back before the flag is set
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(done + (size_t)p * units + u, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

using G70 = Geo<70, 70, 20>;

}  // namespace fx

// the compiled geometry, or false when the grid is another one
static inline bool is_g70(const Grid& g) {
    return g.nvx == 70 && g.nvz == 70 && g.pad == 22 && g.region == 24 && g.region_adj == 26;
}

}  // namespace wg
