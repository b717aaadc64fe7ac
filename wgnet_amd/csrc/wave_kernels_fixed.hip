// Resident forward / adjoint kernels specialised for the reference grid
// (70 x 70 model, 20-cell C-PML: 114 x 114 padded), one workgroup per unit,
// and their time-split work queues; the device templates are in
// fixed_kernels.h.
#include "fixed_kernels.h"

namespace wg {

namespace fx {
template <class G, bool STORE, bool BORN>
__global__ __launch_bounds__(NT) void k_forward_fx(FwdParams P) {
    extern __shared__ float lds[];
    forward_run<G, STORE, BORN>(P, lds, blockIdx.x, 0, P.g.nt, nullptr, nullptr);
}

template <class G, bool DWAV>
__global__ __launch_bounds__(NT) void k_adjoint_fx(AdjParams P) {
    extern __shared__ float lds[];
    adjoint_run<G, DWAV>(P, lds, blockIdx.x, P.g.nt, 0, nullptr, nullptr);
}

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
__device__ __forceinline__ void wait_done(const int* flag, int* err, int limit) {
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
__device__ __forceinline__ bool next_task(int* next, const int* done, int ntask, int units, int* err, int limit,
                                          int* ts, int& p, int& u) {
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

__device__ __forceinline__ void task_done(int* done, int units, int p, int u) {
    // every wave waits until its stores are acknowledged by the L2, then one
    // agent-scope release writes the L2 back before the flag is set
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(done + (size_t)p * units + u, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

template <class G>
__global__ __launch_bounds__(NT) void k_queue_fwd_fx(QueueParams Q) {
    extern __shared__ float lds[];
    __shared__ int task_s;
    const int parts = Q.parts, units = Q.units, nt = Q.f.g.nt;
    int p, u;
    while (next_task(Q.next_f, Q.done_f, parts * units, units, Q.err, Q.spin_limit, &task_s, p, u)) {
        float* st = Q.fstate + (size_t)u * FState<G>::FLOATS;
        const int t0 = (int)((long long)nt * p / parts), t1 = (int)((long long)nt * (p + 1) / parts);
        forward_run<G, true, false>(Q.f, lds, u, t0, t1, p > 0 ? st : nullptr, p < parts - 1 ? st : nullptr);
        task_done(Q.done_f, units, p, u);
    }
}

template <class G>
__global__ __launch_bounds__(NT) void k_queue_adj_fx(QueueParams Q) {
    extern __shared__ float lds[];
    __shared__ int task_s;
    const int parts = Q.parts, units = Q.units, nt = Q.a.g.nt;
    int j, u;
    while (next_task(Q.next_a, Q.done_a, parts * units, units, Q.err, Q.spin_limit, &task_s, j, u)) {
        float* st = Q.astate + (size_t)u * AState<G>::FLOATS;
        const int sg = parts - 1 - j;
        const int t_lo = (int)((long long)nt * sg / parts), t_hi = (int)((long long)nt * (sg + 1) / parts);
        adjoint_run<G, false>(Q.a, lds, u, t_hi, t_lo, j > 0 ? st : nullptr, j < parts - 1 ? st : nullptr);
        task_done(Q.done_a, units, j, u);
    }
}

using G70 = Geo<70, 70, 20>;

}  // namespace fx

// ---------------------------------------------------------------------------
// Dispatch: the compiled geometry, or false when the grid is another one.
// ---------------------------------------------------------------------------
static bool is_g70(const Grid& g) {
    return g.nvx == 70 && g.nvz == 70 && g.pad == 22 && g.region == 24 && g.region_adj == 26;
}

size_t forward_lds_bytes_fx(const Grid& g) { return is_g70(g) ? 4ull * fx::G70::F_TOTAL : 0; }

size_t adjoint_lds_bytes_fx(const Grid& g, const Geometry& geo) {
    if (!is_g70(g)) return 0;
    return 4ull * (fx::G70::A_TAIL + g.nrec + 8 + geo.band_rows * fx::G70::NX + 1 + 2 * geo.n_ent);
}

bool launch_forward_fx(const FwdParams& p, int units, bool store, bool born, hipStream_t st, hipError_t* err) {
    if (!is_g70(p.g)) return false;
    const size_t lds = forward_lds_bytes_fx(p.g);
    void (*fn)(FwdParams) = born ? fx::k_forward_fx<fx::G70, false, true>
                                 : (store ? fx::k_forward_fx<fx::G70, true, false> : fx::k_forward_fx<fx::G70, false, false>);
    hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fn, dim3(units), dim3(fx::NT), lds, st, p);
        e = hipGetLastError();
    }
    *err = e;
    return true;
}

bool launch_adjoint_fx(const AdjParams& p, int units, hipStream_t st, hipError_t* err) {
    if (!is_g70(p.g)) return false;
    const size_t lds = adjoint_lds_bytes_fx(p.g, p.geo);
    void (*fn)(AdjParams) = p.dwav ? fx::k_adjoint_fx<fx::G70, true> : fx::k_adjoint_fx<fx::G70, false>;
    hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fn, dim3(units), dim3(fx::NT), lds, st, p);
        e = hipGetLastError();
    }
    *err = e;
    return true;
}

bool queue_state_fx(const Grid& g, size_t* fwd_floats, size_t* adj_floats) {
    if (!is_g70(g)) return false;
    *fwd_floats = fx::FState<fx::G70>::FLOATS;
    *adj_floats = fx::AState<fx::G70>::FLOATS;
    return true;
}

bool launch_queue_fx(const QueueParams& q, bool adjoint, int workgroups, hipStream_t st, hipError_t* err) {
    if (!is_g70(q.f.g)) return false;
    void (*fn)(QueueParams) = adjoint ? fx::k_queue_adj_fx<fx::G70> : fx::k_queue_fwd_fx<fx::G70>;
    const size_t lds = adjoint ? adjoint_lds_bytes_fx(q.a.g, q.a.geo) : forward_lds_bytes_fx(q.f.g);
    hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess && workgroups > 0) {
        hipLaunchKernelGGL(fn, dim3(workgroups), dim3(fx::NT), lds, st, q);
        e = hipGetLastError();
    }
    *err = e;
    return true;
}

}  // namespace wg
