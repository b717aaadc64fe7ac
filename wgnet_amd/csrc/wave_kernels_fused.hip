// The uneven batch as ONE persistent launch (k_fused), on the compiled
// whole-unit and row-band bodies of fixed_kernels.h.  Its own translation
// unit: it is built with the machine scheduler's max-ilp strategy, which the
// band kernels of small batches do not want.
#include "band_split.h"

namespace wg {

namespace fx {

// ===========================================================================
// An uneven batch in ONE launch.  Task = one 70 x 70 unit, or one band of a
// 100 x 100 or 141 x 141 unit (a unit's bands are consecutive tasks); a task
// belongs to the last group whose task0 it reaches (the table is in task
// order, longest tasks first).  A workgroup takes a task from a device
// counter and runs, for that unit or band: the forward (gathers, snapshots),
// then (the unit's top band) the unit's residual (pred - obs) * scale, then
// the adjoint (its exchanges continue the forward's step numbers).  The
// snapshots go to a slot the workgroup takes from a flag table for the
// task's lifetime, so the snapshot memory is a slot per resident workgroup,
// not per unit of the batch.  Tasks are taken in start order, so a unit's
// bands start one after another and every unit before the newest one is
// complete: no workgroup waits for one that cannot start.  The group's
// parameters are copied out of the table (uniform: scalar loads).
// ===========================================================================
struct FusedSlots {
    float* base;          // ns slots of `floats` floats
    unsigned long long floats;
    int* flags;           // ns flags, 0 = free
    int ns;
    int* err;             // set when no slot was free (then results are flagged invalid)
};

// A free slot (every thread gets the same), or -1.
static __device__ __noinline__ int take_slot(const FusedSlots& S, int start, int* s_slot) {
    if (threadIdx.x == 0) {
        int got = -1, s = start % S.ns;
        for (int k = 0; k < 2 * S.ns && got < 0; ++k) {
            int zero = 0;
            if (__hip_atomic_compare_exchange_strong(S.flags + s, &zero, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT))
                got = s;
            s = s + 1 == S.ns ? 0 : s + 1;
        }
        if (got < 0) __hip_atomic_fetch_or(S.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *s_slot = got;
    }
    __syncthreads();
    return *s_slot;
}

// The residual (pred - obs) * scale of one unit (k_residual's per-element
// expression), after its forward; visible to the unit's adjoint after the
// fence and barrier.
__device__ __forceinline__ void unit_residual(const MultiGroup& M, int unit) {
    const Grid& g = M.f.g;
    const size_t n = (size_t)g.nt * g.nrec, o0 = (size_t)unit * n;
    const float sc = M.scale[unit / g.shots];
    const float* pr = M.f.traces + o0;
    const float* ob = M.obs + o0;
    float* rs = M.resid + o0;
    constexpr int RU = 8;  // eight elements a thread in flight
    for (size_t i0 = threadIdx.x; i0 < n; i0 += (size_t)RU * NT) {
        float vp[RU], vo[RU];
#pragma unroll
        for (int k = 0; k < RU; ++k) {
            const size_t i = i0 + (size_t)k * NT;
            vp[k] = i < n ? pr[i] : 0.0f;
            vo[k] = i < n ? ob[i] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < RU; ++k) {
            const size_t i = i0 + (size_t)k * NT;
            const float d = vp[k] - vo[k];
            if (i < n) rs[i] = d * sc;
        }
    }
}

__global__ __launch_bounds__(NT) void k_fused(const MultiGroup* __restrict__ T, int n, int tasks, int* next,
                                              FusedSlots SL) {
    extern __shared__ float lds[];
    __shared__ int task_s, fail_s, slot_s;
    const int task = __builtin_amdgcn_readfirstlane(band_task(next, &task_s));
    if (task >= tasks) return;
    int gi = 0;
    for (int k = 1; k < n; ++k)
        if (task >= T[k].task0) gi = k;
    gi = __builtin_amdgcn_readfirstlane(gi);
    const MultiGroup* M = T + gi;
    const int kind = __builtin_amdgcn_readfirstlane(M->kind);
    const int local = task - __builtin_amdgcn_readfirstlane(M->task0);
    const int P = kind == 1 ? S100::P : (kind == 2 ? S141::P : 1);
    const int u = local / P, q = local - u * P;
    const int slot = __builtin_amdgcn_readfirstlane(take_slot(SL, task, &slot_s));
    if (slot < 0) return;
    // the slot's lines a task that held it before left in this compute
    // unit's L1 are dropped before its snapshots are written and read again
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const BandLaunch b = M->b;
    const Xchg X{b.xbuf, b.err, b.P, b.spin};
    // the slot holds this unit's snapshots; the parameters are read in place
    // (scalar loads from the table)
    float* sbase = SL.base + (size_t)slot * SL.floats;
    const FwdParams& f = M->f;
    if (kind == 0) forward_run<G70, true, false>(f, lds, u, 0, f.g.nt, nullptr, nullptr, nullptr, 0, nullptr, sbase);
    else if (kind == 1) fwd_band<S100, 0, true>(f, lds, u, q, X, &fail_s, sbase);
    else fwd_band<S141, 0, true>(f, lds, u, q, X, &fail_s, sbase);
    __syncthreads();  // the gathers and snapshots are read back by this workgroup only
    const int unit = f.unit0 + u;
    if (q == 0 && unit_active(f.geo, unit)) {
        unit_residual(*M, unit);
        __syncthreads();
    }
    const AdjParams& a = M->a;
    if (kind == 0) adjoint_run<G70, false>(a, lds, u, a.g.nt, 0, nullptr, nullptr, nullptr, 0, nullptr, 0, sbase);
    else if (kind == 1) adj_band<S100, 0>(a, lds, u, q, X, &fail_s, a.g.nt, sbase);
    else adj_band<S141, 0>(a, lds, u, q, X, &fail_s, a.g.nt, sbase);
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(SL.flags + slot, 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_multi_put(MultiGroup* table, int i, MultiGroup m) {
    if (threadIdx.x == 0) table[i] = m;
}

}  // namespace fx

int multi_kind(const Grid& g, int P, const int* bounds) {
    BandLaunch b{};
    b.P = P;
    for (int i = 0; i <= P && i <= kMaxBands; ++i) b.bounds[i] = bounds[i];
    if (is_grid(g, 70) && P <= 1) return 0;
    if (is_grid(g, 100) && same_split<fx::S100>(b)) return 1;
    if (is_grid(g, 141) && same_split<fx::S141>(b)) return 2;
    return -1;
}

int multi_tasks_per_unit(int kind) { return kind == 1 ? fx::S100::P : (kind == 2 ? fx::S141::P : 1); }

size_t multi_lds_bytes(int kind, const Grid& g, const Geometry& geo, bool adjoint) {
    const int nx = kind == 0 ? fx::G70::NX : (kind == 1 ? fx::G100::NX : fx::G141::NX);
    if (!adjoint) {
        const int f[3] = {fx::G70::F_TOTAL, fx::band_fwd_floats<fx::S100>(), fx::band_fwd_floats<fx::S141>()};
        return 4ull * f[kind];
    }
    const size_t tail = (size_t)g.nrec + 8 + (size_t)geo.band_rows * nx + 1 + 2ull * geo.n_ent;
    const int a[3] = {fx::G70::A_TAIL, fx::band_adj_floats<fx::S100>(), fx::band_adj_floats<fx::S141>()};
    return 4ull * (a[kind] + tail);
}

hipError_t launch_multi_put(MultiGroup* table, int i, const MultiGroup& m, hipStream_t st) {
    if (i < 0 || i >= kMultiMaxGroups) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fx::k_multi_put, dim3(1), dim3(64), 0, st, table, i, m);
    return hipGetLastError();
}

hipError_t launch_fused(const MultiGroup* table, int n, int tasks, int* next, float* slots,
                        unsigned long long slot_floats, int* slot_flags, int ns, int* err, size_t lds, hipStream_t st) {
    if (n < 1 || n > kMultiMaxGroups || tasks < 1 || ns < 1) return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute((const void*)fx::k_fused, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const fx::FusedSlots S{slots, slot_floats, slot_flags, ns, err};
    hipLaunchKernelGGL(fx::k_fused, dim3(tasks), dim3(fx::NT), lds, st, table, n, tasks, next, S);
    return hipGetLastError();
}

}  // namespace wg
