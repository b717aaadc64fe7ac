// Row-band kernels of the compiled grids, on the device templates of
// fixed_kernels.h: the 100 x 100 and 141 x 141 models (20-cell C-PML:
// 144 x 144 and 185 x 185 padded), too large for one workgroup, as 2 and 3
// row bands (the 141 x 141 one as 4 in a call that ends sooner that way);
// and the reference's 70 x 70 model (114 x 114) as 2 or 4 bands when a batch
// has so few units that whole-unit workgroups would leave most compute units
// idle.
#include "fixed_kernels.h"

namespace wg {

namespace fx {

// ===========================================================================
// Row bands of the compiled grids.  A unit of a grid too large for one
// workgroup runs as P bands with compile-time row bounds, one workgroup each,
// exchanging edge rows after every step (band_xchg.h).  A workgroup takes its
// (unit, band) from a device counter; each band's body is the compiled
// kernel's for the rows it holds, so the P bodies are inlined side by side
// and selected by the band index.
// ===========================================================================
template <class G, int LO, int HI>
using BandG = Geo<G::NVX, G::NVZ, G::PML, LO, HI>;

// A compiled split: band p holds the inner rows [B[p], B[p + 1]).
template <class G, int... Bs>
struct Split {
    using Grid = G;
    static constexpr int P = sizeof...(Bs) + 1;
    static constexpr int B[P + 1] = {0, Bs..., G::NZI};
};

template <class S, int p, bool STORE>
__device__ __forceinline__ void fwd_band(const FwdParams& P, float* lds, int u, int q, const Xchg& X, int* fail_s,
                                         float* snaps_at = nullptr) {
    if constexpr (p < S::P) {
        if (q == p)
            forward_run<BandG<typename S::Grid, S::B[p], S::B[p + 1]>, STORE, false>(P, lds, u, 0, P.g.nt, nullptr,
                                                                                    nullptr, &X, p, fail_s, snaps_at);
        else
            fwd_band<S, p + 1, STORE>(P, lds, u, q, X, fail_s, snaps_at);
    }
}

template <class S, int p>
__device__ __forceinline__ void adj_band(const AdjParams& P, float* lds, int u, int q, const Xchg& X, int* fail_s,
                                         int tag0 = 0, const float* snaps_at = nullptr) {
    if constexpr (p < S::P) {
        if (q == p)
            adjoint_run<BandG<typename S::Grid, S::B[p], S::B[p + 1]>, false>(P, lds, u, P.g.nt, 0, nullptr, nullptr,
                                                                             &X, p, fail_s, tag0, snaps_at);
        else
            adj_band<S, p + 1>(P, lds, u, q, X, fail_s, tag0, snaps_at);
    }
}

template <class S, bool STORE>
__global__ __launch_bounds__(NT) void k_forward_bfx(FwdParams P, BandLaunch B, int units) {
    extern __shared__ float lds[];
    __shared__ int task_s, fail_s;
    const Xchg X{B.xbuf, B.err, S::P, B.spin};
    const int task = __builtin_amdgcn_readfirstlane(band_task(B.next, &task_s));
    if (task >= units * S::P) return;
    const int u = task / S::P, p = task - u * S::P;
    fwd_band<S, 0, STORE>(P, lds, u, p, X, &fail_s);
}

template <class S>
__global__ __launch_bounds__(NT) void k_adjoint_bfx(AdjParams P, BandLaunch B, int units) {
    extern __shared__ float lds[];
    __shared__ int task_s, fail_s;
    const Xchg X{B.xbuf, B.err, S::P, B.spin};
    const int task = __builtin_amdgcn_readfirstlane(band_task(B.next, &task_s));
    if (task >= units * S::P) return;
    const int u = task / S::P, p = task - u * S::P;
    adj_band<S, 0>(P, lds, u, p, X, &fail_s);
}

// LDS of the largest band of a split (floats; the adjoint without its
// receiver tables).
template <class S, int p = 0>
constexpr int band_fwd_floats() {
    if constexpr (p >= S::P) {
        return 0;
    } else {
        constexpr int a = BandG<typename S::Grid, S::B[p], S::B[p + 1]>::F_TOTAL, b = band_fwd_floats<S, p + 1>();
        return a > b ? a : b;
    }
}
template <class S, int p = 0>
constexpr int band_adj_floats() {
    if constexpr (p >= S::P) {
        return 0;
    } else {
        constexpr int a = BandG<typename S::Grid, S::B[p], S::B[p + 1]>::A_TAIL, b = band_adj_floats<S, p + 1>();
        return a > b ? a : b;
    }
}

using G70 = Geo<70, 70, 20>;
using G100 = Geo<100, 100, 20>;
using G141 = Geo<141, 141, 20>;
using S70x2 = Split<G70, 55>;           // 114 x 114 padded: two bands of 55 rows
using S70x4 = Split<G70, 28, 55, 82>;   // four bands of 28, 27, 27 and 28 rows
using S100 = Split<G100, 70>;             // 144 x 144 padded: two bands of 70 rows
using S141 = Split<G141, 59, 122>;        // 185 x 185 padded: 59, 63 and 59 rows
using S141x4 = Split<G141, 44, 90, 136>;  // 44, 46, 46 and 45 rows, for some calls (band_split_call_fx)


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

static bool is_grid(const Grid& g, int nv) {
    return g.nvx == nv && g.nvz == nv && g.pad == 22 && g.region == 24 && g.region_adj == 26;
}

namespace {

template <class S>
bool same_split(const BandLaunch& b) {
    if (b.P != S::P) return false;
    for (int p = 0; p <= S::P; ++p)
        if (b.bounds[p] != S::B[p]) return false;
    return true;
}

template <class S>
void split_of(int* P, int bounds[kMaxBands + 1]) {
    *P = S::P;
    for (int p = 0; p <= S::P; ++p) bounds[p] = S::B[p];
}

template <class S>
hipError_t fwd_band_fx_t(const FwdParams& p, int units, bool store, const BandLaunch& b, hipStream_t st) {
    const size_t lds = 4ull * fx::band_fwd_floats<S>();
    void (*fn)(FwdParams, BandLaunch, int) = store ? fx::k_forward_bfx<S, true> : fx::k_forward_bfx<S, false>;
    hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, dim3(units * S::P), dim3(fx::NT), lds, st, p, b, units);
    return hipGetLastError();
}

template <class S>
hipError_t adj_band_fx_t(const AdjParams& p, int units, const BandLaunch& b, hipStream_t st) {
    using G = typename S::Grid;
    const size_t lds =
        4ull * (fx::band_adj_floats<S>() + p.g.nrec + 8 + p.geo.band_rows * G::NX + 1 + 2 * p.geo.n_ent);
    void (*fn)(AdjParams, BandLaunch, int) = fx::k_adjoint_bfx<S>;
    hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, dim3(units * S::P), dim3(fx::NT), lds, st, p, b, units);
    return hipGetLastError();
}

}  // namespace

// One receiver per thread, every receiver row and residual CSR row in the
// top band [0, top).
static bool receivers_fit(const Grid& g, const Geometry& geo, int top) {
    if (g.nrec > fx::NT) return false;
    return geo.band_rows <= 0 || (geo.band_z0 - 2 >= 0 && geo.band_z0 + geo.band_rows - 1 - 2 < top);
}

bool band_split_fx(const Grid& g, const Geometry& geo, int* P, int bounds[kMaxBands + 1]) {
    int q = 0, b[kMaxBands + 1] = {};
    if (is_grid(g, 100)) split_of<fx::S100>(&q, b);
    else if (is_grid(g, 141)) split_of<fx::S141>(&q, b);
    else return false;
    if (!receivers_fit(g, geo, b[1])) return false;
    *P = q;
    for (int p = 0; p <= q; ++p) bounds[p] = b[p];
    return true;
}

// Rounds of band tasks a call of `units` units in P bands a unit takes on
// `cus` compute units, one band workgroup each.  Workgroups go to the 8 XCDs
// in turn (workgroup b to XCD b mod 8) and start in order within an XCD.
// Where P divides cus every round runs whole units.  Otherwise a unit can
// straddle two rounds: its first bands start early and hold their compute
// units until its last band has started and the unit ends, so, in the worst
// case, every later round of the fullest XCD has one compute unit fewer.
static long long band_rounds(long long units, int P, int cus) {
    const long long n = units * P;
    if (n <= cus) return 1;
    if (cus % P == 0) return (n + cus - 1) / cus;
    const int xcds = cus % 8 == 0 ? 8 : 1, per = cus / xcds;
    if (per < 2) return n;
    const long long nx = (n + xcds - 1) / xcds;  // workgroups of the fullest XCD
    return (nx - 1 + per - 2) / (per - 1);
}

// The split of one call over `units` units: the grid's own split, or for the
// 141 x 141 grid 4 bands where that call ends sooner.  A unit's bands run in
// lockstep and all tasks of a call are alike, so a call takes band_rounds()
// rounds of one band task each; a band step (forward + adjoint) takes
// 15.2 us in 3 bands and 12.8 us in 4 (measured on the MI355X).  3 bands cost
// a unit 11 % less, through fewer halo exchanges per cell, but 4 fill the
// compute units in fewer rounds at some unit counts (up to 64 units, and for
// example 86-128 and 169-192, on 256 compute units).
bool band_split_call_fx(const Grid& g, const Geometry& geo, int units, int cus, int* P,
                        int bounds[kMaxBands + 1]) {
    if (is_grid(g, 141) && units >= 1 && cus >= 1) {
        const double t3 = 15.2 * band_rounds(units, fx::S141::P, cus);
        const double t4 = 12.8 * band_rounds(units, fx::S141x4::P, cus);
        int q = 0, b[kMaxBands + 1] = {};
        split_of<fx::S141x4>(&q, b);
        if (t4 < t3 && receivers_fit(g, geo, b[1])) {
            *P = q;
            for (int p = 0; p <= q; ++p) bounds[p] = b[p];
            return true;
        }
    }
    return band_split_fx(g, geo, P, bounds);
}

bool band_split_small_fx(const Grid& g, const Geometry& geo, int units, int cus, int* P,
                         int bounds[kMaxBands + 1]) {
    if (!is_grid(g, 70) || units < 1) return false;
    int q = 0, b[kMaxBands + 1] = {};
    if ((long long)units * fx::S70x4::P <= cus) split_of<fx::S70x4>(&q, b);
    else if ((long long)units * fx::S70x2::P <= cus) split_of<fx::S70x2>(&q, b);
    else return false;
    if (!receivers_fit(g, geo, b[1])) return false;
    *P = q;
    for (int p = 0; p <= q; ++p) bounds[p] = b[p];
    return true;
}

bool launch_forward_band_fx(const FwdParams& p, int units, bool store, const BandLaunch& b, hipStream_t st,
                            hipError_t* err) {
    if (is_grid(p.g, 100) && same_split<fx::S100>(b)) *err = fwd_band_fx_t<fx::S100>(p, units, store, b, st);
    else if (is_grid(p.g, 141) && same_split<fx::S141>(b)) *err = fwd_band_fx_t<fx::S141>(p, units, store, b, st);
    else if (is_grid(p.g, 141) && same_split<fx::S141x4>(b))
        *err = fwd_band_fx_t<fx::S141x4>(p, units, store, b, st);
    else if (is_grid(p.g, 70) && same_split<fx::S70x4>(b)) *err = fwd_band_fx_t<fx::S70x4>(p, units, store, b, st);
    else if (is_grid(p.g, 70) && same_split<fx::S70x2>(b)) *err = fwd_band_fx_t<fx::S70x2>(p, units, store, b, st);
    else return false;
    return true;
}

bool launch_adjoint_band_fx(const AdjParams& p, int units, const BandLaunch& b, hipStream_t st, hipError_t* err) {
    if (p.dwav) return false;
    if (is_grid(p.g, 100) && same_split<fx::S100>(b)) *err = adj_band_fx_t<fx::S100>(p, units, b, st);
    else if (is_grid(p.g, 141) && same_split<fx::S141>(b)) *err = adj_band_fx_t<fx::S141>(p, units, b, st);
    else if (is_grid(p.g, 141) && same_split<fx::S141x4>(b)) *err = adj_band_fx_t<fx::S141x4>(p, units, b, st);
    else if (is_grid(p.g, 70) && same_split<fx::S70x4>(b)) *err = adj_band_fx_t<fx::S70x4>(p, units, b, st);
    else if (is_grid(p.g, 70) && same_split<fx::S70x2>(b)) *err = adj_band_fx_t<fx::S70x2>(p, units, b, st);
    else return false;
    return true;
}

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
