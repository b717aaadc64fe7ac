// Row bands of a resident unit and the halo exchange between them, shared by
// the runtime-geometry kernels (wave_kernels.hip) and the compiled-geometry
// ones (wave_kernels_fixed.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "wave_kernels.h"

namespace wg {

// ---------------------------------------------------------------------------
// The rows a workgroup holds.  The whole grid is the band [0, nzi) with both
// z-strip sides and no neighbours; band p of a split unit holds the inner
// rows [lo, hi) plus two halo rows on each side (padded rows [lo, hi + 4)),
// and the z-strip side its rows reach (band boundaries lie in the interior,
// at least one strip width from either edge).
// ---------------------------------------------------------------------------
struct Band {
    int lo, hi;     // inner rows owned
    int zb0, rows;  // padded rows held in LDS: [zb0, zb0 + rows)
    int top, bot;   // z-strip sides held (rows [0, sw), rows [nz - sw, nz))
    int up, down;   // a neighbour band above / below
};

__host__ __device__ inline Band whole_band(const Grid& g) {
    return Band{0, g.nzi, 0, g.nz, 1, 1, 0, 0};
}

__host__ __device__ inline Band split_band(const Grid& g, const int* bounds, int P, int p) {
    Band b;
    b.lo = bounds[p];
    b.hi = bounds[p + 1];
    b.zb0 = b.lo;  // padded row of inner row lo - 2
    b.rows = b.hi - b.lo + 4;
    b.top = p == 0;
    b.bot = p == P - 1;
    b.up = p > 0;
    b.down = p < P - 1;
    return b;
}

// ---------------------------------------------------------------------------
// Halo exchange between the bands of a split unit.  After a step each band
// stores its two edge rows of the field into its slot of `buf` (two slots,
// by step parity), each float in a 64-bit word whose high half is the step
// number; then it loads each neighbour's edge rows into its halo rows,
// polling every word until it carries the step number.  Words move with
// relaxed agent-scope 64-bit atomic stores and loads: single-copy atomic,
// past the compute unit's L1, coherent across the device.  The step number
// in every word is what orders the data, so there is no flag to publish after
// the rows and no wait for their stores to complete, and a stale word (an
// older step of the same parity, or the zeros the slots start with) is never
// taken.  (Release and acquire fences would order the rows too, but they
// write back and invalidate the whole L2 of the XCD at every step: several
// microseconds more per step.)  A neighbour overwrites a slot two steps
// later, after loading this band's rows of the step in between, which this
// band stores only after it has loaded the slot (every thread's loads end
// before the barrier at the end of the exchange, and its next stores come
// after the next step's barriers).
// ---------------------------------------------------------------------------
struct Xchg {
    unsigned long long* buf;  // [launch units][P][2 sides (top rows, bottom rows)][2 parities][2 rows][nx]
    int* err;                 // set by a wait that gives up
    int P;
    int spin;                 // polls before a wait gives up (< 0: at once)
};

__device__ __forceinline__ unsigned long long* xslot(const Xchg& X, int nx, int u, int p, int side, int par) {
    return X.buf + ((((size_t)u * X.P + p) * 2 + side) * 2 + par) * 2 * nx;
}

// Exchange of field F (LDS rows of band bd, nx floats each) for unit u (its
// position in the launch), band p, after step number k >= 1.  Every thread
// of the workgroup calls it and moves at most one word of each neighbour's
// rows per round (2 nx words a side, so one round for nx <= 256 with 512
// threads): its own edge words go out first, then its loads from both
// neighbours are issued together and each is polled until it carries step k.
// So the exchange costs about one round trip through the memory system, not
// one per word a thread handles.  Returns false (in every thread) when a
// wait gave up.
__device__ __forceinline__ bool band_exchange(const Xchg& X, float* F, const Band& bd, int nx, int u, int p, int k,
                                              int* fail_s) {
    const int par = k & 1, n2 = 2 * nx, nth = blockDim.x;
    const unsigned long long tag = (unsigned long long)(unsigned)k << 32;
    unsigned long long* up_out = xslot(X, nx, u, p, 0, par);    // first two own rows
    unsigned long long* down_out = xslot(X, nx, u, p, 1, par);  // last two own rows
    const unsigned long long* up_in = bd.up ? xslot(X, nx, u, p - 1, 1, par) : nullptr;
    const unsigned long long* down_in = bd.down ? xslot(X, nx, u, p + 1, 0, par) : nullptr;
    float* own_first = F + 2 * nx;
    float* own_last = F + (bd.hi - bd.lo) * nx;
    float* halo_up = F;
    float* halo_down = F + (bd.hi - bd.lo + 2) * nx;
    for (int j = threadIdx.x; j < n2; j += nth) {
        if (bd.up)
            __hip_atomic_store(up_out + j, tag | __float_as_uint(own_first[j]), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        if (bd.down)
            __hip_atomic_store(down_out + j, tag | __float_as_uint(own_last[j]), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
    bool ok = true;
    for (int j = threadIdx.x; j < n2; j += nth) {
        unsigned long long wu = 0, wd = 0;
        if (bd.up) wu = __hip_atomic_load(up_in + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (bd.down) wd = __hip_atomic_load(down_in + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int c = 0; ok && bd.up && (unsigned)(wu >> 32) != (unsigned)k; ++c) {
            if (c > X.spin) ok = false;
            else {
                __builtin_amdgcn_s_sleep(1);
                wu = __hip_atomic_load(up_in + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        for (int c = 0; ok && bd.down && (unsigned)(wd >> 32) != (unsigned)k; ++c) {
            if (c > X.spin) ok = false;
            else {
                __builtin_amdgcn_s_sleep(1);
                wd = __hip_atomic_load(down_in + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (!ok) break;
        if (bd.up) halo_up[j] = __uint_as_float((unsigned)wu);
        if (bd.down) halo_down[j] = __uint_as_float((unsigned)wd);
    }
    const int failed = __syncthreads_or(ok ? 0 : 1);
    if (failed && threadIdx.x == 0) __hip_atomic_fetch_or(X.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    (void)fail_s;
    return failed == 0;
}

// Task of a band workgroup: one (unit, band) per workgroup, numbered in the
// order the workgroups start (a device counter), so a unit's P bands go to P
// workgroups that start one after another.  All units before the newest are
// complete and run to their end, freeing compute units for the bands still
// to start, so bands that wait for each other are always running together.
// (A loop over tasks in one workgroup would cost registers: the compiler keeps
// values live across tasks.  Out of line, the fetch costs the band kernels no
// registers either: inlined, it cost them 40.)
static __device__ __noinline__ int band_task(int* next, int* ts) {
    if (threadIdx.x == 0) *ts = __hip_atomic_fetch_add(next, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return *ts;
}

}  // namespace wg
