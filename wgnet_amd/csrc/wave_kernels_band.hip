// Row-band kernels of the compiled grids, on the device templates of
// fixed_kernels.h: the 100 x 100 and 141 x 141 models (20-cell C-PML:
// 144 x 144 and 185 x 185 padded), too large for one workgroup, as 2 and 3
// row bands (the 141 x 141 one as 4 in a call that ends sooner that way);
// and the reference's 70 x 70 model (114 x 114) as 2 or 4 bands when a batch
// has so few units that whole-unit workgroups would leave most compute units
// idle.
#include "band_split.h"

namespace wg {

namespace fx {

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

}  // namespace fx

namespace {

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

}  // namespace wg
