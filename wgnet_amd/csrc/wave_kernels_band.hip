// Row-band kernels of the compiled grids, on the device templates of
// fixed_kernels.h: the 100 x 100 and 141 x 141 models (20-cell C-PML:
// 144 x 144 and 185 x 185 padded), too large for one workgroup, as 2 and 4
// row bands; and the reference's 70 x 70 model (114 x 114) as 2 or 4 bands
// when a batch has so few units that whole-unit workgroups would leave most
// compute units idle.
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
__device__ __forceinline__ void fwd_band(const FwdParams& P, float* lds, int u, int q, const Xchg& X, int* fail_s) {
    if constexpr (p < S::P) {
        if (q == p)
            forward_run<BandG<typename S::Grid, S::B[p], S::B[p + 1]>, STORE, false>(P, lds, u, 0, P.g.nt, nullptr,
                                                                                    nullptr, &X, p, fail_s);
        else
            fwd_band<S, p + 1, STORE>(P, lds, u, q, X, fail_s);
    }
}

template <class S, int p>
__device__ __forceinline__ void adj_band(const AdjParams& P, float* lds, int u, int q, const Xchg& X, int* fail_s) {
    if constexpr (p < S::P) {
        if (q == p)
            adjoint_run<BandG<typename S::Grid, S::B[p], S::B[p + 1]>, false>(P, lds, u, P.g.nt, 0, nullptr, nullptr,
                                                                             &X, p, fail_s);
        else
            adj_band<S, p + 1>(P, lds, u, q, X, fail_s);
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
using S100 = Split<G100, 70>;           // 144 x 144 padded: two bands of 70 rows
using S141 = Split<G141, 44, 90, 136>;  // 185 x 185 padded: 44, 46, 46 and 45 rows


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
    else if (is_grid(p.g, 70) && same_split<fx::S70x4>(b)) *err = fwd_band_fx_t<fx::S70x4>(p, units, store, b, st);
    else if (is_grid(p.g, 70) && same_split<fx::S70x2>(b)) *err = fwd_band_fx_t<fx::S70x2>(p, units, store, b, st);
    else return false;
    return true;
}

bool launch_adjoint_band_fx(const AdjParams& p, int units, const BandLaunch& b, hipStream_t st, hipError_t* err) {
    if (p.dwav) return false;
    if (is_grid(p.g, 100) && same_split<fx::S100>(b)) *err = adj_band_fx_t<fx::S100>(p, units, b, st);
    else if (is_grid(p.g, 141) && same_split<fx::S141>(b)) *err = adj_band_fx_t<fx::S141>(p, units, b, st);
    else if (is_grid(p.g, 70) && same_split<fx::S70x4>(b)) *err = adj_band_fx_t<fx::S70x4>(p, units, b, st);
    else if (is_grid(p.g, 70) && same_split<fx::S70x2>(b)) *err = adj_band_fx_t<fx::S70x2>(p, units, b, st);
    else return false;
    return true;
}

}  // namespace wg
