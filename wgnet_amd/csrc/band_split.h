// Compiled row-band splits of the 100 x 100, 141 x 141 and 70 x 70 grids
// (20-cell C-PML) on the device templates of fixed_kernels.h, shared by the
// band kernels (wave_kernels_band.hip) and the one-launch uneven batch
// (wave_kernels_fused.hip).
#pragma once

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

}  // namespace fx

inline bool is_grid(const Grid& g, int nv) {
    return g.nvx == nv && g.nvz == nv && g.pad == 22 && g.region == 24 && g.region_adj == 26;
}


template <class S>
inline bool same_split(const BandLaunch& b) {
    if (b.P != S::P) return false;
    for (int p = 0; p <= S::P; ++p)
        if (b.bounds[p] != S::B[p]) return false;
    return true;
}

}  // namespace wg
