// HBM-streamed forward / adjoint steps for grids whose state does not fit one
// workgroup (e.g. 144 x 144, 185 x 185, 1044 x 344 padded).
//
// One launch advances every unit of a chunk by one time step.  The state lives
// in HBM: forward u (double-buffered: the next value overwrites u^{t-1}), psi
// (double-buffered, neighbours read it) and zeta (own cell only, in place);
// adjoint lambda (double-buffered likewise), P and Z (double-buffered by the
// step parity, as splatAdjoint.wgsl:70-73) and the gradient.  A 256-thread
// workgroup computes a TX x TZ tile of inner cells (64 x 16 by default;
// 128 x 8, 32 x 32 and 64 x 32 are the other shapes an engine can select),
// TX * TZ / 256 cells per thread (rows lz, lz + RZ, ... of column lx,
// RZ = 256 / TX); the neighbour values it
// needs (u or V = v^2 dt^2 lambda, a*psi, T, U1) are staged in LDS with a
// 2-cell halo.  Tiles with no C-PML cell along an axis skip that axis's strips
// and memory variables (they are zero there).  Per-cell arithmetic is
// stencil.h's, in the reference's order.
//
// Most launches are framed: a lean kernel (no memory variables, no per-cell
// C-PML or bounds tests, the clamped velocity from a padded array) runs the
// whole tiles of a rectangle that starts region_adj cells in from each side,
// and a general kernel runs the four bands around it, each band tile
// clipped to its band.  On 1000 x 300 the rectangle holds 78 % of the cells.
#include <algorithm>

#include "stencil.h"
#include "wave_kernels.h"

namespace wg {
namespace st {

constexpr int NT = 256;

template <int TX_, int TZ_>
struct Tile {
    static constexpr int TX = TX_, TZ = TZ_;
    static constexpr int RZ = NT / TX;   // rows per pass
    static constexpr int CPT = TZ / RZ;  // cells per thread
    static constexpr int HX = TX + 4, HZ = TZ + 4;
    static_assert(NT % TX == 0 && TZ % RZ == 0, "tile must tile the workgroup");
};

// The tile a workgroup computes.  The dispatcher hands workgroups to the
// XCDs round-robin by linear id, so neighbouring tiles, which read each
// other's halos, would sit behind different L2 caches.  With P.xcd_order the
// linear ids are renumbered so that each XCD gets one contiguous run of
// tiles in (x, z, unit) order: the halos a tile reads were mostly fetched
// into its own XCD's L2 by its neighbours.  Tiles of one launch are
// independent, so the order changes no result.  At 1000 x 300 (374 tiles a
// unit) this cuts the steps' HBM reads by a third and their time by 2-4 %; on
// grids of a few dozen tiles a unit (100 x 100, 141 x 141) the round-robin
// order is the faster one, so the launchers turn it on from kXcdMinTiles
// tiles a unit.
constexpr unsigned kXcds = 8;  // MI355X
constexpr unsigned kXcdMinTiles = 128;

struct TileId {
    int x, z, unit;
};

// n / d with the launcher's multiplier m = ceil(2^32 / d) (exact for the n a
// launch divides; m = 0: a plain division).  The scalar unit has no divider:
// a division by a value known only at run time is a long scalar sequence in
// every wave of every tile.
__device__ __forceinline__ unsigned udiv(unsigned n, unsigned d, unsigned m) { return m ? __umulhi(n, m) : n / d; }

__device__ __forceinline__ TileId tile_id(const StepParams& P) {
    if (!P.xcd_order) return TileId{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z};
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned n = gx * gy * gridDim.z;
    const unsigned lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned k = lin % kXcds, j = lin / kXcds, q = n / kXcds, r = n % kXcds;
    const unsigned t = k * q + (k < r ? k : r) + j;  // a bijection of [0, n)
    const unsigned u = udiv(t, gx * gy, P.m_gxy), w = t - u * gx * gy;
    const unsigned tz = udiv(w, gx, P.m_gx);
    return TileId{(int)(w - tz * gx), (int)tz, (int)u};
}

// Model and source slot of a unit.
__device__ __forceinline__ void unit_model(const StepParams& P, int unit, int& b, int& s) {
    b = (int)udiv((unsigned)unit, (unsigned)P.g.shots, P.m_shots);
    s = src_slot(P.geo, unit, unit - b * P.g.shots);
}

// Storage index (z * sp + x) of grid cell `cell` = z * nx + x (receivers and
// the wavelet gradient; the steps compare source corners by grid cell).
__device__ __forceinline__ int sidx(const StepParams& P, int cell) {
    const int z = (int)udiv((unsigned)cell, (unsigned)P.g.nx, P.m_nx);
    return cell + z * (P.sp - P.g.nx);
}

// The cells a workgroup of a general launch stores: the tile whose first
// cell is inner cell (xi0, zi0), clipped to inner columns [cx0, cx1) and rows
// [cz0, cz1).  Unframed (bnd_n <= 0): tile (x, z) of the full tiling, clipped
// to the inner grid.  Framed: tile i of the four bands around the lean
// rectangle (rows above it, rows below it, columns left and right of it),
// clipped to its band, so that every cell is stored by one workgroup.
struct TileAt {
    int xi0, zi0, unit;
    int cx0, cx1, cz0, cz1;
};

template <int TX, int TZ>
__device__ __forceinline__ TileAt tile_frame(const StepParams& P) {
    const TileId t = tile_id(P);
    const Grid& g = P.g;
    if (P.bnd_n <= 0) return TileAt{t.x * TX, t.z * TZ, t.unit, 0, g.ni, 0, g.nzi};
    const int lx1 = P.lx1, lz1 = P.lz1;
    const int top = P.gxa * P.gzt, bot = P.gxa * P.gzb, left = P.gxl * P.gzs;
    int i = t.x;
    if (i < top) {
        const int z = (int)udiv((unsigned)i, (unsigned)P.gxa, P.m_gxa);
        return TileAt{(i - z * P.gxa) * TX, z * TZ, t.unit, 0, g.ni, 0, P.lz0};
    }
    i -= top;
    if (i < bot) {
        const int z = (int)udiv((unsigned)i, (unsigned)P.gxa, P.m_gxa);
        return TileAt{(i - z * P.gxa) * TX, lz1 + z * TZ, t.unit, 0, g.ni, lz1, g.nzi};
    }
    i -= bot;
    if (i < left) {
        const int z = (int)udiv((unsigned)i, (unsigned)P.gxl, P.m_gxl);
        return TileAt{(i - z * P.gxl) * TX, P.lz0 + z * TZ, t.unit, 0, P.lx0, P.lz0, lz1};
    }
    i -= left;
    const int z = (int)udiv((unsigned)i, (unsigned)P.gxr, P.m_gxr);
    return TileAt{lx1 + (i - z * P.gxr) * TX, P.lz0 + z * TZ, t.unit, lx1, g.ni, P.lz0, lz1};
}

#define WG_TILE_CONSTANTS                                              \
    constexpr int TX = TL::TX, TZ = TL::TZ, RZ = TL::RZ, CPT = TL::CPT; \
    constexpr int HX = TL::HX, HZ = TL::HZ;

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ float clampv(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ float vel_at(const float* vvis, const Grid& g, int x, int z) {
    const int vz = clampi(z - g.pad, 0, g.nvz - 1), vx = clampi(x - g.pad, 0, g.nvx - 1);
    return clampv(vvis[vz * g.nvx + vx], g.vmin, g.vmax);
}

__device__ __forceinline__ float vis_raw(const float* vvis, const Grid& g, int x, int z) {
    const int vz = clampi(z - g.pad, 0, g.nvz - 1), vx = clampi(x - g.pad, 0, g.nvx - 1);
    return vvis[vz * g.nvx + vx];
}

// --------------------------------------------------------------------------
// Lean paths of the single steps for tiles with no C-PML cell, receiver or
// source (most of a large grid): the tile and its halo lie inside the grid,
// so there are no bounds or C-PML tests; the clamped velocity comes from the
// padded array P.vpad; the staging needs no division by the halo width (the
// tile's own columns are staged by the column map of its cells, the four halo
// columns separately).  Per-cell arithmetic is the general path's, operation
// for operation.
// --------------------------------------------------------------------------
template <class TL>
__device__ __forceinline__ void stage_interior(float* sF, const float* F, int nx, int xi0, int zi0) {
    WG_TILE_CONSTANTS
    const int tid = threadIdx.x, lx = tid % TX, r0 = tid / TX;
    constexpr int NP = (HZ + RZ - 1) / RZ;  // passes over the tile's own columns
    constexpr int NHC = 4 * HZ;             // halo-column words
    constexpr int NH = (NHC + NT - 1) / NT;
    float v[NP], h[NH];
    const float* src = F + (size_t)zi0 * nx + xi0;  // halo origin (padded row zi0, column xi0)
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int r = r0 + k * RZ;
        v[k] = r < HZ ? src[r * nx + 2 + lx] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < NH; ++k) {
        const int j = tid + k * NT;
        const int r = j >> 2, c = j & 3, hx = c < 2 ? c : TX + c;
        h[k] = j < NHC ? src[r * nx + hx] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int r = r0 + k * RZ;
        if (r < HZ) sF[r * HX + 2 + lx] = v[k];
    }
#pragma unroll
    for (int k = 0; k < NH; ++k) {
        const int j = tid + k * NT;
        const int r = j >> 2, c = j & 3, hx = c < 2 ? c : TX + c;
        if (j < NHC) sF[r * HX + hx] = h[k];
    }
}

// Does a source corner of slot s lie in the tile's cells?  The forward
// injects only corners of nonzero weight (nonzero_only); the adjoint adds the
// gradient term of every corner.
// Does a source corner of slot s lie in padded columns [x0, x1) and rows
// [z0, z1)?
__device__ __forceinline__ bool box_has_source(const StepParams& P, int s, int x0, int x1, int z0, int z1,
                                               bool nonzero_only) {
    bool any = false;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int cell = P.geo.src_cell[s * 4 + c];
        const int cz = (int)udiv((unsigned)cell, (unsigned)P.g.nx, P.m_nx), cx = cell - cz * P.g.nx;
        any |= (!nonzero_only || P.geo.src_w[s * 4 + c] != 0.0f) && cx >= x0 && cx < x1 && cz >= z0 && cz < z1;
    }
    return any;
}
__device__ __forceinline__ bool tile_has_source(const StepParams& P, int s, int xi0, int zi0, int TX, int TZ,
                                                bool nonzero_only) {
    return box_has_source(P, s, xi0 + 2, xi0 + TX + 2, zi0 + 2, zi0 + TZ + 2, nonzero_only);
}

// src: a source corner of slot s lies in the tile (injected after the
// update, lib.rs:209-215, with wavelet sample f_t).
template <class TL, bool STORE>
__device__ __forceinline__ void fwd_interior(const StepParams& P, float* sU, int ul, int unit, int b, int xi0,
                                             int zi0, bool src = false, int s = 0, float f_t = 0.0f) {
    WG_TILE_CONSTANTS
    const Grid& g = P.g;
    const size_t cells = P.sc;
    const int nx = P.sp;  // row pitch of the arrays
    const int tid = threadIdx.x, lx = tid % TX, lz0 = tid / TX;
    const float* Ucur = P.state + ((size_t)fwd_u_arr(P.u_cur) * P.units + ul) * cells;
    float* Unxt = P.state + ((size_t)fwd_u_arr(P.u_prev) * P.units + ul) * cells;
    const float* vp = P.vpad + (size_t)(b - P.vpad_b0) * cells;
    const int xo = xi0 + lx + 2;
    const int i0 = (zi0 + lz0 + 2) * nx + xo;  // the thread's first cell
    float pu[CPT], pv[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        pu[k] = Unxt[i0 + k * RZ * nx];
        pv[k] = vp[i0 + k * RZ * nx];
    }
    stage_interior<TL>(sU, Ucur, nx, xi0, zi0);
    __syncthreads();
    float* snap = STORE ? P.snaps + ((size_t)(P.t - P.snap_t0) * P.units + ul) * P.sinner : nullptr;
    const int q0 = (zi0 + lz0) * P.sni + xi0 + lx;  // snapshot index of the first cell
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int lz = lz0 + RZ * k;
        const int li = (lz + 2) * HX + (lx + 2);
        const float u = sU[li];
        Pml none{0, 0, 0, 0};
        float dummy = 0.0f;
        const float w = fwd_w<false, false>(sU, li, HX, u, g.rh, g.rh2, nullptr, none, dummy, dummy, nullptr, TX, none,
                                            dummy, dummy);
        const float v = pv[k];
        if (STORE) snap[q0 + k * RZ * P.sni] = 2.0f * v * g.dt2 * w;
        float next = v * v * g.dt2 * w + 2.0f * u - pu[k];
        if (src) {
            const int gi = (zi0 + lz + 2) * g.nx + xo;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const float wc = P.geo.src_w[s * 4 + c];
                if (P.geo.src_cell[s * 4 + c] == gi && wc != 0.0f) next += -(v * v * g.dt2) * wc * f_t;
            }
        }
        Unxt[i0 + k * RZ * nx] = next;
    }
}

// Adjoint: V = v^2 dt^2 lambda^t staged from lambda^t and the padded
// velocity, then the interior update and the gradient term.
template <class TL>
__device__ __forceinline__ void adj_interior(const StepParams& P, float* sV, int ul, int unit, int b, int xi0,
                                             int zi0) {
    WG_TILE_CONSTANTS
    const Grid& g = P.g;
    const size_t cells = P.sc;
    const int nx = P.sp;  // row pitch of the arrays
    const int tid = threadIdx.x, lx = tid % TX, r0 = tid / TX;
    auto arr = [&](int a) { return P.state + ((size_t)a * P.units + ul) * cells; };
    const float* Lcur = arr(P.lam_cur < 2 ? P.lam_cur : P.lam_cur + 8);
    float* Lnxt = arr(P.lam_prev < 2 ? P.lam_prev : P.lam_prev + 8);
    float* grad = P.grad + (size_t)unit * cells;
    const float* vp = P.vpad + (size_t)(b - P.vpad_b0) * cells;
    const int xo = xi0 + lx + 2;
    const int i0 = (zi0 + r0 + 2) * nx + xo;
    const float* snap = P.snaps + ((size_t)(P.t - P.snap_t0) * P.units + ul) * P.sinner;
    const int q0 = (zi0 + r0) * P.sni + xi0 + lx;
    float pl[CPT], pn[CPT], ps[CPT], pg[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int i = i0 + k * RZ * nx;
        pl[k] = Lcur[i];
        pn[k] = Lnxt[i];
        ps[k] = snap[q0 + k * RZ * P.sni];
        pg[k] = grad[i];
    }
    {
        constexpr int NP = (HZ + RZ - 1) / RZ;
        constexpr int NHC = 4 * HZ;
        constexpr int NH = (NHC + NT - 1) / NT;
        float tl[NP], tv[NP], hl[NH], hv[NH];
        const size_t o = (size_t)zi0 * nx + xi0;
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            const int r = r0 + k * RZ;
            tl[k] = r < HZ ? Lcur[o + r * nx + 2 + lx] : 0.0f;
            tv[k] = r < HZ ? vp[o + r * nx + 2 + lx] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < NH; ++k) {
            const int j = tid + k * NT;
            const int r = j >> 2, c = j & 3, hx = c < 2 ? c : TX + c;
            hl[k] = j < NHC ? Lcur[o + r * nx + hx] : 0.0f;
            hv[k] = j < NHC ? vp[o + r * nx + hx] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            const int r = r0 + k * RZ;
            if (r < HZ) sV[r * HX + 2 + lx] = tv[k] * tv[k] * g.dt2 * tl[k];
        }
#pragma unroll
        for (int k = 0; k < NH; ++k) {
            const int j = tid + k * NT;
            const int r = j >> 2, c = j & 3, hx = c < 2 ? c : TX + c;
            if (j < NHC) sV[r * HX + hx] = hv[k] * hv[k] * g.dt2 * hl[k];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int lz = r0 + RZ * k;
        const int li = (lz + 2) * HX + (lx + 2);
        const float v0 = sV[li];
        const float rh = g.rh;
        float w = adj_axis_int(sV, li, 1, v0, rh);
        w += adj_axis_int(sV, li, HX, v0, rh);
        const float lam = pl[k];
        const float next = w + 2.0f * lam - pn[k];
        const float gk = lam * ps[k];
        const int i = i0 + k * RZ * nx;
        grad[i] = pg[k] + gk;
        Lnxt[i] = next;
    }
}

// --------------------------------------------------------------------------
// Forward step on the tiles of the lean rectangle (a framed launch, beside
// k_fwd_step over the bands around it).
// --------------------------------------------------------------------------
template <class TL, bool STORE>
__global__ __launch_bounds__(NT) void k_fwd_step_in(StepParams P) {
    WG_TILE_CONSTANTS
    __shared__ float sU[HZ * HX];
    const TileId tl = tile_id(P);
    const int unit = P.unit0 + tl.unit;
    int b, s;
    unit_model(P, unit, b, s);
    if (!unit_active(P.geo, unit)) return;
    const int xi0 = P.lx0 + tl.x * TX, zi0 = P.lz0 + tl.z * TZ;
    const bool src = tile_has_source(P, s, xi0, zi0, TX, TZ, true);
    const float f_t = P.wavelet[(size_t)unit * P.wav_stride + P.t];
    fwd_interior<TL, STORE>(P, sU, tl.unit, unit, b, xi0, zi0, src, s, f_t);
}

// --------------------------------------------------------------------------
// Forward step (lib.rs:145-220 for one t, every unit of the chunk).
// --------------------------------------------------------------------------
template <class TL, bool STORE, bool BORN>
__global__ __launch_bounds__(NT) void k_fwd_step(StepParams P) {
    WG_TILE_CONSTANTS
    __shared__ float sU[HZ * HX];
    __shared__ float sAX[TZ * HX];
    __shared__ float sAZ[HZ * TX];
    const Grid& g = P.g;
    const TileAt tile = tile_frame<TX, TZ>(P);
    const int ul = tile.unit;              // unit within the chunk
    const int unit = P.unit0 + ul;
    int b, s;
    unit_model(P, unit, b, s);
    const size_t cells = P.sc;
    const int t = P.t, h = P.ps_half;
    auto arr = [&](int a) { return P.state + ((size_t)a * P.units + ul) * cells; };
    float* Ucur = arr(fwd_u_arr(P.u_cur));
    float* Unxt = arr(fwd_u_arr(P.u_prev));  // u^{t-1}, overwritten with u^{t+1}
    float* PXc = arr(2 + h);
    float* PXn = arr(3 - h);
    float* PZc = arr(4 + h);
    float* PZn = arr(5 - h);
    float* ZX = arr(fwd_zx_arr(P.ze_half));
    float* ZZ = arr(fwd_zz_arr(P.ze_half));
    const float* vvis = P.vel + (size_t)b * g.nvx * g.nvz;
    // the clamped velocity on the padded grid, or null (then clamped from vvis)
    const float* vpg = P.vpad ? P.vpad + (size_t)(b - P.vpad_b0) * cells : nullptr;
    const float* pax = P.geo.prof;
    const float* pbx = pax + g.nx;
    const float* pdbx = pbx + g.nx;
    const float* paz = P.geo.prof + 3 * g.nx;
    const float* pbz = paz + g.nz;
    const float* pdbz = pbz + g.nz;
    const int tid = threadIdx.x;
    const int xi0 = tile.xi0, zi0 = tile.zi0;  // tile origin, inner coords
    const bool rec0 = xi0 == 0 && zi0 == 0;  // the workgroup recording the gathers

    if (!unit_active(P.geo, unit)) {  // padding unit of an uneven batch
        if (P.traces && rec0)
            for (int r = tid; r < g.nrec; r += NT) P.traces[((size_t)unit * g.nt + t) * g.nrec + r] = 0.0f;
        return;
    }
    // Receivers record u^t before the update (lib.rs:146-149): one tile per unit.
    if (P.traces && rec0) {
        float* tr = P.traces + ((size_t)unit * g.nt + t) * g.nrec;
        for (int r = tid; r < g.nrec; r += NT) {
            float acc = 0.0f;
#pragma unroll
            for (int c = 0; c < 4; ++c) acc += P.geo.rec_w[r * 4 + c] * Ucur[sidx(P, P.geo.rec_cell[r * 4 + c])];
            tr[r] = acc;
        }
    }
    // does the tile hold C-PML cells along x / z (padded coords [xi0+2, xi0+TX+2))?
    const bool tpx = xi0 + 2 < g.region || xi0 + TX + 2 > g.nx - g.region;
    const bool tpz = zi0 + 2 < g.region || zi0 + TZ + 2 > g.nz - g.region;
    if constexpr (!BORN) {
        // (the C-PML region is wider than the 2-cell halo, so such a tile and
        // its halo lie inside the grid; tile (0, 0), which records the
        // gathers, is in the region)
        if (P.bnd_n <= 0 && P.vpad && !tpx && !tpz && !tile_has_source(P, s, xi0, zi0, TX, TZ, true)) {
            fwd_interior<TL, STORE>(P, sU, ul, unit, b, xi0, zi0);
            return;
        }
    }
    // Everything a thread reads from HBM for its own cells is loaded before
    // the staging so that its latency overlaps it: u^{t-1}, the velocity,
    // the C-PML memory variables and profile coefficients, the unit's
    // source corners.
    const int lx = tid % TX;
    const int xo = xi0 + lx + 2;  // the thread's column (padded)
    const bool pxo = tpx && (xo < g.region || xo >= g.nx - g.region);
    Pml mx{0, 0, 0, 0};
    if (pxo && xo < g.nx - 2) {
        const float bb = pbx[xo];
        mx = Pml{pax[xo], 1.0f + bb, bb, pdbx[xo]};
    }
    int scell[4];
    float sw[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        scell[c] = P.geo.src_cell[s * 4 + c];  // grid cell z * nx + x
        sw[c] = P.geo.src_w[s * 4 + c];
    }
    float pu[CPT], pv[CPT], ppx[CPT], pzx[CPT], ppz[CPT], pzz[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int z = zi0 + tid / TX + RZ * k + 2;
        pu[k] = pv[k] = ppx[k] = pzx[k] = ppz[k] = pzz[k] = 0.0f;
        if (xo < g.nx - 2 && z < g.nz - 2) {
            const int i = z * P.sp + xo;
            pu[k] = Unxt[i];
            if constexpr (BORN) pv[k] = vis_raw(vvis, g, xo, z);  // (the Born mask needs the raw value)
            else pv[k] = vpg ? vpg[i] : vis_raw(vvis, g, xo, z);
            if (pxo) {
                ppx[k] = PXc[i];
                pzx[k] = ZX[i];
            }
            if (tpz && (z < g.region || z >= g.nz - g.region)) {
                ppz[k] = PZc[i];
                pzz[k] = ZZ[i];
            }
        }
    }
    // Stage u (tile + 2-cell halo), a*psi_x (x halo) and a*psi_z (z halo),
    // each thread's loads issued together before its LDS writes.
    {
        constexpr int NS = (HZ * HX + NT - 1) / NT;
        float tu[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int j = tid + k * NT;
            const int x = xi0 + (j % HX), z = zi0 + (j / HX);  // padded coords: halo starts at inner - 2
            tu[k] = (j < HZ * HX && x < g.nx && z < g.nz) ? Ucur[z * P.sp + x] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < NS; ++k)
            if (tid + k * NT < HZ * HX) sU[tid + k * NT] = tu[k];
    }
    if (tpx) {
        constexpr int NS = (TZ * HX + NT - 1) / NT;
        float ta[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int j = tid + k * NT;
            const int x = xi0 + (j % HX), z = zi0 + 2 + (j / HX);
            ta[k] = (j < TZ * HX && x < g.nx && z < g.nz) ? pax[x] * PXc[z * P.sp + x] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < NS; ++k)
            if (tid + k * NT < TZ * HX) sAX[tid + k * NT] = ta[k];
    }
    if (tpz) {
        constexpr int NS = (HZ * TX + NT - 1) / NT;
        float ta[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int j = tid + k * NT;
            const int x = xi0 + 2 + (j % TX), z = zi0 + (j / TX);
            ta[k] = (j < HZ * TX && x < g.nx && z < g.nz) ? paz[z] * PZc[z * P.sp + x] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < NS; ++k)
            if (tid + k * NT < HZ * TX) sAZ[tid + k * NT] = ta[k];
    }
    __syncthreads();
    const float f_t = P.wavelet[(size_t)unit * P.wav_stride + t];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int lz = tid / TX + RZ * k;
        const int x = xo, z = zi0 + lz + 2;
        // (the clip lies inside the inner grid)
        if (x - 2 < tile.cx0 || x - 2 >= tile.cx1 || z - 2 < tile.cz0 || z - 2 >= tile.cz1) continue;
        const int i = z * P.sp + x;
        const int li = (lz + 2) * HX + (lx + 2);
        const float u = sU[li];
        const bool px = x < g.region || x >= g.nx - g.region;
        const bool pz = z < g.region || z >= g.nz - g.region;
        Pml mz{0, 0, 0, 0};
        if (pz) {
            const float bb = pbz[z];
            mz = Pml{paz[z], 1.0f + bb, bb, pdbz[z]};
        }
        float psx = ppx[k], zex = pzx[k], psz = ppz[k], zez = pzz[k];
        const float* AXc = sAX + lz * HX + (lx + 2);
        const float* AZc = sAZ + (lz + 2) * TX + lx;
        float w;
        if (px && pz)
            w = fwd_w<true, true>(sU, li, HX, u, g.rh, g.rh2, AXc, mx, psx, zex, AZc, TX, mz, psz, zez);
        else if (px)
            w = fwd_w<true, false>(sU, li, HX, u, g.rh, g.rh2, AXc, mx, psx, zex, AZc, TX, mz, psz, zez);
        else if (pz)
            w = fwd_w<false, true>(sU, li, HX, u, g.rh, g.rh2, AXc, mx, psx, zex, AZc, TX, mz, psz, zez);
        else
            w = fwd_w<false, false>(sU, li, HX, u, g.rh, g.rh2, AXc, mx, psx, zex, AZc, TX, mz, psz, zez);
        const float vr = pv[k];
        const float v = clampv(vr, g.vmin, g.vmax);
        const size_t q = (size_t)(z - 2) * P.sni + (x - 2);
        if (STORE) P.snaps[((size_t)(t - P.snap_t0) * P.units + ul) * P.sinner + q] = 2.0f * v * g.dt2 * w;
        float next = v * v * g.dt2 * w + 2.0f * u - pu[k];
        if (BORN) {
            const float m = (vr > g.vmin && vr < g.vmax) ? 1.0f : 0.0f;
            next += (vis_raw(P.dvel + (size_t)b * g.nvx * g.nvz, g, x, z) * m) *
                    P.born_snaps[((size_t)(t - P.snap_t0) * P.units + ul) * P.sinner + q];
        }
        // point source after the update (lib.rs:209-215); Born: its derivative
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float wc = sw[c];
            if (scell[c] == z * g.nx + x && wc != 0.0f) {
                if (BORN) {
                    const float m = (vr > g.vmin && vr < g.vmax) ? 1.0f : 0.0f;
                    const float dv = vis_raw(P.dvel + (size_t)b * g.nvx * g.nvz, g, x, z) * m;
                    next += -(2.0f * v * dv * g.dt2) * wc * f_t;
                } else {
                    next += -(v * v * g.dt2) * wc * f_t;
                }
            }
        }
        Unxt[i] = next;
        if (px) {
            PXn[i] = psx;
            ZX[i] = zex;
        }
        if (pz) {
            PZn[i] = psz;
            ZZ[i] = zez;
        }
    }
}

// --------------------------------------------------------------------------
// Two forward steps per launch (temporal blocking; lib.rs:145-220 for t and
// t + 1).  Step A advances the tile and the 2-cell ring around it (the cells
// step B's stencils read; the ring's corners are not needed) from u^t on the
// tile plus a 4-cell halo, u^{t-1} and the memory variables, which are read
// from HBM once per two steps; step B advances the tile from step A's values,
// kept in registers and LDS.  Ring cells are recomputed by the neighbouring
// tiles with the same operations, so the result is bit-identical to two
// single steps.  u^{t+1} and u^{t+2} go to the two wavefield buffers not
// read, psi and zeta to the other halves, so no tile reads what another
// writes in the same launch.  Gathers: step t's from u^t, and with
// P.trace_prev step t - 1's from u^{t-1}; step t + 1's are recorded by the
// next launch.
// --------------------------------------------------------------------------
template <class TL, bool STORE>
__global__ __launch_bounds__(NT) void k_fwd_step2(StepParams P) {
    WG_TILE_CONSTANTS
    constexpr int AX = TX + 8, AZ = TZ + 8;  // tile + 4-cell halo
    constexpr int BX = TX + 4, BZ = TZ + 4;  // tile + 2-cell halo
    constexpr int RING = 4 * TX + 4 * TZ;    // 2-cell ring without its corners
    constexpr int RPT = (RING + NT - 1) / NT;
    // step A: u^t on tile + 4 (pitch AX); a psi_x^t on rows tile + 2, columns
    // tile + 4; a psi_z^t on rows tile + 4, columns tile + 2.  Step B reuses
    // them for u^{t+1} on tile + 2 (pitch BX), a psi_x^{t+1} on rows tile,
    // columns tile + 2, and a psi_z^{t+1} on rows tile + 2, columns tile.
    __shared__ float sU[AZ * AX];
    __shared__ float sAX[BZ * AX];
    __shared__ float sAZ[AZ * BX];
    const Grid& g = P.g;
    const TileAt tile = tile_frame<TX, TZ>(P);
    const int ul = tile.unit;
    const int unit = P.unit0 + ul;
    int b, s;
    unit_model(P, unit, b, s);
    const size_t cells = P.sc;
    const int t = P.t, h = P.ps_half;
    auto arr = [&](int a) { return P.state + ((size_t)a * P.units + ul) * cells; };
    const float* Ucur = arr(fwd_u_arr(P.u_cur));    // u^t
    const float* Uprev = arr(fwd_u_arr(P.u_prev));  // u^{t-1}
    float* O0 = arr(fwd_u_arr(P.u_out0));           // u^{t+1}
    float* O1 = arr(fwd_u_arr(P.u_out1));           // u^{t+2}
    const float* PXr = arr(2 + h);
    float* PXw = arr(3 - h);
    const float* PZr = arr(4 + h);
    float* PZw = arr(5 - h);
    const float* ZXr = arr(fwd_zx_arr(P.ze_half));
    float* ZXw = arr(fwd_zx_arr(1 - P.ze_half));
    const float* ZZr = arr(fwd_zz_arr(P.ze_half));
    float* ZZw = arr(fwd_zz_arr(1 - P.ze_half));
    const float* vvis = P.vel + (size_t)b * g.nvx * g.nvz;
    // the clamped velocity on the padded grid, or null (then clamped from vvis)
    const float* vpg = P.vpad ? P.vpad + (size_t)(b - P.vpad_b0) * cells : nullptr;
    const float* pax = P.geo.prof;
    const float* pbx = pax + g.nx;
    const float* pdbx = pbx + g.nx;
    const float* paz = P.geo.prof + 3 * g.nx;
    const float* pbz = paz + g.nz;
    const float* pdbz = pbz + g.nz;
    const int tid = threadIdx.x;
    const int xi0 = tile.xi0, zi0 = tile.zi0;  // tile origin; cell (rx, rz) is padded (xi0+2+rx, zi0+2+rz)
    const bool rec0 = xi0 == 0 && zi0 == 0;    // the workgroup recording the gathers

    if (!unit_active(P.geo, unit)) {  // padding unit of an uneven batch
        if (P.traces && rec0)
            for (int r = tid; r < g.nrec; r += NT) {
                P.traces[((size_t)unit * g.nt + t) * g.nrec + r] = 0.0f;
                if (P.trace_prev) P.traces[((size_t)unit * g.nt + t - 1) * g.nrec + r] = 0.0f;
            }
        return;
    }
    // Receivers record u^t (and u^{t-1}) before the update (lib.rs:146-149).
    if (P.traces && rec0) {
        for (int r = tid; r < g.nrec; r += NT) {
            float acc = 0.0f;
#pragma unroll
            for (int c = 0; c < 4; ++c) acc += P.geo.rec_w[r * 4 + c] * Ucur[sidx(P, P.geo.rec_cell[r * 4 + c])];
            P.traces[((size_t)unit * g.nt + t) * g.nrec + r] = acc;
            if (P.trace_prev) {
                float ap = 0.0f;
#pragma unroll
                for (int c = 0; c < 4; ++c) ap += P.geo.rec_w[r * 4 + c] * Uprev[sidx(P, P.geo.rec_cell[r * 4 + c])];
                P.traces[((size_t)unit * g.nt + t - 1) * g.nrec + r] = ap;
            }
        }
    }
    auto inner_cell = [&](int rx, int rz) {
        const int x = xi0 + 2 + rx, z = zi0 + 2 + rz;
        return x >= 2 && z >= 2 && x < g.nx - 2 && z < g.nz - 2;
    };
    // the tile cells this workgroup stores (step A still advances all of
    // them, as step B reads them)
    auto own = [&](int rx, int rz) {
        const int x = xi0 + rx, z = zi0 + rz;
        return x >= tile.cx0 && x < tile.cx1 && z >= tile.cz0 && z < tile.cz1;
    };
    auto pml_x = [&](int x) { return x < g.region || x >= g.nx - g.region; };
    auto pml_z = [&](int z) { return z < g.region || z >= g.nz - g.region; };
    int scell[4];
    float sw[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        scell[c] = P.geo.src_cell[s * 4 + c];  // grid cell z * nx + x
        sw[c] = P.geo.src_w[s * 4 + c];
    }
    // own state of this thread's cells, loaded before the staging so that its
    // latency overlaps it: CPT tile cells (rows tid / TX + k RZ of column
    // tid % TX) and up to RPT ring cells (top and bottom rows, then the left
    // and right columns): u^{t-1}, the velocity, psi and zeta
    const int lx = tid % TX;
    float pu[CPT], pv[CPT], ppx[CPT], pzx[CPT], ppz[CPT], pzz[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int lz = tid / TX + RZ * k;
        pu[k] = pv[k] = ppx[k] = pzx[k] = ppz[k] = pzz[k] = 0.0f;
        if (inner_cell(lx, lz)) {
            const int x = xi0 + 2 + lx, z = zi0 + 2 + lz;
            const int i = z * P.sp + x;
            pu[k] = Uprev[i];
            pv[k] = vpg ? vpg[i] : vis_raw(vvis, g, x, z);
            if (pml_x(x)) {
                ppx[k] = PXr[i];
                pzx[k] = ZXr[i];
            }
            if (pml_z(z)) {
                ppz[k] = PZr[i];
                pzz[k] = ZZr[i];
            }
        }
    }
    int rrx[RPT], rrz[RPT];
    float ru[RPT], rv[RPT], rpx[RPT], rzx[RPT], rpz[RPT], rzz[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int j = tid + NT * r;
        int rx = 0, rz = -100;  // rz = -100: no ring cell for this thread
        if (j < 2 * TX) {
            rz = -2 + j / TX;
            rx = j % TX;
        } else if (j < 4 * TX) {
            rz = TZ + (j - 2 * TX) / TX;
            rx = (j - 2 * TX) % TX;
        } else if (j < 4 * TX + 2 * TZ) {
            rz = (j - 4 * TX) / 2;
            rx = -2 + ((j - 4 * TX) & 1);
        } else if (j < RING) {
            rz = (j - 4 * TX - 2 * TZ) / 2;
            rx = TX + ((j - 4 * TX - 2 * TZ) & 1);
        }
        rrx[r] = rx;
        rrz[r] = rz;
        ru[r] = rv[r] = rpx[r] = rzx[r] = rpz[r] = rzz[r] = 0.0f;
        if (rz != -100 && inner_cell(rx, rz)) {
            const int x = xi0 + 2 + rx, z = zi0 + 2 + rz;
            const int i = z * P.sp + x;
            ru[r] = Uprev[i];
            rv[r] = vpg ? vpg[i] : vis_raw(vvis, g, x, z);
            if (pml_x(x)) {
                rpx[r] = PXr[i];
                rzx[r] = ZXr[i];
            }
            if (pml_z(z)) {
                rpz[r] = PZr[i];
                rzz[r] = ZZr[i];
            }
        }
    }
    // Stage u^t (tile + 4), a*psi_x^t and a*psi_z^t, each thread's loads
    // issued together before its LDS writes; 0 outside the grid.
    {
        constexpr int NS = (AZ * AX + NT - 1) / NT;
        float tu[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int j = tid + k * NT;
            const int x = xi0 - 2 + (j % AX), z = zi0 - 2 + (j / AX);
            tu[k] = (j < AZ * AX && x >= 0 && z >= 0 && x < g.nx && z < g.nz) ? Ucur[z * P.sp + x] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < NS; ++k)
            if (tid + k * NT < AZ * AX) sU[tid + k * NT] = tu[k];
    }
    // does step A's region (the tile and its ring, padded [xi0, xi0 + TX + 4))
    // hold C-PML cells along x / z?
    const bool tpx2 = xi0 < g.region || xi0 + TX + 4 > g.nx - g.region;
    const bool tpz2 = zi0 < g.region || zi0 + TZ + 4 > g.nz - g.region;
    if (tpx2) {
        constexpr int NS = (BZ * AX + NT - 1) / NT;
        float ta[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int j = tid + k * NT;
            const int x = xi0 - 2 + (j % AX), z = zi0 + (j / AX);
            ta[k] = (j < BZ * AX && x >= 0 && x < g.nx && z < g.nz) ? pax[x] * PXr[z * P.sp + x] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < NS; ++k)
            if (tid + k * NT < BZ * AX) sAX[tid + k * NT] = ta[k];
    }
    if (tpz2) {
        constexpr int NS = (AZ * BX + NT - 1) / NT;
        float ta[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int j = tid + k * NT;
            const int x = xi0 + (j % BX), z = zi0 - 2 + (j / BX);
            ta[k] = (j < AZ * BX && z >= 0 && x < g.nx && z < g.nz) ? paz[z] * PZr[z * P.sp + x] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < NS; ++k)
            if (tid + k * NT < AZ * BX) sAZ[tid + k * NT] = ta[k];
    }
    __syncthreads();

    // Step A at inner cell (rx, rz): returns u^{t+1}, updates psi / zeta,
    // gives u^t in uc and stores step t's snapshot when `snap`.
    const float f_a = P.wavelet[(size_t)unit * P.wav_stride + t];
    auto step_a = [&](int rx, int rz, float up, float vr, float& psx, float& zex, float& psz, float& zez, float& uc,
                      bool snap) -> float {
        const int x = xi0 + 2 + rx, z = zi0 + 2 + rz;
        const int li = (rz + 4) * AX + (rx + 4);
        const float u = sU[li];
        uc = u;
        const bool px = pml_x(x), pz = pml_z(z);
        Pml mx{0, 0, 0, 0}, mz{0, 0, 0, 0};
        if (px) {
            const float bb = pbx[x];
            mx = Pml{pax[x], 1.0f + bb, bb, pdbx[x]};
        }
        if (pz) {
            const float bb = pbz[z];
            mz = Pml{paz[z], 1.0f + bb, bb, pdbz[z]};
        }
        const float* AXc = sAX + (rz + 2) * AX + (rx + 4);
        const float* AZc = sAZ + (rz + 4) * BX + (rx + 2);
        float w;
        if (px && pz)
            w = fwd_w<true, true>(sU, li, AX, u, g.rh, g.rh2, AXc, mx, psx, zex, AZc, BX, mz, psz, zez);
        else if (px)
            w = fwd_w<true, false>(sU, li, AX, u, g.rh, g.rh2, AXc, mx, psx, zex, AZc, BX, mz, psz, zez);
        else if (pz)
            w = fwd_w<false, true>(sU, li, AX, u, g.rh, g.rh2, AXc, mx, psx, zex, AZc, BX, mz, psz, zez);
        else
            w = fwd_w<false, false>(sU, li, AX, u, g.rh, g.rh2, AXc, mx, psx, zex, AZc, BX, mz, psz, zez);
        const float v = clampv(vr, g.vmin, g.vmax);
        if (STORE && snap) {
            const size_t q = (size_t)(z - 2) * P.sni + (x - 2);
            P.snaps[((size_t)(t - P.snap_t0) * P.units + ul) * P.sinner + q] = 2.0f * v * g.dt2 * w;
        }
        float next = v * v * g.dt2 * w + 2.0f * u - up;
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (scell[c] == z * g.nx + x && sw[c] != 0.0f) next += -(v * v * g.dt2) * sw[c] * f_a;
        return next;
    };
    float ut[CPT], un[CPT];  // u^t and u^{t+1} of the tile cells
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int lz = tid / TX + RZ * k;
        ut[k] = un[k] = 0.0f;
        if (inner_cell(lx, lz))
            un[k] = step_a(lx, lz, pu[k], pv[k], ppx[k], pzx[k], ppz[k], pzz[k], ut[k], own(lx, lz));
    }
    float rn[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        rn[r] = 0.0f;
        float uc;
        if (rrz[r] != -100 && inner_cell(rrx[r], rrz[r]))
            rn[r] = step_a(rrx[r], rrz[r], ru[r], rv[r], rpx[r], rzx[r], rpz[r], rzz[r], uc, false);
    }
    __syncthreads();  // step A's reads of the LDS are done

    // u^{t+1} on tile + 2 and the step B strips a*psi^{t+1}, from step A's
    // registers (psi is 0 where a cell is outside the region or the grid)
    const bool tpx = xi0 + 2 < g.region || xi0 + TX + 2 > g.nx - g.region;
    const bool tpz = zi0 + 2 < g.region || zi0 + TZ + 2 > g.nz - g.region;
    auto put = [&](int rx, int rz, float u1, float psx, float psz) {
        const int x = xi0 + 2 + rx, z = zi0 + 2 + rz;
        const bool ing = x < g.nx && z < g.nz;
        sU[(rz + 2) * BX + (rx + 2)] = u1;
        if (tpx && rz >= 0 && rz < TZ) sAX[rz * BX + (rx + 2)] = ing ? pax[x] * psx : 0.0f;
        if (tpz && rx >= 0 && rx < TX) sAZ[(rz + 2) * TX + rx] = ing ? paz[z] * psz : 0.0f;
    };
#pragma unroll
    for (int k = 0; k < CPT; ++k) put(lx, tid / TX + RZ * k, un[k], ppx[k], ppz[k]);
#pragma unroll
    for (int r = 0; r < RPT; ++r)
        if (rrz[r] != -100) put(rrx[r], rrz[r], rn[r], rpx[r], rpz[r]);
    __syncthreads();

    // Step B (time t + 1) on the tile
    const int t1 = t + 1;
    const float f_b = P.wavelet[(size_t)unit * P.wav_stride + t1];
    const int xo = xi0 + 2 + lx;
    const bool pxo = pml_x(xo);
    Pml mx{0, 0, 0, 0};
    if (pxo && xo < g.nx - 2) {
        const float bb = pbx[xo];
        mx = Pml{pax[xo], 1.0f + bb, bb, pdbx[xo]};
    }
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int lz = tid / TX + RZ * k;
        if (!own(lx, lz)) continue;  // (the clip lies inside the inner grid)
        const int x = xo, z = zi0 + 2 + lz;
        const int i = z * P.sp + x;
        const int li = (lz + 2) * BX + (lx + 2);
        const float u = sU[li];
        const bool pz = pml_z(z);
        Pml mz{0, 0, 0, 0};
        if (pz) {
            const float bb = pbz[z];
            mz = Pml{paz[z], 1.0f + bb, bb, pdbz[z]};
        }
        float psx = ppx[k], zex = pzx[k], psz = ppz[k], zez = pzz[k];
        const float* AXc = sAX + lz * BX + (lx + 2);
        const float* AZc = sAZ + (lz + 2) * TX + lx;
        float w;
        if (pxo && pz)
            w = fwd_w<true, true>(sU, li, BX, u, g.rh, g.rh2, AXc, mx, psx, zex, AZc, TX, mz, psz, zez);
        else if (pxo)
            w = fwd_w<true, false>(sU, li, BX, u, g.rh, g.rh2, AXc, mx, psx, zex, AZc, TX, mz, psz, zez);
        else if (pz)
            w = fwd_w<false, true>(sU, li, BX, u, g.rh, g.rh2, AXc, mx, psx, zex, AZc, TX, mz, psz, zez);
        else
            w = fwd_w<false, false>(sU, li, BX, u, g.rh, g.rh2, AXc, mx, psx, zex, AZc, TX, mz, psz, zez);
        const float v = clampv(pv[k], g.vmin, g.vmax);
        if (STORE) {
            const size_t q = (size_t)(z - 2) * P.sni + (x - 2);
            P.snaps[((size_t)(t1 - P.snap_t0) * P.units + ul) * P.sinner + q] = 2.0f * v * g.dt2 * w;
        }
        float next = v * v * g.dt2 * w + 2.0f * u - ut[k];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (scell[c] == z * g.nx + x && sw[c] != 0.0f) next += -(v * v * g.dt2) * sw[c] * f_b;
        O0[i] = u;
        O1[i] = next;
        if (pxo) {
            PXw[i] = psx;
            ZXw[i] = zex;
        }
        if (pz) {
            PZw[i] = psz;
            ZZw[i] = zez;
        }
    }
}

// Gathers of step P.t from wavefield buffer P.u_cur, one workgroup a unit.
__global__ __launch_bounds__(NT) void k_fwd_record(StepParams P) {
    const Grid& g = P.g;
    const int ul = blockIdx.x, unit = P.unit0 + ul;
    const size_t cells = P.sc;
    const float* U = P.state + ((size_t)fwd_u_arr(P.u_cur) * P.units + ul) * cells;
    float* tr = P.traces + ((size_t)unit * g.nt + P.t) * g.nrec;
    const bool on = unit_active(P.geo, unit);
    for (int r = threadIdx.x; r < g.nrec; r += NT) {
        float acc = 0.0f;
        if (on) {
#pragma unroll
            for (int c = 0; c < 4; ++c) acc += P.geo.rec_w[r * 4 + c] * U[sidx(P, P.geo.rec_cell[r * 4 + c])];
        }
        tr[r] = acc;
    }
}

// --------------------------------------------------------------------------
// Adjoint step (splatAdjoint.wgsl:58-131 for one t, every unit of the chunk).
// --------------------------------------------------------------------------
template <class TL, bool DWAV>
__global__ __launch_bounds__(NT) void k_adj_step(StepParams P) {
    WG_TILE_CONSTANTS
    __shared__ float sV[HZ * HX];
    __shared__ float sTX[TZ * HX], sUX[TZ * HX];
    __shared__ float sTZ[HZ * TX], sUZ[HZ * TX];
    const Grid& g = P.g;
    const TileId tile = tile_id(P);
    const int ul = tile.unit;
    const int unit = P.unit0 + ul;
    int b, s;
    unit_model(P, unit, b, s);
    const size_t cells = P.sc;
    const int t = P.t, q = P.half;
    // arrays: lam0 lam1 | Px0 Px1 | Zx0 Zx1 | Pz0 Pz1 | Zz0 Zz1 | lam2 lam3; the
    // memory variables are read from half q and written to half 1 - q, and
    // lambda^{t-1} overwrites lambda^{t+1} in place (own cell only)
    auto arr = [&](int a) { return P.state + ((size_t)a * P.units + ul) * cells; };
    float* Lcur = arr(P.lam_cur < 2 ? P.lam_cur : P.lam_cur + 8);
    float* Lnxt = arr(P.lam_prev < 2 ? P.lam_prev : P.lam_prev + 8);
    const float* PXr = arr(2 + q);
    float* PXw = arr(3 - q);
    const float* ZXr = arr(4 + q);
    float* ZXw = arr(5 - q);
    const float* PZr = arr(6 + q);
    float* PZw = arr(7 - q);
    const float* ZZr = arr(8 + q);
    float* ZZw = arr(9 - q);
    float* grad = P.grad + (size_t)unit * cells;
    const float* vvis = P.vel + (size_t)b * g.nvx * g.nvz;
    // the clamped velocity on the padded grid, or null (then clamped from vvis)
    const float* vpg = P.vpad ? P.vpad + (size_t)(b - P.vpad_b0) * cells : nullptr;
    const float* pax = P.geo.prof;
    const float* pbx = pax + g.nx;
    const float* pdbx = pbx + g.nx;
    const float* paz = P.geo.prof + 3 * g.nx;
    const float* pbz = paz + g.nz;
    const float* pdbz = pbz + g.nz;
    const int tid = threadIdx.x;
    const int xi0 = tile.x * TX, zi0 = tile.z * TZ;

    if (!unit_active(P.geo, unit)) {  // padding unit; its gradient rows are zeroed by the host
        if (DWAV && tile.x == 0 && tile.z == 0 && tid == 0) P.dwav[(size_t)unit * g.nt + t] = 0.0f;
        return;
    }
    if (DWAV && tile.x == 0 && tile.z == 0 && tid == 0) {
        float acc = 0.0f;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int cell = P.geo.src_cell[s * 4 + c];
            const int cz = cell / g.nx, cx = cell - cz * g.nx;
            const float v = vel_at(vvis, g, cx, cz);
            acc += (-(v * v * g.dt2) * P.geo.src_w[s * 4 + c]) * Lcur[sidx(P, cell)];
        }
        P.dwav[(size_t)unit * g.nt + t] = acc;
    }
    // does the tile hold cells of the adjoint C-PML region along x / z?
    const bool tax = xi0 + 2 < g.region_adj || xi0 + TX + 2 > g.nx - g.region_adj;
    const bool taz = zi0 + 2 < g.region_adj || zi0 + TZ + 2 > g.nz - g.region_adj;
    if constexpr (!DWAV) {
        // receiver rows [band_z0, band_z0 + band_rows) and source corners stay
        // on the general path
        const bool rec = zi0 + 2 < P.geo.band_z0 + P.geo.band_rows && zi0 + TZ + 2 > P.geo.band_z0;
        if (P.vpad && !tax && !taz && !rec && !tile_has_source(P, s, xi0, zi0, TX, TZ, false)) {
            adj_interior<TL>(P, sV, ul, unit, b, xi0, zi0);
            return;
        }
    }
    // Everything a thread reads from HBM for its own cells is loaded before
    // the staging so that its latency overlaps it: lambda^t, lambda^{t+1},
    // the snapshot, the gradient, the C-PML memory variables and the unit's
    // source corners.
    const int lx = tid % TX;
    const int xo = xi0 + lx + 2;  // the thread's column (padded)
    const bool pxo = xo < g.region_adj || xo >= g.nx - g.region_adj;
    int scell[4];
    float sw[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        scell[c] = P.geo.src_cell[s * 4 + c];  // grid cell z * nx + x
        sw[c] = P.geo.src_w[s * 4 + c];
    }
    float pl[CPT], pn[CPT], ps[CPT], pg[CPT], ppx[CPT], pzx[CPT], ppz[CPT], pzz[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int z = zi0 + tid / TX + RZ * k + 2;
        pl[k] = pn[k] = ps[k] = pg[k] = ppx[k] = pzx[k] = ppz[k] = pzz[k] = 0.0f;
        if (xo < g.nx - 2 && z < g.nz - 2) {
            const int i = z * P.sp + xo;
            pl[k] = Lcur[i];
            pn[k] = Lnxt[i];
            ps[k] = P.snaps[((size_t)(t - P.snap_t0) * P.units + ul) * P.sinner + (size_t)(z - 2) * P.sni + (xo - 2)];
            pg[k] = grad[i];
            if (pxo) {
                ppx[k] = PXr[i];
                pzx[k] = ZXr[i];
            }
            if (z < g.region_adj || z >= g.nz - g.region_adj) {
                ppz[k] = PZr[i];
                pzz[k] = ZZr[i];
            }
        }
    }
    // V on the tile + halo (vw(), wgsl:53-56); 0 on the ring (lambda is 0
    // there); each thread's loads issued together before its LDS writes
    {
        constexpr int NS = (HZ * HX + NT - 1) / NT;
        float tl[NS], tv[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int j = tid + k * NT;
            const int x = xi0 + (j % HX), z = zi0 + (j / HX);
            tl[k] = tv[k] = 0.0f;
            if (j < HZ * HX && x < g.nx && z < g.nz) {
                tv[k] = (vpg ? vpg[z * P.sp + x] : vel_at(vvis, g, x, z));
                tl[k] = Lcur[z * P.sp + x];
            }
        }
#pragma unroll
        for (int k = 0; k < NS; ++k)
            if (tid + k * NT < HZ * HX) sV[tid + k * NT] = tv[k] * tv[k] * g.dt2 * tl[k];
    }
    __syncthreads();
    // T = (1+b)V + bZ, U1 = db T + bP of the x-neighbours (x halo) and z-neighbours
    if (tax)
        for (int j = tid; j < TZ * HX; j += NT) {
            const int hx = j % HX, hz = j / HX;
            const int x = xi0 + hx, z = zi0 + 2 + hz;
            float T = 0.0f, U1 = 0.0f;
            if (x < g.nx && z < g.nz) {
                const int i = z * P.sp + x;
                const float bb = pbx[x];
                T = (1.0f + bb) * sV[(hz + 2) * HX + hx] + bb * ZXr[i];
                U1 = pdbx[x] * T + bb * PXr[i];
            }
            sTX[j] = T;
            sUX[j] = U1;
        }
    if (taz)
        for (int j = tid; j < HZ * TX; j += NT) {
            const int hx = j % TX, hz = j / TX;
            const int x = xi0 + 2 + hx, z = zi0 + hz;
            float T = 0.0f, U1 = 0.0f;
            if (x < g.nx && z < g.nz) {
                const int i = z * P.sp + x;
                const float bb = pbz[z];
                T = (1.0f + bb) * sV[hz * HX + hx + 2] + bb * ZZr[i];
                U1 = pdbz[z] * T + bb * PZr[i];
            }
            sTZ[j] = T;
            sUZ[j] = U1;
        }
    __syncthreads();
    const float f_t = P.wavelet[(size_t)unit * P.wav_stride + t];
    const float* res = P.resid + ((size_t)unit * g.nt + t) * g.nrec;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int lz = tid / TX + RZ * k;
        const int x = xo, z = zi0 + lz + 2;
        if (x >= g.nx - 2 || z >= g.nz - 2) continue;
        const int i = z * P.sp + x;
        const int li = (lz + 2) * HX + (lx + 2);
        const float v0 = sV[li];
        const float rh = g.rh;
        float w;  // x term, then += z term (0 + x skipped: equal up to the sign of zero)
        if (pxo) {
            const float opb[5] = {1.0f + pbx[x - 2], 1.0f + pbx[x - 1], 1.0f + pbx[x], 1.0f + pbx[x + 1], 1.0f + pbx[x + 2]};
            float Pv = ppx[k], Zv = pzx[k];
            w = adj_axis_pml(sTX + lz * HX + (lx + 2), sUX + lz * HX + (lx + 2), 1, opb, pax[x], v0, rh, Pv, Zv);
            PXw[i] = Pv;
            ZXw[i] = Zv;
        } else {
            w = adj_axis_int(sV, li, 1, v0, rh);
        }
        if (z < g.region_adj || z >= g.nz - g.region_adj) {
            const float opb[5] = {1.0f + pbz[z - 2], 1.0f + pbz[z - 1], 1.0f + pbz[z], 1.0f + pbz[z + 1], 1.0f + pbz[z + 2]};
            float Pv = ppz[k], Zv = pzz[k];
            w += adj_axis_pml(sTZ + (lz + 2) * TX + lx, sUZ + (lz + 2) * TX + lx, TX, opb, paz[z], v0, rh, Pv, Zv);
            PZw[i] = Pv;
            ZZw[i] = Zv;
        } else {
            w += adj_axis_int(sV, li, HX, v0, rh);
        }
        const float lam = pl[k];
        float next = w + 2.0f * lam - pn[k];
        float gk = lam * ps[k];
        if (z >= P.geo.band_z0 && z < P.geo.band_z0 + P.geo.band_rows) {
            const int o = (z - P.geo.band_z0) * g.nx + x;
            for (int e = P.geo.band_off[o]; e < P.geo.band_off[o + 1]; ++e) next += P.geo.ent_w[e] * res[P.geo.ent_r[e]];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (scell[c] == z * g.nx + x) {
                const float v = (vpg ? vpg[z * P.sp + x] : vel_at(vvis, g, x, z));
                gk += lam * (-2.0f * v * g.dt2 * sw[c] * f_t);
            }
        grad[i] = pg[k] + gk;
        Lnxt[i] = next;
    }
}


// --------------------------------------------------------------------------
// Two forward steps on the tiles of the lean rectangle, whose cells and
// 2-cell ring lie outside the C-PML region (the gathers are recorded by the
// band tile at the grid's corner): k_fwd_step2's scheme with no memory
// variables, strips or bounds tests, sources tested only in a tile that
// holds one, and the clamped velocity from P.vpad; launched beside
// k_fwd_step2 over the bands around the rectangle.
// --------------------------------------------------------------------------
template <class TL, bool STORE>
__global__ __launch_bounds__(NT) void k_fwd_step2_in(StepParams P) {
    WG_TILE_CONSTANTS
    constexpr int AX = TX + 8, AZ = TZ + 8;  // tile + 4-cell halo
    constexpr int BX = TX + 4;               // tile + 2-cell halo
    constexpr int RING = 4 * TX + 4 * TZ;    // 2-cell ring without its corners
    constexpr int RPT = (RING + NT - 1) / NT;
    __shared__ float sU[AZ * AX];  // u^t on tile + 4 (pitch AX), then u^{t+1} on tile + 2 (pitch BX)
    const Grid& g = P.g;
    const TileId tl = tile_id(P);
    const int ul = tl.unit;
    const int unit = P.unit0 + ul;
    int b, s;
    unit_model(P, unit, b, s);
    if (!unit_active(P.geo, unit)) return;
    const int xi0 = P.lx0 + tl.x * TX, zi0 = P.lz0 + tl.z * TZ;
    // a source corner in the tile or its ring: the injection after each
    // update (lib.rs:209-215), tested per cell only in such a tile
    const bool src = box_has_source(P, s, xi0, xi0 + TX + 4, zi0, zi0 + TZ + 4, true);
    const float f_a = P.wavelet[(size_t)unit * P.wav_stride + P.t];
    const float f_b = P.wavelet[(size_t)unit * P.wav_stride + P.t + 1];
    auto inject = [&](int gi, float v, float f, float& next) {  // gi: grid cell; corners read only in such a tile
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float wc = P.geo.src_w[s * 4 + c];
            if (P.geo.src_cell[s * 4 + c] == gi && wc != 0.0f) next += -(v * v * g.dt2) * wc * f;
        }
    };
    const size_t cells = P.sc;
    const int nx = P.sp, t = P.t;  // nx: row pitch of the arrays
    auto arr = [&](int a) { return P.state + ((size_t)a * P.units + ul) * cells; };
    const float* Ucur = arr(fwd_u_arr(P.u_cur));    // u^t
    const float* Uprev = arr(fwd_u_arr(P.u_prev));  // u^{t-1}
    float* O0 = arr(fwd_u_arr(P.u_out0));           // u^{t+1}
    float* O1 = arr(fwd_u_arr(P.u_out1));           // u^{t+2}
    const float* vp = P.vpad + (size_t)(b - P.vpad_b0) * cells;
    const int tid = threadIdx.x, lx = tid % TX, r0 = tid / TX;

    const int i0 = (zi0 + 2 + r0) * nx + xi0 + 2 + lx;
    const int g0 = (zi0 + 2 + r0) * g.nx + xi0 + 2 + lx;  // its grid cell
    float pu[CPT], pv[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        pu[k] = Uprev[i0 + k * RZ * nx];
        pv[k] = vp[i0 + k * RZ * nx];
    }
    int rrx[RPT], rrz[RPT];
    float ru[RPT], rv[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int j = tid + NT * r;
        int rx = 0, rz = -100;  // rz = -100: no ring cell for this thread
        if (j < 2 * TX) {
            rz = -2 + j / TX;
            rx = j % TX;
        } else if (j < 4 * TX) {
            rz = TZ + (j - 2 * TX) / TX;
            rx = (j - 2 * TX) % TX;
        } else if (j < 4 * TX + 2 * TZ) {
            rz = (j - 4 * TX) / 2;
            rx = -2 + ((j - 4 * TX) & 1);
        } else if (j < RING) {
            rz = (j - 4 * TX - 2 * TZ) / 2;
            rx = TX + ((j - 4 * TX - 2 * TZ) & 1);
        }
        rrx[r] = rx;
        rrz[r] = rz;
        ru[r] = rv[r] = 0.0f;
        if (rz != -100) {
            const int i = (zi0 + 2 + rz) * nx + (xi0 + 2 + rx);
            ru[r] = Uprev[i];
            rv[r] = vp[i];
        }
    }
    {
        constexpr int NS = (AZ * AX + NT - 1) / NT;
        float tu[NS];
        const size_t o = (size_t)(zi0 - 2) * nx + (xi0 - 2);
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int j = tid + k * NT;
            tu[k] = j < AZ * AX ? Ucur[o + (size_t)(j / AX) * nx + (j % AX)] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < NS; ++k)
            if (tid + k * NT < AZ * AX) sU[tid + k * NT] = tu[k];
    }
    __syncthreads();
    Pml none{0, 0, 0, 0};
    float dummy = 0.0f;
    float* snap0 = STORE ? P.snaps + ((size_t)(t - P.snap_t0) * P.units + ul) * P.sinner : nullptr;
    float* snap1 = STORE ? P.snaps + ((size_t)(t + 1 - P.snap_t0) * P.units + ul) * P.sinner : nullptr;
    const int q0 = (zi0 + r0) * P.sni + xi0 + lx;
    // Step A (time t) on the tile and its ring: u^{t+1}
    float ut[CPT], un[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int li = (r0 + RZ * k + 4) * AX + (lx + 4);
        const float u = sU[li];
        ut[k] = u;
        const float w = fwd_w<false, false>(sU, li, AX, u, g.rh, g.rh2, nullptr, none, dummy, dummy, nullptr, BX,
                                            none, dummy, dummy);
        const float v = pv[k];
        if (STORE) snap0[q0 + k * RZ * P.sni] = 2.0f * v * g.dt2 * w;
        float next = v * v * g.dt2 * w + 2.0f * u - pu[k];
        if (src) inject(g0 + k * RZ * g.nx, v, f_a, next);
        un[k] = next;
    }
    float rn[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        rn[r] = 0.0f;
        if (rrz[r] != -100) {
            const int li = (rrz[r] + 4) * AX + (rrx[r] + 4);
            const float u = sU[li];
            const float w = fwd_w<false, false>(sU, li, AX, u, g.rh, g.rh2, nullptr, none, dummy, dummy, nullptr, BX,
                                                none, dummy, dummy);
            const float v = rv[r];
            float next = v * v * g.dt2 * w + 2.0f * u - ru[r];
            if (src) inject((zi0 + 2 + rrz[r]) * g.nx + (xi0 + 2 + rrx[r]), v, f_a, next);
            rn[r] = next;
        }
    }
    __syncthreads();  // step A's reads of sU are done
#pragma unroll
    for (int k = 0; k < CPT; ++k) sU[(r0 + RZ * k + 2) * BX + (lx + 2)] = un[k];
#pragma unroll
    for (int r = 0; r < RPT; ++r)
        if (rrz[r] != -100) sU[(rrz[r] + 2) * BX + (rrx[r] + 2)] = rn[r];
    __syncthreads();
    // Step B (time t + 1) on the tile
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int li = (r0 + RZ * k + 2) * BX + (lx + 2);
        const float u = sU[li];
        const float w = fwd_w<false, false>(sU, li, BX, u, g.rh, g.rh2, nullptr, none, dummy, dummy, nullptr, TX,
                                            none, dummy, dummy);
        const float v = pv[k];
        if (STORE) snap1[q0 + k * RZ * P.sni] = 2.0f * v * g.dt2 * w;
        const int i = i0 + k * RZ * nx;
        float next = v * v * g.dt2 * w + 2.0f * u - ut[k];
        if (src) inject(g0 + k * RZ * g.nx, v, f_b, next);
        O0[i] = u;
        O1[i] = next;
    }
}

// --------------------------------------------------------------------------
// Two adjoint steps on the tiles of the lean rectangle, whose cells and
// 2-cell ring lie outside the C-PML region and the receiver rows (the bulk
// of a large grid; launched beside k_adj_step2 over the bands around it):
// k_adj_step2's scheme with no memory variables, strips, receivers or bounds
// tests, and the clamped velocity from P.vpad.  Same arithmetic in the same
// order, so the two kernels together equal two single steps bit for bit.
// --------------------------------------------------------------------------
template <class TL>
__global__ __launch_bounds__(NT) void k_adj_step2_in(StepParams P) {
    WG_TILE_CONSTANTS
    constexpr int AX = TX + 8, AZ = TZ + 8;  // tile + 4-cell halo
    constexpr int BX = TX + 4;               // tile + 2-cell halo
    constexpr int RING = 4 * TX + 4 * TZ;    // 2-cell ring without its corners
    constexpr int RPT = (RING + NT - 1) / NT;
    __shared__ float sV[AZ * AX];  // V^t on tile + 4 (pitch AX), then V^{t-1} on tile + 2 (pitch BX)
    const Grid& g = P.g;
    const TileId tl = tile_id(P);
    const int ul = tl.unit;
    const int unit = P.unit0 + ul;
    int b, s;
    unit_model(P, unit, b, s);
    if (!unit_active(P.geo, unit)) return;
    const int xi0 = P.lx0 + tl.x * TX, zi0 = P.lz0 + tl.z * TZ;
    // a source corner in the tile: its gradient term (wgsl:245-249), tested
    // per cell only in such a tile
    const bool src = tile_has_source(P, s, xi0, zi0, TX, TZ, false);
    auto src_grad = [&](int gi, float v, float l, int tt, float& gg) {  // gi: grid cell; corners read only in such a tile
        const float f = P.wavelet[(size_t)unit * P.wav_stride + tt];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (P.geo.src_cell[s * 4 + c] == gi) gg += l * (-2.0f * v * g.dt2 * P.geo.src_w[s * 4 + c] * f);
    };
    const size_t cells = P.sc;
    const int nx = P.sp, t = P.t;  // nx: row pitch of the arrays
    auto lam = [&](int k) { return P.state + ((size_t)(k < 2 ? k : k + 8) * P.units + ul) * cells; };
    const float* Lcur = lam(P.lam_cur);    // lambda^t
    const float* Lprev = lam(P.lam_prev);  // lambda^{t+1}
    float* O0 = lam(P.lam_out0);           // lambda^{t-1}
    float* O1 = lam(P.lam_out1);           // lambda^{t-2}
    float* grad = P.grad + (size_t)unit * cells;
    const float* vp = P.vpad + (size_t)(b - P.vpad_b0) * cells;
    const int tid = threadIdx.x, lx = tid % TX, r0 = tid / TX;
    const float rh = g.rh;

    // the thread's tile cells and ring cells, loaded before the staging
    const int i0 = (zi0 + 2 + r0) * nx + xi0 + 2 + lx;
    const int g0 = (zi0 + 2 + r0) * g.nx + xi0 + 2 + lx;  // its grid cell
    const float* sn0 = P.snaps + ((size_t)(t - P.snap_t0) * P.units + ul) * P.sinner;
    const float* sn1 = P.snaps + ((size_t)(t - 1 - P.snap_t0) * P.units + ul) * P.sinner;
    const int q0 = (zi0 + r0) * P.sni + xi0 + lx;
    float pl[CPT], pp[CPT], ps0[CPT], ps1[CPT], pg[CPT], pv[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int i = i0 + k * RZ * nx, q = q0 + k * RZ * P.sni;
        pl[k] = Lcur[i];
        pp[k] = Lprev[i];
        ps0[k] = sn0[q];
        ps1[k] = sn1[q];
        pg[k] = grad[i];
        pv[k] = vp[i];
    }
    int rrx[RPT], rrz[RPT];
    float rl[RPT], rp[RPT], rv[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int j = tid + NT * r;
        int rx = 0, rz = -100;  // rz = -100: no ring cell for this thread
        if (j < 2 * TX) {
            rz = -2 + j / TX;
            rx = j % TX;
        } else if (j < 4 * TX) {
            rz = TZ + (j - 2 * TX) / TX;
            rx = (j - 2 * TX) % TX;
        } else if (j < 4 * TX + 2 * TZ) {
            rz = (j - 4 * TX) / 2;
            rx = -2 + ((j - 4 * TX) & 1);
        } else if (j < RING) {
            rz = (j - 4 * TX - 2 * TZ) / 2;
            rx = TX + ((j - 4 * TX - 2 * TZ) & 1);
        }
        rrx[r] = rx;
        rrz[r] = rz;
        rl[r] = rp[r] = rv[r] = 0.0f;
        if (rz != -100) {
            const int i = (zi0 + 2 + rz) * nx + (xi0 + 2 + rx);
            rl[r] = Lcur[i];
            rp[r] = Lprev[i];
            rv[r] = vp[i];
        }
    }
    // V^t on tile + 4 (vw(), wgsl:53-56)
    {
        constexpr int NS = (AZ * AX + NT - 1) / NT;
        float tl_[NS], tv[NS];
        const size_t o = (size_t)(zi0 - 2) * nx + (xi0 - 2);
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int j = tid + k * NT;
            const int hx = j % AX, hz = j / AX;
            tl_[k] = tv[k] = 0.0f;
            if (j < AZ * AX) {
                tl_[k] = Lcur[o + (size_t)hz * nx + hx];
                tv[k] = vp[o + (size_t)hz * nx + hx];
            }
        }
#pragma unroll
        for (int k = 0; k < NS; ++k)
            if (tid + k * NT < AZ * AX) sV[tid + k * NT] = tv[k] * tv[k] * g.dt2 * tl_[k];
    }
    __syncthreads();
    // Step A (time t) on the tile and its ring: lambda^{t-1} and V^{t-1}
    auto step_a = [&](int rx, int rz, float l0, float lp, float v, float& l1, float& V1) {
        const int li = (rz + 4) * AX + (rx + 4);
        const float v0 = sV[li];
        float w = adj_axis_int(sV, li, 1, v0, rh);
        w += adj_axis_int(sV, li, AX, v0, rh);
        const float next = w + 2.0f * l0 - lp;
        l1 = next;
        V1 = v * v * g.dt2 * next;
    };
    float lt1[CPT], Vo[CPT], gk[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        step_a(lx, r0 + RZ * k, pl[k], pp[k], pv[k], lt1[k], Vo[k]);
        float gg = pl[k] * ps0[k];
        if (src) src_grad(g0 + k * RZ * g.nx, pv[k], pl[k], t, gg);
        gk[k] = gg;
    }
    float rV[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        float l1;
        rV[r] = 0.0f;
        if (rrz[r] != -100) step_a(rrx[r], rrz[r], rl[r], rp[r], rv[r], l1, rV[r]);
    }
    __syncthreads();  // step A's reads of sV are done
#pragma unroll
    for (int k = 0; k < CPT; ++k) sV[(r0 + RZ * k + 2) * BX + (lx + 2)] = Vo[k];
#pragma unroll
    for (int r = 0; r < RPT; ++r)
        if (rrz[r] != -100) sV[(rrz[r] + 2) * BX + (rrx[r] + 2)] = rV[r];
    __syncthreads();
    // Step B (time t - 1) on the tile
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int lz = r0 + RZ * k;
        const int li = (lz + 2) * BX + (lx + 2);
        const float v0 = sV[li];
        float w = adj_axis_int(sV, li, 1, v0, rh);
        w += adj_axis_int(sV, li, BX, v0, rh);
        const float l1 = lt1[k];
        const float next = w + 2.0f * l1 - pl[k];
        const int i = i0 + k * RZ * nx;
        float gb = l1 * ps1[k];
        if (src) src_grad(g0 + k * RZ * g.nx, pv[k], l1, t - 1, gb);
        float gacc = pg[k];
        gacc += gk[k];
        gacc += gb;
        grad[i] = gacc;
        O0[i] = l1;
        O1[i] = next;
    }
}

// --------------------------------------------------------------------------
// Two adjoint steps, t and t - 1, per launch (temporal blocking).  Step A
// advances the tile and the 2-cell ring around it (the cells step B's
// stencils read; the ring's corners are not needed) from lambda^t,
// lambda^{t+1} and the C-PML memory variables, which are read from HBM once
// per two steps; step B advances the tile from step A's values, kept in
// registers and LDS.  Ring cells are recomputed by the neighbouring tiles
// with the same operations, so the result is bit-identical to two single
// steps.  The gradient is read and written once per two steps, its two terms
// added in the single steps' order.  lambda^{t-1} and lambda^{t-2} go to two
// buffers other than the two read, and the memory variables to the other
// half, so no tile reads what another writes in the same launch.
// --------------------------------------------------------------------------
// Waves a SIMD the compiler must fit k_adj_step2<TL> into: on the 32 x 32
// band tiles it fits 5 (96 VGPRs, no spills: the band call 3 % faster than
// at the 4 of its natural 101 VGPRs); other shapes would spill.
template <class TL>
struct AdjWaves {
    static constexpr int value = (TL::TX == 32 && TL::TZ == 32) ? 5 : 1;
};

template <class TL>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(AdjWaves<TL>::value))) void k_adj_step2(
    StepParams P) {
    WG_TILE_CONSTANTS
    constexpr int AX = TX + 8, AZ = TZ + 8;  // tile + 4-cell halo
    constexpr int BX = TX + 4, BZ = TZ + 4;  // tile + 2-cell halo
    constexpr int RING = 4 * TX + 4 * TZ;    // 2-cell ring without its corners
    constexpr int RPT = (RING + NT - 1) / NT;
    // strips: step A x (rows tile+2, cols tile+4) and z (rows tile+4, cols
    // tile+2), T then U1 each; step B reuses the space for x (rows tile, cols
    // tile+2) and z (rows tile+2, cols tile)
    constexpr int SAX = BZ * AX, SAZ = AZ * BX;
    constexpr int A_XT = 0, A_XU = SAX, A_ZT = 2 * SAX, A_ZU = 2 * SAX + SAZ;
    constexpr int SBX = TZ * BX, SBZ = BZ * TX;
    constexpr int B_XT = 0, B_XU = SBX, B_ZT = 2 * SBX, B_ZU = 2 * SBX + SBZ;
    static_assert(2 * SBX + 2 * SBZ <= 2 * SAX + 2 * SAZ, "step B strips must fit in step A's");
    __shared__ float sV[AZ * AX];  // V^t on tile + 4 (pitch AX), then V^{t-1} on tile + 2 (pitch BX)
    __shared__ float sS[2 * SAX + 2 * SAZ];
    const Grid& g = P.g;
    const TileAt tile = tile_frame<TX, TZ>(P);
    const int ul = tile.unit;
    const int unit = P.unit0 + ul;
    int b, s;
    unit_model(P, unit, b, s);
    const size_t cells = P.sc;
    const int t = P.t, h = P.half;
    auto arr = [&](int a) { return P.state + ((size_t)a * P.units + ul) * cells; };
    auto lam = [&](int k) { return arr(k < 2 ? k : k + 8); };
    const float* Lcur = lam(P.lam_cur);    // lambda^t
    const float* Lprev = lam(P.lam_prev);  // lambda^{t+1}
    float* O0 = lam(P.lam_out0);           // lambda^{t-1}
    float* O1 = lam(P.lam_out1);           // lambda^{t-2}
    const float* PXr = arr(2 + h);
    float* PXw = arr(3 - h);
    const float* ZXr = arr(4 + h);
    float* ZXw = arr(5 - h);
    const float* PZr = arr(6 + h);
    float* PZw = arr(7 - h);
    const float* ZZr = arr(8 + h);
    float* ZZw = arr(9 - h);
    float* grad = P.grad + (size_t)unit * cells;
    const float* vvis = P.vel + (size_t)b * g.nvx * g.nvz;
    // the clamped velocity on the padded grid, or null (then clamped from vvis)
    const float* vpg = P.vpad ? P.vpad + (size_t)(b - P.vpad_b0) * cells : nullptr;
    const float* pax = P.geo.prof;
    const float* pbx = pax + g.nx;
    const float* pdbx = pbx + g.nx;
    const float* paz = P.geo.prof + 3 * g.nx;
    const float* pbz = paz + g.nz;
    const float* pdbz = pbz + g.nz;
    const int tid = threadIdx.x;
    const int xi0 = tile.xi0, zi0 = tile.zi0;  // tile origin; cell (rx, rz) is padded (xi0+2+rx, zi0+2+rz)
    const int ra = g.region_adj;
    const float rh = g.rh;

    if (!unit_active(P.geo, unit)) return;  // padding unit; its gradient rows are zeroed by the host

    // cells of this thread: CPT of the tile (rows tid / TX + k RZ of column
    // tid % TX) and up to RPT of the ring (top and bottom rows, then the left
    // and right columns); their lambda^t and lambda^{t+1} (and for the tile
    // cells the two snapshots and the gradient) are loaded before the staging
    // so that their latency overlaps it
    auto inner_cell = [&](int rx, int rz) {
        const int x = xi0 + 2 + rx, z = zi0 + 2 + rz;
        return x >= 2 && z >= 2 && x < g.nx - 2 && z < g.nz - 2;
    };
    const int lx = tid % TX;
    int scell[4];
    float sw[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        scell[c] = P.geo.src_cell[s * 4 + c];  // grid cell z * nx + x
        sw[c] = P.geo.src_w[s * 4 + c];
    }
    float pl[CPT], pp[CPT], ps0[CPT], ps1[CPT], pg[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int lz = tid / TX + RZ * k;
        pl[k] = pp[k] = ps0[k] = ps1[k] = pg[k] = 0.0f;
        if (inner_cell(lx, lz)) {
            const int x = xi0 + 2 + lx, z = zi0 + 2 + lz;
            const int i = z * P.sp + x;
            const size_t qq = (size_t)(z - 2) * P.sni + (x - 2);
            pl[k] = Lcur[i];
            pp[k] = Lprev[i];
            ps0[k] = P.snaps[((size_t)(t - P.snap_t0) * P.units + ul) * P.sinner + qq];
            ps1[k] = P.snaps[((size_t)(t - 1 - P.snap_t0) * P.units + ul) * P.sinner + qq];
            pg[k] = grad[i];
        }
    }
    int rrx[RPT], rrz[RPT];
    float rl[RPT], rp[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int j = tid + NT * r;
        int rx = 0, rz = -100;  // rz = -100: no ring cell for this thread
        if (j < 2 * TX) {
            rz = -2 + j / TX;
            rx = j % TX;
        } else if (j < 4 * TX) {
            rz = TZ + (j - 2 * TX) / TX;
            rx = (j - 2 * TX) % TX;
        } else if (j < 4 * TX + 2 * TZ) {
            rz = (j - 4 * TX) / 2;
            rx = -2 + ((j - 4 * TX) & 1);
        } else if (j < RING) {
            rz = (j - 4 * TX - 2 * TZ) / 2;
            rx = TX + ((j - 4 * TX - 2 * TZ) & 1);
        }
        rrx[r] = rx;
        rrz[r] = rz;
        rl[r] = rp[r] = 0.0f;
        if (rz != -100 && inner_cell(rx, rz)) {
            const int i = (zi0 + 2 + rz) * P.sp + (xi0 + 2 + rx);
            rl[r] = Lcur[i];
            rp[r] = Lprev[i];
        }
    }

    // V^t on tile + 4 (vw(), wgsl:53-56); 0 on the ring of the grid and
    // outside it; each thread's loads issued together before its LDS writes
    {
        constexpr int NS = (AZ * AX + NT - 1) / NT;
        float tl[NS], tv[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int j = tid + k * NT;
            const int x = xi0 - 2 + (j % AX), z = zi0 - 2 + (j / AX);
            tl[k] = tv[k] = 0.0f;
            if (j < AZ * AX && x >= 0 && z >= 0 && x < g.nx && z < g.nz) {
                tv[k] = (vpg ? vpg[z * P.sp + x] : vel_at(vvis, g, x, z));
                tl[k] = Lcur[z * P.sp + x];
            }
        }
#pragma unroll
        for (int k = 0; k < NS; ++k)
            if (tid + k * NT < AZ * AX) sV[tid + k * NT] = tv[k] * tv[k] * g.dt2 * tl[k];
    }
    __syncthreads();
    // step A strips: T = (1+b)V + bZ, U1 = db T + bP along each axis
    const bool tax2 = xi0 < ra || xi0 + TX + 4 > g.nx - ra;
    const bool taz2 = zi0 < ra || zi0 + TZ + 4 > g.nz - ra;
    // (each thread's HBM loads issued together before it computes and writes)
    if (tax2) {
        constexpr int NS = (SAX + NT - 1) / NT;
        float tb[NS], td[NS], tz[NS], tp[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int j = tid + k * NT;
            const int hx = j % AX, hz = j / AX;
            const int x = xi0 - 2 + hx, z = zi0 + hz;
            tb[k] = td[k] = tz[k] = tp[k] = 0.0f;
            if (j < SAX && x >= 0 && x < g.nx && z < g.nz) {
                const int i = z * P.sp + x;
                tb[k] = pbx[x];
                td[k] = pdbx[x];
                tz[k] = ZXr[i];
                tp[k] = PXr[i];
            }
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int j = tid + k * NT;
            if (j >= SAX) continue;
            const int hx = j % AX, hz = j / AX;  // rx = hx - 4, rz = hz - 2
            const int x = xi0 - 2 + hx, z = zi0 + hz;
            float T = 0.0f, U1 = 0.0f;
            if (x >= 0 && x < g.nx && z < g.nz) {
                const float bb = tb[k];
                T = (1.0f + bb) * sV[(hz + 2) * AX + hx] + bb * tz[k];
                U1 = td[k] * T + bb * tp[k];
            }
            sS[A_XT + j] = T;
            sS[A_XU + j] = U1;
        }
    }
    if (taz2) {
        constexpr int NS = (SAZ + NT - 1) / NT;
        float tb[NS], td[NS], tz[NS], tp[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int j = tid + k * NT;
            const int hx = j % BX, hz = j / BX;
            const int x = xi0 + hx, z = zi0 - 2 + hz;
            tb[k] = td[k] = tz[k] = tp[k] = 0.0f;
            if (j < SAZ && z >= 0 && x < g.nx && z < g.nz) {
                const int i = z * P.sp + x;
                tb[k] = pbz[z];
                td[k] = pdbz[z];
                tz[k] = ZZr[i];
                tp[k] = PZr[i];
            }
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int j = tid + k * NT;
            if (j >= SAZ) continue;
            const int hx = j % BX, hz = j / BX;  // rx = hx - 2, rz = hz - 4
            const int x = xi0 + hx, z = zi0 - 2 + hz;
            float T = 0.0f, U1 = 0.0f;
            if (z >= 0 && x < g.nx && z < g.nz) {
                const float bb = tb[k];
                T = (1.0f + bb) * sV[hz * AX + hx + 2] + bb * tz[k];
                U1 = td[k] * T + bb * tp[k];
            }
            sS[A_ZT + j] = T;
            sS[A_ZU + j] = U1;
        }
    }
    __syncthreads();

    // Step A at cell (rx, rz): lambda^{t-1}, V^{t-1} and the new memory
    // variables (0 where the axis is outside the region, as the never-written
    // HBM entries are); false outside the inner grid (all zero there).
    const float* res_a = P.resid + ((size_t)unit * g.nt + t) * g.nrec;
    auto step_a = [&](int rx, int rz, float l0, float lp, float& l1, float& V1, float& px, float& zx, float& pz,
                      float& zz) -> bool {
        l1 = V1 = px = zx = pz = zz = 0.0f;
        const int x = xi0 + 2 + rx, z = zi0 + 2 + rz;
        if (!inner_cell(rx, rz)) return false;
        const int i = z * P.sp + x;
        const int li = (rz + 4) * AX + (rx + 4);
        const float v0 = sV[li];
        float w;  // x term, then += z term (0 + x skipped: equal up to the sign of zero)
        if (x < ra || x >= g.nx - ra) {
            const float opb[5] = {1.0f + pbx[x - 2], 1.0f + pbx[x - 1], 1.0f + pbx[x], 1.0f + pbx[x + 1],
                                  1.0f + pbx[x + 2]};
            float Pv = PXr[i], Zv = ZXr[i];
            const int o = (rz + 2) * AX + (rx + 4);
            w = adj_axis_pml(sS + A_XT + o, sS + A_XU + o, 1, opb, pax[x], v0, rh, Pv, Zv);
            px = Pv;
            zx = Zv;
        } else {
            w = adj_axis_int(sV, li, 1, v0, rh);
        }
        if (z < ra || z >= g.nz - ra) {
            const float opb[5] = {1.0f + pbz[z - 2], 1.0f + pbz[z - 1], 1.0f + pbz[z], 1.0f + pbz[z + 1],
                                  1.0f + pbz[z + 2]};
            float Pv = PZr[i], Zv = ZZr[i];
            const int o = (rz + 4) * BX + (rx + 2);
            w += adj_axis_pml(sS + A_ZT + o, sS + A_ZU + o, BX, opb, paz[z], v0, rh, Pv, Zv);
            pz = Pv;
            zz = Zv;
        } else {
            w += adj_axis_int(sV, li, AX, v0, rh);
        }
        float next = w + 2.0f * l0 - lp;
        if (z >= P.geo.band_z0 && z < P.geo.band_z0 + P.geo.band_rows) {
            const int o = (z - P.geo.band_z0) * g.nx + x;
            for (int e = P.geo.band_off[o]; e < P.geo.band_off[o + 1]; ++e) next += P.geo.ent_w[e] * res_a[P.geo.ent_r[e]];
        }
        l1 = next;
        const float v = (vpg ? vpg[z * P.sp + x] : vel_at(vvis, g, x, z));
        V1 = v * v * g.dt2 * next;
        return true;
    };

    float lt1[CPT], Vo[CPT], pxo[CPT], zxo[CPT], pzo[CPT], zzo[CPT], gk[CPT];
    bool in[CPT];
    const float f_a = P.wavelet[(size_t)unit * P.wav_stride + t];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int lz = tid / TX + RZ * k;
        in[k] = step_a(lx, lz, pl[k], pp[k], lt1[k], Vo[k], pxo[k], zxo[k], pzo[k], zzo[k]);
        gk[k] = 0.0f;
        if (in[k]) {
            const int x = xi0 + 2 + lx, z = zi0 + 2 + lz;
            float gg = pl[k] * ps0[k];
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (scell[c] == z * g.nx + x) {
                    const float v = (vpg ? vpg[z * P.sp + x] : vel_at(vvis, g, x, z));
                    gg += pl[k] * (-2.0f * v * g.dt2 * sw[c] * f_a);
                }
            gk[k] = gg;
        }
    }
    float rV[RPT], rpx[RPT], rzx[RPT], rpz[RPT], rzz[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        float l1;
        rV[r] = rpx[r] = rzx[r] = rpz[r] = rzz[r] = 0.0f;
        if (rrz[r] != -100) step_a(rrx[r], rrz[r], rl[r], rp[r], l1, rV[r], rpx[r], rzx[r], rpz[r], rzz[r]);
    }
    __syncthreads();  // step A's reads of sV and the strips are done

    // V^{t-1} on tile + 2 and the step B strips, from step A's registers
    const bool tax = xi0 + 2 < ra || xi0 + TX + 2 > g.nx - ra;
    const bool taz = zi0 + 2 < ra || zi0 + TZ + 2 > g.nz - ra;
    auto put = [&](int rx, int rz, float V1, float px, float zx, float pz, float zz) {
        const int x = xi0 + 2 + rx, z = zi0 + 2 + rz;
        const bool ing = x < g.nx && z < g.nz;
        sV[(rz + 2) * BX + (rx + 2)] = V1;
        if (tax && rz >= 0 && rz < TZ) {
            float T = 0.0f, U1 = 0.0f;
            if (ing) {
                const float bb = pbx[x];
                T = (1.0f + bb) * V1 + bb * zx;
                U1 = pdbx[x] * T + bb * px;
            }
            sS[B_XT + rz * BX + (rx + 2)] = T;
            sS[B_XU + rz * BX + (rx + 2)] = U1;
        }
        if (taz && rx >= 0 && rx < TX) {
            float T = 0.0f, U1 = 0.0f;
            if (ing) {
                const float bb = pbz[z];
                T = (1.0f + bb) * V1 + bb * zz;
                U1 = pdbz[z] * T + bb * pz;
            }
            sS[B_ZT + (rz + 2) * TX + rx] = T;
            sS[B_ZU + (rz + 2) * TX + rx] = U1;
        }
    };
#pragma unroll
    for (int k = 0; k < CPT; ++k) put(lx, tid / TX + RZ * k, Vo[k], pxo[k], zxo[k], pzo[k], zzo[k]);
#pragma unroll
    for (int r = 0; r < RPT; ++r)
        if (rrz[r] != -100) put(rrx[r], rrz[r], rV[r], rpx[r], rzx[r], rpz[r], rzz[r]);
    __syncthreads();

    // Step B (time t - 1) on the tile
    const float* res_b = P.resid + ((size_t)unit * g.nt + (t - 1)) * g.nrec;
    const float f_b = P.wavelet[(size_t)unit * P.wav_stride + (t - 1)];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int lz = tid / TX + RZ * k;
        const int x = xi0 + 2 + lx, z = zi0 + 2 + lz;
        // the cells this workgroup stores (the clip lies inside the inner grid)
        if (!in[k] || x - 2 < tile.cx0 || x - 2 >= tile.cx1 || z - 2 < tile.cz0 || z - 2 >= tile.cz1) continue;
        const int i = z * P.sp + x;
        const int li = (lz + 2) * BX + (lx + 2);
        const float v0 = sV[li];
        float w;
        if (x < ra || x >= g.nx - ra) {
            const float opb[5] = {1.0f + pbx[x - 2], 1.0f + pbx[x - 1], 1.0f + pbx[x], 1.0f + pbx[x + 1],
                                  1.0f + pbx[x + 2]};
            float Pv = pxo[k], Zv = zxo[k];
            const int o = lz * BX + (lx + 2);
            w = adj_axis_pml(sS + B_XT + o, sS + B_XU + o, 1, opb, pax[x], v0, rh, Pv, Zv);
            PXw[i] = Pv;
            ZXw[i] = Zv;
        } else {
            w = adj_axis_int(sV, li, 1, v0, rh);
        }
        if (z < ra || z >= g.nz - ra) {
            const float opb[5] = {1.0f + pbz[z - 2], 1.0f + pbz[z - 1], 1.0f + pbz[z], 1.0f + pbz[z + 1],
                                  1.0f + pbz[z + 2]};
            float Pv = pzo[k], Zv = zzo[k];
            const int o = (lz + 2) * TX + lx;
            w += adj_axis_pml(sS + B_ZT + o, sS + B_ZU + o, TX, opb, paz[z], v0, rh, Pv, Zv);
            PZw[i] = Pv;
            ZZw[i] = Zv;
        } else {
            w += adj_axis_int(sV, li, BX, v0, rh);
        }
        const float l1 = lt1[k];
        float next = w + 2.0f * l1 - pl[k];
        float gb = l1 * ps1[k];
        if (z >= P.geo.band_z0 && z < P.geo.band_z0 + P.geo.band_rows) {
            const int o = (z - P.geo.band_z0) * g.nx + x;
            for (int e = P.geo.band_off[o]; e < P.geo.band_off[o + 1]; ++e) next += P.geo.ent_w[e] * res_b[P.geo.ent_r[e]];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (scell[c] == z * g.nx + x) {
                const float v = (vpg ? vpg[z * P.sp + x] : vel_at(vvis, g, x, z));
                gb += l1 * (-2.0f * v * g.dt2 * sw[c] * f_b);
            }
        float gacc = pg[k];
        gacc += gk[k];
        gacc += gb;
        grad[i] = gacc;
        O0[i] = l1;
        O1[i] = next;
    }
}

}  // namespace st

namespace {

// ceil(2^32 / d) when n / d = mulhi(n, m) is exact for every n < n_end
// (n_end <= 2^32 / d), else 0.
unsigned div_magic(unsigned long long d, unsigned long long n_end) {
    if (d < 2 || n_end > (1ull << 32) / d) return 0;
    return (unsigned)(((1ull << 32) + d - 1) / d);
}

// xcd_tiles: tiles a unit of the grid's full tiling, which decides the
// XCD-aware order (0: those of this launch).
StepParams ordered(const StepParams& p, const dim3& grid, unsigned xcd_tiles = 0) {
    StepParams q = p;
    q.xcd_order = (xcd_tiles ? xcd_tiles : grid.x * grid.y) >= st::kXcdMinTiles ? 1 : 0;
    const unsigned long long tiles = (unsigned long long)grid.x * grid.y * grid.z;
    q.m_gx = div_magic(grid.x, tiles);
    q.m_gxy = div_magic((unsigned long long)grid.x * grid.y, tiles);
    q.m_shots = div_magic(p.g.shots, (unsigned long long)p.unit0 + grid.z);
    q.m_nx = div_magic(p.g.nx, (unsigned long long)p.g.nx * p.g.nz);
    return q;
}

// The lean rectangle of a framed launch (StepParams::lx0 ...): whole TX x TZ
// tiles from inner cell (ra, ra), ra the adjoint C-PML region, so that the
// tiles' cells and 2-cell ring (padded [ra, ra + n T + 4)) lie outside every
// step's C-PML region; empty without the padded velocity, when no tile fits,
// or when the receiver rows (the adjoint's residual injection) reach into
// it.  The bands around it get their tile counts and multipliers.  Returns
// whether the launch is framed.
bool frame(StepParams& q, int TX, int TZ, int TXb, int TZb) {
    const Grid& g = q.g;
    const int ra = g.region_adj;
    q.lx0 = q.lz0 = ra;
    q.nlx = (g.nx - 2 * ra - 4) / TX;
    q.nlz = (g.nz - 2 * ra - 4) / TZ;
    q.bnd_n = 0;
    const int lz1 = ra + q.nlz * TZ;  // inner rows [ra, lz1): padded [ra + 2, lz1 + 2)
    // (the lean adjoint also advances the 2-cell ring: padded rows [ra, lz1 + 4))
    const bool rec_in = q.geo.band_rows > 0 && q.geo.band_z0 < lz1 + 4 && q.geo.band_z0 + q.geo.band_rows > ra;
    if (!q.vpad || q.nlx <= 0 || q.nlz <= 0 || rec_in) {
        q.nlx = q.nlz = 0;
        return false;
    }
    const int lx1 = ra + q.nlx * TX;
    q.lx1 = lx1;
    q.lz1 = lz1;
    q.gxa = (g.ni + TXb - 1) / TXb;
    q.gzt = (ra + TZb - 1) / TZb;
    q.gzb = (g.nzi - lz1 + TZb - 1) / TZb;
    q.gxl = (ra + TXb - 1) / TXb;
    q.gxr = (g.ni - lx1 + TXb - 1) / TXb;
    q.gzs = (lz1 - ra + TZb - 1) / TZb;
    q.bnd_n = q.gxa * (q.gzt + q.gzb) + (q.gxl + q.gxr) * q.gzs;
    const unsigned long long n = (unsigned long long)q.bnd_n;
    q.m_gxa = div_magic(q.gxa, n);
    q.m_gxl = div_magic(q.gxl, n);
    q.m_gxr = div_magic(q.gxr, n);
    return true;
}

// A framed launch: the lean kernel `lean` over the rectangle frame() gives
// and the general kernel `full` over the bands around it, or `full` over
// every tile when the rectangle is empty.
template <class T, class TB, class TF = T>
hipError_t framed(const StepParams& p, int units, void (*lean)(StepParams), void (*full)(StepParams),
                  void (*full_b)(StepParams), hipStream_t st) {
    StepParams q = p;
    const unsigned gx = (p.g.ni + T::TX - 1) / T::TX, gz = (p.g.nzi + T::TZ - 1) / T::TZ;
    if (!frame(q, T::TX, T::TZ, TB::TX, TB::TZ)) {  // unframed: the general kernel on TF tiles
        const dim3 grid((p.g.ni + TF::TX - 1) / TF::TX, (p.g.nzi + TF::TZ - 1) / TF::TZ, units);
        hipLaunchKernelGGL(full, grid, dim3(st::NT), 0, st, ordered(q, grid));
        return hipGetLastError();
    }
    const dim3 gin(q.nlx, q.nlz, units);
    hipLaunchKernelGGL(lean, gin, dim3(st::NT), 0, st, ordered(q, gin, gx * gz));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const dim3 gb(q.bnd_n, 1, units);
    hipLaunchKernelGGL(full_b, gb, dim3(st::NT), 0, st, ordered(q, gb, gx * gz));
    return hipGetLastError();
}

// The band kernels' tile, whatever the lean kernels' is: 32 x 32 fits the
// bands, which are region_adj = 26 cells wide (one tile column of a side
// band, one tile row of the top band).  At 1000 x 300 the bands take 93 such
// tiles a unit, 95 k thread slots for their 77 k cells, against 104 tiles
// and 106 k slots of 64 x 16; forward and adjoint calls 11 % and 4 % faster.
template <class T>
using BandTile = st::Tile<32, 32>;

template <class T>
hipError_t fwd_step_t(const StepParams& p, int units, bool store, bool born, hipStream_t st) {
    if (!born)
        return framed<T, BandTile<T>>(p, units, store ? st::k_fwd_step_in<T, true> : st::k_fwd_step_in<T, false>,
                                      store ? st::k_fwd_step<T, true, false> : st::k_fwd_step<T, false, false>,
                                      store ? st::k_fwd_step<BandTile<T>, true, false>
                                            : st::k_fwd_step<BandTile<T>, false, false>,
                                      st);
    const dim3 grid((p.g.ni + T::TX - 1) / T::TX, (p.g.nzi + T::TZ - 1) / T::TZ, units);
    StepParams q = p;
    q.bnd_n = 0;
    void (*fn)(StepParams) = st::k_fwd_step<T, false, true>;
    hipLaunchKernelGGL(fn, grid, dim3(st::NT), 0, st, ordered(q, grid));
    return hipGetLastError();
}

// Lean tiles of T, and the general kernel of an unframed launch on TF tiles.
template <class T, class TF = T>
hipError_t fwd_step2_t(const StepParams& p, int units, bool store, hipStream_t st) {
    return framed<T, BandTile<T>, TF>(p, units, store ? st::k_fwd_step2_in<T, true> : st::k_fwd_step2_in<T, false>,
                                      store ? st::k_fwd_step2<TF, true> : st::k_fwd_step2<TF, false>,
                                      store ? st::k_fwd_step2<BandTile<T>, true> : st::k_fwd_step2<BandTile<T>, false>,
                                      st);
}

template <class T>
hipError_t adj_step_t(const StepParams& p, int units, hipStream_t st) {
    const dim3 grid((p.g.ni + T::TX - 1) / T::TX, (p.g.nzi + T::TZ - 1) / T::TZ, units);
    void (*fn)(StepParams) = p.dwav ? st::k_adj_step<T, true> : st::k_adj_step<T, false>;
    hipLaunchKernelGGL(fn, grid, dim3(st::NT), 0, st, ordered(p, grid));
    return hipGetLastError();
}

template <class T>
hipError_t adj_step2_t(const StepParams& p, int units, hipStream_t st) {
    return framed<T, BandTile<T>>(p, units, st::k_adj_step2_in<T>, st::k_adj_step2<T>, st::k_adj_step2<BandTile<T>>, st);
}

using T0 = st::Tile<64, 16>;
using T1 = st::Tile<128, 8>;
using T2 = st::Tile<32, 32>;
using T3 = st::Tile<64, 32>;

}  // namespace

hipError_t launch_fwd_step(const StepParams& p, int units, bool store, bool born, hipStream_t st) {
    if (units < 1 || units > 65535) return hipErrorInvalidValue;
    switch (p.tile) {
        case 1: return fwd_step_t<T1>(p, units, store, born, st);
        case 2: return fwd_step_t<T2>(p, units, store, born, st);
        case 3: return fwd_step_t<T3>(p, units, store, born, st);
        default: return fwd_step_t<T0>(p, units, store, born, st);
    }
}

hipError_t launch_fwd_step2(const StepParams& p, int units, bool store, hipStream_t st) {
    if (units < 1 || units > 65535 || p.t < 0 || p.t + 1 >= p.g.nt || (p.trace_prev && p.t < 1)) return hipErrorInvalidValue;
    if (store && (p.t < p.snap_t0 || p.t + 1 >= p.snap_t0 + p.seg_len)) return hipErrorInvalidValue;
    switch (p.tile) {
        case 1: return fwd_step2_t<T1>(p, units, store, st);
        case 2: return fwd_step2_t<T2>(p, units, store, st);
        case 3: return fwd_step2_t<T3>(p, units, store, st);
        // the default's lean tiles are 64 x 32 in the two-step forward (its
        // call 0.446 against 0.459 ms at 1000 x 300 x 160 units), 64 x 16
        // for the general kernel of an unframed launch (181-185 against
        // 101-103 VGPRs at 64 x 32)
        default: return fwd_step2_t<T3, T0>(p, units, store, st);
    }
}

hipError_t launch_fwd_record(const StepParams& p, int units, hipStream_t st) {
    if (units < 1 || p.t < 0 || p.t >= p.g.nt || !p.traces) return hipErrorInvalidValue;
    hipLaunchKernelGGL(st::k_fwd_record, dim3(units), dim3(st::NT), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_adj_step2(const StepParams& p, int units, hipStream_t st) {
    if (units < 1 || units > 65535 || p.dwav || p.t < 1) return hipErrorInvalidValue;
    switch (p.tile) {
        case 1: return adj_step2_t<T1>(p, units, st);
        case 2: return adj_step2_t<T2>(p, units, st);
        case 3: return adj_step2_t<T3>(p, units, st);
        default: return adj_step2_t<T0>(p, units, st);
    }
}

namespace st {
__global__ void k_pad_velocity(Grid g, const float* vel, int b0, int nb, float* vpad, CellLayout lay) {
    const size_t cells = (size_t)g.nx * g.nz, n = cells * nb;
    for (size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x; j < n; j += (size_t)gridDim.x * blockDim.x) {
        const int m = (int)(j / cells);
        const int i = (int)(j - (size_t)m * cells);
        const int z = i / g.nx, x = i - z * g.nx;
        vpad[(size_t)m * lay.ustride + lay.off + (size_t)z * lay.pitch + x] =
            vel_at(vel + (size_t)(b0 + m) * g.nvx * g.nvz, g, x, z);
    }
}
}  // namespace st

hipError_t launch_pad_velocity(const Grid& g, const float* vel, int b0, int nb, float* vpad, const CellLayout& lay,
                               hipStream_t st) {
    if (nb < 1 || lay.pitch < g.nx || lay.off < 0 ||
        lay.ustride < (size_t)lay.off + (size_t)lay.pitch * (g.nz - 1) + g.nx)
        return hipErrorInvalidValue;
    const size_t n = (size_t)g.nx * g.nz * nb;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(st::k_pad_velocity, dim3(blocks), dim3(256), 0, st, g, vel, b0, nb, vpad, lay);
    return hipGetLastError();
}

hipError_t launch_adj_step(const StepParams& p, int units, hipStream_t st) {
    if (units < 1 || units > 65535) return hipErrorInvalidValue;
    switch (p.tile) {
        case 1: return adj_step_t<T1>(p, units, st);
        case 2: return adj_step_t<T2>(p, units, st);
        case 3: return adj_step_t<T3>(p, units, st);
        default: return adj_step_t<T0>(p, units, st);
    }
}

}  // namespace wg
