// Device templates of the resident kernels specialised for a compile-time
// grid: the reference's 70 x 70 model with a 20-cell C-PML (114 x 114
// padded) as one workgroup per unit (wave_kernels_fixed.hip), and the
// 100 x 100 and 141 x 141 models as row bands (wave_kernels_band.hip).  The
// two sets of kernels are separate translation units, so that each is
// compiled (inlined, register-allocated) on its own.
//
// Same algorithm and arithmetic as wave_kernels.hip (bit-identical results);
// what changes is the cell-to-thread map.  The inner cells form six
// rectangles per kernel: interior, x-band, z-band top / bottom, corner top /
// bottom (bands of the kernel's C-PML region).  In a rectangle of width W a
// thread owns column tid % W and rows tid / W + k * RP, RP = NT / W, so every
// slot of a thread sits at a compile-time offset from the thread's base cell:
// LDS and snapshot addresses are base + immediate, and the x-axis C-PML
// coefficients are per-thread constants held in registers.
#pragma once

#include <algorithm>

#include "band_xchg.h"
#include "stencil.h"
#include "wave_kernels.h"

namespace wg {

namespace fx {

constexpr int NT = 512;

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ float clampv(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Opaque copies, taken once per phase: they stop the compiler from hoisting
// per-slot addresses and profile loads out of the time loop, which would cost
// one register per slot each; per slot, addresses are base + immediate.
// Exact product of small ints (|a|, |b| < 2^23) on the full-rate 24-bit
// multiplier; the laundered offsets below are recomputed every step.
__device__ __forceinline__ int m24(int a, int b) { return __mul24(a, b); }

// One step's snapshot row (G::INNER floats) as a raw buffer resource: its
// stores and loads take a 32-bit byte offset from the row, so the slots of a
// thread need no 64-bit address arithmetic (accesses past the row are
// dropped by the hardware bounds check).  Word 3: the gfx9 raw-buffer format.
// The row address is the same in every lane; it is moved to scalar registers
// explicitly, since a resource held in vector registers would make the
// compiler loop over lanes at every access.
template <class G>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t snap_rsrc(const float* row) {
    const unsigned long long a = reinterpret_cast<unsigned long long>(row);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    float* base = reinterpret_cast<float*>(((unsigned long long)hi << 32) | lo);
    return __builtin_amdgcn_make_buffer_rsrc(base, 0, G::INNER * 4, 0x00020000);
}

__device__ __forceinline__ int opq(int v) {
    asm volatile("" : "+v"(v));
    return v;
}
// Launders an offset, not a pointer, so LDS pointers keep their address space.
__device__ __forceinline__ const float* opq_s(const float* p) {
    int off = 0;
    asm volatile("" : "+s"(off));
    return p + off;
}

// A compiled grid and the inner rows [LO, HI) one workgroup holds: the whole
// grid (the defaults), or one row band of it.  A band's field is held in LDS
// from padded row ZB0 = LO (inner row LO - 2) for HI - LO + 4 rows, its
// x-strips for its own rows, and the z-strips of the sides it reaches (the
// band boundaries lie at least one adjoint strip width inside the grid).  The
// kernels index the field by padded row - ZB0, the x-strips by inner row -
// LO and a bottom z-strip row z by z - NZ + *ZB: for the whole grid these are
// the plain grid indices.
template <int NVX_, int NVZ_, int PML_, int LO_ = 0, int HI_ = -1>
struct Geo {
    static constexpr int NVX = NVX_, NVZ = NVZ_, PML = PML_;
    static constexpr int PAD = PML + 2;
    static constexpr int NX = NVX + 2 * PAD, NZ = NVZ + 2 * PAD;
    static constexpr int NI = NX - 4, NZI = NZ - 4, INNER = NI * NZI;
    static constexpr int REG = PML + 4, REGA = PML + 6;  // wave_wgpu.rs:291, 295
    static constexpr int SWF = REG + 2, SWA = REGA + 2;
    static constexpr int NPR = 4 * (NX + NZ);  // a, 1+b, b, db for x then z
    // rows held
    static constexpr int LO = LO_, HI = HI_ < 0 ? NZI : HI_;
    static constexpr bool WHOLE = LO == 0 && HI == NZI;
    static constexpr bool TOP = LO == 0, BOT = HI == NZI;  // z-strip sides held
    static constexpr int ZB0 = LO, ROWS = HI - LO + 4;     // padded field rows in LDS
    static constexpr int ZS = (TOP ? 1 : 0) + (BOT ? 1 : 0);
    static_assert(LO >= 0 && HI <= NZI && HI - LO >= 4, "band rows");
    static_assert(WHOLE || ((TOP || LO >= SWA) && (BOT || HI <= NZI - SWA)),
                  "band boundaries must lie at least one adjoint strip width inside the grid");
    // forward LDS (floats)
    static constexpr int F_U = 0, F_AX = ROWS * NX, F_AZ = F_AX + 2 * SWF * (HI - LO), F_PR = F_AZ + ZS * SWF * NI;
    static constexpr int F_TOTAL = F_PR + NPR;
    static constexpr int FZB = TOP ? 2 * SWF : SWF;  // strip row of padded row z >= NZ - SWF: z - NZ + FZB
    // adjoint LDS (floats); receiver tables follow A_TAIL
    static constexpr int A_V = 0, A_TX = ROWS * NX, A_UX = A_TX + 2 * SWA * (HI - LO), A_TZ = A_UX + 2 * SWA * (HI - LO);
    static constexpr int A_UZ = A_TZ + ZS * SWA * NI, A_PR = A_UZ + ZS * SWA * NI, A_TAIL = A_PR + NPR;
    static constexpr int AZB = TOP ? 2 * SWA : SWA;
    // the rows this workgroup owns as a Band (for the halo exchange)
    __device__ static Band band() { return Band{LO, HI, ZB0, ROWS, TOP, BOT, !TOP, !BOT}; }
};

// Profile table offsets inside PR.
template <class G>
struct Prof {
    static constexpr int AX = 0, OPBX = G::NX, BX = 2 * G::NX, DBX = 3 * G::NX;
    static constexpr int AZ = 4 * G::NX, OPBZ = AZ + G::NZ, BZ = AZ + 2 * G::NZ, DBZ = AZ + 3 * G::NZ;
};

// One rectangle of cells.  BX: columns are the two x-bands (else the interior
// columns); rows [ZOFF, ZOFF + H) of the inner grid (H = 0: a band without
// these cells); RB: band width.
template <class G, int RB, bool BX_, bool BZ_, int ZOFF_, int H_>
struct Rect {
    static constexpr bool BX = BX_, BZ = BZ_;
    static constexpr int ZOFF = ZOFF_, H = H_;
    static constexpr int W = BX ? 2 * RB : G::NI - 2 * RB;
    static constexpr int RP = NT / W;
    static constexpr int K = (H + RP - 1) / RP;
    static constexpr int ACTIVE = H > 0 ? RP * W : 0;
    static_assert(W > 0 && RP > 0 && H >= 0, "rectangle must be narrower than a workgroup");
    __device__ static int xi_of(int col) { return BX ? (col < RB ? col : col + G::NI - 2 * RB) : RB + col; }
};

// The rectangles of the rows G holds: interior rows [ILO, IHI) and the z
// bands of the sides it reaches.
template <class G, int RB>
struct Rects {
    static constexpr int ILO = G::LO > RB ? G::LO : RB;
    static constexpr int IHI = G::HI < G::NZI - RB ? G::HI : G::NZI - RB;
    using I = Rect<G, RB, false, false, ILO, IHI - ILO>;
    using X = Rect<G, RB, true, false, ILO, IHI - ILO>;
    using ZT = Rect<G, RB, false, true, 0, G::TOP ? RB : 0>;
    using ZB = Rect<G, RB, false, true, G::NZI - RB, G::BOT ? RB : 0>;
    using CT = Rect<G, RB, true, true, 0, G::TOP ? RB : 0>;
    using CB = Rect<G, RB, true, true, G::NZI - RB, G::BOT ? RB : 0>;
};

// Thread's place in a rectangle.
template <class R>
struct Place {
    bool active, last_ok;
    int col, row0, xi, zi0;
    __device__ void init(int tid) {
        active = tid < R::ACTIVE;
        col = tid % R::W;
        row0 = tid / R::W;
        xi = R::xi_of(col);
        zi0 = R::ZOFF + row0;
        last_ok = active && (row0 + (R::K - 1) * R::RP < R::H);
    }
    __device__ bool valid(int k) const { return k < R::K - 1 ? active : last_ok; }
    // xi and zi0 recomputed from the thread index, so that they need not stay
    // live in registers across the time loop
    __device__ static int xi_at(int tid) { return R::xi_of(tid % R::W); }
    __device__ static int zi0_at(int tid) { return R::ZOFF + tid / R::W; }
    // The slot whose padded row lies in [z_lo, z_hi), or -1 (rows of one
    // thread are RP > 1 apart, so a band of at most RP rows holds one).
    __device__ int slot_in_rows(int z_lo, int z_hi) const {
        int r = -1;
#pragma unroll
        for (int k = 0; k < R::K; ++k) {
            const int z = zi0 + k * R::RP + 2;
            if (valid(k) && z >= z_lo && z < z_hi) r = k;
        }
        return r;
    }
};

// Which rectangle of band width RB holds inner cell (xi, zi), and the owner
// (thread, slot): returns rectangle id 0..5 (I, X, ZT, ZB, CT, CB), or -1
// (tid = -1) when the cell is outside the rows G holds.
template <class G, int RB>
__device__ inline int owner_of(int xi, int zi, int& tid, int& slot) {
    if constexpr (!G::WHOLE) {
        if (zi < G::LO || zi >= G::HI) {
            tid = -1;
            slot = 0;
            return -1;
        }
    }
    const bool bx = xi < RB || xi >= G::NI - RB;
    const bool bz = zi < RB || zi >= G::NZI - RB;
    const int W = bx ? 2 * RB : G::NI - 2 * RB;
    const int col = bx ? (xi < RB ? xi : xi - (G::NI - 2 * RB)) : xi - RB;
    const int zoff = !bz ? Rects<G, RB>::ILO : (zi < RB ? 0 : G::NZI - RB);
    const int row = zi - zoff;
    const int RP = NT / W;
    tid = (row % RP) * W + col;
    slot = row / RP;
    return !bz ? (bx ? 1 : 0) : (zi < RB ? (bx ? 4 : 2) : (bx ? 5 : 3));
}

// ===========================================================================
// Forward
// ===========================================================================
template <int K_, bool PX, bool PZ, bool STORE, bool BORN>
struct FSt {
    static constexpr int K = K_ > 0 ? K_ : 1;  // array sizes (an empty rectangle has no slots)
    float uprev[K], unext[K], c1[K];          // c1 = v^2 dt^2
    float c2[STORE ? K : 1];                  // 2 v dt^2 (snapshot factor)
    float px[PX ? K : 1], zx[PX ? K : 1];
    float pz[PZ ? K : 1], zz[PZ ? K : 1];
    float dv[BORN ? K : 1];
};

// Time-split runs hand their state to the next part through HBM: entry j of
// thread tid at st[j * NT + tid].  SAVE writes the register, else reads it.
template <bool SAVE>
__device__ __forceinline__ void xfer(float& v, float* st, int& j, int tid) {
    float* p = st + (size_t)(j++) * NT + tid;
    if constexpr (SAVE) *p = v;
    else v = *p;
}

// Forward: u^{t-1} and the C-PML memory variables of every slot.
template <bool SAVE, int K, bool PX, bool PZ, bool STORE, bool BORN>
__device__ __forceinline__ void f_xfer(FSt<K, PX, PZ, STORE, BORN>& S, float* st, int& j, int tid) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
        xfer<SAVE>(S.uprev[k], st, j, tid);
        if constexpr (PX) {
            xfer<SAVE>(S.px[k], st, j, tid);
            xfer<SAVE>(S.zx[k], st, j, tid);
        }
        if constexpr (PZ) {
            xfer<SAVE>(S.pz[k], st, j, tid);
            xfer<SAVE>(S.zz[k], st, j, tid);
        }
    }
}
template <class R>
constexpr int f_regs() { return R::K * (1 + (R::BX ? 2 : 0) + (R::BZ ? 2 : 0)); }

struct FCtxF {
    float *U, *AX, *AZ;
    const float* PR;
    float rh, rh2, dt2;
};

template <class G, class R, bool STORE, bool BORN>
__device__ __forceinline__ void f_init(FSt<R::K, R::BX, R::BZ, STORE, BORN>& S, const Place<R>& pl,
                                       const float* vvis, const float* dvis, const Grid& g) {
#pragma unroll
    for (int k = 0; k < R::K; ++k) {
        S.uprev[k] = S.unext[k] = S.c1[k] = 0.0f;
        if constexpr (STORE) S.c2[k] = 0.0f;
        if constexpr (R::BX) { S.px[k] = 0.0f; S.zx[k] = 0.0f; }
        if constexpr (R::BZ) { S.pz[k] = 0.0f; S.zz[k] = 0.0f; }
        if constexpr (BORN) S.dv[k] = 0.0f;
        // clamped model index: every thread loads, valid slots keep the values
        const int x = pl.xi + 2, z = pl.zi0 + k * R::RP + 2;
        const int vz = clampi(z - G::PAD, 0, G::NVZ - 1), vx = clampi(x - G::PAD, 0, G::NVX - 1);
        const float vr = vvis[vz * G::NVX + vx];
        const float v = clampv(vr, g.vmin, g.vmax);
        const bool ok = pl.valid(k);
        S.c1[k] = ok ? v * v * g.dt2 : 0.0f;
        if constexpr (STORE) S.c2[k] = ok ? 2.0f * v * g.dt2 : 0.0f;
        if constexpr (BORN) {
            const float m = (vr > g.vmin && vr < g.vmax) ? 1.0f : 0.0f;
            S.dv[k] = ok ? dvis[vz * G::NVX + vx] * m : 0.0f;
        }
    }
}

template <class G, class R>
__device__ __forceinline__ Pml pml_x(const float* PR, int x) {
    using P = Prof<G>;
    return Pml{PR[P::AX + x], PR[P::OPBX + x], PR[P::BX + x], PR[P::DBX + x]};
}
template <class G>
__device__ __forceinline__ Pml pml_z(const float* PR, int z) {
    using P = Prof<G>;
    return Pml{PR[P::AZ + z], PR[P::OPBZ + z], PR[P::BZ + z], PR[P::DBZ + z]};
}

// lib.rs:151-206 for one rectangle's slots; new values stay in registers
// until every thread has read u^t.
template <class G, class R, bool STORE, bool BORN>
__device__ __forceinline__ void f_compute(FSt<R::K, R::BX, R::BZ, STORE, BORN>& S, const Place<R>& pl,
                                          const FCtxF& C, const Pml& pmx, float* snap_t, const float* bsnap_t) {
    if (!pl.active) return;
    const int xi = opq(pl.xi), zi0 = opq(pl.zi0);
    const float* PR = opq_s(C.PR);
    const int x = xi + 2;
    const int z0 = zi0 + 2;
    const int i0 = m24(z0 - G::ZB0, G::NX) + x;
    const int q0 = m24(zi0, G::NI) + xi;
    constexpr int SW2 = 2 * G::SWF;
    // Base pointers at the lowest address each slot reads (LDS offsets are
    // unsigned immediates): field at i0 - 2 NX - 2, strips at -2 / -2 NI.
    const float* Ub = C.U + (i0 - 2 * G::NX - 2);
    const float* AXb = C.AX + (m24(zi0 - G::LO, SW2) + (x < G::REG ? x : x - G::NX + SW2) - 2);
    const float* AZb = C.AZ + (m24(R::ZOFF == 0 ? z0 : z0 - G::NZ + G::FZB, G::NI) + xi - 2 * G::NI);
    float dummy = 0.0f;
    [[maybe_unused]] const __amdgpu_buffer_rsrc_t srs = snap_rsrc<G>(snap_t);
#pragma unroll
    for (int k = 0; k < R::K; ++k) {
        if (k == R::K - 1 && !pl.last_ok) break;
        const int li = k * R::RP * G::NX + 2 * G::NX + 2;  // slot k relative to Ub
        const float u = Ub[li];
        Pml pmz{0, 0, 0, 0};
        if constexpr (R::BZ) pmz = pml_z<G>(PR, z0 + k * R::RP);
        float& px = R::BX ? S.px[R::BX ? k : 0] : dummy;
        float& zx = R::BX ? S.zx[R::BX ? k : 0] : dummy;
        float& pz = R::BZ ? S.pz[R::BZ ? k : 0] : dummy;
        float& zz = R::BZ ? S.zz[R::BZ ? k : 0] : dummy;
        const float w = fwd_w<R::BX, R::BZ>(Ub, li, G::NX, u, C.rh, C.rh2, AXb + (k * R::RP * SW2 + 2), pmx, px, zx,
                                            AZb + (k * R::RP * G::NI + 2 * G::NI), G::NI, pmz, pz, zz);
        const int q = q0 + k * R::RP * G::NI;
        if constexpr (STORE)  // 2 v dt^2 w (splatForward.wgsl:108-110)
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(S.c2[k] * w), srs, q * 4, 0, 0);
        float next = S.c1[k] * w + 2.0f * u - S.uprev[k];
        if constexpr (BORN) next += S.dv[k] * bsnap_t[q];
        S.unext[k] = next;
        S.uprev[k] = u;
    }
}

template <class G, class R, bool STORE, bool BORN>
__device__ __forceinline__ void f_write(FSt<R::K, R::BX, R::BZ, STORE, BORN>& S, const Place<R>& pl,
                                        const FCtxF& C, const Pml& pmx) {
    if (!pl.active) return;
    const int xi = opq(pl.xi), zi0 = opq(pl.zi0);
    const float* PR = opq_s(C.PR);
    const int x = xi + 2;
    const int z0 = zi0 + 2;
    const int i0 = m24(z0 - G::ZB0, G::NX) + x;
    constexpr int SW2 = 2 * G::SWF;
    float* AX0 = C.AX + (m24(zi0 - G::LO, SW2) + (x < G::REG ? x : x - G::NX + SW2));
    float* AZ0 = C.AZ + (m24(R::ZOFF == 0 ? z0 : z0 - G::NZ + G::FZB, G::NI) + xi);
#pragma unroll
    for (int k = 0; k < R::K; ++k) {
        if (k == R::K - 1 && !pl.last_ok) break;
        C.U[i0 + k * R::RP * G::NX] = S.unext[k];
        // a * psi_next, the product neighbours take in lib.rs:167-170, 190-193
        if constexpr (R::BX) AX0[k * R::RP * SW2] = pmx.a * S.px[k];
        if constexpr (R::BZ) AZ0[k * R::RP * G::NI] = PR[Prof<G>::AZ + z0 + k * R::RP] * S.pz[k];
    }
}

// Per-unit state of a time-split forward: the LDS field and strips [0, F_PR),
// then f_regs of every rectangle per thread.
template <class G>
struct FState {
    using RR = Rects<G, G::REG - 2>;
    static constexpr int REGS = f_regs<typename RR::I>() + f_regs<typename RR::X>() + f_regs<typename RR::ZT>() +
                                f_regs<typename RR::ZB>() + f_regs<typename RR::CT>() + f_regs<typename RR::CB>();
    static constexpr size_t FLOATS = (size_t)G::F_PR + (size_t)REGS * NT;
};

// Steps [t0, t1) of the forward of unit P.unit0 + blk (snapshots at launch
// position blk).  st_in (t0 > 0) restores the state a previous part saved;
// st_out (t1 < nt) saves it for the next part; snaps_at, when set, is where
// this unit's snapshots go (else P.snaps at position blk).  A band (G not the whole
// grid) exchanges its edge rows with band bp's neighbours through X after
// every step; the top band records the gathers, NaN when a wait gave up.
template <class G, bool STORE, bool BORN>
__device__ __forceinline__ void forward_run(const FwdParams& P, float* lds, int blk, int t0, int t1,
                                            const float* st_in, float* st_out, const Xchg* X = nullptr,
                                            int bp = 0, int* fail_s = nullptr, float* snaps_at = nullptr) {
    using RR = Rects<G, G::REG - 2>;
    const Grid& g = P.g;
    FCtxF C;
    C.U = lds + G::F_U;
    C.AX = lds + G::F_AX;
    C.AZ = lds + G::F_AZ;
    float* PR = lds + G::F_PR;
    C.PR = PR;
    C.rh = g.rh;
    C.rh2 = g.rh2;
    C.dt2 = g.dt2;

    const int tid = threadIdx.x;
    const int unit = P.unit0 + blk;
    const int b = unit / g.shots;
    const int s = src_slot(P.geo, unit, unit - b * g.shots);
    if (!unit_active(P.geo, unit)) {  // padding unit of an uneven batch
        if (G::TOP) zero_unit_outputs(P.traces + (size_t)unit * g.nt * g.nrec, g.nt * g.nrec, nullptr, 0);
        return;
    }

    // the LDS field and strips: restored from the previous part (eight loads
    // in flight per thread, not one load and wait at a time) or cleared
    {
        constexpr int RU = 8;
        for (int j0 = tid; j0 < G::F_PR; j0 += RU * NT) {
            float v[RU];
#pragma unroll
            for (int u = 0; u < RU; ++u) {
                const int j = j0 + u * NT;
                v[u] = (st_in && j < G::F_PR) ? st_in[j] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < RU; ++u) {
                const int j = j0 + u * NT;
                if (j < G::F_PR) lds[j] = v[u];
            }
        }
    }
    // profile tables: a, 1 + b, b, db for x then z (geometry prof: ax bx dbx az bz dbz)
    for (int j = tid; j < G::NX; j += NT) {
        const float bb = P.geo.prof[G::NX + j];
        PR[Prof<G>::AX + j] = P.geo.prof[j];
        PR[Prof<G>::OPBX + j] = 1.0f + bb;
        PR[Prof<G>::BX + j] = bb;
        PR[Prof<G>::DBX + j] = P.geo.prof[2 * G::NX + j];
    }
    for (int j = tid; j < G::NZ; j += NT) {
        const float* pz = P.geo.prof + 3 * G::NX;
        const float bb = pz[G::NZ + j];
        PR[Prof<G>::AZ + j] = pz[j];
        PR[Prof<G>::OPBZ + j] = 1.0f + bb;
        PR[Prof<G>::BZ + j] = bb;
        PR[Prof<G>::DBZ + j] = pz[2 * G::NZ + j];
    }

    const float* vvis = P.vel + (size_t)b * G::NVX * G::NVZ;
    const float* dvis = BORN ? P.dvel + (size_t)b * G::NVX * G::NVZ : nullptr;
    const float* wav = P.wavelet + (size_t)unit * P.wav_stride;
    float* traces = P.traces + (size_t)unit * g.nt * g.nrec;
    float* snaps = STORE ? (snaps_at ? snaps_at : P.snaps + (size_t)blk * g.nt * G::INNER) : nullptr;
    const float* bsnaps = BORN ? P.born_snaps + (size_t)blk * g.nt * G::INNER : nullptr;

    Place<typename RR::I> pI;
    Place<typename RR::X> pX;
    Place<typename RR::ZT> pZT;
    Place<typename RR::ZB> pZB;
    Place<typename RR::CT> pCT;
    Place<typename RR::CB> pCB;
    pI.init(tid); pX.init(tid); pZT.init(tid); pZB.init(tid); pCT.init(tid); pCB.init(tid);
    FSt<RR::I::K, false, false, STORE, BORN> sI;
    FSt<RR::X::K, true, false, STORE, BORN> sX;
    FSt<RR::ZT::K, false, true, STORE, BORN> sZT;
    FSt<RR::ZB::K, false, true, STORE, BORN> sZB;
    FSt<RR::CT::K, true, true, STORE, BORN> sCT;
    FSt<RR::CB::K, true, true, STORE, BORN> sCB;
    f_init<G>(sI, pI, vvis, dvis, g);
    f_init<G>(sX, pX, vvis, dvis, g);
    f_init<G>(sZT, pZT, vvis, dvis, g);
    f_init<G>(sZB, pZB, vvis, dvis, g);
    f_init<G>(sCT, pCT, vvis, dvis, g);
    f_init<G>(sCB, pCB, vvis, dvis, g);
    if (st_in) {
        float* sr = const_cast<float*>(st_in) + G::F_PR;
        int j = 0;
        f_xfer<false>(sI, sr, j, tid);
        f_xfer<false>(sX, sr, j, tid);
        f_xfer<false>(sZT, sr, j, tid);
        f_xfer<false>(sZB, sr, j, tid);
        f_xfer<false>(sCT, sr, j, tid);
        f_xfer<false>(sCB, sr, j, tid);
    }

    // Point-source corners owned by this thread: added after the update by
    // the owner, -(v^2 dt^2) w f[t] (lib.rs:209-215); Born adds its derivative.
    int scell[4];
    float scoef[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        scell[c] = -1;
        scoef[c] = 0.0f;
        const int cell = P.geo.src_cell[s * 4 + c];
        const float w = P.geo.src_w[s * 4 + c];
        const int cz = cell / G::NX, cx = cell - cz * G::NX;
        if (w != 0.0f && cx >= 2 && cx < G::NX - 2 && cz >= 2 && cz < G::NZ - 2) {
            int otid, oslot;
            owner_of<G, G::REG - 2>(cx - 2, cz - 2, otid, oslot);
            if (otid == tid) {
                const int vz = clampi(cz - G::PAD, 0, G::NVZ - 1), vx = clampi(cx - G::PAD, 0, G::NVX - 1);
                const float vr = vvis[vz * G::NVX + vx];
                const float v = clampv(vr, g.vmin, g.vmax);
                scell[c] = cell - G::ZB0 * G::NX;  // LDS index
                if (BORN) {
                    const float m = (vr > g.vmin && vr < g.vmax) ? 1.0f : 0.0f;
                    scoef[c] = -(2.0f * v * (dvis[vz * G::NVX + vx] * m) * g.dt2) * w;
                } else {
                    scoef[c] = -(v * v * g.dt2) * w;
                }
            }
        }
    }
    int rc[4];
    float rw[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        rc[c] = 0;
        rw[c] = 0.0f;
        if (G::TOP && tid < g.nrec) {  // the receiver rows lie in the top band
            rc[c] = P.geo.rec_cell[tid * 4 + c];
            rw[c] = P.geo.rec_w[tid * 4 + c];
        }
    }
    __syncthreads();
    // x-axis coefficients of this thread's band column (shared by X, CT, CB)
    const Pml pmx = pml_x<G, typename RR::X>(PR, pX.xi + 2);

    bool failed = false;
    for (int t = t0; t < t1; ++t) {
        if (G::TOP && tid < g.nrec) {  // receivers record u^t before the update (lib.rs:146-149)
            float acc = 0.0f;
#pragma unroll
            for (int c = 0; c < 4; ++c) acc += rw[c] * C.U[rc[c]];
            traces[(size_t)t * g.nrec + tid] = acc;
        }
        float* snap_t = STORE ? snaps + (size_t)t * G::INNER : nullptr;
        const float* bsnap_t = BORN ? bsnaps + (size_t)t * G::INNER : nullptr;
        f_compute<G>(sI, pI, C, pmx, snap_t, bsnap_t);
        f_compute<G>(sX, pX, C, pmx, snap_t, bsnap_t);
        f_compute<G>(sZT, pZT, C, pmx, snap_t, bsnap_t);
        f_compute<G>(sZB, pZB, C, pmx, snap_t, bsnap_t);
        f_compute<G>(sCT, pCT, C, pmx, snap_t, bsnap_t);
        f_compute<G>(sCB, pCB, C, pmx, snap_t, bsnap_t);
        __syncthreads();
        f_write<G>(sI, pI, C, pmx);
        f_write<G>(sX, pX, C, pmx);
        f_write<G>(sZT, pZT, C, pmx);
        f_write<G>(sZB, pZB, C, pmx);
        f_write<G>(sCT, pCT, C, pmx);
        f_write<G>(sCB, pCB, C, pmx);
        const float f_t = wav[t];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (scell[c] >= 0) C.U[scell[c]] += scoef[c] * f_t;
        __syncthreads();
        if constexpr (!G::WHOLE) {
            if (!band_exchange(*X, C.U, G::band(), G::NX, blk, bp, t + 1, fail_s)) {
                failed = true;
                break;
            }
        }
    }
    if (!G::WHOLE && failed && G::TOP)  // a neighbour never arrived: the gathers are not valid
        for (int j = tid; j < g.nt * g.nrec; j += NT) traces[j] = __builtin_nanf("");
    if (st_out) {
        for (int j = tid; j < G::F_PR; j += NT) st_out[j] = lds[j];
        float* sr = st_out + G::F_PR;
        int j = 0;
        f_xfer<true>(sI, sr, j, tid);
        f_xfer<true>(sX, sr, j, tid);
        f_xfer<true>(sZT, sr, j, tid);
        f_xfer<true>(sZB, sr, j, tid);
        f_xfer<true>(sCT, sr, j, tid);
        f_xfer<true>(sCB, sr, j, tid);
    }
}


// ===========================================================================
// Adjoint (splatAdjoint.wgsl:58-131), phases as in wave_kernels.hip.
// ===========================================================================
template <int K_, bool PX, bool PZ>
struct ASt {
    static constexpr int K = K_ > 0 ? K_ : 1;  // array sizes (an empty rectangle has no slots)
    float lam[K], lamp[K], gr[K], c1[K];  // c1 = v^2 dt^2
    float Px[PX ? K : 1], Zx[PX ? K : 1];
    float Pz[PZ ? K : 1], Zz[PZ ? K : 1];
};

// Adjoint: lambda^t, lambda^{t+1}, the gradient sums and the memory variables.
template <bool SAVE, int K, bool PX, bool PZ>
__device__ __forceinline__ void a_xfer(ASt<K, PX, PZ>& S, float* st, int& j, int tid) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
        xfer<SAVE>(S.lam[k], st, j, tid);
        xfer<SAVE>(S.lamp[k], st, j, tid);
        xfer<SAVE>(S.gr[k], st, j, tid);
        if constexpr (PX) {
            xfer<SAVE>(S.Px[k], st, j, tid);
            xfer<SAVE>(S.Zx[k], st, j, tid);
        }
        if constexpr (PZ) {
            xfer<SAVE>(S.Pz[k], st, j, tid);
            xfer<SAVE>(S.Zz[k], st, j, tid);
        }
    }
}
template <class R>
constexpr int a_regs() { return R::K * (3 + (R::BX ? 2 : 0) + (R::BZ ? 2 : 0)); }

struct ACtxF {
    float *V, *TX, *UX, *TZ, *UZ;
    const float* PR;
    const float* RES;
    float* SRC;
    const int* BOFF;
    const int* ER;
    const float* EW;
    int bz0, bz1;
    float rh;
};

struct SrcOwnF {
    int slot, corner;
    float coef;
};

// x-axis data of a band column: a, b, db and (1 + b) at x - 2 .. x + 2.
struct ColX {
    float a, b, db;
    float opb[5];
};

template <class G, class R>
__device__ __forceinline__ void a_init(ASt<R::K, R::BX, R::BZ>& S, const Place<R>& pl, const float* vvis,
                                       const Grid& g) {
#pragma unroll
    for (int k = 0; k < R::K; ++k) {
        S.lam[k] = S.lamp[k] = S.gr[k] = S.c1[k] = 0.0f;
        if constexpr (R::BX) { S.Px[k] = 0.0f; S.Zx[k] = 0.0f; }
        if constexpr (R::BZ) { S.Pz[k] = 0.0f; S.Zz[k] = 0.0f; }
        // The model index is clamped, so every thread may load (the loads of
        // all slots then overlap); only valid slots keep the factor.
        const int x = pl.xi + 2, z = pl.zi0 + k * R::RP + 2;
        const int vz = clampi(z - G::PAD, 0, G::NVZ - 1), vx = clampi(x - G::PAD, 0, G::NVX - 1);
        const float v = clampv(vvis[vz * G::NVX + vx], g.vmin, g.vmax);
        S.c1[k] = pl.valid(k) ? v * v * g.dt2 : 0.0f;  // the factor of vw(), wgsl:53-56
    }
}

// Phase 1: publish V, and T / U1 on the strips (T = V just outside the region).
// With DWAV the step's lambda at the source corners goes to LDS for the
// wavelet gradient; the plain adjoint needs it nowhere.
template <class G, class R, bool DWAV>
__device__ __forceinline__ void a_phase1(ASt<R::K, R::BX, R::BZ>& S, const Place<R>& pl, const ACtxF& C,
                                         const ColX& cx, const SrcOwnF& so) {
    if (!pl.active) return;
    constexpr int SW2 = 2 * G::SWA;
    const int tq = opq((int)threadIdx.x);
    const int xi = Place<R>::xi_at(tq), zi0 = Place<R>::zi0_at(tq);
    const float* PR = opq_s(C.PR);
    const int sk = DWAV ? opq(so.slot) : -1;
    const int x = xi + 2;
    const int z0 = zi0 + 2;
    const int i0 = m24(z0 - G::ZB0, G::NX) + x;
    const bool xstrip = x < G::SWA || x >= G::NX - G::SWA;
    const int xcol = x < G::SWA ? x : x - G::NX + SW2;
    const int tx0 = m24(zi0 - G::LO, SW2) + xcol;
    const int tz0 = m24(R::ZOFF == 0 ? z0 : z0 - G::NZ + G::AZB, G::NI) + xi;
#pragma unroll
    for (int k = 0; k < R::K; ++k) {
        if (k == R::K - 1 && !pl.last_ok) break;
        const float Vc = S.c1[k] * S.lam[k];
        C.V[i0 + k * R::RP * G::NX] = Vc;
        const int tx = tx0 + k * R::RP * SW2;
        if constexpr (R::BX) {
            const float T = (cx.opb[2]) * Vc + cx.b * S.Zx[k];
            const float U1 = cx.db * T + cx.b * S.Px[k];
            C.TX[tx] = T;
            C.UX[tx] = U1;
        } else {
            if (xstrip) C.TX[tx] = Vc;
        }
        const int z = z0 + k * R::RP;
        if constexpr (R::BZ) {
            const float bb = PR[Prof<G>::BZ + z];
            const float T = PR[Prof<G>::OPBZ + z] * Vc + bb * S.Zz[k];
            const float U1 = PR[Prof<G>::DBZ + z] * T + bb * S.Pz[k];
            const int tz = tz0 + k * R::RP * G::NI;
            C.TZ[tz] = T;
            C.UZ[tz] = U1;
        } else {
            // Slot k's rows are [ZOFF + 2 + k RP, ZOFF + 1 + (k + 1) RP]: only
            // the slots whose range meets a z strip test for it (folded at
            // compile time after unrolling).
            constexpr int zlo = R::ZOFF + 2;
            if (zlo + k * R::RP < G::SWA && z < G::SWA) C.TZ[m24(z, G::NI) + xi] = Vc;
            else if (zlo + (k + 1) * R::RP - 1 >= G::NZ - G::SWA && z >= G::NZ - G::SWA)
                C.TZ[m24(z - G::NZ + G::AZB, G::NI) + xi] = Vc;
        }
        if constexpr (DWAV) {
            if (k == sk) C.SRC[so.corner] = S.lam[k];
        }
    }
}

template <class G, class R>
__device__ __forceinline__ void a_phase2(ASt<R::K, R::BX, R::BZ>& S, const Place<R>& pl, const ACtxF& C,
                                         const ColX& cx, const SrcOwnF& so, int rslot, const float* snap_t,
                                         float f_t) {
    if (!pl.active) return;
    constexpr int SW2 = 2 * G::SWA;
    const int tq = opq((int)threadIdx.x);
    const int xi = Place<R>::xi_at(tq), zi0 = Place<R>::zi0_at(tq);
    const float* PR = opq_s(C.PR);
    const int sk = opq(so.slot);
    const int rk = opq(rslot);
    const int x = xi + 2;
    const int z0 = zi0 + 2;
    const int i0 = m24(z0 - G::ZB0, G::NX) + x;
    const int q0 = m24(zi0, G::NI) + xi;
    // Base pointers at the lowest address each slot reads (LDS offsets are
    // unsigned immediates): V at i0 - 2 NX - 2, x strips at -2, z strips at -2 NI.
    const float* Vb = C.V + (i0 - 2 * G::NX - 2);
    const int txb = m24(zi0 - G::LO, SW2) + (x < G::REGA ? x : x - G::NX + SW2) - 2;
    const float* TXb = C.TX + txb;
    const float* UXb = C.UX + txb;
    const int tzb = m24(R::ZOFF == 0 ? z0 : z0 - G::NZ + G::AZB, G::NI) + xi - 2 * G::NI;
    const float* TZb = C.TZ + tzb;
    const float* UZb = C.UZ + tzb;
    const __amdgpu_buffer_rsrc_t srs = snap_rsrc<G>(snap_t);
    const float rh = C.rh;
#pragma unroll
    for (int k = 0; k < R::K; ++k) {
        if (k == R::K - 1 && !pl.last_ok) break;
        const int li = k * R::RP * G::NX + 2 * G::NX + 2;  // slot k relative to Vb
        const int z = z0 + k * R::RP;
        const float Sn = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(srs, (q0 + k * R::RP * G::NI) * 4, 0, 0));
        const float v0 = Vb[li];
        float w;  // x term, then += z term (0 + x skipped: equal up to the sign of zero)
        if constexpr (R::BX) {
            w = adj_axis_pml(TXb + (k * R::RP * SW2 + 2), UXb + (k * R::RP * SW2 + 2), 1, cx.opb, cx.a, v0, rh,
                             S.Px[k], S.Zx[k]);
        } else {
            w = adj_axis_int(Vb, li, 1, v0, rh);
        }
        if constexpr (R::BZ) {
            const float opb[5] = {PR[Prof<G>::OPBZ + z - 2], PR[Prof<G>::OPBZ + z - 1], PR[Prof<G>::OPBZ + z],
                                  PR[Prof<G>::OPBZ + z + 1], PR[Prof<G>::OPBZ + z + 2]};
            const int tz = k * R::RP * G::NI + 2 * G::NI;
            w += adj_axis_pml(TZb + tz, UZb + tz, G::NI, opb, PR[Prof<G>::AZ + z], v0, rh, S.Pz[k], S.Zz[k]);
        } else {
            w += adj_axis_int(Vb, li, G::NX, v0, rh);
        }
        float next = w + 2.0f * S.lam[k] - S.lamp[k];
        float gk = S.lam[k] * Sn;  // store_every = 1 (wgsl:234-236)
        if (k == rk) {  // receiver residuals, CSR order (wgsl:239-242)
            const int o = m24(z - C.bz0, G::NX) + x;
            for (int e = C.BOFF[o]; e < C.BOFF[o + 1]; ++e) next += C.EW[e] * C.RES[C.ER[e]];
        }
        if (k == sk) gk += S.lam[k] * (so.coef * f_t);  // wgsl:245-249
        S.gr[k] += gk;
        S.lamp[k] = S.lam[k];
        S.lam[k] = next;
    }
}

template <class G, class R>
__device__ __forceinline__ void a_store(ASt<R::K, R::BX, R::BZ>& S, const Place<R>& pl, float* gout) {
    if (!pl.active) return;
    const int i0 = (pl.zi0 + 2) * G::NX + pl.xi + 2;
#pragma unroll
    for (int k = 0; k < R::K; ++k) {
        if (k == R::K - 1 && !pl.last_ok) break;
        gout[i0 + k * R::RP * G::NX] = S.gr[k];
    }
}

// Per-unit state of a time-split adjoint: a_regs of every rectangle per
// thread (the LDS content is rebuilt from it at every step).
template <class G>
struct AState {
    using RR = Rects<G, G::REGA - 2>;
    static constexpr int REGS = a_regs<typename RR::I>() + a_regs<typename RR::X>() + a_regs<typename RR::ZT>() +
                                a_regs<typename RR::ZB>() + a_regs<typename RR::CT>() + a_regs<typename RR::CB>();
    static constexpr size_t FLOATS = (size_t)REGS * NT;
};

// Steps t_hi - 1 down to t_lo of the adjoint of unit P.unit0 + blk
// (snapshots at launch position blk); st_in (t_hi < nt) restores, st_out
// (t_lo > 0) saves the state; the gradient is written when t_lo = 0.  A band
// (G not the whole grid) exchanges V with band bp's neighbours through X
// after every phase 1 (step numbers tag0 + 1, tag0 + 2, ...: a band that ran
// its forward on the same exchange words continues that numbering) and
// writes the gradient of its own rows, NaN when a wait gave up.
template <class G, bool DWAV>
__device__ __forceinline__ void adjoint_run(const AdjParams& P, float* lds, int blk, int t_hi, int t_lo,
                                            const float* st_in, float* st_out, const Xchg* X = nullptr,
                                            int bp = 0, int* fail_s = nullptr, int tag0 = 0,
                                            const float* snaps_at = nullptr) {
    static_assert(G::WHOLE || !DWAV, "bands run the adjoint without the wavelet gradient");
    using RR = Rects<G, G::REGA - 2>;
    const Grid& g = P.g;
    float* PR = lds + G::A_PR;
    float* RES = lds + G::A_TAIL;
    float* SRC = RES + g.nrec;  // [0..4) lambda at corners, [4..8) -v^2 dt^2 w
    int* BOFF = reinterpret_cast<int*>(SRC + 8);
    int* ER = BOFF + P.geo.band_rows * G::NX + 1;
    float* EW = reinterpret_cast<float*>(ER + P.geo.n_ent);
    ACtxF C;
    C.V = lds + G::A_V;
    C.TX = lds + G::A_TX;
    C.UX = lds + G::A_UX;
    C.TZ = lds + G::A_TZ;
    C.UZ = lds + G::A_UZ;
    C.PR = PR;
    C.RES = RES;
    C.SRC = SRC;
    C.BOFF = BOFF;
    C.ER = ER;
    C.EW = EW;
    C.bz0 = P.geo.band_z0;
    C.bz1 = P.geo.band_z0 + P.geo.band_rows;
    C.rh = g.rh;

    const int tid = threadIdx.x;
    const int unit = P.unit0 + blk;
    const int b = unit / g.shots;
    const int s = src_slot(P.geo, unit, unit - b * g.shots);
    if (!unit_active(P.geo, unit)) {  // padding unit of an uneven batch
        if (G::TOP)
            zero_unit_outputs(P.grad + (size_t)unit * g.nx * g.nz, g.nx * g.nz,
                              DWAV ? P.dwav + (size_t)unit * g.nt : nullptr, DWAV ? g.nt : 0);
        return;
    }

    for (int j = tid; j < G::A_PR; j += NT) lds[j] = 0.0f;
    for (int j = tid; j < G::NX; j += NT) {
        const float bb = P.geo.prof[G::NX + j];
        PR[Prof<G>::AX + j] = P.geo.prof[j];
        PR[Prof<G>::OPBX + j] = 1.0f + bb;
        PR[Prof<G>::BX + j] = bb;
        PR[Prof<G>::DBX + j] = P.geo.prof[2 * G::NX + j];
    }
    for (int j = tid; j < G::NZ; j += NT) {
        const float* pz = P.geo.prof + 3 * G::NX;
        const float bb = pz[G::NZ + j];
        PR[Prof<G>::AZ + j] = pz[j];
        PR[Prof<G>::OPBZ + j] = 1.0f + bb;
        PR[Prof<G>::BZ + j] = bb;
        PR[Prof<G>::DBZ + j] = pz[2 * G::NZ + j];
    }
    for (int j = tid; j < P.geo.band_rows * G::NX + 1; j += NT) BOFF[j] = P.geo.band_off[j];
    for (int j = tid; j < P.geo.n_ent; j += NT) {
        ER[j] = P.geo.ent_r[j];
        EW[j] = P.geo.ent_w[j];
    }

    const float* vvis = P.vel + (size_t)b * G::NVX * G::NVZ;
    const float* wav = P.wavelet + (size_t)unit * P.wav_stride;
    const float* resid = P.resid + (size_t)unit * g.nt * g.nrec;
    const float* snaps = snaps_at ? snaps_at : P.snaps + (size_t)blk * g.nt * G::INNER;

    if (tid < 4) {
        const int cell = P.geo.src_cell[s * 4 + tid];
        const int cz = cell / G::NX, cx = cell - cz * G::NX;
        const int vz = clampi(cz - G::PAD, 0, G::NVZ - 1), vx = clampi(cx - G::PAD, 0, G::NVX - 1);
        const float v = clampv(vvis[vz * G::NVX + vx], g.vmin, g.vmax);
        SRC[tid] = 0.0f;
        SRC[4 + tid] = -(v * v * g.dt2) * P.geo.src_w[s * 4 + tid];
    }

    Place<typename RR::I> pI;
    Place<typename RR::X> pX;
    Place<typename RR::ZT> pZT;
    Place<typename RR::ZB> pZB;
    Place<typename RR::CT> pCT;
    Place<typename RR::CB> pCB;
    pI.init(tid); pX.init(tid); pZT.init(tid); pZB.init(tid); pCT.init(tid); pCB.init(tid);

    // Source corners (wgsl:245-249): the owning (rectangle, slot) of each.
    SrcOwnF so[6];
#pragma unroll
    for (int r = 0; r < 6; ++r) so[r] = SrcOwnF{-1, 0, 0.0f};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int cell = P.geo.src_cell[s * 4 + c];
        const float w = P.geo.src_w[s * 4 + c];
        const int cz = cell / G::NX, cx = cell - cz * G::NX;
        if (cx >= 2 && cx < G::NX - 2 && cz >= 2 && cz < G::NZ - 2) {
            int otid, oslot;
            const int rect = owner_of<G, G::REGA - 2>(cx - 2, cz - 2, otid, oslot);
            if (otid == tid) {
                const int vz = clampi(cz - G::PAD, 0, G::NVZ - 1), vx = clampi(cx - G::PAD, 0, G::NVX - 1);
                const float v = clampv(vvis[vz * G::NVX + vx], g.vmin, g.vmax);
                const SrcOwnF o{oslot, c, -2.0f * v * g.dt2 * w};
#pragma unroll
                for (int r = 0; r < 6; ++r)
                    if (r == rect) so[r] = o;
            }
        }
    }

    ASt<RR::I::K, false, false> sI;
    ASt<RR::X::K, true, false> sX;
    ASt<RR::ZT::K, false, true> sZT;
    ASt<RR::ZB::K, false, true> sZB;
    ASt<RR::CT::K, true, true> sCT;
    ASt<RR::CB::K, true, true> sCB;
    a_init<G>(sI, pI, vvis, g);
    a_init<G>(sX, pX, vvis, g);
    a_init<G>(sZT, pZT, vvis, g);
    a_init<G>(sZB, pZB, vvis, g);
    a_init<G>(sCT, pCT, vvis, g);
    a_init<G>(sCB, pCB, vvis, g);
    if (st_in) {
        float* sr = const_cast<float*>(st_in);
        int j = 0;
        a_xfer<false>(sI, sr, j, tid);
        a_xfer<false>(sX, sr, j, tid);
        a_xfer<false>(sZT, sr, j, tid);
        a_xfer<false>(sZB, sr, j, tid);
        a_xfer<false>(sCT, sr, j, tid);
        a_xfer<false>(sCB, sr, j, tid);
    }
    __syncthreads();
    ColX cx;
    {
        const int x = pX.xi + 2;
        cx.a = PR[Prof<G>::AX + x];
        cx.b = PR[Prof<G>::BX + x];
        cx.db = PR[Prof<G>::DBX + x];
#pragma unroll
        for (int d = 0; d < 5; ++d) cx.opb[d] = PR[Prof<G>::OPBX + x - 2 + d];
    }

    int rs[6];  // slot holding the receiver rows, per rectangle
    rs[0] = pI.slot_in_rows(C.bz0, C.bz1);
    rs[1] = pX.slot_in_rows(C.bz0, C.bz1);
    rs[2] = pZT.slot_in_rows(C.bz0, C.bz1);
    rs[3] = pZB.slot_in_rows(C.bz0, C.bz1);
    rs[4] = pCT.slot_in_rows(C.bz0, C.bz1);
    rs[5] = pCB.slot_in_rows(C.bz0, C.bz1);
    // The residual row and the wavelet sample of each step are loaded one
    // step ahead (at most NT receivers: one per thread), so their latency
    // overlaps a whole step instead of stalling its start.
    const bool has_rec = G::TOP && tid < g.nrec;  // the receiver rows lie in the top band
    float res_next = has_rec ? resid[(size_t)(t_hi - 1) * g.nrec + tid] : 0.0f;
    float f_next = wav[t_hi - 1];
    bool failed = false;
    for (int t = t_hi - 1; t >= t_lo; --t) {
        if (has_rec) RES[tid] = res_next;
        const float f_t = f_next;
        if (t > t_lo) {
            if (has_rec) res_next = resid[(size_t)(t - 1) * g.nrec + tid];
            f_next = wav[t - 1];
        }
        a_phase1<G, typename RR::I, DWAV>(sI, pI, C, cx, so[0]);
        a_phase1<G, typename RR::X, DWAV>(sX, pX, C, cx, so[1]);
        a_phase1<G, typename RR::ZT, DWAV>(sZT, pZT, C, cx, so[2]);
        a_phase1<G, typename RR::ZB, DWAV>(sZB, pZB, C, cx, so[3]);
        a_phase1<G, typename RR::CT, DWAV>(sCT, pCT, C, cx, so[4]);
        a_phase1<G, typename RR::CB, DWAV>(sCB, pCB, C, cx, so[5]);
        __syncthreads();
        if constexpr (!G::WHOLE) {
            if (!band_exchange(*X, C.V, G::band(), G::NX, blk, bp, g.nt - t + tag0, fail_s)) {
                failed = true;
                break;
            }
        }
        if (DWAV && tid == 0) {
            float acc = 0.0f;
#pragma unroll
            for (int c = 0; c < 4; ++c) acc += SRC[4 + c] * SRC[c];
            P.dwav[(size_t)unit * g.nt + t] = acc;
        }
        const float* snap_t = snaps + (size_t)t * G::INNER;
        a_phase2<G>(sI, pI, C, cx, so[0], rs[0], snap_t, f_t);
        a_phase2<G>(sX, pX, C, cx, so[1], rs[1], snap_t, f_t);
        a_phase2<G>(sZT, pZT, C, cx, so[2], rs[2], snap_t, f_t);
        a_phase2<G>(sZB, pZB, C, cx, so[3], rs[3], snap_t, f_t);
        a_phase2<G>(sCT, pCT, C, cx, so[4], rs[4], snap_t, f_t);
        a_phase2<G>(sCB, pCB, C, cx, so[5], rs[5], snap_t, f_t);
        __syncthreads();
    }
    if (st_out) {
        int j = 0;
        a_xfer<true>(sI, st_out, j, tid);
        a_xfer<true>(sX, st_out, j, tid);
        a_xfer<true>(sZT, st_out, j, tid);
        a_xfer<true>(sZB, st_out, j, tid);
        a_xfer<true>(sCT, st_out, j, tid);
        a_xfer<true>(sCB, st_out, j, tid);
    }
    if (t_lo > 0) return;
    float* gout = P.grad + (size_t)unit * G::NX * G::NZ;
    // the padded rows this workgroup owns: its inner rows, and the top or
    // bottom ring rows of the grid when it holds that side
    constexpr int OWN0 = G::TOP ? 0 : G::LO + 2, OWN1 = G::BOT ? G::NZ : G::HI + 2;
    if (!G::WHOLE && failed) {  // a neighbour never arrived: the gradient is not valid
        for (int j = OWN0 * G::NX + tid; j < OWN1 * G::NX; j += NT) gout[j] = __builtin_nanf("");
        return;
    }
    for (int j = OWN0 * G::NX + tid; j < OWN1 * G::NX; j += NT) {
        const int z = j / G::NX, x = j - z * G::NX;
        if (x < 2 || x >= G::NX - 2 || z < 2 || z >= G::NZ - 2) gout[j] = 0.0f;
    }
    a_store<G>(sI, pI, gout);
    a_store<G>(sX, pX, gout);
    a_store<G>(sZT, pZT, gout);
    a_store<G>(sZB, pZB, gout);
    a_store<G>(sCT, pCT, gout);
    a_store<G>(sCB, pCB, gout);
}

}  // namespace fx
}  // namespace wg
