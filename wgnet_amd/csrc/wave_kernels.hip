// gfx950 kernels of the wave engine: the 4th-order C-PML acoustic forward and
// its exact discrete adjoint, for grids whose state fits one workgroup.
//
// Resident design
//   One workgroup of NT = 512 threads (8 waves) owns one (model, shot) unit for
//   all nt steps.  The inner cells are split into four classes by the C-PML
//   region of the kernel (forward: pml + 4, adjoint: pml + 6, wave_wgpu.rs:291,
//   295): interior (I), x-strip only (X), z-strip only (Z) and corner (C).
//   Thread tid owns slot k of class c: the cell q_c = tid + k*NT of that class
//   in row-major order.  Every class has its own compile-time slot count and
//   branch-free code, so no wave executes a C-PML branch it does not need and
//   interior cells carry no C-PML registers.
//   VGPRs: previous-step field, C-PML memory variables, velocity, gradient.
//   LDS:   the current field and the strips of per-cell C-PML quantities that
//          neighbours read; C-PML profiles; receiver tables.
//   HBM:   gathers, and snapshots 2 v dt^2 w in inner row-major order.
//
// Row bands
//   A unit too large for one workgroup (the 100 x 100 and 141 x 141 models:
//   144 x 144 and 185 x 185 padded) is cut into P row bands, one workgroup
//   each, all P running at once.  A band holds its own rows and two halo rows
//   above and below in LDS; its cells' state stays in its registers.  The
//   band boundaries lie in the interior rows, away from the z C-PML strips,
//   so only the field (forward u, adjoint V = v^2 dt^2 lambda) crosses a
//   boundary: after each step a band sends its two edge rows to its
//   neighbours through global memory and takes theirs (see Xchg).  The
//   arithmetic per cell is the whole-grid kernel's, so results are the same
//   bits.  Bands run from a device work queue, one workgroup per compute
//   unit, taking a unit's P bands in a row, so the bands that wait for each
//   other are always running together.
//
// Arithmetic follows the reference operation by operation with no FMA
// contraction (-ffp-contract=off and the pragma below), so the float32 gathers
// equal the CPU solver's bit for bit:
//   forward  crates/geology-core/src/lib.rs:145-220 (== splatForward.wgsl:48-121)
//   adjoint  crates/wave-gauss-rs/src/shaders/splatAdjoint.wgsl:58-131
#include <algorithm>

#include "band_xchg.h"
#include "wave_kernels.h"

#pragma clang fp contract(off)

namespace wg {

constexpr int NT = 512;

// lib.rs:139-143 (forward) and splatAdjoint.wgsl:163-167 (adjoint), float32.
constexpr float C1 = 0.6666666666666666f;
constexpr float C2 = 0.08333333333333333f;
constexpr float D0 = -2.5f;
constexpr float D1 = 1.3333333333333333f;
constexpr float D2 = 0.08333333333333333f;
constexpr float D1A = 0.6666666666666666f;
constexpr float D1B = -0.08333333333333333f;
constexpr float D2A = -2.5f;
constexpr float D2B = 1.3333333333333333f;
constexpr float D2C = -0.08333333333333333f;

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Rust's f32::clamp keeps NaN (wave_wgpu.rs:730).
__device__ __forceinline__ float clampv(float v, float lo, float hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

// Edge-replicating pad of the visible model (wave_wgpu.rs:714-734).
__device__ __forceinline__ float vis_at(const float* vvis, const Grid& g, int x, int z) {
    const int vz = clampi(z - g.pad, 0, g.nvz - 1);
    const int vx = clampi(x - g.pad, 0, g.nvx - 1);
    return vvis[vz * g.nvx + vx];
}

// Opaque copy: keeps the compiler from hoisting per-slot coordinates and
// predicates out of the time loop (one register per slot each); they are
// re-derived from the class start every step instead.
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

// ---------------------------------------------------------------------------
// Cell classes.  In inner coordinates (xi, zi in [0, ni) x [0, nzi)) the C-PML
// band of a region R is r = R - 2 cells wide on each side.  A class is a set
// of columns times a set of rows, each either the band (two runs, or all
// columns when 2r >= n) or the interior run.  Column col of the class maps to
// xi = col + (col < sx ? ox0 : ox1); rows likewise.
// ---------------------------------------------------------------------------
struct Cls {
    int w, h;
    int sx, ox0, ox1;
    int sz, oz0, oz1;
    int dX, dZ;  // NT = dZ * w + dX
};

__host__ __device__ inline void axis_set(int n, int r, bool band, int& len, int& split, int& o0, int& o1) {
    if (band) {
        if (2 * r >= n) {
            len = n; split = n; o0 = 0; o1 = 0;
        } else {
            len = 2 * r; split = r; o0 = 0; o1 = n - 2 * r;
        }
    } else {
        len = 2 * r >= n ? 0 : n - 2 * r;
        split = len; o0 = r; o1 = r;
    }
}

// The rows of a class restricted to the inner rows [lo, hi) of a band: the
// band rows ([0, r) and [n - r, n)) or the interior run, each cut to [lo, hi).
// For lo = 0, hi = n it is axis_set.
__host__ __device__ inline void rows_set(int n, int r, bool band, int lo, int hi, int& len, int& split, int& o0,
                                         int& o1) {
    if (band && 2 * r >= n) {
        const int a = lo > 0 ? lo : 0, b = hi < n ? hi : n;
        len = b > a ? b - a : 0; split = len; o0 = a; o1 = a;
    } else if (band) {
        const int a0 = lo > 0 ? lo : 0, a1 = hi < r ? hi : r;
        const int b0 = lo > n - r ? lo : n - r, b1 = hi < n ? hi : n;
        const int la = a1 > a0 ? a1 - a0 : 0, lb = b1 > b0 ? b1 - b0 : 0;
        len = la + lb; split = la; o0 = a0; o1 = b0 - la;
    } else {
        const int a = lo > r ? lo : r, b = hi < n - r ? hi : n - r;
        len = (2 * r >= n || b <= a) ? 0 : b - a; split = len; o0 = a; o1 = a;
    }
}

__host__ __device__ inline Cls make_cls(const Grid& g, int r, bool bx, bool bz, int lo = 0, int hi = -1) {
    Cls c;
    axis_set(g.ni, r, bx, c.w, c.sx, c.ox0, c.ox1);
    rows_set(g.nzi, r, bz, lo, hi < 0 ? g.nzi : hi, c.h, c.sz, c.oz0, c.oz1);
    if (c.w == 0) { c.w = 1; c.h = 0; }
    c.dZ = NT / c.w;
    c.dX = NT - c.dZ * c.w;
    return c;
}

__host__ __device__ inline int cls_count(const Cls& c) { return c.w * c.h; }

// Class and in-class index of inner cell (xi, zi) for region band r, among
// the cells of the inner rows [lo, hi) (the cell's own band).
__host__ __device__ inline void cell_class(const Grid& g, int r, int xi, int zi, int& cls, int& q, int lo = 0,
                                           int hi = -1) {
    const bool bx = xi < r || xi >= g.ni - r;
    const bool bz = zi < r || zi >= g.nzi - r;
    cls = (bx ? 1 : 0) + (bz ? 2 : 0);  // 0 I, 1 X, 2 Z, 3 C
    const Cls c = make_cls(g, r, bx, bz, lo, hi);
    const int col = xi - (xi < c.sx + c.ox0 ? c.ox0 : c.ox1);
    const int row = zi - (zi < c.sz + c.oz0 ? c.oz0 : c.oz1);
    q = row * c.w + col;
}

__device__ __forceinline__ void cls_step(const Cls& c, int& col, int& row) {
    col += c.dX;
    row += c.dZ;
    if (col >= c.w) {
        col -= c.w;
        row += 1;
    }
}

__device__ __forceinline__ int cls_xi(const Cls& c, int col) { return col + (col < c.sx ? c.ox0 : c.ox1); }
__device__ __forceinline__ int cls_zi(const Cls& c, int row) { return row + (row < c.sz ? c.oz0 : c.oz1); }

// ---------------------------------------------------------------------------
// LDS layouts (floats).
//   field: the band's padded rows [zb0, zb0 + rows), nx per row.
//   x-strips: the band's inner rows, columns [0, sw) and [nx-sw, nx): 2*sw per row.
//   z-strips: rows [0, sw) (top side) and [nz-sw, nz) (bottom side), columns
//   [2, nx-2): ni per row; a band holds the sides it reaches.
// A cell lying in both halves (grids narrower than 2*sw) is written to both.
// ---------------------------------------------------------------------------
struct FwdLayout {
    int U, AX, AZ, PR, total;
};
struct AdjLayout {
    int V, TX, UX, TZ, UZ, PR, RES, SRC, BOFF, ER, EW, total;
};

__host__ __device__ inline FwdLayout fwd_layout(const Grid& g, const Band& bd) {
    FwdLayout L;
    L.U = 0;
    L.AX = L.U + g.nx * bd.rows;
    L.AZ = L.AX + 2 * g.swf * (bd.hi - bd.lo);
    L.PR = L.AZ + (bd.top + bd.bot) * g.swf * g.ni;
    L.total = L.PR + 3 * (g.nx + g.nz);
    return L;
}

__host__ __device__ inline AdjLayout adj_layout(const Grid& g, const Geometry& geo, const Band& bd) {
    AdjLayout L;
    L.V = 0;
    L.TX = L.V + g.nx * bd.rows;
    L.UX = L.TX + 2 * g.swa * (bd.hi - bd.lo);
    L.TZ = L.UX + 2 * g.swa * (bd.hi - bd.lo);
    L.UZ = L.TZ + (bd.top + bd.bot) * g.swa * g.ni;
    L.PR = L.UZ + (bd.top + bd.bot) * g.swa * g.ni;
    L.RES = L.PR + 3 * (g.nx + g.nz);
    L.SRC = L.RES + g.nrec;  // 4 lambdas + 4 coefficients
    L.BOFF = L.SRC + 8;
    L.ER = L.BOFF + geo.band_rows * g.nx + 1;
    L.EW = L.ER + geo.n_ent;
    L.total = L.EW + geo.n_ent;
    return L;
}

size_t forward_lds_bytes(const Grid& g) { return 4ull * fwd_layout(g, whole_band(g)).total; }
size_t adjoint_lds_bytes(const Grid& g, const Geometry& geo) {
    return 4ull * adj_layout(g, geo, whole_band(g)).total;
}

// Slot counts per class of a compiled variant.
struct Tuple {
    int kI, kX, kZ, kC;
};
static const Tuple kTuples[] = {{1, 1, 1, 1}, {2, 2, 2, 2}, {4, 3, 3, 3}, {9, 6, 6, 5}};
constexpr int kNumTuples = 4;

static bool tuple_fits(const Grid& g, int r, const Tuple& t) {
    const int need[4] = {cls_count(make_cls(g, r, false, false)), cls_count(make_cls(g, r, true, false)),
                         cls_count(make_cls(g, r, false, true)), cls_count(make_cls(g, r, true, true))};
    const int have[4] = {t.kI, t.kX, t.kZ, t.kC};
    for (int c = 0; c < 4; ++c)
        if (need[c] > have[c] * NT) return false;
    return true;
}

static int pick_tuple(const Grid& g, int r) {
    for (int i = 0; i < kNumTuples; ++i)
        if (tuple_fits(g, r, kTuples[i])) return i;
    return -1;
}

int slots_needed(const Grid& g) {
    const int a = pick_tuple(g, g.region - 2), b = pick_tuple(g, g.region_adj - 2);
    return (a < 0 || b < 0) ? -1 : (a > b ? a : b);
}

// Slot counts of the band kernels, the largest that compile without spills:
// the 100 x 100 model (144 x 144 padded) in two bands, the 141 x 141 model
// (185 x 185) in three.  The forward and the adjoint pick their own (the
// adjoint's C-PML region is two cells wider, and its slots cost more
// registers).
static const Tuple kBandTuples[] = {{9, 5, 5, 3}, {12, 5, 7, 3}, {13, 4, 6, 2}};
constexpr int kNumBandTuples = 3;
static const bool kBandTupleFwd[] = {true, true, true};
static const bool kBandTupleAdj[] = {true, true, false};

// The first band tuple that holds every band's classes of region band r
// (forward: region - 2, adjoint: region_adj - 2), or -1.
static int pick_band_tuple(const Grid& g, int P, const int* bounds, bool adjoint) {
    const int r = adjoint ? g.region_adj - 2 : g.region - 2;
    for (int i = 0; i < kNumBandTuples; ++i) {
        if (!(adjoint ? kBandTupleAdj[i] : kBandTupleFwd[i])) continue;
        const Tuple& t = kBandTuples[i];
        const int have[4] = {t.kI, t.kX, t.kZ, t.kC};
        bool ok = true;
        for (int p = 0; p < P && ok; ++p)
            for (int c = 0; c < 4 && ok; ++c)
                ok = cls_count(make_cls(g, r, c & 1, c & 2, bounds[p], bounds[p + 1])) <= have[c] * NT;
        if (ok) return i;
    }
    return -1;
}

// Boundaries for P bands whose P - 2 middle bands have m rows each and the
// top and bottom bands the rest, evenly.
static void band_bounds(int nzi, int P, int m, int* bounds) {
    const int edge = nzi - (P - 2) * m, top = edge / 2;
    bounds[0] = 0;
    bounds[1] = top;
    for (int p = 2; p < P; ++p) bounds[p] = bounds[p - 1] + m;
    bounds[P] = nzi;
}

// Work of one band per step in interior-cell equivalents: C-PML cells cost
// more than interior ones (about 1.6 times on one axis, 2.2 on both), and each
// exchanged side about 1,500 cells' worth.  A unit's bands wait for each
// other, so a split costs P times its most expensive band.
static double band_cost(const Grid& g, const int* bounds, int P) {
    const double wcls[4] = {1.0, 1.6, 1.6, 2.2};
    double worst = 0.0;
    for (int p = 0; p < P; ++p) {
        double c = 1500.0 * ((p > 0) + (p < P - 1));
        for (int k = 0; k < 4; ++k)
            c += wcls[k] * cls_count(make_cls(g, g.region - 2, k & 1, k & 2, bounds[p], bounds[p + 1]));
        worst = c > worst ? c : worst;
    }
    return worst * P;
}

bool band_plan(const Grid& g, const Geometry& geo, int lds_limit, int* Pout, int bounds[kMaxBands + 1]) {
    if (g.nrec > NT || g.ni >= 511) return false;
    // receiver corners and the residual CSR rows: inner rows of the top band
    int rec_lo = g.nzi, rec_hi = -1;
    if (geo.band_rows > 0) {
        rec_lo = geo.band_z0 - 2;
        rec_hi = geo.band_z0 + geo.band_rows - 1 - 2;
    }
    // every split that fits (LDS, registers, boundaries clear of the z-strips),
    // the cheapest kept
    double best = -1.0;
    int b[kMaxBands + 1];
    for (int P = 2; P <= kMaxBands; ++P) {
        for (int m = (g.nzi + P - 1) / P; m >= 4; --m) {  // middle bands from the even split down
            band_bounds(g.nzi, P, m, b);
            bool ok = true;
            for (int p = 1; p < P && ok; ++p)  // boundaries clear of the z-strips, bands of >= 4 rows
                ok = b[p] >= g.swa && b[p] <= g.nzi - g.swa && b[p] - b[p - 1] >= 4;
            ok = ok && b[P] - b[P - 1] >= 4 && rec_lo >= 0 && rec_hi < b[1];
            for (int p = 0; p < P && ok; ++p) {
                const Band bd = split_band(g, b, P, p);
                const size_t f = 4ull * fwd_layout(g, bd).total, a = 4ull * adj_layout(g, geo, bd).total;
                ok = f + 64 <= (size_t)lds_limit && a + 64 <= (size_t)lds_limit;
            }
            if (ok && pick_band_tuple(g, P, b, false) >= 0 && pick_band_tuple(g, P, b, true) >= 0) {
                const double c = band_cost(g, b, P);
                if (best < 0.0 || c < best) {
                    best = c;
                    *Pout = P;
                    for (int i = 0; i <= P; ++i) bounds[i] = b[i];
                }
            }
            if (P == 2) break;  // two bands have no middle band to shrink
        }
    }
    return best >= 0.0;
}

size_t band_xbuf_words(const Grid& g, int units, int P) { return (size_t)units * P * 2 * 2 * 2 * g.nx; }

// ===========================================================================
// Forward
// ===========================================================================
struct FCtx {
    float* U;
    float* AX;
    float* AZ;
    const float *pax, *pbx, *pdbx, *paz, *pbz, *pdbz;
    int nx, nz, ni, reg, swf, sw2;
    int zb0, lo, azb;  // band: first padded row in LDS, first inner row, bottom z-strip offset
    float rh, rh2, dt2;
};

template <int K, bool PX, bool PZ, bool BORN>
struct FState {
    float uprev[K], unext[K], v[K];
    float px[PX ? K : 1], zx[PX ? K : 1];
    float pz[PZ ? K : 1], zz[PZ ? K : 1];
    float dv[BORN ? K : 1];
};

template <int K, bool PX, bool PZ, bool BORN>
__device__ __forceinline__ void fwd_init(FState<K, PX, PZ, BORN>& S, const Cls& c, int col0, int row0,
                                         const Grid& g, const float* vvis, const float* dvis) {
    int col = col0, row = row0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (k) cls_step(c, col, row);
        S.uprev[k] = 0.0f;
        S.unext[k] = 0.0f;
        S.v[k] = 0.0f;
        if constexpr (PX) { S.px[k] = 0.0f; S.zx[k] = 0.0f; }
        if constexpr (PZ) { S.pz[k] = 0.0f; S.zz[k] = 0.0f; }
        if constexpr (BORN) S.dv[k] = 0.0f;
        if (row < c.h) {
            const int x = cls_xi(c, col) + 2, z = cls_zi(c, row) + 2;
            const float v = vis_at(vvis, g, x, z);
            S.v[k] = clampv(v, g.vmin, g.vmax);
            if constexpr (BORN) {
                const float m = (v > g.vmin && v < g.vmax) ? 1.0f : 0.0f;
                S.dv[k] = vis_at(dvis, g, x, z) * m;
            }
        }
    }
}

// One step of lib.rs:151-206 for the slots of one class; new values are kept
// in registers until every thread has read u^t.
template <int K, bool PX, bool PZ, bool STORE, bool BORN>
__device__ __forceinline__ void fwd_compute(FState<K, PX, PZ, BORN>& S, const Cls& c, int col0, int row0,
                                            const FCtx& C, float* snap_t, const float* bsnap_t) {
    int col = opaque(col0), row = opaque(row0);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (k) cls_step(c, col, row);
        if (row >= c.h) continue;
        const int xi = cls_xi(c, col), zi = cls_zi(c, row);
        const int x = xi + 2, z = zi + 2;
        const int i = (z - C.zb0) * C.nx + x;
        const float* U = C.U;
        const float u = U[i];
        float w = 0.0f;

        const float um1x = U[i - 1], up1x = U[i + 1];
        const float um2x = U[i - 2], up2x = U[i + 2];
        const float d2x = (D0 * u + D1 * (up1x + um1x) - D2 * (up2x + um2x)) * C.rh2;
        if constexpr (PX) {
            const float d1x = (C1 * (up1x - um1x) - C2 * (up2x - um2x)) * C.rh;
            const float a = C.pax[x], bb = C.pbx[x];
            const int base = (zi - C.lo) * C.sw2 + (x < C.reg ? x : x - C.nx + C.sw2);
            const float ap1 = C.AX[base + 1], am1 = C.AX[base - 1];
            const float ap2 = C.AX[base + 2], am2 = C.AX[base - 2];
            const float d1_apsi = (C1 * (ap1 - am1) - C2 * (ap2 - am2)) * C.rh;
            const float tmp = (1.0f + bb) * d2x + C.pdbx[x] * d1x + d1_apsi;
            w += (1.0f + bb) * tmp + a * S.zx[k];
            S.px[k] = bb * d1x + a * S.px[k];
            S.zx[k] = bb * tmp + a * S.zx[k];
        } else {
            w += d2x;
        }

        const int nx = C.nx;
        const float um1z = U[i - nx], up1z = U[i + nx];
        const float um2z = U[i - 2 * nx], up2z = U[i + 2 * nx];
        const float d2z = (D0 * u + D1 * (up1z + um1z) - D2 * (up2z + um2z)) * C.rh2;
        if constexpr (PZ) {
            const float d1z = (C1 * (up1z - um1z) - C2 * (up2z - um2z)) * C.rh;
            const float a = C.paz[z], bb = C.pbz[z];
            const int ni = C.ni;
            const int base = (z < C.reg ? z : z - C.nz + C.swf + C.azb) * ni + xi;
            const float ap1 = C.AZ[base + ni], am1 = C.AZ[base - ni];
            const float ap2 = C.AZ[base + 2 * ni], am2 = C.AZ[base - 2 * ni];
            const float d1_apsi = (C1 * (ap1 - am1) - C2 * (ap2 - am2)) * C.rh;
            const float tmp = (1.0f + bb) * d2z + C.pdbz[z] * d1z + d1_apsi;
            w += (1.0f + bb) * tmp + a * S.zz[k];
            S.pz[k] = bb * d1z + a * S.pz[k];
            S.zz[k] = bb * tmp + a * S.zz[k];
        } else {
            w += d2z;
        }

        const float v = S.v[k];
        const int q = zi * C.ni + xi;
        if constexpr (STORE) snap_t[q] = 2.0f * v * C.dt2 * w;  // splatForward.wgsl:108-110
        float next = v * v * C.dt2 * w + 2.0f * u - S.uprev[k];
        if constexpr (BORN) next += S.dv[k] * bsnap_t[q];
        S.unext[k] = next;
        S.uprev[k] = u;
    }
}

template <int K, bool PX, bool PZ, bool BORN>
__device__ __forceinline__ void fwd_write(FState<K, PX, PZ, BORN>& S, const Cls& c, int col0, int row0,
                                          const FCtx& C) {
    int col = opaque(col0), row = opaque(row0);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (k) cls_step(c, col, row);
        if (row >= c.h) continue;
        const int xi = cls_xi(c, col), zi = cls_zi(c, row);
        const int x = xi + 2, z = zi + 2;
        C.U[(z - C.zb0) * C.nx + x] = S.unext[k];
        // a * psi_next, the product neighbours take in lib.rs:167-170, 190-193.
        if constexpr (PX) {
            const float prod = C.pax[x] * S.px[k];
            const int rb = (zi - C.lo) * C.sw2;
            if (x < C.swf) C.AX[rb + x] = prod;
            if (x >= C.nx - C.swf) C.AX[rb + x - C.nx + C.sw2] = prod;
        }
        if constexpr (PZ) {
            const float prod = C.paz[z] * S.pz[k];
            if (z < C.swf) C.AZ[z * C.ni + xi] = prod;
            if (z >= C.nz - C.swf) C.AZ[(z - C.nz + C.swf + C.azb) * C.ni + xi] = prod;
        }
    }
}

// The forward of one unit (all nt steps), or of one band of it (BAND: the
// inner rows of `bd`, exchanging edge rows with the neighbour bands after
// each step).  pos is the unit's position in the launch (snapshot and
// exchange slots).  The band holding the receiver rows (the top one) records
// the gathers.
template <int KI, int KX, int KZ, int KC, bool STORE, bool BORN, bool BAND>
__device__ __forceinline__ void forward_unit(const FwdParams& P, float* lds, int unit, int pos, const Band& bd,
                                             const Xchg& X, int p, int* fail_s) {
    const Grid& g = P.g;
    const FwdLayout L = fwd_layout(g, bd);
    float* PR = lds + L.PR;
    FCtx C;
    C.U = lds + L.U;
    C.AX = lds + L.AX;
    C.AZ = lds + L.AZ;
    C.pax = PR;
    C.pbx = PR + g.nx;
    C.pdbx = PR + 2 * g.nx;
    C.paz = PR + 3 * g.nx;
    C.pbz = C.paz + g.nz;
    C.pdbz = C.pbz + g.nz;
    C.nx = g.nx;
    C.nz = g.nz;
    C.ni = g.ni;
    C.reg = g.region;
    C.swf = g.swf;
    C.sw2 = 2 * g.swf;
    C.zb0 = bd.zb0;
    C.lo = bd.lo;
    C.azb = bd.top ? g.swf : 0;
    C.rh = g.rh;
    C.rh2 = g.rh2;
    C.dt2 = g.dt2;

    const int tid = threadIdx.x;
    const int b = unit / g.shots;
    const int s = src_slot(P.geo, unit, unit - b * g.shots);
    const bool rec = bd.top;  // the receiver rows (the host checks) lie in the top band
    if (!unit_active(P.geo, unit)) {  // padding unit of an uneven batch
        if (rec) zero_unit_outputs(P.traces + (size_t)unit * g.nt * g.nrec, g.nt * g.nrec, nullptr, 0);
        return;
    }

    for (int j = tid; j < L.PR; j += NT) lds[j] = 0.0f;
    for (int j = tid; j < 3 * (g.nx + g.nz); j += NT) PR[j] = P.geo.prof[j];

    const float* vvis = P.vel + (size_t)b * g.nvx * g.nvz;
    const float* dvis = BORN ? P.dvel + (size_t)b * g.nvx * g.nvz : nullptr;
    const float* wav = P.wavelet + (size_t)unit * P.wav_stride;
    float* traces = P.traces + (size_t)unit * g.nt * g.nrec;
    float* snaps = STORE ? P.snaps + (size_t)pos * g.nt * g.inner : nullptr;
    const float* bsnaps = BORN ? P.born_snaps + (size_t)pos * g.nt * g.inner : nullptr;

    const int r = g.region - 2;
    const Cls cI = make_cls(g, r, false, false, bd.lo, bd.hi);
    const Cls cX = make_cls(g, r, true, false, bd.lo, bd.hi);
    const Cls cZ = make_cls(g, r, false, true, bd.lo, bd.hi);
    const Cls cC = make_cls(g, r, true, true, bd.lo, bd.hi);
    const int rI = tid / cI.w, kI0 = tid - rI * cI.w;
    const int rX = tid / cX.w, kX0 = tid - rX * cX.w;
    const int rZ = tid / cZ.w, kZ0 = tid - rZ * cZ.w;
    const int rC = tid / cC.w, kC0 = tid - rC * cC.w;

    FState<KI, false, false, BORN> SI;
    FState<KX, true, false, BORN> SX;
    FState<KZ, false, true, BORN> SZ;
    FState<KC, true, true, BORN> SC;
    fwd_init(SI, cI, kI0, rI, g, vvis, dvis);
    fwd_init(SX, cX, kX0, rX, g, vvis, dvis);
    fwd_init(SZ, cZ, kZ0, rZ, g, vvis, dvis);
    fwd_init(SC, cC, kC0, rC, g, vvis, dvis);

    // Source corners owned by this thread: the point source is added after the
    // update by the cell's owner, -(v^2 dt^2) w f[t] (lib.rs:209-215); Born
    // adds the derivative -(2 v dv dt^2) w f[t].  scell is the LDS index.
    int scell[4];
    float scoef[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        scell[c] = -1;
        scoef[c] = 0.0f;
        const int cell = P.geo.src_cell[s * 4 + c];
        const float w = P.geo.src_w[s * 4 + c];
        const int cz = cell / g.nx, cx = cell - cz * g.nx;
        if (w != 0.0f && cx >= 2 && cx < g.nx - 2 && cz - 2 >= bd.lo && cz - 2 < bd.hi) {
            int cls, q;
            cell_class(g, r, cx - 2, cz - 2, cls, q, bd.lo, bd.hi);
            if (q % NT == tid) {
                const float vraw = vis_at(vvis, g, cx, cz);
                const float v = clampv(vraw, g.vmin, g.vmax);
                scell[c] = (cz - bd.zb0) * g.nx + cx;
                if (BORN) {
                    const float m = (vraw > g.vmin && vraw < g.vmax) ? 1.0f : 0.0f;
                    scoef[c] = -(2.0f * v * (vis_at(dvis, g, cx, cz) * m) * g.dt2) * w;
                } else {
                    scoef[c] = -(v * v * g.dt2) * w;
                }
            }
        }
    }
    // Receiver r = tid: its 4 corners (the host guarantees nrec <= NT).
    int rc[4];
    float rw[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        rc[c] = 0;
        rw[c] = 0.0f;
        if (rec && tid < g.nrec) {
            rc[c] = P.geo.rec_cell[tid * 4 + c] - bd.zb0 * g.nx;
            rw[c] = P.geo.rec_w[tid * 4 + c];
        }
    }
    __syncthreads();

    bool failed = false;
    for (int t = 0; t < g.nt; ++t) {
        // Receivers record u^t before the update (lib.rs:146-149, wgsl:61-67).
        if (rec && tid < g.nrec) {
            float acc = 0.0f;
#pragma unroll
            for (int c = 0; c < 4; ++c) acc += rw[c] * C.U[rc[c]];
            traces[(size_t)t * g.nrec + tid] = acc;
        }
        float* snap_t = STORE ? snaps + (size_t)t * g.inner : nullptr;
        const float* bsnap_t = BORN ? bsnaps + (size_t)t * g.inner : nullptr;
        fwd_compute<KI, false, false, STORE, BORN>(SI, cI, kI0, rI, C, snap_t, bsnap_t);
        fwd_compute<KX, true, false, STORE, BORN>(SX, cX, kX0, rX, C, snap_t, bsnap_t);
        fwd_compute<KZ, false, true, STORE, BORN>(SZ, cZ, kZ0, rZ, C, snap_t, bsnap_t);
        fwd_compute<KC, true, true, STORE, BORN>(SC, cC, kC0, rC, C, snap_t, bsnap_t);
        __syncthreads();
        fwd_write(SI, cI, kI0, rI, C);
        fwd_write(SX, cX, kX0, rX, C);
        fwd_write(SZ, cZ, kZ0, rZ, C);
        fwd_write(SC, cC, kC0, rC, C);
        const float f_t = wav[t];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (scell[c] >= 0) C.U[scell[c]] += scoef[c] * f_t;
        __syncthreads();
        if constexpr (BAND) {
            if (!band_exchange(X, C.U, bd, g.nx, pos, p, t + 1, fail_s)) {
                failed = true;
                break;
            }
        }
    }
    if (BAND && failed && rec)  // a neighbour never arrived: the gathers are not valid
        for (int j = tid; j < g.nt * g.nrec; j += NT) traces[j] = __builtin_nanf("");
}

template <int KI, int KX, int KZ, int KC, bool STORE, bool BORN>
__global__ __launch_bounds__(NT) void k_forward(FwdParams P) {
    extern __shared__ float lds[];
    const Xchg X{};
    forward_unit<KI, KX, KZ, KC, STORE, BORN, false>(P, lds, P.unit0 + blockIdx.x, blockIdx.x, whole_band(P.g), X,
                                                     0, nullptr);
}

// Band launch: `units` units of P bands each, one workgroup per band.
template <int KI, int KX, int KZ, int KC, bool STORE>
__global__ __launch_bounds__(NT) void k_forward_band(FwdParams P, BandLaunch B, int units) {
    extern __shared__ float lds[];
    __shared__ int task_s, fail_s;
    const Xchg X{B.xbuf, B.err, B.P, B.spin};
    const int task = __builtin_amdgcn_readfirstlane(band_task(B.next, &task_s));
    if (task >= units * B.P) return;
    const int u = task / B.P, p = task - u * B.P;
    forward_unit<KI, KX, KZ, KC, STORE, false, true>(P, lds, P.unit0 + u, u, split_band(P.g, B.bounds, B.P, p), X, p,
                                                     &fail_s);
}

// ===========================================================================
// Adjoint: splatAdjoint.wgsl:58-131 for one unit, t = nt-1 .. 0.
// Phase 1 publishes, for every owned cell j, V(j) = v^2 dt^2 lambda_t(j) and,
// on the strips, T(j) = (1+b)V + bZ and U1(j) = db T + bP: per-cell values the
// wgsl recomputes at each of five neighbours (wgsl:212-220).  Outside the
// region T = V and U1 = 0, which owners of the two columns (rows) beyond the
// region publish as V (U1 stays zero).  Phase 2 forms lambda_{t-1}, the new
// P/Z and the gradient from neighbour values.
// ===========================================================================
struct ACtx {
    float* V;
    float *TX, *UX, *TZ, *UZ;
    const float *pax, *pbx, *pdbx, *paz, *pbz, *pdbz;
    const float* RES;
    float* SRC;
    const int* BOFF;
    const int* ER;
    const float* EW;
    int nx, nz, ni, ra, swa, sw2, bz0, bz1;
    int zb0, lo, azb;  // band: first padded row in LDS, first inner row, bottom z-strip offset
    float rh, dt2;
};

template <int K, bool PX, bool PZ>
struct AState {
    float lam[K], lamp[K], gr[K], v2[K];
    float Px[PX ? K : 1], Zx[PX ? K : 1];
    float Pz[PZ ? K : 1], Zz[PZ ? K : 1];
};

// Source corner owned within one class: slot, corner index and -2 v dt^2 w.
struct SrcOwn {
    int slot, corner;
    float coef;
};

template <int K, bool PX, bool PZ>
__device__ __forceinline__ void adj_init(AState<K, PX, PZ>& S, const Cls& c, int col0, int row0, const Grid& g,
                                         const float* vvis) {
    int col = col0, row = row0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (k) cls_step(c, col, row);
        S.lam[k] = S.lamp[k] = S.gr[k] = S.v2[k] = 0.0f;
        if constexpr (PX) { S.Px[k] = 0.0f; S.Zx[k] = 0.0f; }
        if constexpr (PZ) { S.Pz[k] = 0.0f; S.Zz[k] = 0.0f; }
        if (row < c.h) {
            const float v = clampv(vis_at(vvis, g, cls_xi(c, col) + 2, cls_zi(c, row) + 2), g.vmin, g.vmax);
            S.v2[k] = v * v * g.dt2;  // the factor of vw(), wgsl:53-56
        }
    }
}

template <int K, bool PX, bool PZ>
__device__ __forceinline__ void adj_phase1(AState<K, PX, PZ>& S, const Cls& c, int col0, int row0, const ACtx& C,
                                           const SrcOwn& so) {
    int col = opaque(col0), row = opaque(row0);
    const int sk = opaque(so.slot);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (k) cls_step(c, col, row);
        if (row >= c.h) continue;
        const int xi = cls_xi(c, col), zi = cls_zi(c, row);
        const int x = xi + 2, z = zi + 2;
        const float Vc = S.v2[k] * S.lam[k];
        C.V[(z - C.zb0) * C.nx + x] = Vc;
        const int rb = (zi - C.lo) * C.sw2;
        if constexpr (PX) {
            const float bb = C.pbx[x];
            const float T = (1.0f + bb) * Vc + bb * S.Zx[k];
            const float U1 = C.pdbx[x] * T + bb * S.Px[k];
            if (x < C.swa) {
                C.TX[rb + x] = T;
                C.UX[rb + x] = U1;
            }
            if (x >= C.nx - C.swa) {
                C.TX[rb + x - C.nx + C.sw2] = T;
                C.UX[rb + x - C.nx + C.sw2] = U1;
            }
        } else {
            if (x < C.swa) C.TX[rb + x] = Vc;
            if (x >= C.nx - C.swa) C.TX[rb + x - C.nx + C.sw2] = Vc;
        }
        const int zb = z - C.nz + C.swa + C.azb;  // row of the bottom z-strip side
        if constexpr (PZ) {
            const float bb = C.pbz[z];
            const float T = (1.0f + bb) * Vc + bb * S.Zz[k];
            const float U1 = C.pdbz[z] * T + bb * S.Pz[k];
            if (z < C.swa) {
                C.TZ[z * C.ni + xi] = T;
                C.UZ[z * C.ni + xi] = U1;
            }
            if (z >= C.nz - C.swa) {
                C.TZ[zb * C.ni + xi] = T;
                C.UZ[zb * C.ni + xi] = U1;
            }
        } else {
            if (z < C.swa) C.TZ[z * C.ni + xi] = Vc;
            if (z >= C.nz - C.swa) C.TZ[zb * C.ni + xi] = Vc;
        }
        if (k == sk) C.SRC[so.corner] = S.lam[k];
    }
}

template <int K, bool PX, bool PZ>
__device__ __forceinline__ void adj_phase2(AState<K, PX, PZ>& S, const Cls& c, int col0, int row0, const ACtx& C,
                                           const SrcOwn& so, const float* snap_t, float f_t) {
    int col = opaque(col0), row = opaque(row0);
    const int sk = opaque(so.slot);
    const float rh = C.rh;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (k) cls_step(c, col, row);
        if (row >= c.h) continue;
        const int xi = cls_xi(c, col), zi = cls_zi(c, row);
        const int x = xi + 2, z = zi + 2;
        const int nx = C.nx, ni = C.ni;
        const int i = (z - C.zb0) * nx + x;
        const float Sn = snap_t[zi * ni + xi];
        const float v0 = C.V[i];
        float w = 0.0f;
        if constexpr (PX) {
            const int base = (zi - C.lo) * C.sw2 + (x < C.ra ? x : x - nx + C.sw2);
            const float tm2 = C.TX[base - 2], tm1 = C.TX[base - 1], t00 = C.TX[base];
            const float tp1 = C.TX[base + 1], tp2 = C.TX[base + 2];
            const float um2 = C.UX[base - 2], um1 = C.UX[base - 1];
            const float up1 = C.UX[base + 1], up2 = C.UX[base + 2];
            const float w2m2 = (1.0f + C.pbx[x - 2]) * tm2;
            const float w2m1 = (1.0f + C.pbx[x - 1]) * tm1;
            const float w200 = (1.0f + C.pbx[x]) * t00;
            const float w2p1 = (1.0f + C.pbx[x + 1]) * tp1;
            const float w2p2 = (1.0f + C.pbx[x + 2]) * tp2;
            const float d1_u1 = (D1A * (up1 - um1) + D1B * (up2 - um2)) * rh;
            const float d2_u2 = (D2A * w200 + D2B * (w2p1 + w2m1) + D2C * (w2p2 + w2m2)) * rh * rh;
            const float d1_t = (D1A * (tp1 - tm1) + D1B * (tp2 - tm2)) * rh;
            w += -d1_u1 + d2_u2;
            const float a = C.pax[x];
            S.Px[k] = -a * d1_t + a * S.Px[k];
            S.Zx[k] = a * v0 + a * S.Zx[k];
        } else {
            const float vm1 = C.V[i - 1], vp1 = C.V[i + 1], vm2 = C.V[i - 2], vp2 = C.V[i + 2];
            w += (D2A * v0 + D2B * (vp1 + vm1) + D2C * (vp2 + vm2)) * rh * rh;
        }
        if constexpr (PZ) {
            const int base = (z < C.ra ? z : z - C.nz + C.swa + C.azb) * ni + xi;
            const float tm2 = C.TZ[base - 2 * ni], tm1 = C.TZ[base - ni], t00 = C.TZ[base];
            const float tp1 = C.TZ[base + ni], tp2 = C.TZ[base + 2 * ni];
            const float um2 = C.UZ[base - 2 * ni], um1 = C.UZ[base - ni];
            const float up1 = C.UZ[base + ni], up2 = C.UZ[base + 2 * ni];
            const float w2m2 = (1.0f + C.pbz[z - 2]) * tm2;
            const float w2m1 = (1.0f + C.pbz[z - 1]) * tm1;
            const float w200 = (1.0f + C.pbz[z]) * t00;
            const float w2p1 = (1.0f + C.pbz[z + 1]) * tp1;
            const float w2p2 = (1.0f + C.pbz[z + 2]) * tp2;
            const float d1_u1 = (D1A * (up1 - um1) + D1B * (up2 - um2)) * rh;
            const float d2_u2 = (D2A * w200 + D2B * (w2p1 + w2m1) + D2C * (w2p2 + w2m2)) * rh * rh;
            const float d1_t = (D1A * (tp1 - tm1) + D1B * (tp2 - tm2)) * rh;
            w += -d1_u1 + d2_u2;
            const float a = C.paz[z];
            S.Pz[k] = -a * d1_t + a * S.Pz[k];
            S.Zz[k] = a * v0 + a * S.Zz[k];
        } else {
            const float vm1 = C.V[i - nx], vp1 = C.V[i + nx], vm2 = C.V[i - 2 * nx], vp2 = C.V[i + 2 * nx];
            w += (D2A * v0 + D2B * (vp1 + vm1) + D2C * (vp2 + vm2)) * rh * rh;
        }
        float next = w + 2.0f * S.lam[k] - S.lamp[k];
        float gk = S.lam[k] * Sn;  // store_every = 1 (wgsl:234-236)
        if (z >= C.bz0 && z < C.bz1) {  // receiver residuals, CSR order (wgsl:239-242)
            const int o = (z - C.bz0) * nx + x;
            for (int e = C.BOFF[o]; e < C.BOFF[o + 1]; ++e) next += C.EW[e] * C.RES[C.ER[e]];
        }
        if (k == sk) gk += S.lam[k] * (so.coef * f_t);  // wgsl:245-249
        S.gr[k] += gk;
        S.lamp[k] = S.lam[k];
        S.lam[k] = next;
    }
}

template <int K, bool PX, bool PZ>
__device__ __forceinline__ void adj_store(AState<K, PX, PZ>& S, const Cls& c, int col0, int row0, int nx,
                                          float* gout) {
    int col = col0, row = row0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (k) cls_step(c, col, row);
        if (row >= c.h) continue;
        gout[(cls_zi(c, row) + 2) * nx + cls_xi(c, col) + 2] = S.gr[k];
    }
}

// The adjoint of one unit (t = nt-1 .. 0), or of one band of it (BAND: the
// inner rows of `bd`, exchanging the edge rows of V with the neighbour bands
// after phase 1 of every step).  pos is the unit's position in the launch.
// A band writes the padded gradient of the padded rows it owns (the top and
// bottom bands also the outer FD_PAD rows).
template <int KI, int KX, int KZ, int KC, bool DWAV, bool BAND>
__device__ __forceinline__ void adjoint_unit(const AdjParams& P, float* lds, int unit, int pos, const Band& bd,
                                             const Xchg& X, int p, int* fail_s) {
    const Grid& g = P.g;
    const AdjLayout L = adj_layout(g, P.geo, bd);
    float* PR = lds + L.PR;
    float* RES = lds + L.RES;
    float* SRC = lds + L.SRC;  // [0..4) lambda at corners, [4..8) -v^2 dt^2 w
    int* BOFF = reinterpret_cast<int*>(lds + L.BOFF);
    int* ER = reinterpret_cast<int*>(lds + L.ER);
    float* EW = lds + L.EW;
    ACtx C;
    C.V = lds + L.V;
    C.TX = lds + L.TX;
    C.UX = lds + L.UX;
    C.TZ = lds + L.TZ;
    C.UZ = lds + L.UZ;
    C.pax = PR;
    C.pbx = PR + g.nx;
    C.pdbx = PR + 2 * g.nx;
    C.paz = PR + 3 * g.nx;
    C.pbz = C.paz + g.nz;
    C.pdbz = C.pbz + g.nz;
    C.RES = RES;
    C.SRC = SRC;
    C.BOFF = BOFF;
    C.ER = ER;
    C.EW = EW;
    C.nx = g.nx;
    C.nz = g.nz;
    C.ni = g.ni;
    C.ra = g.region_adj;
    C.swa = g.swa;
    C.sw2 = 2 * g.swa;
    C.bz0 = P.geo.band_z0;
    C.bz1 = P.geo.band_z0 + P.geo.band_rows;
    C.zb0 = bd.zb0;
    C.lo = bd.lo;
    C.azb = bd.top ? g.swa : 0;
    C.rh = g.rh;
    C.dt2 = g.dt2;

    const int tid = threadIdx.x;
    const int b = unit / g.shots;
    const int s = src_slot(P.geo, unit, unit - b * g.shots);
    const bool rec = bd.top;  // the receiver rows (the host checks) lie in the top band
    const int own_z0 = bd.top ? 0 : bd.lo + 2, own_z1 = bd.bot ? g.nz : bd.hi + 2;  // padded rows owned
    float* gout = P.grad + (size_t)unit * g.nx * g.nz;
    if (!unit_active(P.geo, unit)) {  // padding unit of an uneven batch
        zero_unit_outputs(gout + (size_t)own_z0 * g.nx, (own_z1 - own_z0) * g.nx,
                          DWAV ? P.dwav + (size_t)unit * g.nt : nullptr, DWAV ? g.nt : 0);
        return;
    }

    for (int j = tid; j < L.PR; j += NT) lds[j] = 0.0f;
    for (int j = tid; j < 3 * (g.nx + g.nz); j += NT) PR[j] = P.geo.prof[j];
    for (int j = tid; j < P.geo.band_rows * g.nx + 1; j += NT) BOFF[j] = P.geo.band_off[j];
    for (int j = tid; j < P.geo.n_ent; j += NT) {
        ER[j] = P.geo.ent_r[j];
        EW[j] = P.geo.ent_w[j];
    }

    const float* vvis = P.vel + (size_t)b * g.nvx * g.nvz;
    const float* wav = P.wavelet + (size_t)unit * P.wav_stride;
    const float* resid = P.resid + (size_t)unit * g.nt * g.nrec;
    const float* snaps = P.snaps + (size_t)pos * g.nt * g.inner;

    if (tid < 4) {
        const int cell = P.geo.src_cell[s * 4 + tid];
        const int cz = cell / g.nx, cx = cell - cz * g.nx;
        const float v = clampv(vis_at(vvis, g, cx, cz), g.vmin, g.vmax);
        SRC[tid] = 0.0f;
        SRC[4 + tid] = -(v * v * g.dt2) * P.geo.src_w[s * 4 + tid];
    }

    const int r = g.region_adj - 2;
    const Cls cI = make_cls(g, r, false, false, bd.lo, bd.hi);
    const Cls cX = make_cls(g, r, true, false, bd.lo, bd.hi);
    const Cls cZ = make_cls(g, r, false, true, bd.lo, bd.hi);
    const Cls cC = make_cls(g, r, true, true, bd.lo, bd.hi);
    const int rI = tid / cI.w, kI0 = tid - rI * cI.w;
    const int rX = tid / cX.w, kX0 = tid - rX * cX.w;
    const int rZ = tid / cZ.w, kZ0 = tid - rZ * cZ.w;
    const int rC = tid / cC.w, kC0 = tid - rC * cC.w;

    // Source corners (wgsl:245-249): at most one per class per thread (corner
    // cells differ by 1, w or w +- 1 in a class's numbering; the host checks
    // that no class is 511 or more columns wide).  A band owns the corners in
    // its rows.
    SrcOwn so[4];
#pragma unroll
    for (int cl = 0; cl < 4; ++cl) so[cl] = SrcOwn{-1, 0, 0.0f};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int cell = P.geo.src_cell[s * 4 + c];
        const float w = P.geo.src_w[s * 4 + c];
        const int cz = cell / g.nx, cx = cell - cz * g.nx;
        if (cx >= 2 && cx < g.nx - 2 && cz - 2 >= bd.lo && cz - 2 < bd.hi) {
            int cls, q;
            cell_class(g, r, cx - 2, cz - 2, cls, q, bd.lo, bd.hi);
            if (q % NT == tid) {
                const float v = clampv(vis_at(vvis, g, cx, cz), g.vmin, g.vmax);
                const SrcOwn o{q / NT, c, -2.0f * v * g.dt2 * w};
#pragma unroll
                for (int cl = 0; cl < 4; ++cl)
                    if (cl == cls) so[cl] = o;
            }
        }
    }

    AState<KI, false, false> SI;
    AState<KX, true, false> SX;
    AState<KZ, false, true> SZ;
    AState<KC, true, true> SC;
    adj_init(SI, cI, kI0, rI, g, vvis);
    adj_init(SX, cX, kX0, rX, g, vvis);
    adj_init(SZ, cZ, kZ0, rZ, g, vvis);
    adj_init(SC, cC, kC0, rC, g, vvis);
    __syncthreads();

    bool failed = false;
    for (int t = g.nt - 1; t >= 0; --t) {
        if (rec)
            for (int rr = tid; rr < g.nrec; rr += NT) RES[rr] = resid[(size_t)t * g.nrec + rr];
        adj_phase1(SI, cI, kI0, rI, C, so[0]);
        adj_phase1(SX, cX, kX0, rX, C, so[1]);
        adj_phase1(SZ, cZ, kZ0, rZ, C, so[2]);
        adj_phase1(SC, cC, kC0, rC, C, so[3]);
        __syncthreads();
        if constexpr (BAND) {
            if (!band_exchange(X, C.V, bd, g.nx, pos, p, g.nt - t, fail_s)) {
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
        const float f_t = wav[t];
        const float* snap_t = snaps + (size_t)t * g.inner;
        adj_phase2(SI, cI, kI0, rI, C, so[0], snap_t, f_t);
        adj_phase2(SX, cX, kX0, rX, C, so[1], snap_t, f_t);
        adj_phase2(SZ, cZ, kZ0, rZ, C, so[2], snap_t, f_t);
        adj_phase2(SC, cC, kC0, rC, C, so[3], snap_t, f_t);
        __syncthreads();
    }
    // Padded-grid gradient of the owned rows; the FD_PAD ring is zero.
    for (int j = own_z0 * g.nx + tid; j < own_z1 * g.nx; j += NT) {
        const int z = j / g.nx, x = j - z * g.nx;
        if (x < 2 || x >= g.nx - 2 || z < 2 || z >= g.nz - 2) gout[j] = 0.0f;
    }
    if (BAND && failed) {  // a neighbour never arrived: the gradient is not valid
        for (int j = own_z0 * g.nx + tid; j < own_z1 * g.nx; j += NT) gout[j] = __builtin_nanf("");
        return;
    }
    adj_store(SI, cI, kI0, rI, g.nx, gout);
    adj_store(SX, cX, kX0, rX, g.nx, gout);
    adj_store(SZ, cZ, kZ0, rZ, g.nx, gout);
    adj_store(SC, cC, kC0, rC, g.nx, gout);
}

template <int KI, int KX, int KZ, int KC, bool DWAV>
__global__ __launch_bounds__(NT) void k_adjoint(AdjParams P) {
    extern __shared__ float lds[];
    const Xchg X{};
    adjoint_unit<KI, KX, KZ, KC, DWAV, false>(P, lds, P.unit0 + blockIdx.x, blockIdx.x, whole_band(P.g), X, 0,
                                              nullptr);
}

template <int KI, int KX, int KZ, int KC>
__global__ __launch_bounds__(NT) void k_adjoint_band(AdjParams P, BandLaunch B, int units) {
    extern __shared__ float lds[];
    __shared__ int task_s, fail_s;
    const Xchg X{B.xbuf, B.err, B.P, B.spin};
    const int task = __builtin_amdgcn_readfirstlane(band_task(B.next, &task_s));
    if (task >= units * B.P) return;
    const int u = task / B.P, p = task - u * B.P;
    adjoint_unit<KI, KX, KZ, KC, false, true>(P, lds, P.unit0 + u, u, split_band(P.g, B.bounds, B.P, p), X, p,
                                              &fail_s);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <int KI, int KX, int KZ, int KC>
static hipError_t fwd_t(const FwdParams& p, int units, bool store, bool born, hipStream_t st) {
    const size_t lds = forward_lds_bytes(p.g);
    void (*fn)(FwdParams) = born ? k_forward<KI, KX, KZ, KC, false, true>
                                 : (store ? k_forward<KI, KX, KZ, KC, true, false>
                                          : k_forward<KI, KX, KZ, KC, false, false>);
    hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, dim3(units), dim3(NT), lds, st, p);
    return hipGetLastError();
}

template <int KI, int KX, int KZ, int KC>
static hipError_t adj_t(const AdjParams& p, int units, hipStream_t st) {
    const size_t lds = adjoint_lds_bytes(p.g, p.geo);
    void (*fn)(AdjParams) = p.dwav ? k_adjoint<KI, KX, KZ, KC, true> : k_adjoint<KI, KX, KZ, KC, false>;
    hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, dim3(units), dim3(NT), lds, st, p);
    return hipGetLastError();
}

template <int KI, int KX, int KZ, int KC>
static hipError_t fwd_band_t(const FwdParams& p, int units, bool store, const BandLaunch& b, hipStream_t st) {
    size_t lds = 0;
    for (int q = 0; q < b.P; ++q) lds = std::max<size_t>(lds, 4ull * fwd_layout(p.g, split_band(p.g, b.bounds, b.P, q)).total);
    void (*fn)(FwdParams, BandLaunch, int) = store ? k_forward_band<KI, KX, KZ, KC, true>
                                                   : k_forward_band<KI, KX, KZ, KC, false>;
    hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, dim3(units * b.P), dim3(NT), lds, st, p, b, units);
    return hipGetLastError();
}

template <int KI, int KX, int KZ, int KC>
static hipError_t adj_band_t(const AdjParams& p, int units, const BandLaunch& b, hipStream_t st) {
    size_t lds = 0;
    for (int q = 0; q < b.P; ++q)
        lds = std::max<size_t>(lds, 4ull * adj_layout(p.g, p.geo, split_band(p.g, b.bounds, b.P, q)).total);
    void (*fn)(AdjParams, BandLaunch, int) = k_adjoint_band<KI, KX, KZ, KC>;
    hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, dim3(units * b.P), dim3(NT), lds, st, p, b, units);
    return hipGetLastError();
}

hipError_t launch_forward_band(const FwdParams& p, int units, bool store, const BandLaunch& b, hipStream_t st) {
    if (units <= 0) return hipSuccess;
    if (b.P < 2 || b.P > kMaxBands) return hipErrorInvalidValue;
    hipError_t err;
    if (b.fixed && launch_forward_band_fx(p, units, store, b, st, &err)) return err;
    switch (pick_band_tuple(p.g, b.P, b.bounds, false)) {
        case 0: return fwd_band_t<9, 5, 5, 3>(p, units, store, b, st);
        case 1: return fwd_band_t<12, 5, 7, 3>(p, units, store, b, st);
        case 2: return fwd_band_t<13, 4, 6, 2>(p, units, store, b, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_adjoint_band(const AdjParams& p, int units, const BandLaunch& b, hipStream_t st) {
    if (units <= 0) return hipSuccess;
    if (b.P < 2 || b.P > kMaxBands || p.dwav) return hipErrorInvalidValue;
    hipError_t err;
    if (b.fixed && launch_adjoint_band_fx(p, units, b, st, &err)) return err;
    switch (pick_band_tuple(p.g, b.P, b.bounds, true)) {
        case 0: return adj_band_t<9, 5, 5, 3>(p, units, b, st);
        case 1: return adj_band_t<12, 5, 7, 3>(p, units, b, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_forward(const FwdParams& p, int units, bool store, bool born, hipStream_t st) {
    if (units <= 0) return hipSuccess;
    hipError_t err;
    if (launch_forward_fx(p, units, store, born, st, &err)) return err;
    switch (pick_tuple(p.g, p.g.region - 2)) {
        case 0: return fwd_t<1, 1, 1, 1>(p, units, store, born, st);
        case 1: return fwd_t<2, 2, 2, 2>(p, units, store, born, st);
        case 2: return fwd_t<4, 3, 3, 3>(p, units, store, born, st);
        case 3: return fwd_t<9, 6, 6, 5>(p, units, store, born, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_adjoint(const AdjParams& p, int units, hipStream_t st) {
    if (units <= 0) return hipSuccess;
    hipError_t err;
    if (launch_adjoint_fx(p, units, st, &err)) return err;
    switch (pick_tuple(p.g, p.g.region_adj - 2)) {
        case 0: return adj_t<1, 1, 1, 1>(p, units, st);
        case 1: return adj_t<2, 2, 2, 2>(p, units, st);
        case 2: return adj_t<4, 3, 3, 3>(p, units, st);
        case 3: return adj_t<9, 6, 6, 5>(p, units, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace wg
