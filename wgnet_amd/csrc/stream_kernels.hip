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
// Forward step (lib.rs:145-220 for one t, every unit of the chunk).
// --------------------------------------------------------------------------
template <class TL, bool STORE, bool BORN>
__global__ __launch_bounds__(NT) void k_fwd_step(StepParams P) {
    WG_TILE_CONSTANTS
    __shared__ float sU[HZ * HX];
    __shared__ float sAX[TZ * HX];
    __shared__ float sAZ[HZ * TX];
    const Grid& g = P.g;
    const int ul = blockIdx.z;             // unit within the chunk
    const int unit = P.unit0 + ul;
    const int b = unit / g.shots, s = src_slot(P.geo, unit, unit - (unit / g.shots) * g.shots);
    const size_t cells = (size_t)g.nx * g.nz;
    const int t = P.t, par = t & 1;
    float* Ucur = P.state + ((size_t)(0 + par) * P.units + ul) * cells;
    float* Unxt = P.state + ((size_t)(1 - par) * P.units + ul) * cells;
    float* PXc = P.state + ((size_t)(2 + par) * P.units + ul) * cells;
    float* PXn = P.state + ((size_t)(3 - par) * P.units + ul) * cells;
    float* PZc = P.state + ((size_t)(4 + par) * P.units + ul) * cells;
    float* PZn = P.state + ((size_t)(5 - par) * P.units + ul) * cells;
    float* ZX = P.state + ((size_t)6 * P.units + ul) * cells;
    float* ZZ = P.state + ((size_t)7 * P.units + ul) * cells;
    const float* vvis = P.vel + (size_t)b * g.nvx * g.nvz;
    const float* pax = P.geo.prof;
    const float* pbx = pax + g.nx;
    const float* pdbx = pbx + g.nx;
    const float* paz = P.geo.prof + 3 * g.nx;
    const float* pbz = paz + g.nz;
    const float* pdbz = pbz + g.nz;
    const int tid = threadIdx.x;
    const int xi0 = blockIdx.x * TX, zi0 = blockIdx.y * TZ;  // tile origin, inner coords

    if (!unit_active(P.geo, unit)) {  // padding unit of an uneven batch
        if (P.traces && blockIdx.x == 0 && blockIdx.y == 0)
            for (int r = tid; r < g.nrec; r += NT) P.traces[((size_t)unit * g.nt + t) * g.nrec + r] = 0.0f;
        return;
    }
    // Receivers record u^t before the update (lib.rs:146-149): one tile per unit.
    if (P.traces && blockIdx.x == 0 && blockIdx.y == 0) {
        float* tr = P.traces + ((size_t)unit * g.nt + t) * g.nrec;
        for (int r = tid; r < g.nrec; r += NT) {
            float acc = 0.0f;
#pragma unroll
            for (int c = 0; c < 4; ++c) acc += P.geo.rec_w[r * 4 + c] * Ucur[P.geo.rec_cell[r * 4 + c]];
            tr[r] = acc;
        }
    }
    // Stage u (tile + 2-cell halo), a*psi_x (x halo) and a*psi_z (z halo).
    for (int j = tid; j < HZ * HX; j += NT) {
        const int x = xi0 + (j % HX), z = zi0 + (j / HX);  // padded coords: halo starts at inner - 2
        sU[j] = (x < g.nx && z < g.nz) ? Ucur[z * g.nx + x] : 0.0f;
    }
    // does the tile hold C-PML cells along x / z (padded coords [xi0+2, xi0+TX+2))?
    const bool tpx = xi0 + 2 < g.region || xi0 + TX + 2 > g.nx - g.region;
    const bool tpz = zi0 + 2 < g.region || zi0 + TZ + 2 > g.nz - g.region;
    if (tpx)
        for (int j = tid; j < TZ * HX; j += NT) {
            const int x = xi0 + (j % HX), z = zi0 + 2 + (j / HX);
            sAX[j] = (x < g.nx && z < g.nz) ? pax[x] * PXc[z * g.nx + x] : 0.0f;
        }
    if (tpz)
        for (int j = tid; j < HZ * TX; j += NT) {
            const int x = xi0 + 2 + (j % TX), z = zi0 + (j / TX);
            sAZ[j] = (x < g.nx && z < g.nz) ? paz[z] * PZc[z * g.nx + x] : 0.0f;
        }
    __syncthreads();
    const int lx = tid % TX;
    const float f_t = P.wavelet[(size_t)unit * P.wav_stride + t];
    for (int k = 0; k < CPT; ++k) {
        const int lz = tid / TX + RZ * k;
        const int x = xi0 + lx + 2, z = zi0 + lz + 2;
        if (x >= g.nx - 2 || z >= g.nz - 2) continue;
        const int i = z * g.nx + x;
        const int li = (lz + 2) * HX + (lx + 2);
        const float u = sU[li];
        const bool px = x < g.region || x >= g.nx - g.region;
        const bool pz = z < g.region || z >= g.nz - g.region;
        Pml mx{0, 0, 0, 0}, mz{0, 0, 0, 0};
        float psx = 0.0f, zex = 0.0f, psz = 0.0f, zez = 0.0f;
        if (px) {
            const float bb = pbx[x];
            mx = Pml{pax[x], 1.0f + bb, bb, pdbx[x]};
            psx = PXc[i];
            zex = ZX[i];
        }
        if (pz) {
            const float bb = pbz[z];
            mz = Pml{paz[z], 1.0f + bb, bb, pdbz[z]};
            psz = PZc[i];
            zez = ZZ[i];
        }
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
        const float vr = vis_raw(vvis, g, x, z);
        const float v = clampv(vr, g.vmin, g.vmax);
        const size_t q = (size_t)(z - 2) * g.ni + (x - 2);
        if (STORE) P.snaps[((size_t)(t - P.snap_t0) * P.units + ul) * g.inner + q] = 2.0f * v * g.dt2 * w;
        float next = v * v * g.dt2 * w + 2.0f * u - Unxt[i];
        if (BORN) {
            const float m = (vr > g.vmin && vr < g.vmax) ? 1.0f : 0.0f;
            next += (vis_raw(P.dvel + (size_t)b * g.nvx * g.nvz, g, x, z) * m) *
                    P.born_snaps[((size_t)(t - P.snap_t0) * P.units + ul) * g.inner + q];
        }
        // point source after the update (lib.rs:209-215); Born: its derivative
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float wc = P.geo.src_w[s * 4 + c];
            if (P.geo.src_cell[s * 4 + c] == i && wc != 0.0f) {
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
// Adjoint step (splatAdjoint.wgsl:58-131 for one t, every unit of the chunk).
// --------------------------------------------------------------------------
template <class TL, bool DWAV>
__global__ __launch_bounds__(NT) void k_adj_step(StepParams P) {
    WG_TILE_CONSTANTS
    __shared__ float sV[HZ * HX];
    __shared__ float sTX[TZ * HX], sUX[TZ * HX];
    __shared__ float sTZ[HZ * TX], sUZ[HZ * TX];
    const Grid& g = P.g;
    const int ul = blockIdx.z;
    const int unit = P.unit0 + ul;
    const int b = unit / g.shots, s = src_slot(P.geo, unit, unit - (unit / g.shots) * g.shots);
    const size_t cells = (size_t)g.nx * g.nz;
    const int t = P.t, q = (g.nt - 1 - t) & 1;
    // arrays: lam0 lam1 | Px0 Px1 | Zx0 Zx1 | Pz0 Pz1 | Zz0 Zz1 (halves by q)
    auto arr = [&](int a) { return P.state + ((size_t)a * P.units + ul) * cells; };
    float* Lcur = arr(q);
    float* Lnxt = arr(1 - q);
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
    const float* pax = P.geo.prof;
    const float* pbx = pax + g.nx;
    const float* pdbx = pbx + g.nx;
    const float* paz = P.geo.prof + 3 * g.nx;
    const float* pbz = paz + g.nz;
    const float* pdbz = pbz + g.nz;
    const int tid = threadIdx.x;
    const int xi0 = blockIdx.x * TX, zi0 = blockIdx.y * TZ;

    if (!unit_active(P.geo, unit)) {  // padding unit; its gradient rows are zeroed by the host
        if (DWAV && blockIdx.x == 0 && blockIdx.y == 0 && tid == 0) P.dwav[(size_t)unit * g.nt + t] = 0.0f;
        return;
    }
    if (DWAV && blockIdx.x == 0 && blockIdx.y == 0 && tid == 0) {
        float acc = 0.0f;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int cell = P.geo.src_cell[s * 4 + c];
            const int cz = cell / g.nx, cx = cell - cz * g.nx;
            const float v = vel_at(vvis, g, cx, cz);
            acc += (-(v * v * g.dt2) * P.geo.src_w[s * 4 + c]) * Lcur[cell];
        }
        P.dwav[(size_t)unit * g.nt + t] = acc;
    }
    // V on the tile + halo (vw(), wgsl:53-56); 0 on the ring (lambda is 0 there)
    for (int j = tid; j < HZ * HX; j += NT) {
        const int x = xi0 + (j % HX), z = zi0 + (j / HX);
        float V = 0.0f;
        if (x < g.nx && z < g.nz) {
            const float v = vel_at(vvis, g, x, z);
            V = v * v * g.dt2 * Lcur[z * g.nx + x];
        }
        sV[j] = V;
    }
    __syncthreads();
    // T = (1+b)V + bZ, U1 = db T + bP of the x-neighbours (x halo) and z-neighbours
    // does the tile hold cells of the adjoint C-PML region along x / z?
    const bool tax = xi0 + 2 < g.region_adj || xi0 + TX + 2 > g.nx - g.region_adj;
    const bool taz = zi0 + 2 < g.region_adj || zi0 + TZ + 2 > g.nz - g.region_adj;
    if (tax)
        for (int j = tid; j < TZ * HX; j += NT) {
            const int hx = j % HX, hz = j / HX;
            const int x = xi0 + hx, z = zi0 + 2 + hz;
            float T = 0.0f, U1 = 0.0f;
            if (x < g.nx && z < g.nz) {
                const int i = z * g.nx + x;
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
                const int i = z * g.nx + x;
                const float bb = pbz[z];
                T = (1.0f + bb) * sV[hz * HX + hx + 2] + bb * ZZr[i];
                U1 = pdbz[z] * T + bb * PZr[i];
            }
            sTZ[j] = T;
            sUZ[j] = U1;
        }
    __syncthreads();
    const int lx = tid % TX;
    const float f_t = P.wavelet[(size_t)unit * P.wav_stride + t];
    const float* res = P.resid + ((size_t)unit * g.nt + t) * g.nrec;
    for (int k = 0; k < CPT; ++k) {
        const int lz = tid / TX + RZ * k;
        const int x = xi0 + lx + 2, z = zi0 + lz + 2;
        if (x >= g.nx - 2 || z >= g.nz - 2) continue;
        const int i = z * g.nx + x;
        const int li = (lz + 2) * HX + (lx + 2);
        const float v0 = sV[li];
        const float rh = g.rh;
        float w;  // x term, then += z term (0 + x skipped: equal up to the sign of zero)
        if (x < g.region_adj || x >= g.nx - g.region_adj) {
            const float opb[5] = {1.0f + pbx[x - 2], 1.0f + pbx[x - 1], 1.0f + pbx[x], 1.0f + pbx[x + 1], 1.0f + pbx[x + 2]};
            float Pv = PXr[i], Zv = ZXr[i];
            w = adj_axis_pml(sTX + lz * HX + (lx + 2), sUX + lz * HX + (lx + 2), 1, opb, pax[x], v0, rh, Pv, Zv);
            PXw[i] = Pv;
            ZXw[i] = Zv;
        } else {
            w = adj_axis_int(sV, li, 1, v0, rh);
        }
        if (z < g.region_adj || z >= g.nz - g.region_adj) {
            const float opb[5] = {1.0f + pbz[z - 2], 1.0f + pbz[z - 1], 1.0f + pbz[z], 1.0f + pbz[z + 1], 1.0f + pbz[z + 2]};
            float Pv = PZr[i], Zv = ZZr[i];
            w += adj_axis_pml(sTZ + (lz + 2) * TX + lx, sUZ + (lz + 2) * TX + lx, TX, opb, paz[z], v0, rh, Pv, Zv);
            PZw[i] = Pv;
            ZZw[i] = Zv;
        } else {
            w += adj_axis_int(sV, li, HX, v0, rh);
        }
        const float lam = Lcur[i];
        float next = w + 2.0f * lam - Lnxt[i];
        const size_t qq = (size_t)(z - 2) * g.ni + (x - 2);
        float gk = lam * P.snaps[((size_t)(t - P.snap_t0) * P.units + ul) * g.inner + qq];
        if (z >= P.geo.band_z0 && z < P.geo.band_z0 + P.geo.band_rows) {
            const int o = (z - P.geo.band_z0) * g.nx + x;
            for (int e = P.geo.band_off[o]; e < P.geo.band_off[o + 1]; ++e) next += P.geo.ent_w[e] * res[P.geo.ent_r[e]];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (P.geo.src_cell[s * 4 + c] == i) {
                const float v = vel_at(vvis, g, x, z);
                gk += lam * (-2.0f * v * g.dt2 * P.geo.src_w[s * 4 + c] * f_t);
            }
        grad[i] += gk;
        Lnxt[i] = next;
    }
}

}  // namespace st

namespace {

template <class T>
hipError_t fwd_step_t(const StepParams& p, int units, bool store, bool born, hipStream_t st) {
    const dim3 grid((p.g.ni + T::TX - 1) / T::TX, (p.g.nzi + T::TZ - 1) / T::TZ, units);
    void (*fn)(StepParams) =
        born ? st::k_fwd_step<T, false, true> : (store ? st::k_fwd_step<T, true, false> : st::k_fwd_step<T, false, false>);
    hipLaunchKernelGGL(fn, grid, dim3(st::NT), 0, st, p);
    return hipGetLastError();
}

template <class T>
hipError_t adj_step_t(const StepParams& p, int units, hipStream_t st) {
    const dim3 grid((p.g.ni + T::TX - 1) / T::TX, (p.g.nzi + T::TZ - 1) / T::TZ, units);
    void (*fn)(StepParams) = p.dwav ? st::k_adj_step<T, true> : st::k_adj_step<T, false>;
    hipLaunchKernelGGL(fn, grid, dim3(st::NT), 0, st, p);
    return hipGetLastError();
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
