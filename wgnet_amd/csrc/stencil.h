// Per-cell arithmetic of the 4th-order C-PML acoustic step and its exact
// adjoint, shared by the kernels.  Operation order is the reference's, with no
// FMA contraction, so results are bit-identical to the float32 CPU solver:
//   forward  crates/geology-core/src/lib.rs:155-205
//   adjoint  crates/wave-gauss-rs/src/shaders/splatAdjoint.wgsl:196-232
// (1 + b) is read from a table; it is the same float32 sum the reference forms.
#pragma once

#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

namespace wg {

// lib.rs:139-143 (forward) and splatAdjoint.wgsl:163-167 (adjoint), float32.
constexpr float kC1 = 0.6666666666666666f;
constexpr float kC2 = 0.08333333333333333f;
constexpr float kD0 = -2.5f;
constexpr float kD1 = 1.3333333333333333f;
constexpr float kD2 = 0.08333333333333333f;
constexpr float kD1A = 0.6666666666666666f;
constexpr float kD1B = -0.08333333333333333f;
constexpr float kD2A = -2.5f;
constexpr float kD2B = 1.3333333333333333f;
constexpr float kD2C = -0.08333333333333333f;

// C-PML coefficients of one cell along one axis.
struct Pml {
    float a, opb, b, db;  // a, 1 + b, b, db/dx
};

// Forward: w = sum over axes of the C-PML Laplacian term at one cell, and the
// update of that cell's psi / zeta (lib.rs:155-202).  U points at the field,
// i is the cell, sx / sz the strides of the x and z neighbours; A*c point at
// this cell's entry of the a*psi strips (neighbours at +-1, +-2 strides).
template <bool PX, bool PZ>
__device__ __forceinline__ float fwd_w(const float* U, int i, int sz, float u, float rh, float rh2,
                                       const float* AXc, const Pml& px_, float& psx, float& zex,
                                       const float* AZc, int az_stride, const Pml& pz_, float& psz, float& zez) {
    float w;  // the x term, then += the z term (0 + x skipped: equal up to the sign of zero)
    const float um1x = U[i - 1], up1x = U[i + 1];
    const float um2x = U[i - 2], up2x = U[i + 2];
    const float d2x = (kD0 * u + kD1 * (up1x + um1x) - kD2 * (up2x + um2x)) * rh2;
    if constexpr (PX) {
        const float d1x = (kC1 * (up1x - um1x) - kC2 * (up2x - um2x)) * rh;
        const float ap1 = AXc[1], am1 = AXc[-1];
        const float ap2 = AXc[2], am2 = AXc[-2];
        const float d1_apsi = (kC1 * (ap1 - am1) - kC2 * (ap2 - am2)) * rh;
        const float tmp = px_.opb * d2x + px_.db * d1x + d1_apsi;
        w = px_.opb * tmp + px_.a * zex;
        psx = px_.b * d1x + px_.a * psx;
        zex = px_.b * tmp + px_.a * zex;
    } else {
        w = d2x;
    }
    const float um1z = U[i - sz], up1z = U[i + sz];
    const float um2z = U[i - 2 * sz], up2z = U[i + 2 * sz];
    const float d2z = (kD0 * u + kD1 * (up1z + um1z) - kD2 * (up2z + um2z)) * rh2;
    if constexpr (PZ) {
        const float d1z = (kC1 * (up1z - um1z) - kC2 * (up2z - um2z)) * rh;
        const float ap1 = AZc[az_stride], am1 = AZc[-az_stride];
        const float ap2 = AZc[2 * az_stride], am2 = AZc[-2 * az_stride];
        const float d1_apsi = (kC1 * (ap1 - am1) - kC2 * (ap2 - am2)) * rh;
        const float tmp = pz_.opb * d2z + pz_.db * d1z + d1_apsi;
        w += pz_.opb * tmp + pz_.a * zez;
        psz = pz_.b * d1z + pz_.a * psz;
        zez = pz_.b * tmp + pz_.a * zez;
    } else {
        w += d2z;
    }
    return w;
}

// Adjoint, one axis inside the C-PML region (wgsl:206-227): T and U1 of the
// five neighbours along the axis are read from strips (Tc / Uc point at this
// cell, stride st), U2 = (1 + b) T with (1 + b) of each neighbour in opb[].
// Returns -d1(U1) + d2(U2) and updates this cell's P / Z.
__device__ __forceinline__ float adj_axis_pml(const float* Tc, const float* Uc, int st, const float opb[5],
                                              float a, float v0, float rh, float& P, float& Z) {
    const float tm2 = Tc[-2 * st], tm1 = Tc[-st], t00 = Tc[0], tp1 = Tc[st], tp2 = Tc[2 * st];
    const float um2 = Uc[-2 * st], um1 = Uc[-st], up1 = Uc[st], up2 = Uc[2 * st];
    const float w2m2 = opb[0] * tm2;
    const float w2m1 = opb[1] * tm1;
    const float w200 = opb[2] * t00;
    const float w2p1 = opb[3] * tp1;
    const float w2p2 = opb[4] * tp2;
    const float d1_u1 = (kD1A * (up1 - um1) + kD1B * (up2 - um2)) * rh;
    const float d2_u2 = (kD2A * w200 + kD2B * (w2p1 + w2m1) + kD2C * (w2p2 + w2m2)) * rh * rh;
    const float d1_t = (kD1A * (tp1 - tm1) + kD1B * (tp2 - tm2)) * rh;
    P = -a * d1_t + a * P;
    Z = a * v0 + a * Z;
    return -d1_u1 + d2_u2;
}

// Adjoint, one axis outside the region (wgsl:228-230).
__device__ __forceinline__ float adj_axis_int(const float* V, int i, int st, float v0, float rh) {
    const float vm1 = V[i - st], vp1 = V[i + st], vm2 = V[i - 2 * st], vp2 = V[i + 2 * st];
    return (kD2A * v0 + kD2B * (vp1 + vm1) + kD2C * (vp2 + vm2)) * rh * rh;
}

}  // namespace wg
