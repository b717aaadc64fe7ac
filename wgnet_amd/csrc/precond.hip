// Learned gradient preconditioner: a 4-layer network of dilated 3x3
// convolutions (2 -> 32 -> 32 -> 32 -> 1 channels, dilations 4, 16, 8, 1,
// GELU after the first three layers, zero padding equal to the dilation, so
// each layer keeps the image size), inference and training on MFMA.  The
// dilations give a 59 x 59 receptive field on the 70 x 70 models.
//
// No reference counterpart: the reference refines with plain projected L-BFGS
// (python/wgnet/refine.py:92-169); the network maps (velocity-space gradient,
// normalised velocity) [2, H, W] to a model-update direction [H, W] and is
// applied at the gradient hook of refine.py:105-106.
//
// Precision: weights and activations are bfloat16, every product is
// accumulated in float32 by v_mfma_f32_16x16x16_bf16; pre-activations are kept
// in float32 for the GELU derivative, master weights and AdamW moments are
// float32.
//
// Layouts: activations NHWC, [P = N*H*W pixels][C]; layer weights
// [COUTP][KP] with k = (ky*3 + kx)*CIN + ci (tap (ky, kx) reads the pixel
// offset by ((ky - 1) D, (kx - 1) D) for dilation D), COUTP and KP rounded up to 16
// (padding stays zero); their transposes [CINP][K2P] with k2 = tap*COUT + co
// for the data gradient.  Convolutions are implicit GEMMs: M = pixels,
// N = output channels, K = 9*CIN (forward); M = pixels, N = CIN, K = 9*COUT
// (data gradient); M = COUT, N = 9*CIN, K = pixels (weight gradient, per-block
// partial sums reduced in a fixed order, so results are deterministic).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/wgnet_hip.h"

namespace wg {
namespace pc {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NT = 256;  // threads per block (4 waves)

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return 0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_d(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

__host__ __device__ constexpr int rup16(int v) { return (v + 15) / 16 * 16; }

// --------------------------------------------------------------------------
// Forward: pre[p][co] = bias[co] + sum_k W[co][k] * in[nbr(p, tap(k), D)][ci(k)];
// act = bf16(gelu(pre)) (or bf16(pre) without GELU).  One wave computes 16
// pixels x COUTP channels.
// --------------------------------------------------------------------------
template <int CIN, int COUT, bool GELU>
__global__ __launch_bounds__(NT) void k_fwd(const uint16_t* __restrict__ in, int P, int H, int W, int D,
                                            const uint16_t* __restrict__ wt, const float* __restrict__ bias,
                                            float* __restrict__ pre, uint16_t* __restrict__ act) {
    constexpr int K = 9 * CIN, KP = rup16(K), COUTP = rup16(COUT), NTL = COUTP / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int p0 = (blockIdx.x * 4 + wave) * 16;
    if (p0 >= P) return;  // wave-uniform; the kernel has no barriers
    const int r = lane & 15, kq = lane >> 4;
    const int p = p0 + r;
    const int HW = H * W;
    const int img = p / HW, rem = p - img * HW, y = rem / W, x = rem - y * W;
    const bool pv = p < P;
    f32x4 acc[NTL];
#pragma unroll
    for (int t = 0; t < NTL; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int kb = 0; kb < KP; kb += 16) {
        bf16x4 a;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = kb + 4 * kq + j;
            uint16_t v = 0;
            if (pv && k < K) {
                const int tap = k / CIN, ci = k - tap * CIN;
                const int yy = y + (tap / 3 - 1) * D, xx = x + (tap % 3 - 1) * D;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W) v = in[((size_t)img * HW + yy * W + xx) * CIN + ci];
            }
            a[j] = (short)v;
        }
#pragma unroll
        for (int t = 0; t < NTL; ++t) {
            const uint16_t* wr = wt + (size_t)(t * 16 + r) * KP + kb + 4 * kq;
            bf16x4 b;
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = (short)wr[j];
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
    }
    // lane holds D[row 4*kq + j][col r] of each 16 x 16 tile
#pragma unroll
    for (int t = 0; t < NTL; ++t) {
        const int co = t * 16 + r;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int pp = p0 + 4 * kq + j;
            if (pp < P && co < COUT) {
                const float v = acc[t][j] + bias[co];
                if (pre) pre[(size_t)pp * COUT + co] = v;
                if (act) act[(size_t)pp * COUT + co] = f2bf(GELU ? gelu(v) : v);
            }
        }
    }
}

// --------------------------------------------------------------------------
// Data gradient: d_in[p][ci] = sum_{tap, co} dpre[p - D off(tap)][co] * W[co][tap*CIN + ci];
// dpre_prev = bf16(d_in * gelu'(pre_prev)).  wtT is [CINP][K2P], k2 = tap*COUT + co.
// --------------------------------------------------------------------------
template <int CIN, int COUT>
__global__ __launch_bounds__(NT) void k_bwd_data(const uint16_t* __restrict__ dpre, int P, int H, int W, int D,
                                                 const uint16_t* __restrict__ wtT, const float* __restrict__ pre_prev,
                                                 uint16_t* __restrict__ dpre_prev) {
    constexpr int K2 = 9 * COUT, K2P = rup16(K2), CINP = rup16(CIN), NTL = CINP / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int p0 = (blockIdx.x * 4 + wave) * 16;
    if (p0 >= P) return;
    const int r = lane & 15, kq = lane >> 4;
    const int p = p0 + r;
    const int HW = H * W;
    const int img = p / HW, rem = p - img * HW, y = rem / W, x = rem - y * W;
    const bool pv = p < P;
    f32x4 acc[NTL];
#pragma unroll
    for (int t = 0; t < NTL; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int kb = 0; kb < K2P; kb += 16) {
        bf16x4 a;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = kb + 4 * kq + j;
            uint16_t v = 0;
            if (pv && k < K2) {
                const int tap = k / COUT, co = k - tap * COUT;
                const int yy = y - (tap / 3 - 1) * D, xx = x - (tap % 3 - 1) * D;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W) v = dpre[((size_t)img * HW + yy * W + xx) * COUT + co];
            }
            a[j] = (short)v;
        }
#pragma unroll
        for (int t = 0; t < NTL; ++t) {
            const uint16_t* wr = wtT + (size_t)(t * 16 + r) * K2P + kb + 4 * kq;
            bf16x4 b;
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = (short)wr[j];
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < NTL; ++t) {
        const int ci = t * 16 + r;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int pp = p0 + 4 * kq + j;
            if (pp < P && ci < CIN) {
                const size_t o = (size_t)pp * CIN + ci;
                dpre_prev[o] = f2bf(acc[t][j] * gelu_d(pre_prev[o]));
            }
        }
    }
}

// --------------------------------------------------------------------------
// Weight and bias gradient partials over a chunk of `chunk` pixels per block:
// part[blk][co*KP + k] = sum_p dpre[p][co] * in[nbr(p, tap(k), D)][ci(k)],
// part[blk][COUTP*KP + co] = sum_p dpre[p][co].
// --------------------------------------------------------------------------
template <int CIN, int COUT>
__global__ __launch_bounds__(NT) void k_bwd_w(const uint16_t* __restrict__ dpre, const uint16_t* __restrict__ in,
                                              int P, int H, int W, int D, int chunk, float* __restrict__ part) {
    constexpr int K = 9 * CIN, KP = rup16(K), COUTP = rup16(COUT), MT = COUTP / 16, NTL = KP / 16;
    constexpr int NPW = (NTL + 3) / 4;  // column tiles per wave
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, kq = lane >> 4;
    const int HW = H * W;
    const int pbeg = blockIdx.x * chunk, pend = min(P, pbeg + chunk);
    f32x4 acc[MT][NPW];
    float bsum[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        bsum[m] = 0.0f;
#pragma unroll
        for (int u = 0; u < NPW; ++u) acc[m][u] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    for (int pb = pbeg; pb < pend; pb += 16) {
        bf16x4 a[MT];
        int py[4], px[4], pimg[4];
        bool ok[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int p = pb + 4 * kq + j;
            ok[j] = p < pend;
            pimg[j] = p / HW;
            const int rem = p - pimg[j] * HW;
            py[j] = rem / W;
            px[j] = rem - py[j] * W;
        }
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int co = m * 16 + r;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint16_t v = (ok[j] && co < COUT) ? dpre[(size_t)(pb + 4 * kq + j) * COUT + co] : (uint16_t)0;
                a[m][j] = (short)v;
                bsum[m] += bf2f(v);
            }
        }
#pragma unroll
        for (int u = 0; u < NPW; ++u) {
            const int t = wave + 4 * u;
            if (t < NTL) {
                const int k = t * 16 + r;
                const int tap = k / CIN, ci = k - tap * CIN;
                bf16x4 b;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    uint16_t v = 0;
                    if (ok[j] && k < K) {
                        const int yy = py[j] + (tap / 3 - 1) * D, xx = px[j] + (tap % 3 - 1) * D;
                        if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                            v = in[((size_t)pimg[j] * HW + yy * W + xx) * CIN + ci];
                    }
                    b[j] = (short)v;
                }
#pragma unroll
                for (int m = 0; m < MT; ++m)
                    acc[m][u] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[m], b, acc[m][u], 0, 0, 0);
            }
        }
    }
    float* out = part + (size_t)blockIdx.x * (COUTP * KP + COUTP);
#pragma unroll
    for (int u = 0; u < NPW; ++u) {
        const int t = wave + 4 * u;
        if (t < NTL) {
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int j = 0; j < 4; ++j) out[(size_t)(m * 16 + 4 * kq + j) * KP + t * 16 + r] = acc[m][u][j];
        }
    }
    if (wave == 0) {
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            float s = bsum[m];
            s += __shfl_xor(s, 16);
            s += __shfl_xor(s, 32);
            if (kq == 0) out[COUTP * KP + m * 16 + r] = s;
        }
    }
}

// grad[i] = sum over blocks (fixed order) of part[blk][i]
__global__ void k_reduce(const float* __restrict__ part, int nblk, int n, float* __restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int b = 0; b < nblk; ++b) s += part[(size_t)b * n + i];
    grad[i] = s;
}

// x [N][2][H][W] float32 -> bf16 NHWC [P][2]
__global__ void k_prep(const float* __restrict__ x, int N, int HW, uint16_t* __restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N * HW) return;
    const int img = p / HW, q = p - img * HW;
    out[2 * (size_t)p] = f2bf(x[((size_t)img * 2 + 0) * HW + q]);
    out[2 * (size_t)p + 1] = f2bf(x[((size_t)img * 2 + 1) * HW + q]);
}

// Mean squared error: d_out = bf16(2 (out - target) / P); part[blk] = sum of squares (double).
__global__ __launch_bounds__(NT) void k_mse(const float* __restrict__ out, const float* __restrict__ target, int P,
                                            uint16_t* __restrict__ dout, double* __restrict__ part) {
    __shared__ double red[NT];
    const int p = blockIdx.x * NT + threadIdx.x;
    double sq = 0.0;
    if (p < P) {
        const float d = out[p] - target[p];
        dout[p] = f2bf(2.0f * d / (float)P);
        sq = (double)d * (double)d;
    }
    red[threadIdx.x] = sq;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_mse_final(const double* __restrict__ part, int n, int P, float* __restrict__ loss) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += part[i];
    *loss = (float)(s / (double)P);
}

// AdamW (Loshchilov & Hutter) on the flat float32 master parameters.
__global__ void k_adamw(float* __restrict__ th, float* __restrict__ m, float* __restrict__ v,
                        const float* __restrict__ g, int n, float lr, float b1, float b2, float eps, float wd,
                        float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / bc1, vh = vi / bc2;
    th[i] = th[i] - lr * (mh / (sqrtf(vh) + eps) + wd * th[i]);
}

// bf16 copies of one layer's weights: wt [COUTP][KP] and wtT [CINP][K2P]
__global__ void k_cast_layer(const float* __restrict__ th, int cin, int cout, uint16_t* __restrict__ wt,
                             uint16_t* __restrict__ wtT) {
    const int K = 9 * cin, KP = rup16(K), COUTP = rup16(cout), CINP = rup16(cin), K2 = 9 * cout, K2P = rup16(K2);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < COUTP * KP) wt[i] = f2bf(th[i]);
    if (wtT && i < CINP * K2P) {
        const int ci = i / K2P, k2 = i - ci * K2P;
        float w = 0.0f;
        if (ci < cin && k2 < K2) {
            const int tap = k2 / cout, co = k2 - tap * cout;
            w = th[(size_t)co * KP + tap * cin + ci];
        }
        wtT[i] = f2bf(w);
    }
}

}  // namespace pc
}  // namespace wg

// ===========================================================================
// Host side
// ===========================================================================
namespace {

using namespace wg::pc;

thread_local std::string pc_err;

int pc_fail(int code, const std::string& m) {
    pc_err = m;
    return code;
}

#define PCK(call)                                                                   \
    do {                                                                            \
        hipError_t _e = (call);                                                     \
        if (_e != hipSuccess) return pc_fail(WG_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

struct Layer {
    int cin, cout, dil;
    int KP() const { return rup16(9 * cin); }
    int COUTP() const { return rup16(cout); }
    int CINP() const { return rup16(cin); }
    int K2P() const { return rup16(9 * cout); }
    int flat() const { return COUTP() * KP() + COUTP(); }  // [W padded | b padded]
    int nparams() const { return cout * cin * 9 + cout; }   // unpadded
};

constexpr int kLayers = 4;
const Layer kNet[kLayers] = {{2, 32, 4}, {32, 32, 16}, {32, 32, 8}, {32, 1, 1}};
constexpr int kChunk = 4096;  // pixels per weight-gradient block

}  // namespace

struct wg_pc {
    int max_n = 0, H = 70, W = 70;
    int device = 0;
    int off[kLayers + 1] = {};  // offsets of each layer in the flat buffers
    float *th = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    uint16_t *wt[kLayers] = {}, *wtT[kLayers] = {};
    uint16_t* x16 = nullptr;
    float* pre[3] = {};
    uint16_t* act[3] = {};
    uint16_t* dpre[3] = {};
    float* out = nullptr;
    uint16_t* dout = nullptr;
    float* part = nullptr;
    size_t part_floats = 0;
    double* mse_part = nullptr;
    float* loss = nullptr;
    int64_t step = 0;
    std::mutex mu;
};

namespace {

template <class T>
int dalloc(T** p, size_t n) {
    PCK(hipMalloc((void**)p, n * sizeof(T) + 16));
    return WG_OK;
}

int cast_all(wg_pc* pc, hipStream_t st) {
    for (int l = 0; l < kLayers; ++l) {
        const Layer& L = kNet[l];
        const int n = std::max(L.COUTP() * L.KP(), L.CINP() * L.K2P());
        hipLaunchKernelGGL(k_cast_layer, dim3((n + 255) / 256), dim3(256), 0, st, pc->th + pc->off[l], L.cin, L.cout,
                           pc->wt[l], l > 0 ? pc->wtT[l] : nullptr);
        PCK(hipGetLastError());
    }
    return WG_OK;
}

// torch Conv2d order (W[co][ci][ky][kx], then b[co]) <-> padded flat layout
void flat_index(std::vector<int>& idx) {
    idx.clear();
    int base = 0;
    for (int l = 0; l < kLayers; ++l) {
        const Layer& L = kNet[l];
        for (int co = 0; co < L.cout; ++co)
            for (int ci = 0; ci < L.cin; ++ci)
                for (int tap = 0; tap < 9; ++tap) idx.push_back(base + co * L.KP() + tap * L.cin + ci);
        for (int co = 0; co < L.cout; ++co) idx.push_back(base + L.COUTP() * L.KP() + co);
        base += L.flat();
    }
}

int forward(wg_pc* pc, int N, const float* x, hipStream_t st) {
    const int HW = pc->H * pc->W, P = N * HW;
    const dim3 g4((P + 63) / 64);
    hipLaunchKernelGGL(k_prep, dim3((P + 255) / 256), dim3(256), 0, st, x, N, HW, pc->x16);
    hipLaunchKernelGGL((k_fwd<2, 32, true>), g4, dim3(NT), 0, st, pc->x16, P, pc->H, pc->W, kNet[0].dil,
                       pc->wt[0], pc->th + pc->off[0] + kNet[0].COUTP() * kNet[0].KP(), pc->pre[0], pc->act[0]);
    hipLaunchKernelGGL((k_fwd<32, 32, true>), g4, dim3(NT), 0, st, pc->act[0], P, pc->H, pc->W, kNet[1].dil,
                       pc->wt[1], pc->th + pc->off[1] + kNet[1].COUTP() * kNet[1].KP(), pc->pre[1], pc->act[1]);
    hipLaunchKernelGGL((k_fwd<32, 32, true>), g4, dim3(NT), 0, st, pc->act[1], P, pc->H, pc->W, kNet[2].dil,
                       pc->wt[2], pc->th + pc->off[2] + kNet[2].COUTP() * kNet[2].KP(), pc->pre[2], pc->act[2]);
    hipLaunchKernelGGL((k_fwd<32, 1, false>), g4, dim3(NT), 0, st, pc->act[2], P, pc->H, pc->W, kNet[3].dil,
                       pc->wt[3], pc->th + pc->off[3] + kNet[3].COUTP() * kNet[3].KP(), pc->out, (uint16_t*)nullptr);
    PCK(hipGetLastError());
    return WG_OK;
}

int backward(wg_pc* pc, int N, const float* target, hipStream_t st) {
    const int HW = pc->H * pc->W, P = N * HW;
    const dim3 g4((P + 63) / 64);
    const int nb = (P + NT - 1) / NT;
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(NT), 0, st, pc->out, target, P, pc->dout, pc->mse_part);
    hipLaunchKernelGGL(k_mse_final, dim3(1), dim3(64), 0, st, pc->mse_part, nb, P, pc->loss);
    const int nchunk = (P + kChunk - 1) / kChunk;
    // layer 4 (32 -> 1)
    float* p4 = pc->part;
    hipLaunchKernelGGL((k_bwd_w<32, 1>), dim3(nchunk), dim3(NT), 0, st, pc->dout, pc->act[2], P, pc->H, pc->W,
                       kNet[3].dil, kChunk, p4);
    hipLaunchKernelGGL((k_reduce), dim3((kNet[3].flat() + 255) / 256), dim3(256), 0, st, p4, nchunk, kNet[3].flat(),
                       pc->grad + pc->off[3]);
    hipLaunchKernelGGL((k_bwd_data<32, 1>), g4, dim3(NT), 0, st, pc->dout, P, pc->H, pc->W, kNet[3].dil,
                       pc->wtT[3], pc->pre[2], pc->dpre[2]);
    // layer 3 (32 -> 32)
    hipLaunchKernelGGL((k_bwd_w<32, 32>), dim3(nchunk), dim3(NT), 0, st, pc->dpre[2], pc->act[1], P, pc->H, pc->W,
                       kNet[2].dil, kChunk, pc->part);
    hipLaunchKernelGGL((k_reduce), dim3((kNet[2].flat() + 255) / 256), dim3(256), 0, st, pc->part, nchunk,
                       kNet[2].flat(), pc->grad + pc->off[2]);
    hipLaunchKernelGGL((k_bwd_data<32, 32>), g4, dim3(NT), 0, st, pc->dpre[2], P, pc->H, pc->W, kNet[2].dil,
                       pc->wtT[2], pc->pre[1], pc->dpre[1]);
    // layer 2
    hipLaunchKernelGGL((k_bwd_w<32, 32>), dim3(nchunk), dim3(NT), 0, st, pc->dpre[1], pc->act[0], P, pc->H, pc->W,
                       kNet[1].dil, kChunk, pc->part);
    hipLaunchKernelGGL((k_reduce), dim3((kNet[1].flat() + 255) / 256), dim3(256), 0, st, pc->part, nchunk,
                       kNet[1].flat(), pc->grad + pc->off[1]);
    hipLaunchKernelGGL((k_bwd_data<32, 32>), g4, dim3(NT), 0, st, pc->dpre[1], P, pc->H, pc->W, kNet[1].dil,
                       pc->wtT[1], pc->pre[0], pc->dpre[0]);
    // layer 1 (2 -> 32)
    hipLaunchKernelGGL((k_bwd_w<2, 32>), dim3(nchunk), dim3(NT), 0, st, pc->dpre[0], pc->x16, P, pc->H, pc->W,
                       kNet[0].dil, kChunk, pc->part);
    hipLaunchKernelGGL((k_reduce), dim3((kNet[0].flat() + 255) / 256), dim3(256), 0, st, pc->part, nchunk,
                       kNet[0].flat(), pc->grad + pc->off[0]);
    PCK(hipGetLastError());
    return WG_OK;
}

int check_n(wg_pc* pc, int N) {
    if (N < 1 || N > pc->max_n) return pc_fail(WG_E_INVALID, "batch size must be in [1, max_batch]");
    return WG_OK;
}

}  // namespace

extern "C" {

const char* wg_pc_last_error(void) { return pc_err.c_str(); }

int32_t wg_pc_create(int32_t max_batch, int32_t height, int32_t width, uint64_t seed, wg_pc** out) {
    if (!out || max_batch < 1 || height < 1 || width < 1) return pc_fail(WG_E_INVALID, "invalid argument");
    if ((int64_t)max_batch * height * width > (1 << 26)) return pc_fail(WG_E_TOO_LARGE, "batch too large");
    wg_pc* pc = new wg_pc();
    pc->max_n = max_batch;
    pc->H = height;
    pc->W = width;
    PCK(hipGetDevice(&pc->device));
    for (int l = 0; l < kLayers; ++l) pc->off[l + 1] = pc->off[l] + kNet[l].flat();
    const int nf = pc->off[kLayers];
    const size_t P = (size_t)max_batch * height * width;
    int rc = 0;
    const int nchunk = (int)((P + kChunk - 1) / kChunk);
    pc->part_floats = (size_t)nchunk * kNet[1].flat();
    if (!rc) rc = dalloc(&pc->th, nf);
    if (!rc) rc = dalloc(&pc->m, nf);
    if (!rc) rc = dalloc(&pc->v, nf);
    if (!rc) rc = dalloc(&pc->grad, nf);
    for (int l = 0; l < kLayers && !rc; ++l) {
        rc = dalloc(&pc->wt[l], (size_t)kNet[l].COUTP() * kNet[l].KP());
        if (!rc) rc = dalloc(&pc->wtT[l], (size_t)kNet[l].CINP() * kNet[l].K2P());
    }
    if (!rc) rc = dalloc(&pc->x16, 2 * P);
    for (int l = 0; l < 3 && !rc; ++l) {
        rc = dalloc(&pc->pre[l], 32 * P);
        if (!rc) rc = dalloc(&pc->act[l], 32 * P);
        if (!rc) rc = dalloc(&pc->dpre[l], 32 * P);
    }
    if (!rc) rc = dalloc(&pc->out, P);
    if (!rc) rc = dalloc(&pc->dout, P);
    if (!rc) rc = dalloc(&pc->part, pc->part_floats);
    if (!rc) rc = dalloc(&pc->mse_part, (P + NT - 1) / NT);
    if (!rc) rc = dalloc(&pc->loss, 1);
    if (rc) {
        wg_pc_destroy(pc);
        return rc;
    }
    // Deterministic initialisation (PyTorch Conv2d default: U(-1/sqrt(fan_in), 1/sqrt(fan_in))).
    std::vector<float> h((size_t)nf, 0.0f);
    uint64_t s = seed ? seed : 0x9e3779b97f4a7c15ull;
    auto next = [&]() {  // splitmix64 -> [0, 1)
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    std::vector<int> idx;
    flat_index(idx);
    size_t k = 0;
    for (int l = 0; l < kLayers; ++l) {
        const double bound = 1.0 / std::sqrt(9.0 * kNet[l].cin);
        for (int i = 0; i < kNet[l].nparams(); ++i) h[idx[k++]] = (float)((2.0 * next() - 1.0) * bound);
    }
    PCK(hipMemcpy(pc->th, h.data(), nf * sizeof(float), hipMemcpyHostToDevice));
    PCK(hipMemset(pc->m, 0, nf * sizeof(float)));
    PCK(hipMemset(pc->v, 0, nf * sizeof(float)));
    if ((rc = cast_all(pc, nullptr))) {
        wg_pc_destroy(pc);
        return rc;
    }
    PCK(hipDeviceSynchronize());
    *out = pc;
    return WG_OK;
}

void wg_pc_destroy(wg_pc* pc) {
    if (!pc) return;
    (void)hipDeviceSynchronize();
    void* bufs[] = {pc->th, pc->m, pc->v, pc->grad, pc->x16, pc->out, pc->dout, pc->part, pc->mse_part, pc->loss};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (int l = 0; l < kLayers; ++l) {
        if (pc->wt[l]) (void)hipFree(pc->wt[l]);
        if (pc->wtT[l]) (void)hipFree(pc->wtT[l]);
    }
    for (int l = 0; l < 3; ++l) {
        if (pc->pre[l]) (void)hipFree(pc->pre[l]);
        if (pc->act[l]) (void)hipFree(pc->act[l]);
        if (pc->dpre[l]) (void)hipFree(pc->dpre[l]);
    }
    delete pc;
}

int32_t wg_pc_num_layers(void) { return kLayers; }

int32_t wg_pc_layer_shape(int32_t layer, int32_t* cin, int32_t* cout, int32_t* dilation) {
    if (layer < 0 || layer >= kLayers || !cin || !cout || !dilation) return pc_fail(WG_E_INVALID, "invalid argument");
    *cin = kNet[layer].cin;
    *cout = kNet[layer].cout;
    *dilation = kNet[layer].dil;
    return WG_OK;
}

int32_t wg_pc_num_params(void) {
    int n = 0;
    for (int l = 0; l < kLayers; ++l) n += kNet[l].nparams();
    return n;
}

int32_t wg_pc_get_params(wg_pc* pc, float* host, int32_t n) {
    if (!pc || !host || n != wg_pc_num_params()) return pc_fail(WG_E_INVALID, "invalid argument");
    std::lock_guard<std::mutex> lk(pc->mu);
    std::vector<float> h((size_t)pc->off[kLayers]);
    PCK(hipDeviceSynchronize());
    PCK(hipMemcpy(h.data(), pc->th, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    std::vector<int> idx;
    flat_index(idx);
    for (int i = 0; i < n; ++i) host[i] = h[idx[i]];
    return WG_OK;
}

int32_t wg_pc_set_params(wg_pc* pc, const float* host, int32_t n) {
    if (!pc || !host || n != wg_pc_num_params()) return pc_fail(WG_E_INVALID, "invalid argument");
    std::lock_guard<std::mutex> lk(pc->mu);
    std::vector<float> h((size_t)pc->off[kLayers], 0.0f);
    std::vector<int> idx;
    flat_index(idx);
    for (int i = 0; i < n; ++i) h[idx[i]] = host[i];
    PCK(hipDeviceSynchronize());
    PCK(hipMemcpy(pc->th, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    PCK(hipMemset(pc->m, 0, h.size() * sizeof(float)));
    PCK(hipMemset(pc->v, 0, h.size() * sizeof(float)));
    pc->step = 0;
    int rc = cast_all(pc, nullptr);
    if (rc) return rc;
    PCK(hipDeviceSynchronize());
    return WG_OK;
}

int32_t wg_pc_forward_dev(wg_pc* pc, int32_t N, const float* x_dev, float* y_dev, void* stream) {
    if (!pc || !x_dev || !y_dev) return pc_fail(WG_E_INVALID, "invalid argument");
    std::lock_guard<std::mutex> lk(pc->mu);
    int rc;
    if ((rc = check_n(pc, N))) return rc;
    hipStream_t st = (hipStream_t)stream;
    if ((rc = forward(pc, N, x_dev, st))) return rc;
    PCK(hipMemcpyAsync(y_dev, pc->out, (size_t)N * pc->H * pc->W * sizeof(float), hipMemcpyDeviceToDevice, st));
    return WG_OK;
}

int32_t wg_pc_loss_grad(wg_pc* pc, int32_t N, const float* x_dev, const float* target_dev, float* loss_host,
                        float* grad_host, void* stream) {
    if (!pc || !x_dev || !target_dev || !loss_host || !grad_host) return pc_fail(WG_E_INVALID, "invalid argument");
    std::lock_guard<std::mutex> lk(pc->mu);
    int rc;
    if ((rc = check_n(pc, N))) return rc;
    hipStream_t st = (hipStream_t)stream;
    if ((rc = forward(pc, N, x_dev, st))) return rc;
    if ((rc = backward(pc, N, target_dev, st))) return rc;
    PCK(hipStreamSynchronize(st));
    PCK(hipMemcpy(loss_host, pc->loss, sizeof(float), hipMemcpyDeviceToHost));
    std::vector<float> h((size_t)pc->off[kLayers]);
    PCK(hipMemcpy(h.data(), pc->grad, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    std::vector<int> idx;
    flat_index(idx);
    for (int i = 0; i < wg_pc_num_params(); ++i) grad_host[i] = h[idx[i]];
    return WG_OK;
}

int32_t wg_pc_train_step_dev(wg_pc* pc, int32_t N, const float* x_dev, const float* target_dev, float lr,
                             float weight_decay, float* loss_dev, void* stream) {
    if (!pc || !x_dev || !target_dev || !(lr >= 0.0f) || !(weight_decay >= 0.0f))
        return pc_fail(WG_E_INVALID, "invalid argument");
    std::lock_guard<std::mutex> lk(pc->mu);
    int rc;
    if ((rc = check_n(pc, N))) return rc;
    hipStream_t st = (hipStream_t)stream;
    if ((rc = forward(pc, N, x_dev, st))) return rc;
    if ((rc = backward(pc, N, target_dev, st))) return rc;
    pc->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = (float)(1.0 - std::pow((double)b1, (double)pc->step));
    const float bc2 = (float)(1.0 - std::pow((double)b2, (double)pc->step));
    const int nf = pc->off[kLayers];
    hipLaunchKernelGGL(k_adamw, dim3((nf + 255) / 256), dim3(256), 0, st, pc->th, pc->m, pc->v, pc->grad, nf, lr, b1,
                       b2, 1e-8f, weight_decay, bc1, bc2);
    PCK(hipGetLastError());
    if ((rc = cast_all(pc, st))) return rc;
    if (loss_dev) PCK(hipMemcpyAsync(loss_dev, pc->loss, sizeof(float), hipMemcpyDeviceToDevice, st));
    return WG_OK;
}

}  // extern "C"
