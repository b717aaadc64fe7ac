// Misfit and gradient-fold kernels of the wave engine (gfx950).
//   residual / loss  wave_wgpu.rs:13-54 (RESIDUAL_WGSL), 852-854, 960-965
//   shot sum, clamp mask and fold  wave_wgpu.rs:974-1005
// These are small HBM-bound passes over the gathers [B, shots, nt, nrec] and
// the per-unit padded gradients; the float32 tree of the reference's 256-wide
// residual workgroup is reproduced exactly so the loss matches bit for bit.
#include <algorithm>
#include "wave_kernels.h"

#pragma clang fp contract(off)

namespace wg {

// ---------------------------------------------------------------------------
// Misfit pieces.
// ---------------------------------------------------------------------------
// Per-model observed energy in f64 (wave_wgpu.rs:852; summation order is a
// tree, not the reference's sequential loop) and scale = f32(1 / energy).
constexpr int kEnergyThreads = 1024;

__global__ __launch_bounds__(kEnergyThreads) void k_energy(const float* obs, long long per, double* energy,
                                                           float* scale) {
    constexpr int NTH = kEnergyThreads;
    __shared__ double part[NTH];
    const float* o = obs + (size_t)blockIdx.x * per;
    double acc = 0.0;
    long long i = threadIdx.x;
    // eight loads in flight, added in the same order as one at a time
    for (; i + 7 * NTH < per; i += 8 * NTH) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = o[i + u * NTH];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += (double)v[u] * (double)v[u];
    }
    for (; i < per; i += NTH) acc += (double)o[i] * (double)o[i];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int st = NTH / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) part[threadIdx.x] += part[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double e = part[0];
        energy[blockIdx.x] = e;
        scale[blockIdx.x] = e > 0.0 ? (float)(1.0 / e) : 0.0f;
    }
}

// RESIDUAL_WGSL (wave_wgpu.rs:28-53): one 256-wide workgroup per chunk of
// 256 elements of one model's [shots, nt, nrec] gathers; residual (p-d)*scale
// (not written when resid is null) and the float32 tree sum of 0.5 d^2 scale.
__global__ __launch_bounds__(256) void k_residual(const float* pred, const float* obs,
                                                  const float* scale, long long per,
                                                  long long chunks, float* resid,
                                                  float* partials) {
    __shared__ float wg_sums[256];
    const long long blk = blockIdx.x;
    const long long b = blk / chunks;
    const long long c = blk - b * chunks;
    const long long i = c * 256 + threadIdx.x;
    const float sc = scale[b];
    float e = 0.0f;
    if (i < per) {
        const size_t o = (size_t)b * per + i;
        const float d = pred[o] - obs[o];
        if (resid) resid[o] = d * sc;  // null: the loss alone
        e = 0.5f * d * d * sc;
    }
    wg_sums[threadIdx.x] = e;
    __syncthreads();
    for (int st = 128; st > 0; st /= 2) {
        if ((int)threadIdx.x < st) wg_sums[threadIdx.x] += wg_sums[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blk] = wg_sums[0];
}

// Host-side f64 sum of the partials (wave_wgpu.rs:960-965), one thread per
// model, chunks in order.  NaN for zero observed energy.
__global__ void k_loss_sum(const float* partials, long long chunks, int B, const double* energy,
                           float* loss) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    double acc = 0.0;
    const float* pb = partials + (size_t)b * chunks;
    long long c = 0;
    for (; c + 8 <= chunks; c += 8) {  // eight loads in flight, added in order
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = pb[c + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += (double)v[u];
    }
    for (; c < chunks; ++c) acc += (double)pb[c];
    loss[b] = energy[b] > 0.0 ? (float)acc : __builtin_nanf("");
}

// k_loss_sum for up to kLossLdsMax partials per model: one workgroup per
// model copies its partials to LDS, then one thread adds them in chunk order
// (k_loss_sum's sum bit for bit), reading LDS instead of waiting on HBM.
constexpr size_t kLossLdsMax = 64 * 1024;

__global__ __launch_bounds__(256) void k_loss_sum_lds(const float* partials, long long chunks,
                                                      const double* energy, float* loss) {
    extern __shared__ float part_s[];
    const int b = blockIdx.x;
    const float* pb = partials + (size_t)b * chunks;
    for (long long c = threadIdx.x; c < chunks; c += 256) part_s[c] = pb[c];
    __syncthreads();
    if (threadIdx.x != 0) return;
    double acc = 0.0;
    long long c = 0;
    for (; c + 8 <= chunks; c += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = part_s[c + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += (double)v[u];
    }
    for (; c < chunks; ++c) acc += (double)part_s[c];
    loss[b] = energy[b] > 0.0 ? (float)acc : __builtin_nanf("");
}

// wave_wgpu.rs:974-981 then fold_node_gradient (:993-1005): one thread per
// visible node; the padded cells folding onto it are visited in row-major
// order and summed in f64, as the reference's sequential fold does.
__global__ void k_fold(Grid g, const float* grad_unit, const float* vel, const double* energy,
                       int B, float* out, CellLayout lay) {
    const int nvis = g.nvx * g.nvz;
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)B * nvis) return;
    const int b = (int)(gid / nvis);
    const int vi = (int)(gid - (long long)b * nvis);
    const int vz = vi / g.nvx, vx = vi - (vi / g.nvx) * g.nvx;
    const float v = vel[(size_t)b * nvis + vi];
    const float mask = (v > g.vmin && v < g.vmax) ? 1.0f : 0.0f;
    const int zlo = vz == 0 ? 0 : vz + g.pad;
    const int zhi = vz == g.nvz - 1 ? g.nz - 1 : vz + g.pad;
    const int xlo = vx == 0 ? 0 : vx + g.pad;
    const int xhi = vx == g.nvx - 1 ? g.nx - 1 : vx + g.pad;
    const size_t cells = lay.ustride;
    const float* gb = grad_unit + (size_t)b * g.shots * cells + lay.off;
    double acc = 0.0;
    for (int z = zlo; z <= zhi; ++z) {
        int x = xlo;
        // four cells' shot sums formed before they are added, in order
        for (; x + 3 <= xhi; x += 4) {
            float sum_s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int s = 0; s < g.shots; ++s) {
                const float* r = gb + (size_t)s * cells + z * lay.pitch + x;
#pragma unroll
                for (int u = 0; u < 4; ++u) sum_s[u] += r[u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) acc += (double)(sum_s[u] * mask);
        }
        for (; x <= xhi; ++x) {
            const int c = z * lay.pitch + x;
            float sum_s = 0.0f;
            for (int s = 0; s < g.shots; ++s) sum_s += gb[(size_t)s * cells + c];
            acc += (double)(sum_s * mask);
        }
    }
    out[gid] = energy && !(energy[b] > 0.0) ? __builtin_nanf("") : (float)acc;
}

// k_fold for grids whose padded cells fit in LDS: one workgroup per model.
// The shot sums of every padded cell (from 0.0f, shots in order) go to LDS
// first, then each thread folds whole visible cells from LDS in k_fold's
// (z, x) order, so the result is k_fold's bit for bit.  k_fold gives each
// border cell (up to (PAD + 1)^2 padded cells times the shots) to one thread
// reading HBM; here the long folds read LDS and the HBM reads are spread
// over the workgroup.
constexpr int kFoldThreads = 1024;
constexpr size_t kFoldLdsMax = 160 * 1024;

__global__ __launch_bounds__(kFoldThreads) void k_fold_lds(Grid g, const float* grad_unit, const float* vel,
                                                           const double* energy, float* out) {
    extern __shared__ float ssum[];
    const int b = blockIdx.x;
    const int cells = g.nx * g.nz;
    const float* gb = grad_unit + (size_t)b * g.shots * cells;
    constexpr int U = 4;
    for (int c0 = threadIdx.x; c0 < cells; c0 += U * kFoldThreads) {
        float sum_s[U] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 5
        for (int s = 0; s < g.shots; ++s) {
            const float* r = gb + (size_t)s * cells;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int c = c0 + u * kFoldThreads;
                if (c < cells) sum_s[u] += r[c];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int c = c0 + u * kFoldThreads;
            if (c < cells) ssum[c] = sum_s[u];
        }
    }
    __syncthreads();
    const int nvis = g.nvx * g.nvz;
    const bool bad = energy && !(energy[b] > 0.0);
    for (int vi = threadIdx.x; vi < nvis; vi += kFoldThreads) {
        const int vz = vi / g.nvx, vx = vi - vz * g.nvx;
        const float v = vel[(size_t)b * nvis + vi];
        const float mask = (v > g.vmin && v < g.vmax) ? 1.0f : 0.0f;
        const int zlo = vz == 0 ? 0 : vz + g.pad;
        const int zhi = vz == g.nvz - 1 ? g.nz - 1 : vz + g.pad;
        const int xlo = vx == 0 ? 0 : vx + g.pad;
        const int xhi = vx == g.nvx - 1 ? g.nx - 1 : vx + g.pad;
        double acc = 0.0;
        for (int z = zlo; z <= zhi; ++z) {
            const float* row = ssum + z * g.nx;
            int x = xlo;
            for (; x + 3 <= xhi; x += 4) {
                const float a0 = row[x], a1 = row[x + 1], a2 = row[x + 2], a3 = row[x + 3];
                acc += (double)(a0 * mask);
                acc += (double)(a1 * mask);
                acc += (double)(a2 * mask);
                acc += (double)(a3 * mask);
            }
            for (; x <= xhi; ++x) acc += (double)(row[x] * mask);
        }
        out[(size_t)b * nvis + vi] = bad ? __builtin_nanf("") : (float)acc;
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_energy(const float* obs, long long per, int B, double* energy, float* scale,
                         hipStream_t st) {
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_energy, dim3(B), dim3(kEnergyThreads), 0, st, obs, per, energy, scale);
    return hipGetLastError();
}

hipError_t launch_residual(const float* pred, const float* obs, const float* scale, long long per,
                           int B, float* resid, float* partials, hipStream_t st) {
    if (B <= 0) return hipSuccess;
    const long long chunks = (per + 255) / 256;
    hipLaunchKernelGGL(k_residual, dim3((unsigned)(chunks * B)), dim3(256), 0, st, pred, obs, scale,
                       per, chunks, resid, partials);
    return hipGetLastError();
}

hipError_t launch_loss_sum(const float* partials, long long chunks, int B, const double* energy,
                           float* loss, hipStream_t st) {
    if (B <= 0) return hipSuccess;
    const size_t lds = (size_t)chunks * sizeof(float);
    if (chunks > 0 && lds <= kLossLdsMax) {
        hipError_t e = hipFuncSetAttribute((const void*)k_loss_sum_lds, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_loss_sum_lds, dim3((unsigned)B), dim3(256), lds, st, partials, chunks, energy, loss);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_loss_sum, dim3((B + 63) / 64), dim3(64), 0, st, partials, chunks, B, energy,
                       loss);
    return hipGetLastError();
}

hipError_t launch_fold(const Grid& g, const float* grad_unit, const float* vel, const double* energy,
                       int B, float* grad_vis, hipStream_t st) {
    return launch_fold(g, grad_unit, vel, energy, B, grad_vis, plain_layout(g), st);
}

hipError_t launch_fold(const Grid& g, const float* grad_unit, const float* vel, const double* energy,
                       int B, float* grad_vis, const CellLayout& lay, hipStream_t st) {
    if (B <= 0) return hipSuccess;
    const size_t lds = (size_t)g.nx * g.nz * sizeof(float);
    const bool plain = lay.pitch == g.nx && lay.off == 0 && lay.ustride == (size_t)g.nx * g.nz;
    if (plain && lds <= kFoldLdsMax) {
        hipError_t e = hipFuncSetAttribute((const void*)k_fold_lds, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_fold_lds, dim3((unsigned)B), dim3(kFoldThreads), lds, st, g, grad_unit, vel, energy,
                           grad_vis);
        return hipGetLastError();
    }
    const long long n = (long long)B * g.nvx * g.nvz;
    hipLaunchKernelGGL(k_fold, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, grad_unit, vel,
                       energy, B, grad_vis, lay);
    return hipGetLastError();
}

// Zero fill (grid-stride), used instead of memset nodes so that the engine's
// work can be captured in a HIP graph.
__global__ void k_zero(float* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

hipError_t launch_zero(float* p, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

// Marks every loss NaN when a device-side wait gave up (err != 0).
// A work-queue wait that gave up (its predecessor's state never arrived) sets
// *err; then every model's loss and folded gradient of that call are NaN, so
// no caller can take a gradient built on an unfinished state for a good one.
__global__ void k_flag_nan(const int* err, float* loss, int B, float* grad, long long nvis) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (*err == 0) return;
    if (i < B) loss[i] = __builtin_nanf("");
    if (grad && i < (long long)B * nvis) grad[i] = __builtin_nanf("");
}

hipError_t launch_flag_nan(const int* err, float* loss, int B, float* grad, long long nvis, hipStream_t st) {
    if (B <= 0) return hipSuccess;
    const long long n = std::max<long long>(B, grad ? (long long)B * nvis : 0);
    hipLaunchKernelGGL(k_flag_nan, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, err, loss, B, grad, nvis);
    return hipGetLastError();
}

// Causal low-pass along time (python/wgnet/physics.py:48-57: conv1d of each
// receiver trace with the taps h, left-padded by L - 1 zeros):
//   y[n][t][r] = sum_{k = max(0, L-1-t)}^{L-1} h[k] x[n][t + k - (L - 1)][r].
// One thread per output sample; threads of a wave run along r, so every tap
// reads consecutive floats.  HBM-bound (x is re-read from L2 per tap).
__global__ __launch_bounds__(256) void k_causal_fir(const float* __restrict__ x, long long total, int nt, int nrec,
                                                    const float* __restrict__ h, int L, float* __restrict__ y) {
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= total) return;
    const int r = (int)(idx % nrec);
    const long long q = idx / nrec;
    const int t = (int)(q % nt);
    const long long n = q / nt;
    const float* xs = x + (size_t)n * nt * nrec + r;
    float acc = 0.0f;
    for (int k = (L - 1 - t > 0 ? L - 1 - t : 0); k < L; ++k) acc += h[k] * xs[(size_t)(t + k - (L - 1)) * nrec];
    y[idx] = acc;
}

hipError_t launch_causal_fir(const float* x, long long nseries, int nt, int nrec, const float* h, int L, float* y,
                             hipStream_t st) {
    const long long total = nseries * nt * nrec;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_causal_fir, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, x, total, nt, nrec, h, L,
                       y);
    return hipGetLastError();
}

// Source illumination of each visible cell: sum over the model's shots and
// all steps of the squared snapshot 2 v dt^2 w at the cell's own padded
// position (the pseudo-Hessian diagonal that splatAdjoint.wgsl:113-117
// accumulates as stored^2 and never reads back), float32 in (shot, t) order.
// snaps [units of the chunk, nt, inner]; one thread per visible cell.
__global__ void k_illum(Grid g, const float* __restrict__ snaps, int nb, float* __restrict__ out) {
    const int nvis = g.nvx * g.nvz;
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)nb * nvis) return;
    const int b = (int)(gid / nvis);
    const int vi = (int)(gid - (long long)b * nvis);
    const int vz = vi / g.nvx, vx = vi - (vi / g.nvx) * g.nvx;
    const size_t q = (size_t)(vz + g.pad - 2) * g.ni + (vx + g.pad - 2);
    float acc = 0.0f;
    for (int s = 0; s < g.shots; ++s) {
        const float* sn = snaps + ((size_t)b * g.shots + s) * g.nt * g.inner + q;
        for (int t = 0; t < g.nt; ++t) {
            const float v = sn[(size_t)t * g.inner];
            acc += v * v;
        }
    }
    out[gid] = acc;
}

hipError_t launch_illum(const Grid& g, const float* snaps, int nb, float* out, hipStream_t st) {
    if (nb <= 0) return hipSuccess;
    const long long n = (long long)nb * g.nvx * g.nvz;
    hipLaunchKernelGGL(k_illum, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, snaps, nb, out);
    return hipGetLastError();
}

}  // namespace wg
