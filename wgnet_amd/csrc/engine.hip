// Host runtime of the wave engine: the C-ABI declared in include/wgnet_hip.h.
// Replaces NativeWaveEngine (crates/wave-gauss-rs/src/wave_wgpu.rs) and the
// PyO3 surface of crates/wave-gauss-rs/src/lib.rs.
//
// Differences from the reference engine, by design:
//  - one kernel launch per batch of (model, shot) units runs all nt steps
//    (the reference dispatches every step, wave_wgpu.rs:820-840);
//  - models are batched (the reference loops over models in Python,
//    physics.py:62-75) and the device path is asynchronous on a caller stream;
//  - misfit, shot sum, clamp mask and fold run on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/wgnet_hip.h"
#include "wave_kernels.h"

#pragma clang fp contract(off)

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCK(call)                                                                      \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(WG_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kFdPad = 2;  // geology-core lib.rs:9

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
};

// Rust's saturating `f32 as usize` (NaN and negatives -> 0).
int64_t sat_floor(float g) {
    const float f = std::floor(g);
    if (!(f > 0.0f)) return 0;
    if (f > 9.0e18f) return INT64_MAX;
    return (int64_t)f;
}

// Bilinear corners of a station (wave_wgpu.rs:217-230 and 766-780).
void corners(float x, float z, float h, int pad, int nx, int nz, int cell[4], float w[4]) {
    const float gx = x / h + (float)pad;
    const float gz = z / h + (float)pad;
    int64_t x0 = std::min<int64_t>(sat_floor(gx), nx - 2);
    int64_t z0 = std::min<int64_t>(sat_floor(gz), nz - 2);
    const float fx = gx - (float)x0;
    const float fz = gz - (float)z0;
    cell[0] = (int)(z0 * nx + x0);
    cell[1] = (int)(z0 * nx + x0 + 1);
    cell[2] = (int)((z0 + 1) * nx + x0);
    cell[3] = (int)((z0 + 1) * nx + x0 + 1);
    w[0] = (1.0f - fx) * (1.0f - fz);
    w[1] = fx * (1.0f - fz);
    w[2] = (1.0f - fx) * fz;
    w[3] = fx * fz;
}

}  // namespace

// Fixed device scratch of wg_loss_noalloc_dev (the per-model loss scale, the
// observed energy when the caller passes none, and the residual tree's
// partial sums).  Each device holds its own copy for the life of the process;
// host code reaches it only through the address hipGetSymbolAddress returns
// for the current device (scratch_setup below), never through this name.
constexpr size_t kScratchBytes = (size_t)4 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};  // device address of `scratch` per device

// Looks up (once per device, with that device current) the device address of
// `scratch`.
int scratch_setup(int dev, void** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (dev < 0 || dev >= kMaxDevices) return fail(WG_E_DEVICE, "device index out of range for the scratch table");
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HIPCK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(WG_E_DEVICE, "hipGetSymbolAddress returned a null scratch address");
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return WG_OK;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace

struct TimedSpan {
    hipEvent_t a, b;
    int64_t n;  // kernel launches between a and b
};

struct wg_engine {
    wg_config cfg{};
    wg::Grid g{};
    int device = 0;
    std::vector<float> src_xz, wavelet, prof;
    std::vector<int> src_cell, rec_cell, band_off, ent_r;
    std::vector<float> src_w, rec_w, ent_w;
    int band_z0 = 0, band_rows = 0;
    // device tables
    float* d_prof = nullptr;
    int* d_src_cell = nullptr;
    float* d_src_w = nullptr;
    int* d_rec_cell = nullptr;
    float* d_rec_w = nullptr;
    int* d_band_off = nullptr;
    int* d_ent_r = nullptr;
    float* d_ent_w = nullptr;
    float* d_wavelet = nullptr;
    // workspace, grown on demand
    DevBuf snaps, traces, resid, gunit, partials, scale, energy, vel, obs, loss, grad;
    // HBM-streamed path (grids too large for one workgroup): step state,
    // checkpoints every seg_len steps, one segment of snapshots.
    bool streamed = false;
    DevBuf fstate, astate, ckpt, vpad;  // vpad: clamped velocity on the padded grid (streamed steps)
    // storage layout of the streamed per-unit arrays (wg::StepParams::sp ...)
    // and of the snapshots; plain (nx, 0, nx * nz) / (ni, inner) otherwise
    wg::CellLayout slay{};
    int s_ni = 0, s_soff = 0;  // snapshot row pitch; offset of inner cell (0, 0) in a snapshot block
    size_t s_inner = 0;
    hipStream_t stream = nullptr;  // host-pointer calls
    hipStream_t side = nullptr;    // second stream of the overlapped chunk pipeline
    hipEvent_t ev_pipe[2] = {nullptr, nullptr};
    int cus = 256;                 // compute units of the device
    bool overlap = true;           // overlap chunk forwards with the previous chunk's adjoint
    int time_block = 2;            // adjoint steps per streamed launch (1 or 2)
    int fwd_block = 2;             // forward steps per streamed launch (1 or 2)
    int qparts = 4;                // time parts per unit of the resident work queues (0: chunk pipeline)
    int qspin = 1 << 24;           // polls before a work-queue wait gives up
    // row bands: grids too large for one workgroup whose units split into
    // bands that fit run the band kernels (one workgroup per band, edge rows
    // exchanged every step) instead of the per-step HBM-streamed kernels
    int bands = 0;                 // bands per unit (0: the grid does not split)
    int band_bounds[wg::kMaxBands + 1] = {};
    int rt_bands = 0;              // the runtime-geometry kernels' split (0: none fits)
    int rt_bounds[wg::kMaxBands + 1] = {};
    int band_on = 1;               // 0: streamed kernels, 1: band kernels (compiled where the grid has them), 2: runtime-geometry band kernels
    DevBuf bxbuf[2], bqflags[2];   // per band-launch slot (one per stream in flight): exchange words; [next, err]
    DevBuf mtable, mnext, mslots;  // one-launch batches led by this engine: group table, [counter, slot flags..., err], snapshot slots
    DevBuf qstate, qflags;         // work queues: per-unit part state; counters, error flag, done flags
    hipStream_t last = nullptr;    // stream of the last enqueued call
    bool last_valid = false;
    size_t snap_budget = 0;
    // per-(model, shot) wavelets [B, shots, nt] of the call in progress, or
    // null for the engine wavelet (set by WavScope under the engine mutex)
    const float* wav_dev = nullptr;
    int tile = 0;  // tile shape of the HBM-streamed kernels (wg::StepParams::tile)
    void* scratch_ptr = nullptr;  // this device's address of the fixed scratch array
    // per-unit source corners and activity mask of the call in progress
    // (uneven batches), or null (set by UnitScope under the engine mutex)
    const int* unit_src_cell = nullptr;
    const float* unit_src_w = nullptr;
    const unsigned char* unit_active = nullptr;
    std::mutex mu;
    // kernel timing: (start, stop, launches) records of forward / adjoint launches
    bool timing = false;
    std::vector<TimedSpan> ev_fwd, ev_adj;

    wg::Geometry geo() const {
        wg::Geometry G;
        G.prof = d_prof;
        G.src_cell = d_src_cell;
        G.src_w = d_src_w;
        G.rec_cell = d_rec_cell;
        G.rec_w = d_rec_w;
        G.band_off = d_band_off;
        G.ent_r = d_ent_r;
        G.ent_w = d_ent_w;
        G.band_z0 = band_z0;
        G.band_rows = band_rows;
        G.n_ent = (int)ent_r.size();
        G.src_by_unit = 0;
        G.active = unit_active;
        if (unit_src_cell) {
            G.src_cell = unit_src_cell;
            G.src_w = unit_src_w;
            G.src_by_unit = 1;
        }
        return G;
    }
};

namespace {

// Waits for the previous call's stream when a call arrives on another one, so
// the shared workspace is never reused while still in flight.
int fence(wg_engine* e, hipStream_t st) {
    if (e->last_valid && e->last != st) HIPCK(hipStreamSynchronize(e->last));
    e->last = st;
    e->last_valid = true;
    return WG_OK;
}

int ensure(wg_engine* e, DevBuf& b, size_t bytes) {
    if (b.n >= bytes) return WG_OK;
    if (e->last_valid) HIPCK(hipStreamSynchronize(e->last));
    if (b.p) HIPCK(hipFree(b.p));
    b.p = nullptr;
    b.n = 0;
    HIPCK(hipMalloc(&b.p, bytes));
    b.n = bytes;
    return WG_OK;
}

template <typename T>
int upload(std::vector<T>& host, T** dev) {
    if (*dev) HIPCK(hipFree(*dev));
    *dev = nullptr;
    const size_t n = std::max<size_t>(host.size(), 1);
    HIPCK(hipMalloc((void**)dev, n * sizeof(T)));
    if (!host.empty()) HIPCK(hipMemcpy(*dev, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
    return WG_OK;
}

// Source corner tables from the current positions (wave_wgpu.rs:761-785).
void build_sources(wg_engine* e) {
    const int S = e->cfg.shots;
    e->src_cell.assign(4 * S, 0);
    e->src_w.assign(4 * S, 0.0f);
    for (int s = 0; s < S; ++s)
        corners(e->src_xz[2 * s], e->src_xz[2 * s + 1], e->cfg.h, e->g.pad, e->g.nx, e->g.nz, &e->src_cell[4 * s],
                &e->src_w[4 * s]);
}

// Receivers at (k h, h), k < nvx (wave_wgpu.rs:215), their corners and the
// per-cell (receiver, weight) lists with zero weights dropped, ordered by
// (cell, receiver) (wave_wgpu.rs:232-270), stored for the rows they touch.
void build_receivers(wg_engine* e) {
    const wg::Grid& g = e->g;
    const int R = g.nrec;
    e->rec_cell.assign(4 * R, 0);
    e->rec_w.assign(4 * R, 0.0f);
    struct Ent {
        int cell, r;
        float w;
    };
    std::vector<Ent> ents;
    for (int r = 0; r < R; ++r) {
        corners((float)r * e->cfg.h, e->cfg.h, e->cfg.h, g.pad, g.nx, g.nz, &e->rec_cell[4 * r], &e->rec_w[4 * r]);
        for (int k = 0; k < 4; ++k)
            if (e->rec_w[4 * r + k] != 0.0f) ents.push_back({e->rec_cell[4 * r + k], r, e->rec_w[4 * r + k]});
    }
    std::stable_sort(ents.begin(), ents.end(),
                     [](const Ent& a, const Ent& b) { return a.cell != b.cell ? a.cell < b.cell : a.r < b.r; });
    e->ent_r.clear();
    e->ent_w.clear();
    if (ents.empty()) {
        e->band_z0 = 0;
        e->band_rows = 0;
        e->band_off.assign(1, 0);
        return;
    }
    const int z0 = ents.front().cell / g.nx, z1 = ents.back().cell / g.nx;
    e->band_z0 = z0;
    e->band_rows = z1 - z0 + 1;
    e->band_off.assign((size_t)e->band_rows * g.nx + 1, 0);
    for (const Ent& en : ents) {
        e->band_off[en.cell - z0 * g.nx + 1]++;
        e->ent_r.push_back(en.r);
        e->ent_w.push_back(en.w);
    }
    for (size_t i = 1; i < e->band_off.size(); ++i) e->band_off[i] += e->band_off[i - 1];
}

int check_fits(const wg::Grid& g, int lds_limit, std::string& why) {
    if (wg::slots_needed(g) < 0) {
        why = "grid has more inner cells than the resident kernels hold";
        return WG_E_TOO_LARGE;
    }
    if (g.nrec > 512) {
        why = "more than 512 receivers";
        return WG_E_TOO_LARGE;
    }
    if (g.ni >= 511 || g.nzi >= 511) {
        why = "grid wider than 510 inner cells";
        return WG_E_TOO_LARGE;
    }
    wg::Geometry dummy{};
    dummy.band_rows = 2;
    dummy.n_ent = 4 * g.nrec;
    const size_t f = wg::forward_lds_bytes(g), a = wg::adjoint_lds_bytes(g, dummy);
    if (f > (size_t)lds_limit || a > (size_t)lds_limit) {
        why = "grid needs " + std::to_string(std::max(f, a)) + " B of LDS per workgroup, the device has " +
              std::to_string(lds_limit);
        return WG_E_TOO_LARGE;
    }
    return WG_OK;
}

const float* wav_ptr(const wg_engine* e) { return e->wav_dev ? e->wav_dev : e->d_wavelet; }
long long wav_stride(const wg_engine* e) { return e->wav_dev ? e->g.nt : 0; }

struct WavScope {
    wg_engine* e;
    WavScope(wg_engine* en, const float* w) : e(en) { e->wav_dev = w; }
    ~WavScope() { e->wav_dev = nullptr; }
};

struct UnitScope {
    wg_engine* e;
    UnitScope(wg_engine* en, const int* cell, const float* w, const unsigned char* act) : e(en) {
        e->unit_src_cell = cell;
        e->unit_src_w = w;
        e->unit_active = act;
    }
    ~UnitScope() {
        e->unit_src_cell = nullptr;
        e->unit_src_w = nullptr;
        e->unit_active = nullptr;
    }
};

wg::FwdParams fwd_params(wg_engine* e, const float* vel, float* traces) {
    wg::FwdParams p{};
    p.g = e->g;
    p.geo = e->geo();
    p.vel = vel;
    p.wavelet = wav_ptr(e);
    p.wav_stride = wav_stride(e);
    p.traces = traces;
    return p;
}

wg::AdjParams adj_params(wg_engine* e, const float* vel, const float* resid, float* grad) {
    wg::AdjParams p{};
    p.g = e->g;
    p.geo = e->geo();
    p.vel = vel;
    p.wavelet = wav_ptr(e);
    p.wav_stride = wav_stride(e);
    p.resid = resid;
    p.grad = grad;
    return p;
}

size_t snap_bytes_per_unit(const wg_engine* e) { return (size_t)e->g.nt * e->g.inner * sizeof(float); }

int models_per_chunk(wg_engine* e, int B) {
    const size_t per_model = snap_bytes_per_unit(e) * e->cfg.shots;
    size_t m = per_model ? e->snap_budget / per_model : (size_t)B;
    if (m < 1) m = 1;
    return (int)std::min<size_t>(m, (size_t)B);
}

#define LAUNCH(call)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) return fail(WG_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Forward (+ snapshots) of models [b0, b0 + nb) into traces, then optionally
// the adjoint with residual `resid` ([B*S, nt, R] absolute) into gunit.
// Records a (start, stop) event pair around `n` back-to-back kernel launches
// when timing is enabled.
struct Timed {
    wg_engine* e;
    std::vector<TimedSpan>* list;
    hipStream_t st;
    int64_t n;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(wg_engine* en, std::vector<TimedSpan>* l, hipStream_t s, int64_t launches = 1)
        : e(en), list(l), st(s), n(launches) {
        if (!e->timing) return;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
            e->timing = false;
            return;
        }
        (void)hipEventRecord(a, st);
    }
    ~Timed() {
        if (!a || !b) return;
        (void)hipEventRecord(b, st);
        list->push_back(TimedSpan{a, b, n});
    }
};

int stream_simulate(wg_engine* e, int b0, int nb, const float* vel, const float* wav, long long wav_stride,
                    float* traces, hipStream_t st);

bool use_bands(const wg_engine* e) { return e->streamed && e->bands > 1 && e->band_on; }

// Band split for a call over `units` units of a grid that fits one workgroup
// (the compiled 70 x 70 one): 4 or 2 bands when so few units would leave
// most compute units idle, else 0 (whole-unit kernels).
int small_bands(const wg_engine* e, int units, int bounds[wg::kMaxBands + 1]) {
    if (e->streamed || e->band_on == 0) return 0;
    wg::Geometry gh{};
    gh.band_z0 = e->band_z0;
    gh.band_rows = e->band_rows;
    int P = 0;
    return wg::band_split_small_fx(e->g, gh, units, e->cus, &P, bounds) ? P : 0;
}

// Buffers and counters of a band launch over `units` units in split (P,
// bounds) (P = 0: the engine's split for that many units), in buffer slot
// `slot` (band launches that may run at the same time use different slots):
// exchange words (zeroed: no step has tag 0) and [next, err].  Clears err
// when `clear_err`.  The split of a call depends only on its unit count, so a
// chunk's forward and adjoint run the same one (the results do not depend on
// the split).
int band_setup(wg_engine* e, int units, bool clear_err, hipStream_t st, wg::BandLaunch* bl, int P = 0,
               const int* bounds = nullptr, int slot = 0) {
    int call_bounds[wg::kMaxBands + 1];
    if (P == 0) {
        P = e->bands;
        bounds = e->band_bounds;
        wg::Geometry gh{};
        gh.band_z0 = e->band_z0;
        gh.band_rows = e->band_rows;
        int q = 0;
        if (e->band_on == 2 && e->rt_bands > 1) {
            P = e->rt_bands;
            bounds = e->rt_bounds;
        } else if (e->band_on == 1 && wg::band_split_call_fx(e->g, gh, units, e->cus, &q, call_bounds)) {
            P = q;
            bounds = call_bounds;
        }
    }
    DevBuf& xb = e->bxbuf[slot];
    DevBuf& qf = e->bqflags[slot];
    const size_t words = wg::band_xbuf_words(e->g, units, P);
    int rc;
    if ((rc = ensure(e, xb, words * sizeof(unsigned long long)))) return rc;
    if ((rc = ensure(e, qf, 2 * sizeof(int)))) return rc;
    int* q = (int*)qf.p;
    LAUNCH(wg::launch_zero((float*)xb.p, 2 * words, st));  // 0 bits
    LAUNCH(wg::launch_zero((float*)q, clear_err ? 2 : 1, st));
    bl->P = P;
    for (int i = 0; i <= wg::kMaxBands; ++i) bl->bounds[i] = i <= P ? bounds[i] : 0;
    bl->xbuf = (unsigned long long*)xb.p;
    bl->next = q;
    bl->err = q + 1;
    bl->spin = e->qspin < 0 ? -1 : std::min(e->qspin, 1 << 22);
    bl->fixed = e->band_on == 1 ? 1 : 0;
    return WG_OK;
}

// Forward of models [b0, b0 + nb) on the band kernels: gathers, and with
// `store` every snapshot (indexed by the unit's position in the chunk, from
// unit snap_unit_off of the snapshot buffer).  P, bounds, slot: band_setup's.
int band_forward(wg_engine* e, int b0, int nb, const float* vel, const float* wav, long long wav_stride,
                 float* traces, bool store, bool clear_err, hipStream_t st, int P = 0, const int* bounds = nullptr,
                 int slot = 0, size_t snap_unit_off = 0) {
    const int units = nb * e->cfg.shots;
    wg::BandLaunch bl{};
    int rc;
    if ((rc = band_setup(e, units, clear_err, st, &bl, P, bounds, slot))) return rc;
    wg::FwdParams p = fwd_params(e, vel, traces);
    p.wavelet = wav;
    p.wav_stride = wav_stride;
    p.unit0 = b0 * e->cfg.shots;
    p.snaps = store ? (float*)e->snaps.p + snap_unit_off * e->g.nt * e->g.inner : nullptr;
    Timed tm(e, &e->ev_fwd, st);
    LAUNCH(wg::launch_forward_band(p, units, store, bl, st));
    return WG_OK;
}

int band_adjoint(wg_engine* e, int b0, int nb, const float* vel, const float* resid, float* gunit, hipStream_t st,
                 int P = 0, const int* bounds = nullptr, int slot = 0, size_t snap_unit_off = 0) {
    const int units = nb * e->cfg.shots;
    wg::BandLaunch bl{};
    int rc;
    if ((rc = band_setup(e, units, false, st, &bl, P, bounds, slot))) return rc;
    wg::AdjParams p = adj_params(e, vel, resid, gunit);
    p.unit0 = b0 * e->cfg.shots;
    p.snaps = (const float*)e->snaps.p + snap_unit_off * e->g.nt * e->g.inner;
    Timed tm(e, &e->ev_adj, st);
    LAUNCH(wg::launch_adjoint_band(p, units, bl, st));
    return WG_OK;
}

// Non-storing band forward of models [b0, b0 + nb); a wait that gave up
// leaves NaN gathers in the units concerned.
int band_simulate(wg_engine* e, int b0, int nb, const float* vel, const float* wav, long long wav_stride,
                  float* traces, hipStream_t st) {
    return band_forward(e, b0, nb, vel, wav, wav_stride, traces, false, true, st);
}

// Forward of models [b0, b0 + nb) of a grid that fits one workgroup (and
// the non-storing forward of any grid); few units run as bands (buffer slot
// `slot`: callers running two at once pass different slots).
int run_chunk_forward(wg_engine* e, int b0, int nb, const float* vel, float* traces, bool store, hipStream_t st,
                      size_t snap_unit_off = 0, int slot = 0) {
    if (use_bands(e) && !store) return band_simulate(e, b0, nb, vel, wav_ptr(e), wav_stride(e), traces, st);
    if (e->streamed && !store) return stream_simulate(e, b0, nb, vel, wav_ptr(e), wav_stride(e), traces, st);
    int bounds[wg::kMaxBands + 1];
    if (const int P = small_bands(e, nb * e->cfg.shots, bounds))
        return band_forward(e, b0, nb, vel, wav_ptr(e), wav_stride(e), traces, store, true, st, P, bounds, slot,
                            snap_unit_off);
    wg::FwdParams p = fwd_params(e, vel, traces);
    p.unit0 = b0 * e->cfg.shots;
    p.snaps = store ? (float*)e->snaps.p + snap_unit_off * e->g.nt * e->g.inner : nullptr;
    Timed tm(e, &e->ev_fwd, st);
    LAUNCH(wg::launch_forward(p, nb * e->cfg.shots, store, false, st));
    return WG_OK;
}

int run_chunk_adjoint(wg_engine* e, int b0, int nb, const float* vel, const float* resid, float* gunit, float* dwav,
                      hipStream_t st, size_t snap_unit_off = 0, int slot = 0) {
    int bounds[wg::kMaxBands + 1];
    if (!dwav)
        if (const int P = small_bands(e, nb * e->cfg.shots, bounds))
            return band_adjoint(e, b0, nb, vel, resid, gunit, st, P, bounds, slot, snap_unit_off);
    wg::AdjParams p = adj_params(e, vel, resid, gunit);
    p.unit0 = b0 * e->cfg.shots;
    p.snaps = (const float*)e->snaps.p + snap_unit_off * e->g.nt * e->g.inner;
    p.dwav = dwav;
    Timed tm(e, &e->ev_adj, st);
    LAUNCH(wg::launch_adjoint(p, nb * e->cfg.shots, st));
    return WG_OK;
}


// ---------------------------------------------------------------------------
// HBM-streamed path: one launch per time step, state in HBM.
// ---------------------------------------------------------------------------
struct StreamPlan {
    int models;   // models per chunk
    int seg_len;  // steps per snapshot segment
    int nseg;     // segments (1: all snapshots kept, no recompute)
    bool snaps;   // snapshots and adjoint state are needed
};

constexpr int kCheckpointEvery = 100;  // steps between forward checkpoints when snapshots do not fit
constexpr int kCkptArrays = 6;         // a checkpoint: u^t, u^{t-1}, psi_x, psi_z, zeta_x, zeta_z
constexpr int kMaxStepUnits = 65535;   // units of one step launch (grid z dimension)

// Storage layout of the streamed arrays.  On grids at least kAlignMinNx
// cells wide the row pitch is rounded up to 32 floats and the arrays start a
// few floats into a unit's block, so that the first column of the lean
// rectangle (padded column region_adj + 2, where the lean tiles' columns
// start) and every row lie on 128-byte boundaries: the lean tiles' loads and
// stores are then whole cache lines (the lean 1000 x 300 step: 160 -> 131 us
// in a probe).  Snapshot rows are rounded to 32 floats and offset likewise.
// Row pitches and per-unit strides are odd multiples of 32 floats: the XCDs
// stream different units at once, and strides with a larger power of two put
// their accesses on the same HBM channels.  Narrower grids keep the plain
// layout (the rounding would cost up to a fifth more memory there).
constexpr int kAlignMinNx = 256;

// The smallest n' >= n that is an odd multiple of 32.
size_t odd32(size_t n) {
    size_t q = (n + 31) / 32;
    if (!(q & 1)) ++q;
    return q * 32;
}

void set_stream_layout(wg_engine* e) {
    const wg::Grid& g = e->g;
    e->slay = wg::plain_layout(g);
    e->s_ni = g.ni;
    e->s_inner = (size_t)g.inner;
    e->s_soff = 0;
    if (!e->streamed || g.nx < kAlignMinNx) return;
    const int off = (32 - (g.region_adj + 2) % 32) % 32;  // padded column region_adj + 2 at a multiple of 32
    const int pitch = (int)odd32((size_t)off + g.nx);
    e->slay = wg::CellLayout{pitch, off, odd32((size_t)pitch * g.nz)};
    e->s_soff = (32 - g.region_adj % 32) % 32;  // inner column region_adj likewise
    e->s_ni = (int)odd32((size_t)g.ni);
    e->s_inner = odd32((size_t)e->s_soff + (size_t)e->s_ni * g.nzi);
}

StreamPlan plan_stream(wg_engine* e, int B, bool need_snaps) {
    const wg::Grid& g = e->g;
    const size_t S = e->cfg.shots, cells = e->slay.ustride, inner = e->s_inner;
    const size_t state = (wg::kFwdArrays + 12) * cells * sizeof(float);
    StreamPlan P{1, g.nt, 1, need_snaps};
    const size_t full = state + (need_snaps ? (size_t)g.nt * inner * sizeof(float) : 0);
    if (!need_snaps || full * S <= e->snap_budget) {
        P.models = (int)std::max<size_t>(1, std::min<size_t>(B, e->snap_budget / std::max<size_t>(full * S, 1)));
        P.models = std::min(P.models, std::max(1, kMaxStepUnits / (int)S));
        return P;
    }
    P.seg_len = std::min(g.nt, kCheckpointEvery);
    P.nseg = (g.nt + P.seg_len - 1) / P.seg_len;
    const size_t per = state + (size_t)P.seg_len * inner * sizeof(float) + (size_t)P.nseg * kCkptArrays * cells * sizeof(float);
    P.models = (int)std::max<size_t>(1, std::min<size_t>(B, e->snap_budget / (per * S)));
    P.models = std::min(P.models, std::max(1, kMaxStepUnits / (int)S));
    return P;
}

wg::StepParams step_params(wg_engine* e, const float* vel, int units, int unit0) {
    wg::StepParams p{};
    p.g = e->g;
    p.geo = e->geo();
    p.vel = vel;
    p.wavelet = wav_ptr(e);
    p.wav_stride = wav_stride(e);
    p.units = units;
    p.unit0 = unit0;
    p.tile = e->tile;
    p.sp = e->slay.pitch;
    p.sc = e->slay.ustride;
    p.sni = e->s_ni;
    p.sinner = e->s_inner;
    return p;
}

int stream_alloc(wg_engine* e, const StreamPlan& P, int units) {
    const size_t cells = e->slay.ustride;
    int rc;
    if ((rc = ensure(e, e->fstate, wg::kFwdArrays * units * cells * sizeof(float)))) return rc;
    if (!P.snaps) return WG_OK;
    if ((rc = ensure(e, e->astate, 12 * units * cells * sizeof(float)))) return rc;
    if ((rc = ensure(e, e->snaps, (size_t)units * P.seg_len * e->s_inner * sizeof(float)))) return rc;
    if (P.nseg > 1 && (rc = ensure(e, e->ckpt, (size_t)P.nseg * kCkptArrays * units * cells * sizeof(float))))
        return rc;
    return WG_OK;
}

// Where the streamed forward state is: the wavefield buffers holding u^t
// and u^{t-1}, the half psi is read from and the half holding zeta.
struct FwdRot {
    int cur = 0, prev = 1, ps = 0, ze = 0;
};

// Forward steps [t0, t1) of `units` units from the state `rot` describes
// (updated), two per launch where the engine's forward block is 2; gathers
// into p.traces when set.  A two-step launch cannot record its second step's
// gathers (u^{t+1} is computed, not read), so the next two-step launch
// records them, or a record launch before a single step or at the end.
int fwd_steps(wg_engine* e, wg::StepParams& p, int units, int t0, int t1, bool store, FwdRot& rot, hipStream_t st) {
    bool pending = false;
    auto record_prev = [&](int t) {
        wg::StepParams q = p;
        q.t = t - 1;
        q.u_cur = rot.prev;
        return wg::launch_fwd_record(q, units, st);
    };
    for (int t = t0; t < t1;) {
        p.t = t;
        p.u_cur = rot.cur;
        p.u_prev = rot.prev;
        p.ps_half = rot.ps;
        p.ze_half = rot.ze;
        if (e->fwd_block == 2 && t + 1 < t1) {
            int o[2], n = 0;
            for (int k = 0; k < 4; ++k)
                if (k != rot.cur && k != rot.prev) o[n++] = k;
            p.u_out0 = o[0];
            p.u_out1 = o[1];
            p.trace_prev = pending ? 1 : 0;
            LAUNCH(wg::launch_fwd_step2(p, units, store, st));
            rot.prev = o[0];
            rot.cur = o[1];
            rot.ps ^= 1;
            rot.ze ^= 1;
            pending = p.traces != nullptr;
            t += 2;
        } else {
            if (pending) LAUNCH(record_prev(t));
            pending = false;
            p.trace_prev = 0;
            LAUNCH(wg::launch_fwd_step(p, units, store, false, st));
            std::swap(rot.cur, rot.prev);
            rot.ps ^= 1;
            t += 1;
        }
    }
    if (pending) LAUNCH(record_prev(t1));
    p.trace_prev = 0;
    return WG_OK;
}

// Launches of fwd_steps over [t0, t1) (for the timing average).
int fwd_launches(const wg_engine* e, int n) { return e->fwd_block == 2 ? n / 2 + n % 2 : n; }

// The live arrays of the forward state, in checkpoint order.
void ckpt_arrays(const FwdRot& rot, int out[kCkptArrays]) {
    const int a[kCkptArrays] = {wg::fwd_u_arr(rot.cur), wg::fwd_u_arr(rot.prev), 2 + rot.ps, 4 + rot.ps,
                                wg::fwd_zx_arr(rot.ze), wg::fwd_zz_arr(rot.ze)};
    for (int k = 0; k < kCkptArrays; ++k) out[k] = a[k];
}

// Copies the live forward state into checkpoint `slot` (save) or checkpoint
// `slot` into the arrays of the state FwdRot{} describes (restore).
int ckpt_copy(wg_engine* e, int units, int slot, const FwdRot& rot, bool save, hipStream_t st) {
    const size_t abytes = (size_t)units * e->slay.ustride * sizeof(float);
    int arrs[kCkptArrays];
    ckpt_arrays(save ? rot : FwdRot{}, arrs);
    for (int k = 0; k < kCkptArrays; ++k) {
        char* state = (char*)e->fstate.p + (size_t)arrs[k] * abytes;
        char* ck = (char*)e->ckpt.p + ((size_t)slot * kCkptArrays + k) * abytes;
        HIPCK(hipMemcpyAsync(save ? ck : state, save ? state : ck, abytes, hipMemcpyDeviceToDevice, st));
    }
    return WG_OK;
}

// Forward of units [unit0, unit0 + units) over all nt steps from rest:
// records traces (if non-null), stores all snapshots when P.nseg == 1 and
// `store`, and checkpoints the state every P.seg_len steps when P.nseg > 1.
// The clamped velocity of the models of units [unit0, unit0 + units) on the
// padded grid (e->vpad), for the step kernels' lean interior path.
int stream_pad_velocity(wg_engine* e, const float* vel, int units, int unit0, hipStream_t st, wg::StepParams* p) {
    const int S = e->cfg.shots, b0 = unit0 / S, nb = (unit0 + units - 1) / S + 1 - b0;
    const size_t cells = e->slay.ustride;
    int rc;
    if ((rc = ensure(e, e->vpad, (size_t)nb * cells * sizeof(float)))) return rc;
    LAUNCH(wg::launch_pad_velocity(e->g, vel, b0, nb, (float*)e->vpad.p, e->slay, st));
    p->vpad = (const float*)e->vpad.p + e->slay.off;
    p->vpad_b0 = b0;
    return WG_OK;
}

int stream_forward(wg_engine* e, const StreamPlan& P, int units, int unit0, const float* vel, const float* wav,
                   long long wav_stride, float* traces, bool store, hipStream_t st) {
    const size_t cells = e->slay.ustride;
    LAUNCH(wg::launch_zero((float*)e->fstate.p, (size_t)wg::kFwdArrays * units * cells, st));
    wg::StepParams p = step_params(e, vel, units, unit0);
    int rc0;
    if ((rc0 = stream_pad_velocity(e, vel, units, unit0, st, &p))) return rc0;
    p.wavelet = wav;
    p.wav_stride = wav_stride;
    p.traces = traces;
    p.snaps = (float*)e->snaps.p + e->s_soff;
    p.snap_t0 = 0;
    p.seg_len = P.seg_len;
    p.state = (float*)e->fstate.p + e->slay.off;
    const bool keep = store && P.nseg == 1;
    const bool ckpt = store && P.nseg > 1;
    const int nt = e->g.nt, seg = ckpt ? P.seg_len : nt;
    int launches = 0;
    for (int t0 = 0; t0 < nt; t0 += seg) launches += fwd_launches(e, std::min(nt, t0 + seg) - t0);
    Timed tm(e, &e->ev_fwd, st, launches);
    FwdRot rot;
    int rc;
    for (int t0 = 0; t0 < nt; t0 += seg) {
        if (ckpt && (rc = ckpt_copy(e, units, t0 / seg, rot, true, st))) return rc;
        if ((rc = fwd_steps(e, p, units, t0, std::min(nt, t0 + seg), keep, rot, st))) return rc;
    }
    return WG_OK;
}

// Adjoint of units [unit0, unit0 + units) with residual `resid` (all units),
// gradient accumulated into gunit (zeroed here for these units); segments are
// recomputed from the checkpoints when P.nseg > 1.
int stream_adjoint(wg_engine* e, const StreamPlan& P, int units, int unit0, const float* vel, const float* resid,
                   float* gunit, float* dwav, hipStream_t st) {
    const wg::Grid& g = e->g;
    const size_t cells = e->slay.ustride;
    LAUNCH(wg::launch_zero((float*)e->astate.p, 12 * units * cells, st));
    LAUNCH(wg::launch_zero(gunit + (size_t)unit0 * cells, units * cells, st));
    wg::StepParams pf = step_params(e, vel, units, unit0);
    pf.snaps = (float*)e->snaps.p + e->s_soff;
    pf.seg_len = P.seg_len;
    pf.state = (float*)e->fstate.p + e->slay.off;
    wg::StepParams pa = step_params(e, vel, units, unit0);
    int rc0;
    if ((rc0 = stream_pad_velocity(e, vel, units, unit0, st, &pf))) return rc0;
    pa.vpad = pf.vpad;
    pa.vpad_b0 = pf.vpad_b0;
    pa.snaps = (float*)e->snaps.p + e->s_soff;
    pa.seg_len = P.seg_len;
    pa.state = (float*)e->astate.p + e->slay.off;
    pa.resid = resid;
    pa.grad = gunit + e->slay.off;
    pa.dwav = dwav;
    // lambda buffers (0..3) holding lambda^t and lambda^{t+1}, and the half the
    // memory variables are read from; a single step updates lambda^{t+1} in
    // place, a two-step launch writes the two other buffers
    const bool two = e->time_block == 2 && !dwav;
    int cur = 0, prev = 1, half = 0;
    for (int seg = P.nseg - 1; seg >= 0; --seg) {
        const int t0 = seg * P.seg_len, t1 = std::min(g.nt, t0 + P.seg_len);
        if (P.nseg > 1) {
            int rc;
            FwdRot rot;
            if ((rc = ckpt_copy(e, units, seg, rot, false, st))) return rc;
            pf.snap_t0 = t0;
            Timed tm(e, &e->ev_fwd, st, fwd_launches(e, t1 - t0));
            if ((rc = fwd_steps(e, pf, units, t0, t1, true, rot, st))) return rc;
        }
        pa.snap_t0 = P.nseg > 1 ? t0 : 0;
        const int launches = two ? (t1 - t0 + 1) / 2 : t1 - t0;
        Timed tm(e, &e->ev_adj, st, launches);
        for (int t = t1 - 1; t >= t0;) {
            pa.t = t;
            pa.lam_cur = cur;
            pa.lam_prev = prev;
            pa.half = half;
            if (two && t - 1 >= t0) {
                int o[2], n = 0;
                for (int k = 0; k < 4; ++k)
                    if (k != cur && k != prev) o[n++] = k;
                pa.lam_out0 = o[0];
                pa.lam_out1 = o[1];
                LAUNCH(wg::launch_adj_step2(pa, units, st));
                prev = o[0];
                cur = o[1];
                t -= 2;
            } else {
                LAUNCH(wg::launch_adj_step(pa, units, st));
                std::swap(cur, prev);
                t -= 1;
            }
            half ^= 1;
        }
    }
    return WG_OK;
}

// Non-storing streamed forward of models [b0, b0 + nb), chunked by memory.
int stream_simulate(wg_engine* e, int b0, int nb, const float* vel, const float* wav, long long wav_stride,
                    float* traces, hipStream_t st) {
    const int S = e->cfg.shots;
    const StreamPlan P = plan_stream(e, nb, false);
    int rc;
    if ((rc = stream_alloc(e, P, P.models * S))) return rc;
    for (int c0 = b0; c0 < b0 + nb; c0 += P.models) {
        const int n = std::min(P.models, b0 + nb - c0);
        if ((rc = stream_forward(e, P, n * S, c0 * S, vel, wav, wav_stride, traces, false, st))) return rc;
    }
    return WG_OK;
}

// Loss and gradient of all B models on the resident work queues (compiled-
// grid kernels): every unit's forward with snapshots, then the residual and
// loss partials, then every unit's adjoint, the forward and the adjoint each
// cut into e->qparts time parts.  Sets *used = false, doing nothing, when the
// queues do not apply (another grid, disabled, no more units than CUs, or the
// snapshots and part states of all units over the snapshot budget).
int queue_loss_grad(wg_engine* e, int B, const float* vel, const float* obs, float* traces, const float* scale,
                    hipStream_t st, bool* used) {
    *used = false;
    const wg::Grid& g = e->g;
    const int S = e->cfg.shots, U = B * S, parts = e->qparts;
    size_t ff = 0, af = 0;
    // With at most one unit per CU there is nothing to balance: every unit's
    // parts would run one after another, so the whole-unit kernels are faster.
    if (parts <= 0 || U <= e->cus || !wg::queue_state_fx(g, &ff, &af)) return WG_OK;
    const size_t per_unit = snap_bytes_per_unit(e) + (ff + af) * sizeof(float);
    if (per_unit * (size_t)U > e->snap_budget) return WG_OK;
    const long long per = (long long)S * g.nt * g.nrec;
    int rc;
    if ((rc = ensure(e, e->snaps, snap_bytes_per_unit(e) * U))) return rc;
    if ((rc = ensure(e, e->qstate, (ff + af) * sizeof(float) * U))) return rc;
    const size_t nflag = 3 + (size_t)2 * parts * U;
    if ((rc = ensure(e, e->qflags, nflag * sizeof(int)))) return rc;
    LAUNCH(wg::launch_zero((float*)e->qflags.p, nflag, st));  // int 0 = float +0 bits
    wg::QueueParams q{};
    q.f = fwd_params(e, vel, traces);
    q.f.snaps = (float*)e->snaps.p;
    q.a = adj_params(e, vel, (const float*)e->resid.p, (float*)e->gunit.p);
    q.a.snaps = (const float*)e->snaps.p;
    q.fstate = (float*)e->qstate.p;
    q.astate = q.fstate + ff * U;
    int* fl = (int*)e->qflags.p;
    q.next_f = fl;
    q.next_a = fl + 1;
    q.err = fl + 2;
    q.done_f = fl + 3;
    q.done_a = q.done_f + (size_t)parts * U;
    q.units = U;
    q.parts = parts;
    q.spin_limit = e->qspin;
    const int wgs = std::min(e->cus, parts * U);
    hipError_t err = hipSuccess;
    {
        Timed tm(e, &e->ev_fwd, st);
        wg::launch_queue_fx(q, false, wgs, st, &err);
    }
    if (err != hipSuccess) return fail(WG_E_DEVICE, std::string("forward queue launch: ") + hipGetErrorString(err));
    LAUNCH(wg::launch_residual(traces, obs, scale, per, B, (float*)e->resid.p, (float*)e->partials.p, st));
    {
        Timed tm(e, &e->ev_adj, st);
        wg::launch_queue_fx(q, true, wgs, st, &err);
    }
    if (err != hipSuccess) return fail(WG_E_DEVICE, std::string("adjoint queue launch: ") + hipGetErrorString(err));
    *used = true;
    return WG_OK;
}

// The side stream (non-blocking, default priority) and the two pipeline
// events, created on first use.
int ensure_pipeline(wg_engine* e) {
    if (!e->side) HIPCK(hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking));
    for (auto& ev : e->ev_pipe)
        if (!ev) HIPCK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    return WG_OK;
}

// Loss and gradient of B models on stream st.  scale/energy may be preset by
// the caller (host path) or computed here (device path).
int loss_grad_core(wg_engine* e, int B, const float* vel, const float* obs, float* loss, float* grad,
                   float* traces_out, double* energy, float* scale, bool compute_energy, hipStream_t st) {
    const wg::Grid& g = e->g;
    const int S = e->cfg.shots;
    const long long per = (long long)S * g.nt * g.nrec;
    const long long chunks = (per + 255) / 256;
    int rc;
    if ((rc = ensure(e, e->resid, (size_t)B * per * sizeof(float)))) return rc;
    if ((rc = ensure(e, e->partials, (size_t)B * chunks * sizeof(float)))) return rc;
    if ((rc = ensure(e, e->gunit, (size_t)B * S * e->slay.ustride * sizeof(float)))) return rc;
    float* traces = traces_out;
    if (!traces) {
        if ((rc = ensure(e, e->traces, (size_t)B * per * sizeof(float)))) return rc;
        traces = (float*)e->traces.p;
    }
    if (compute_energy) LAUNCH(wg::launch_energy(obs, per, B, energy, scale, st));
    if (use_bands(e) && snap_bytes_per_unit(e) * S <= e->snap_budget) {
        // every snapshot of a chunk of models in memory, the units of the
        // chunk as bands: forward, residual, adjoint
        const int mc = models_per_chunk(e, B);
        if ((rc = ensure(e, e->snaps, snap_bytes_per_unit(e) * S * mc))) return rc;
        for (int b0 = 0; b0 < B; b0 += mc) {
            const int nb = std::min(mc, B - b0);
            if ((rc = band_forward(e, b0, nb, vel, wav_ptr(e), wav_stride(e), traces, true, b0 == 0, st))) return rc;
            LAUNCH(wg::launch_residual(traces + b0 * per, obs + b0 * per, scale + b0, per, nb,
                                       (float*)e->resid.p + b0 * per, (float*)e->partials.p + b0 * chunks, st));
            if ((rc = band_adjoint(e, b0, nb, vel, (const float*)e->resid.p, (float*)e->gunit.p, st))) return rc;
        }
        LAUNCH(wg::launch_loss_sum((const float*)e->partials.p, chunks, B, energy, loss, st));
        LAUNCH(wg::launch_fold(g, (const float*)e->gunit.p, vel, energy, B, grad, st));
        LAUNCH(wg::launch_flag_nan((const int*)e->bqflags[0].p + 1, loss, B, grad, (long long)g.nvx * g.nvz, st));
        return WG_OK;
    }
    if (e->streamed) {
        const StreamPlan P = plan_stream(e, B, true);
        if ((rc = stream_alloc(e, P, P.models * S))) return rc;
        for (int b0 = 0; b0 < B; b0 += P.models) {
            const int nb = std::min(P.models, B - b0);
            if ((rc = stream_forward(e, P, nb * S, b0 * S, vel, wav_ptr(e), wav_stride(e), traces, true, st))) return rc;
            LAUNCH(wg::launch_residual(traces + b0 * per, obs + b0 * per, scale + b0, per, nb,
                                       (float*)e->resid.p + b0 * per, (float*)e->partials.p + b0 * chunks, st));
            if ((rc = stream_adjoint(e, P, nb * S, b0 * S, vel, (const float*)e->resid.p, (float*)e->gunit.p,
                                     nullptr, st)))
                return rc;
        }
        LAUNCH(wg::launch_loss_sum((const float*)e->partials.p, chunks, B, energy, loss, st));
        LAUNCH(wg::launch_fold(g, (const float*)e->gunit.p, vel, energy, B, grad, e->slay, st));
        return WG_OK;
    }
    bool queued = false;
    if ((rc = queue_loss_grad(e, B, vel, obs, traces, scale, st, &queued))) return rc;
    if (queued) {
        LAUNCH(wg::launch_loss_sum((const float*)e->partials.p, chunks, B, energy, loss, st));
        LAUNCH(wg::launch_fold(g, (const float*)e->gunit.p, vel, energy, B, grad, st));
        LAUNCH(wg::launch_flag_nan((const int*)e->qflags.p + 2, loss, B, grad, (long long)g.nvx * g.nvz, st));
        return WG_OK;
    }
    const int mc = models_per_chunk(e, B);
    // Overlapped pipeline: chunks of one device's worth of units (one unit per
    // CU), chunk i on stream i % 2 with snapshot region i % 2; chunk i's
    // forward waits for chunk i - 1's forward, so it runs beside chunk i - 1's
    // adjoint instead of after it.  Per-unit arithmetic is unchanged.
    const int wave = std::max(1, e->cus / S);
    const size_t per_model = snap_bytes_per_unit(e) * S;
    const size_t fit = per_model ? e->snap_budget / per_model : (size_t)B;  // models whose snapshots fit
    if (e->overlap && B > wave && fit >= (size_t)std::min(B, 2 * wave)) {
        if ((rc = ensure_pipeline(e))) return rc;
        if ((rc = ensure(e, e->snaps, per_model * std::min(B, 2 * wave)))) return rc;
        hipStream_t sts[2] = {st, e->side};
        for (int i = 0, b0 = 0; b0 < B; ++i, b0 += wave) {
            const int nb = std::min(wave, B - b0);
            hipStream_t cs = sts[i & 1];
            const size_t off = (size_t)(i & 1) * wave * S;
            if (i > 0) HIPCK(hipStreamWaitEvent(cs, e->ev_pipe[(i - 1) & 1], 0));
            if ((rc = run_chunk_forward(e, b0, nb, vel, traces, true, cs, off, i & 1))) return rc;
            HIPCK(hipEventRecord(e->ev_pipe[i & 1], cs));
            LAUNCH(wg::launch_residual(traces + b0 * per, obs + b0 * per, scale + b0, per, nb,
                                       (float*)e->resid.p + b0 * per, (float*)e->partials.p + b0 * chunks, cs));
            if ((rc = run_chunk_adjoint(e, b0, nb, vel, (const float*)e->resid.p, (float*)e->gunit.p, nullptr, cs,
                                        off, i & 1)))
                return rc;
        }
        // join: the caller's stream waits for everything queued on the side stream
        HIPCK(hipEventRecord(e->ev_pipe[1], e->side));
        HIPCK(hipStreamWaitEvent(st, e->ev_pipe[1], 0));
        LAUNCH(wg::launch_loss_sum((const float*)e->partials.p, chunks, B, energy, loss, st));
        LAUNCH(wg::launch_fold(g, (const float*)e->gunit.p, vel, energy, B, grad, st));
        return WG_OK;
    }
    if ((rc = ensure(e, e->snaps, snap_bytes_per_unit(e) * S * mc))) return rc;
    for (int b0 = 0; b0 < B; b0 += mc) {
        const int nb = std::min(mc, B - b0);
        if ((rc = run_chunk_forward(e, b0, nb, vel, traces, true, st))) return rc;
        LAUNCH(wg::launch_residual(traces + b0 * per, obs + b0 * per, scale + b0, per, nb,
                                   (float*)e->resid.p + b0 * per, (float*)e->partials.p + b0 * chunks, st));
        if ((rc = run_chunk_adjoint(e, b0, nb, vel, (const float*)e->resid.p, (float*)e->gunit.p, nullptr, st)))
            return rc;
    }
    LAUNCH(wg::launch_loss_sum((const float*)e->partials.p, chunks, B, energy, loss, st));
    LAUNCH(wg::launch_fold(g, (const float*)e->gunit.p, vel, energy, B, grad, st));
    return WG_OK;
}

struct Guard {
    wg_engine* e;
    std::lock_guard<std::mutex> lk;
    explicit Guard(wg_engine* en) : e(en), lk(en->mu) {}
};

}  // namespace

extern "C" {

void wg_config_default(wg_config* c) {
    if (!c) return;
    c->shots = 5;
    c->nt = 2001;
    c->dt = 0.0005f;
    c->peak_hz = 15.0f;
    c->pml = 20;
    c->vmin = 1000.0f;
    c->vmax = 5000.0f;
    c->nvx = 70;
    c->nvz = 70;
    c->h = 10.0f;
}

const char* wg_last_error(void) { return g_err.c_str(); }

// geology-core lib.rs:12-21 (float32, expf).
void wg_ricker_wavelet(int32_t samples, float dt, float peak_hz, float* out) {
    const float delay = 1.5f / peak_hz;
    for (int32_t i = 0; i < samples; ++i) {
        const float t = (float)i * dt - delay;
        const float u = 3.14159265358979323846f * peak_hz * t;
        const float a = u * u;
        out[i] = (1.0f - 2.0f * a) * expf(-a);
    }
}

// geology-core lib.rs:25-68 (double inside, float32 out).
void wg_cpml_profiles(int32_t n, float h, float dt, int32_t width, float max_vel, float freq, float* ao, float* bo,
                      float* dbo) {
    std::vector<double> a(n, 0.0), b(n, 0.0), db(n, 0.0);
    const double start0 = (double)(kFdPad + width);
    const double start1 = (double)(n - 1 - kFdPad - width);
    const double wf = (double)width, hf = (double)h, dtf = (double)dt;
    const double sigma0 = -(3.0 * (double)max_vel * std::log(0.001)) / (2.0 * wf * hf);
    for (int x = 0; x < n; ++x) {
        const double xf = (double)x;
        const double f0 = (start0 - xf) / wf, f1 = (xf - start1) / wf;
        const double frac = std::min(std::max(std::max(f0, f1), 0.0), 1.0);
        const double sigma = sigma0 * frac * frac;
        const double alpha = 3.14159265358979323846 * (double)freq * (1.0 - frac);
        const double axv = std::exp(-(sigma + alpha) * dtf);
        b[x] = (sigma / (sigma + alpha)) * (axv - 1.0);
        a[x] = frac == 0.0 ? 0.0 : axv;
    }
    for (int x = 2; x < n - 2; ++x)
        db[x] = ((8.0 / 12.0) * (b[x + 1] - b[x - 1]) - (1.0 / 12.0) * (b[x + 2] - b[x - 2])) / hf;
    for (int i = 0; i < n; ++i) {
        ao[i] = (float)a[i];
        bo[i] = (float)b[i];
        dbo[i] = (float)db[i];
    }
}

int32_t wg_engine_create(const wg_config* cfg, wg_engine** out) {
    if (!cfg || !out) return fail(WG_E_INVALID, "null argument");
    *out = nullptr;
    const wg_config c = *cfg;
    if (c.shots < 1 || c.nt < 1 || c.pml < 0 || c.nvx < 1 || c.nvz < 1)
        return fail(WG_E_INVALID, "shots, nt, nvx, nvz must be >= 1 and pml >= 0");
    if (!(c.h > 0.0f) || !(c.dt > 0.0f) || !(c.peak_hz > 0.0f))
        return fail(WG_E_INVALID, "h, dt and peak_hz must be positive");
    if (!(c.vmin <= c.vmax)) return fail(WG_E_INVALID, "vmin must not exceed vmax");

    wg::Grid g{};
    g.pad = kFdPad + c.pml;
    g.nx = c.nvx + 2 * g.pad;
    g.nz = c.nvz + 2 * g.pad;
    g.ni = g.nx - 4;
    g.nzi = g.nz - 4;
    g.inner = g.ni * g.nzi;
    g.region = c.pml + 2 * kFdPad;
    g.region_adj = c.pml + 3 * kFdPad;
    g.swf = g.region + 2;
    g.swa = g.region_adj + 2;
    g.nvx = c.nvx;
    g.nvz = c.nvz;
    g.nt = c.nt;
    g.nrec = c.nvx;
    g.shots = c.shots;
    g.rh = 1.0f / c.h;
    g.rh2 = g.rh * g.rh;
    g.dt2 = c.dt * c.dt;
    g.vmin = c.vmin;
    g.vmax = c.vmax;

    int dev = 0;
    HIPCK(hipGetDevice(&dev));
    int lds_limit = 0;
    HIPCK(hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    std::string why;
    const bool resident = check_fits(g, lds_limit, why) == WG_OK;
    if (!resident && (size_t)g.nx * g.nz > ((size_t)1 << 28))
        return fail(WG_E_TOO_LARGE, "grid of " + std::to_string((size_t)g.nx * g.nz) + " cells is too large");

    void* scr = nullptr;
    int src = scratch_setup(dev, &scr);
    if (src) return src;

    wg_engine* e = new wg_engine();
    e->scratch_ptr = scr;
    e->streamed = !resident;
    e->cfg = c;
    e->g = g;
    set_stream_layout(e);
    e->device = dev;
    // Default sources: x_k = xmax k / (S-1), or xmax / 2 for one shot, at z = h
    // (wave_wgpu.rs:203-214).
    const float xmax = (float)(c.nvx - 1) * c.h;
    e->src_xz.resize(2 * c.shots);
    for (int k = 0; k < c.shots; ++k) {
        e->src_xz[2 * k] = c.shots == 1 ? xmax * 0.5f : (xmax * (float)k) / (float)(c.shots - 1);
        e->src_xz[2 * k + 1] = c.h;
    }
    build_sources(e);
    build_receivers(e);
    if (!resident) {
        wg::Geometry gh{};
        gh.band_z0 = e->band_z0;
        gh.band_rows = e->band_rows;
        gh.n_ent = (int)e->ent_r.size();
        int P = 0;
        if (wg::band_plan(g, gh, lds_limit, &P, e->rt_bounds)) e->rt_bands = P;
        if (wg::band_split_fx(g, gh, &P, e->band_bounds)) {
            e->bands = P;
        } else if (e->rt_bands > 1) {
            e->bands = e->rt_bands;
            for (int q = 0; q <= e->rt_bands; ++q) e->band_bounds[q] = e->rt_bounds[q];
        }
    }
    e->prof.resize(3 * (g.nx + g.nz));
    wg_cpml_profiles(g.nx, c.h, c.dt, c.pml, c.vmax, c.peak_hz, e->prof.data(), e->prof.data() + g.nx,
                     e->prof.data() + 2 * g.nx);
    float* pz = e->prof.data() + 3 * g.nx;
    wg_cpml_profiles(g.nz, c.h, c.dt, c.pml, c.vmax, c.peak_hz, pz, pz + g.nz, pz + 2 * g.nz);
    e->wavelet.resize(c.nt);
    wg_ricker_wavelet(c.nt, c.dt, c.peak_hz, e->wavelet.data());

    int rc = WG_OK;
    if (!rc) rc = upload(e->prof, &e->d_prof);
    if (!rc) rc = upload(e->src_cell, &e->d_src_cell);
    if (!rc) rc = upload(e->src_w, &e->d_src_w);
    if (!rc) rc = upload(e->rec_cell, &e->d_rec_cell);
    if (!rc) rc = upload(e->rec_w, &e->d_rec_w);
    if (!rc) rc = upload(e->band_off, &e->d_band_off);
    if (!rc) rc = upload(e->ent_r, &e->d_ent_r);
    if (!rc) rc = upload(e->ent_w, &e->d_ent_w);
    if (!rc) rc = upload(e->wavelet, &e->d_wavelet);
    if (!rc) {
        hipError_t he = hipStreamCreate(&e->stream);
        if (he != hipSuccess) rc = fail(WG_E_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(he));
    }
    if (!rc) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device) == hipSuccess && cus > 0)
            e->cus = cus;
    }
    if (!rc) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = (size_t)16 << 30;
        e->snap_budget = std::min<size_t>(fr / 2, (size_t)96 << 30);
    }
    if (rc) {
        wg_engine_destroy(e);
        return rc;
    }
    *out = e;
    return WG_OK;
}

void wg_engine_destroy(wg_engine* e) {
    if (!e) return;
    if (e->last_valid) (void)hipStreamSynchronize(e->last);
    for (auto* l : {&e->ev_fwd, &e->ev_adj})
        for (auto& pr : *l) {
            (void)hipEventDestroy(pr.a);
            (void)hipEventDestroy(pr.b);
        }
    for (void* p : {(void*)e->d_prof, (void*)e->d_src_cell, (void*)e->d_src_w, (void*)e->d_rec_cell,
                    (void*)e->d_rec_w, (void*)e->d_band_off, (void*)e->d_ent_r, (void*)e->d_ent_w,
                    (void*)e->d_wavelet})
        if (p) (void)hipFree(p);
    for (DevBuf* b : {&e->snaps, &e->traces, &e->resid, &e->gunit, &e->partials, &e->scale, &e->energy, &e->vel,
                      &e->obs, &e->loss, &e->grad, &e->fstate, &e->astate, &e->ckpt, &e->vpad, &e->qstate, &e->qflags,
                      &e->bxbuf[0], &e->bqflags[0], &e->bxbuf[1], &e->bqflags[1], &e->mtable, &e->mnext,
                      &e->mslots})
        if (b->p) (void)hipFree(b->p);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    if (e->side) {
        (void)hipStreamSynchronize(e->side);
        (void)hipStreamDestroy(e->side);
    }
    for (auto& ev : e->ev_pipe)
        if (ev) (void)hipEventDestroy(ev);
    delete e;
}

int32_t wg_engine_shots(const wg_engine* e) { return e ? e->cfg.shots : -1; }
int32_t wg_engine_nt(const wg_engine* e) { return e ? e->cfg.nt : -1; }
int32_t wg_engine_nrec(const wg_engine* e) { return e ? e->g.nrec : -1; }

int32_t wg_engine_dims(const wg_engine* e, int32_t* nx, int32_t* nz, int32_t* pad) {
    if (!e) return fail(WG_E_INVALID, "null engine");
    if (nx) *nx = e->g.nx;
    if (nz) *nz = e->g.nz;
    if (pad) *pad = e->g.pad;
    return WG_OK;
}

int32_t wg_get_sources(const wg_engine* e, float* xz, int32_t n) {
    if (!e || !xz) return fail(WG_E_INVALID, "null argument");
    if (n != e->cfg.shots) return fail(WG_E_INVALID, "sources must have shape [shots, 2]");
    std::memcpy(xz, e->src_xz.data(), sizeof(float) * 2 * n);
    return WG_OK;
}

int32_t wg_set_sources(wg_engine* e, const float* xz, int32_t n) {
    if (!e || !xz) return fail(WG_E_INVALID, "null argument");
    if (n != e->cfg.shots) return fail(WG_E_INVALID, "sources must have shape [shots, 2]");
    Guard G(e);
    HIPCK(hipSetDevice(e->device));
    if (e->last_valid) HIPCK(hipStreamSynchronize(e->last));
    e->src_xz.assign(xz, xz + 2 * n);
    build_sources(e);
    HIPCK(hipMemcpy(e->d_src_cell, e->src_cell.data(), sizeof(int) * e->src_cell.size(), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(e->d_src_w, e->src_w.data(), sizeof(float) * e->src_w.size(), hipMemcpyHostToDevice));
    return WG_OK;
}

int32_t wg_get_wavelet(const wg_engine* e, float* w, int32_t n) {
    if (!e || !w) return fail(WG_E_INVALID, "null argument");
    if (n != e->cfg.nt) return fail(WG_E_INVALID, "wavelet length must equal nt");
    std::memcpy(w, e->wavelet.data(), sizeof(float) * n);
    return WG_OK;
}

int32_t wg_set_wavelet(wg_engine* e, const float* w, int32_t n) {
    if (!e || !w) return fail(WG_E_INVALID, "null argument");
    if (n != e->cfg.nt) return fail(WG_E_INVALID, "wavelet length must equal nt");
    Guard G(e);
    HIPCK(hipSetDevice(e->device));
    if (e->last_valid) HIPCK(hipStreamSynchronize(e->last));
    e->wavelet.assign(w, w + n);
    HIPCK(hipMemcpy(e->d_wavelet, w, sizeof(float) * n, hipMemcpyHostToDevice));
    return WG_OK;
}

int32_t wg_simulate(wg_engine* e, const float* vel_host, float* traces_host) {
    if (!e || !vel_host || !traces_host) return fail(WG_E_INVALID, "null argument");
    Guard G(e);
    HIPCK(hipSetDevice(e->device));
    int rc;
    if ((rc = fence(e, e->stream))) return rc;
    const size_t nvis = (size_t)e->g.nvx * e->g.nvz;
    const size_t ntr = (size_t)e->cfg.shots * e->g.nt * e->g.nrec;
    if ((rc = ensure(e, e->vel, nvis * sizeof(float)))) return rc;
    if ((rc = ensure(e, e->traces, ntr * sizeof(float)))) return rc;
    HIPCK(hipMemcpyAsync(e->vel.p, vel_host, nvis * sizeof(float), hipMemcpyHostToDevice, e->stream));
    if ((rc = run_chunk_forward(e, 0, 1, (const float*)e->vel.p, (float*)e->traces.p, false, e->stream))) return rc;
    HIPCK(hipMemcpyAsync(traces_host, e->traces.p, ntr * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    HIPCK(hipStreamSynchronize(e->stream));
    return WG_OK;
}

int32_t wg_evaluate_adjoint(wg_engine* e, const float* vel_host, const float* obs_host, float* loss_host,
                            float* grad_host) {
    if (!e || !vel_host || !obs_host || !loss_host || !grad_host) return fail(WG_E_INVALID, "null argument");
    Guard G(e);
    HIPCK(hipSetDevice(e->device));
    const size_t nvis = (size_t)e->g.nvx * e->g.nvz;
    const size_t ntr = (size_t)e->cfg.shots * e->g.nt * e->g.nrec;
    // wave_wgpu.rs:852-854: energy in f64 on the host, in order.
    double energy = 0.0;
    for (size_t i = 0; i < ntr; ++i) energy += (double)obs_host[i] * (double)obs_host[i];
    if (!(energy > 0.0)) return fail(WG_E_ZERO_ENERGY, "Observed traces have zero energy");
    const float scale = (float)(1.0 / energy);
    int rc;
    if ((rc = fence(e, e->stream))) return rc;
    if ((rc = ensure(e, e->vel, nvis * sizeof(float)))) return rc;
    if ((rc = ensure(e, e->obs, ntr * sizeof(float)))) return rc;
    if ((rc = ensure(e, e->scale, 256))) return rc;
    if ((rc = ensure(e, e->energy, 256))) return rc;
    if ((rc = ensure(e, e->loss, 256))) return rc;
    if ((rc = ensure(e, e->grad, nvis * sizeof(float)))) return rc;
    HIPCK(hipMemcpyAsync(e->vel.p, vel_host, nvis * sizeof(float), hipMemcpyHostToDevice, e->stream));
    HIPCK(hipMemcpyAsync(e->obs.p, obs_host, ntr * sizeof(float), hipMemcpyHostToDevice, e->stream));
    HIPCK(hipMemcpyAsync(e->scale.p, &scale, sizeof(float), hipMemcpyHostToDevice, e->stream));
    HIPCK(hipMemcpyAsync(e->energy.p, &energy, sizeof(double), hipMemcpyHostToDevice, e->stream));
    if ((rc = loss_grad_core(e, 1, (const float*)e->vel.p, (const float*)e->obs.p, (float*)e->loss.p,
                             (float*)e->grad.p, nullptr, (double*)e->energy.p, (float*)e->scale.p, false, e->stream)))
        return rc;
    HIPCK(hipMemcpyAsync(loss_host, e->loss.p, sizeof(float), hipMemcpyDeviceToHost, e->stream));
    HIPCK(hipMemcpyAsync(grad_host, e->grad.p, nvis * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    HIPCK(hipStreamSynchronize(e->stream));
    return WG_OK;
}

int32_t wg_simulate_dev(wg_engine* e, int32_t B, const float* vel_dev, float* traces_dev, void* stream) {
    if (!e || !vel_dev || !traces_dev || B < 0) return fail(WG_E_INVALID, "invalid argument");
    Guard G(e);
    HIPCK(hipSetDevice(e->device));
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = fence(e, st))) return rc;
    if (B == 0) return WG_OK;
    return run_chunk_forward(e, 0, B, vel_dev, traces_dev, false, st);
}

static int32_t loss_grad_entry(wg_engine* e, int32_t B, const float* vel_dev, const float* wav_dev,
                               const float* obs_dev, float* loss_dev, float* grad_dev, float* traces_dev,
                               double* energy_dev, void* stream, const int* src_cell = nullptr,
                               const float* src_w = nullptr, const unsigned char* active = nullptr) {
    if (!e || !vel_dev || !obs_dev || !loss_dev || !grad_dev || B < 0) return fail(WG_E_INVALID, "invalid argument");
    Guard G(e);
    WavScope W(e, wav_dev);
    UnitScope U(e, src_cell, src_w, active);
    HIPCK(hipSetDevice(e->device));
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = fence(e, st))) return rc;
    if (B == 0) return WG_OK;
    if ((rc = ensure(e, e->scale, (size_t)B * sizeof(float)))) return rc;
    double* energy = energy_dev;
    if (!energy) {
        if ((rc = ensure(e, e->energy, (size_t)B * sizeof(double)))) return rc;
        energy = (double*)e->energy.p;
    }
    return loss_grad_core(e, B, vel_dev, obs_dev, loss_dev, grad_dev, traces_dev, energy, (float*)e->scale.p, true,
                          st);
}

int32_t wg_loss_grad_dev(wg_engine* e, int32_t B, const float* vel_dev, const float* obs_dev, float* loss_dev,
                         float* grad_dev, float* traces_dev, double* energy_dev, void* stream) {
    return loss_grad_entry(e, B, vel_dev, nullptr, obs_dev, loss_dev, grad_dev, traces_dev, energy_dev, stream);
}

int32_t wg_loss_grad_true_dev(wg_engine* e, int32_t B, const float* vel_dev, const float* true_dev, float* obs_dev,
                              float* loss_dev, float* grad_dev, float* traces_dev, double* energy_dev, void* stream) {
    if (!e || !vel_dev || !true_dev || !obs_dev || !loss_dev || !grad_dev || B < 0)
        return fail(WG_E_INVALID, "invalid argument");
    Guard G(e);
    HIPCK(hipSetDevice(e->device));
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = fence(e, st))) return rc;
    if (B == 0) return WG_OK;
    if ((rc = ensure(e, e->scale, (size_t)B * sizeof(float)))) return rc;
    double* energy = energy_dev;
    if (!energy) {
        if ((rc = ensure(e, e->energy, (size_t)B * sizeof(double)))) return rc;
        energy = (double*)e->energy.p;
    }
    const wg::Grid& g = e->g;
    const int S = e->cfg.shots;
    // One resident round of units, snapshots within budget: the observed
    // forward (no snapshots) runs on the side stream beside the storing
    // forward, and the residual waits for both.  Per-unit arithmetic is that
    // of wg_simulate_dev followed by wg_loss_grad_dev.
    const bool beside = !e->streamed && B * S <= e->cus && models_per_chunk(e, B) >= B;
    if (!beside) {
        if ((rc = run_chunk_forward(e, 0, B, true_dev, obs_dev, false, st))) return rc;
        return loss_grad_core(e, B, vel_dev, obs_dev, loss_dev, grad_dev, traces_dev, energy, (float*)e->scale.p,
                              true, st);
    }
    const long long per = (long long)S * g.nt * g.nrec;
    const long long chunks = (per + 255) / 256;
    if ((rc = ensure(e, e->resid, (size_t)B * per * sizeof(float)))) return rc;
    if ((rc = ensure(e, e->partials, (size_t)B * chunks * sizeof(float)))) return rc;
    if ((rc = ensure(e, e->gunit, (size_t)B * S * g.nx * g.nz * sizeof(float)))) return rc;
    if ((rc = ensure(e, e->snaps, snap_bytes_per_unit(e) * S * B))) return rc;
    float* traces = traces_dev;
    if (!traces) {
        if ((rc = ensure(e, e->traces, (size_t)B * per * sizeof(float)))) return rc;
        traces = (float*)e->traces.p;
    }
    if ((rc = ensure_pipeline(e))) return rc;
    HIPCK(hipEventRecord(e->ev_pipe[0], st));
    HIPCK(hipStreamWaitEvent(e->side, e->ev_pipe[0], 0));
    if ((rc = run_chunk_forward(e, 0, B, true_dev, obs_dev, false, e->side, 0, 1))) return rc;
    HIPCK(hipEventRecord(e->ev_pipe[1], e->side));
    if ((rc = run_chunk_forward(e, 0, B, vel_dev, traces, true, st))) return rc;
    HIPCK(hipStreamWaitEvent(st, e->ev_pipe[1], 0));
    float* scale = (float*)e->scale.p;
    LAUNCH(wg::launch_energy(obs_dev, per, B, energy, scale, st));
    LAUNCH(wg::launch_residual(traces, obs_dev, scale, per, B, (float*)e->resid.p, (float*)e->partials.p, st));
    if ((rc = run_chunk_adjoint(e, 0, B, vel_dev, (const float*)e->resid.p, (float*)e->gunit.p, nullptr, st)))
        return rc;
    LAUNCH(wg::launch_loss_sum((const float*)e->partials.p, chunks, B, energy, loss_dev, st));
    LAUNCH(wg::launch_fold(g, (const float*)e->gunit.p, vel_dev, energy, B, grad_dev, st));
    return WG_OK;
}

int32_t wg_loss_grad_wavelets_dev(wg_engine* e, int32_t B, const float* vel_dev, const float* wav_dev,
                                  const float* obs_dev, float* loss_dev, float* grad_dev, float* traces_dev,
                                  double* energy_dev, void* stream) {
    if (!wav_dev) return fail(WG_E_INVALID, "invalid argument");
    return loss_grad_entry(e, B, vel_dev, wav_dev, obs_dev, loss_dev, grad_dev, traces_dev, energy_dev, stream);
}

int32_t wg_loss_dev(wg_engine* e, int32_t B, const float* vel_dev, const float* wav_dev, const float* obs_dev,
                    float* loss_dev, float* traces_dev, double* energy_dev, void* stream) {
    if (!e || !vel_dev || !obs_dev || !loss_dev || B < 0) return fail(WG_E_INVALID, "invalid argument");
    Guard G(e);
    WavScope W(e, wav_dev);
    HIPCK(hipSetDevice(e->device));
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = fence(e, st))) return rc;
    if (B == 0) return WG_OK;
    const wg::Grid& g = e->g;
    const long long per = (long long)e->cfg.shots * g.nt * g.nrec;
    const long long chunks = (per + 255) / 256;
    if ((rc = ensure(e, e->scale, (size_t)B * sizeof(float)))) return rc;
    if ((rc = ensure(e, e->resid, (size_t)B * per * sizeof(float)))) return rc;
    if ((rc = ensure(e, e->partials, (size_t)B * chunks * sizeof(float)))) return rc;
    double* energy = energy_dev;
    if (!energy) {
        if ((rc = ensure(e, e->energy, (size_t)B * sizeof(double)))) return rc;
        energy = (double*)e->energy.p;
    }
    float* traces = traces_dev;
    if (!traces) {
        if ((rc = ensure(e, e->traces, (size_t)B * per * sizeof(float)))) return rc;
        traces = (float*)e->traces.p;
    }
    float* scale = (float*)e->scale.p;
    LAUNCH(wg::launch_energy(obs_dev, per, B, energy, scale, st));
    // the forward without snapshots gives the storing forward's gathers bit for
    // bit, so the loss is the one wg_loss_grad_dev returns
    if ((rc = run_chunk_forward(e, 0, B, vel_dev, traces, false, st))) return rc;
    LAUNCH(wg::launch_residual(traces, obs_dev, scale, per, B, (float*)e->resid.p, (float*)e->partials.p, st));
    LAUNCH(wg::launch_loss_sum((const float*)e->partials.p, chunks, B, energy, loss_dev, st));
    return WG_OK;
}

int32_t wg_loss_noalloc_dev(wg_engine* e, int32_t B, const float* vel_dev, const float* wav_dev,
                            const float* obs_dev, float* loss_dev, float* traces_dev, double* energy_dev,
                            void* stream) {
    if (!e || !vel_dev || !obs_dev || !loss_dev || !traces_dev || B < 0) return fail(WG_E_INVALID, "invalid argument");
    Guard G(e);
    WavScope W(e, wav_dev);
    HIPCK(hipSetDevice(e->device));
    if (e->streamed)
        return fail(WG_E_TOO_LARGE, "wg_loss_noalloc_dev runs the resident kernels only (grid too large for one "
                                    "workgroup)");
    if (!e->scratch_ptr) return fail(WG_E_DEVICE, "the fixed scratch of this device was not set up");
    const wg::Grid& g = e->g;
    const int S = e->cfg.shots;
    const long long per = (long long)S * g.nt * g.nrec;
    const long long chunks = (per + 255) / 256;
    // scratch layout: scale [B] f32 | energy [B] f64 (when the caller passes
    // none) | partials [B, chunks] f32, each part 256-byte aligned
    const size_t off_e = align256((size_t)B * sizeof(float));
    const size_t off_p = off_e + (energy_dev ? 0 : align256((size_t)B * sizeof(double)));
    const size_t need = off_p + (size_t)B * chunks * sizeof(float);
    if (need > kScratchBytes) {
        const size_t per_model = sizeof(float) + (energy_dev ? 0 : sizeof(double)) + chunks * sizeof(float);
        return fail(WG_E_TOO_LARGE, "wg_loss_noalloc_dev: " + std::to_string(B) + " models need " +
                                        std::to_string(need) + " bytes of scratch; the fixed scratch holds " +
                                        std::to_string(kScratchBytes) + " (about " +
                                        std::to_string(kScratchBytes / per_model - 1) + " models of this engine)");
    }
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = fence(e, st))) return rc;
    if (B == 0) return WG_OK;
    char* base = (char*)e->scratch_ptr;
    float* scale = (float*)base;
    double* energy = energy_dev ? energy_dev : (double*)(base + off_e);
    float* partials = (float*)(base + off_p);
    LAUNCH(wg::launch_energy(obs_dev, per, B, energy, scale, st));
    wg::FwdParams p = fwd_params(e, vel_dev, traces_dev);
    LAUNCH(wg::launch_forward(p, B * S, false, false, st));
    LAUNCH(wg::launch_residual(traces_dev, obs_dev, scale, per, B, nullptr, partials, st));
    LAUNCH(wg::launch_loss_sum(partials, chunks, B, energy, loss_dev, st));
    return WG_OK;
}

int32_t wg_source_corners(const wg_engine* e, int32_t n, const float* xz, int32_t* cell, float* w) {
    if (!e || n < 0 || (n > 0 && (!xz || !cell || !w))) return fail(WG_E_INVALID, "invalid argument");
    for (int k = 0; k < n; ++k) {
        if (!std::isfinite(xz[2 * k]) || !std::isfinite(xz[2 * k + 1]))
            return fail(WG_E_INVALID, "source positions must be finite");
        corners(xz[2 * k], xz[2 * k + 1], e->cfg.h, e->g.pad, e->g.nx, e->g.nz, &cell[4 * k], &w[4 * k]);
    }
    return WG_OK;
}

int32_t wg_loss_grad_units_dev(wg_engine* e, int32_t B, const int32_t* src_cell_dev, const float* src_w_dev,
                               const uint8_t* active_dev, const float* vel_dev, const float* wav_dev,
                               const float* obs_dev, float* loss_dev, float* grad_dev, float* traces_dev,
                               double* energy_dev, void* stream) {
    if (!src_cell_dev || !src_w_dev || !wav_dev) return fail(WG_E_INVALID, "invalid argument");
    return loss_grad_entry(e, B, vel_dev, wav_dev, obs_dev, loss_dev, grad_dev, traces_dev, energy_dev, stream,
                           src_cell_dev, src_w_dev, active_dev);
}


// The loss and gradient of several engines' unit batches (wg_loss_grad_units_dev
// arguments per group) in ONE launch, when every group's grid has compiled
// kernels for it (70 x 70 as whole units, 100 x 100 and 141 x 141 in their
// compiled band splits); otherwise the groups run one after another.  Each
// task (a unit, or a band of one) runs forward, residual and adjoint with its
// snapshots in a slot held for the task's lifetime: slots for the resident
// workgroups (one per compute unit) and a margin, so the snapshot memory does
// not grow with the batch.  Tasks are laid out longest first (the groups'
// step counts times the kind's step time).
int32_t wg_loss_grad_groups_dev(int32_t n, const wg_group_args* ga, void* stream) {
    if (n < 1 || !ga) return fail(WG_E_INVALID, "invalid argument");
    if (n > wg::kMultiMaxGroups) return fail(WG_E_INVALID, "at most 64 groups in one call");
    for (int i = 0; i < n; ++i) {
        const wg_group_args& a = ga[i];
        if (!a.engine || a.B < 0 || !a.src_cell_dev || !a.src_w_dev || !a.vel_dev || !a.wav_dev || !a.obs_dev ||
            !a.loss_dev || !a.grad_dev)
            return fail(WG_E_INVALID, "invalid argument");
    }
    std::vector<wg_engine*> order(n);
    for (int i = 0; i < n; ++i) order[i] = ga[i].engine;
    std::sort(order.begin(), order.end());
    for (int i = 1; i < n; ++i)
        if (order[i] == order[i - 1]) return fail(WG_E_INVALID, "an engine appears in two groups");
    for (int i = 1; i < n; ++i)
        if (order[i]->device != order[0]->device) return fail(WG_E_INVALID, "the engines are on different devices");
    std::vector<std::unique_lock<std::mutex>> locks;  // in address order: callers cannot deadlock
    for (wg_engine* e : order) locks.emplace_back(e->mu);
    std::vector<std::unique_ptr<WavScope>> wscopes;
    std::vector<std::unique_ptr<UnitScope>> uscopes;
    for (int i = 0; i < n; ++i) {
        wscopes.emplace_back(new WavScope(ga[i].engine, ga[i].wav_dev));
        uscopes.emplace_back(new UnitScope(ga[i].engine, ga[i].src_cell_dev, ga[i].src_w_dev, ga[i].active_dev));
    }
    HIPCK(hipSetDevice(order[0]->device));
    hipStream_t st = (hipStream_t)stream;
    int rc;
    for (int i = 0; i < n; ++i)
        if ((rc = fence(ga[i].engine, st))) return rc;
    auto energy_of = [&](int i, double** out) -> int {
        wg_engine* e = ga[i].engine;
        int r;
        if ((r = ensure(e, e->scale, (size_t)ga[i].B * sizeof(float)))) return r;
        *out = ga[i].energy_dev;
        if (!*out) {
            if ((r = ensure(e, e->energy, (size_t)ga[i].B * sizeof(double)))) return r;
            *out = (double*)e->energy.p;
        }
        return WG_OK;
    };
    // per group: its kind in the one launch, or -1 (then every group runs alone)
    std::vector<int> kind(n, -1);
    bool fused = true;
    for (int i = 0; i < n && fused; ++i) {
        wg_engine* e = ga[i].engine;
        if (ga[i].B == 0) continue;
        int one[wg::kMaxBands + 1] = {0, e->g.nzi};
        if (!e->streamed) kind[i] = wg::multi_kind(e->g, 1, one);
        else if (use_bands(e) && e->band_on == 1) kind[i] = wg::multi_kind(e->g, e->bands, e->band_bounds);
        fused = kind[i] >= 0;
    }
    if (!fused) {
        for (int i = 0; i < n; ++i) {
            if (ga[i].B == 0) continue;
            wg_engine* e = ga[i].engine;
            double* energy;
            if ((rc = energy_of(i, &energy))) return rc;
            if ((rc = loss_grad_core(e, ga[i].B, ga[i].vel_dev, ga[i].obs_dev, ga[i].loss_dev, ga[i].grad_dev,
                                     ga[i].traces_dev, energy, (float*)e->scale.p, true, st)))
                return rc;
        }
        return WG_OK;
    }
    // longest tasks first: step count times the kind's forward + adjoint step
    // time (us, measured: 70 x 70 whole unit, 100 x 100 band, 141 x 141 band)
    const double step_us[3] = {17.6, 14.9, 15.2};
    std::vector<int> idx;
    for (int i = 0; i < n; ++i)
        if (ga[i].B > 0) idx.push_back(i);
    if (idx.empty()) return WG_OK;
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) {
        return ga[x].engine->g.nt * step_us[kind[x]] > ga[y].engine->g.nt * step_us[kind[y]];
    });
    wg_engine* lead = ga[0].engine;
    std::vector<wg::MultiGroup> tab(idx.size());
    std::vector<double*> energy(n, nullptr);
    size_t lds = 0, slot_floats = 0;
    int tasks = 0;
    for (size_t k = 0; k < idx.size(); ++k) {
        const int i = idx[k];
        wg_engine* e = ga[i].engine;
        const wg::Grid& g = e->g;
        const int B = ga[i].B, S = e->cfg.shots, units = B * S;
        const long long per = (long long)S * g.nt * g.nrec;
        const long long chunks = (per + 255) / 256;
        if ((rc = energy_of(i, &energy[i]))) return rc;
        if ((rc = ensure(e, e->resid, (size_t)B * per * sizeof(float)))) return rc;
        if ((rc = ensure(e, e->partials, (size_t)B * chunks * sizeof(float)))) return rc;
        if ((rc = ensure(e, e->gunit, (size_t)units * g.nx * g.nz * sizeof(float)))) return rc;
        if (!ga[i].traces_dev && (rc = ensure(e, e->traces, (size_t)B * per * sizeof(float)))) return rc;
        float* traces = ga[i].traces_dev ? ga[i].traces_dev : (float*)e->traces.p;
        int P = 1, bounds[wg::kMaxBands + 1] = {0, g.nzi};
        if (kind[i] != 0) {
            P = e->bands;
            for (int q = 0; q <= P; ++q) bounds[q] = e->band_bounds[q];
        }
        const size_t words = P > 1 ? wg::band_xbuf_words(g, units, P) : 2;
        if ((rc = ensure(e, e->bxbuf[0], words * sizeof(unsigned long long)))) return rc;
        if ((rc = ensure(e, e->bqflags[0], 2 * sizeof(int)))) return rc;
        wg::MultiGroup& m = tab[k];
        m.f = fwd_params(e, ga[i].vel_dev, traces);
        m.a = adj_params(e, ga[i].vel_dev, (const float*)e->resid.p, (float*)e->gunit.p);
        m.b.P = P;
        for (int q = 0; q <= wg::kMaxBands; ++q) m.b.bounds[q] = P > 1 && q <= P ? bounds[q] : 0;
        m.b.xbuf = (unsigned long long*)e->bxbuf[0].p;
        m.b.next = nullptr;
        m.b.err = (int*)e->bqflags[0].p + 1;
        m.b.spin = e->qspin < 0 ? -1 : std::min(e->qspin, 1 << 22);
        m.b.fixed = 1;
        m.obs = ga[i].obs_dev;
        m.scale = (const float*)e->scale.p;
        m.resid = (float*)e->resid.p;
        m.kind = kind[i];
        m.task0 = tasks;
        tasks += units * wg::multi_tasks_per_unit(kind[i]);
        const wg::Geometry geo = e->geo();
        lds = std::max({lds, wg::multi_lds_bytes(kind[i], g, geo, false), wg::multi_lds_bytes(kind[i], g, geo, true)});
        slot_floats = std::max(slot_floats, (size_t)g.nt * g.inner);
    }
    // snapshot slots: one per resident workgroup (one per compute unit: the
    // kernel's LDS and registers allow no second) and a margin
    const int ns = std::min(tasks, lead->cus + 16);
    if ((rc = ensure(lead, lead->mtable, idx.size() * sizeof(wg::MultiGroup)))) return rc;
    if ((rc = ensure(lead, lead->mnext, (size_t)(ns + 2) * sizeof(int)))) return rc;
    if ((rc = ensure(lead, lead->mslots, (size_t)ns * slot_floats * sizeof(float)))) return rc;
    int* next = (int*)lead->mnext.p;  // [counter, ns slot flags, slot error]
    int* slot_err = next + 1 + ns;
    for (size_t k = 0; k < idx.size(); ++k) {
        const int i = idx[k];
        wg_engine* e = ga[i].engine;
        const long long per = (long long)e->cfg.shots * e->g.nt * e->g.nrec;
        LAUNCH(wg::launch_energy(ga[i].obs_dev, per, ga[i].B, energy[i], (float*)e->scale.p, st));
        LAUNCH(wg::launch_zero((float*)e->bxbuf[0].p, e->bxbuf[0].n / sizeof(float), st));
        LAUNCH(wg::launch_zero((float*)e->bqflags[0].p, 2, st));
        LAUNCH(wg::launch_multi_put((wg::MultiGroup*)lead->mtable.p, (int)k, tab[k], st));
    }
    LAUNCH(wg::launch_zero((float*)next, (size_t)ns + 2, st));
    LAUNCH(wg::launch_fused((const wg::MultiGroup*)lead->mtable.p, (int)idx.size(), tasks, next,
                            (float*)lead->mslots.p, slot_floats, next + 1, ns, slot_err, lds, st));
    for (size_t k = 0; k < idx.size(); ++k) {
        const int i = idx[k];
        wg_engine* e = ga[i].engine;
        const wg::Grid& g = e->g;
        const long long per = (long long)e->cfg.shots * g.nt * g.nrec;
        const long long chunks = (per + 255) / 256;
        // the loss partials (the residual again, the same values the tasks used)
        LAUNCH(wg::launch_residual(tab[k].f.traces, ga[i].obs_dev, (const float*)e->scale.p, per, ga[i].B,
                                   (float*)e->resid.p, (float*)e->partials.p, st));
        LAUNCH(wg::launch_loss_sum((const float*)e->partials.p, chunks, ga[i].B, energy[i], ga[i].loss_dev, st));
        LAUNCH(wg::launch_fold(g, (const float*)e->gunit.p, ga[i].vel_dev, energy[i], ga[i].B, ga[i].grad_dev, st));
        const long long nvis = (long long)g.nvx * g.nvz;
        LAUNCH(wg::launch_flag_nan((const int*)e->bqflags[0].p + 1, ga[i].loss_dev, ga[i].B, ga[i].grad_dev, nvis, st));
        LAUNCH(wg::launch_flag_nan(slot_err, ga[i].loss_dev, ga[i].B, ga[i].grad_dev, nvis, st));
    }
    return WG_OK;
}

int32_t wg_adjoint_apply_dev(wg_engine* e, int32_t B, const float* vel_dev, const float* resid_dev, float* grad_dev,
                             float* dwav_dev, float* traces_dev, void* stream) {
    if (!e || !vel_dev || !resid_dev || B < 0) return fail(WG_E_INVALID, "invalid argument");
    Guard G(e);
    HIPCK(hipSetDevice(e->device));
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = fence(e, st))) return rc;
    if (B == 0) return WG_OK;
    const wg::Grid& g = e->g;
    const int S = e->cfg.shots;
    const size_t per = (size_t)S * g.nt * g.nrec;
    float* traces = traces_dev;
    if (!traces) {
        if ((rc = ensure(e, e->traces, B * per * sizeof(float)))) return rc;
        traces = (float*)e->traces.p;
    }
    if ((rc = ensure(e, e->gunit, (size_t)B * S * e->slay.ustride * sizeof(float)))) return rc;
    if (e->streamed) {
        const StreamPlan P = plan_stream(e, B, true);
        if ((rc = stream_alloc(e, P, P.models * S))) return rc;
        for (int b0 = 0; b0 < B; b0 += P.models) {
            const int nb = std::min(P.models, B - b0);
            if ((rc = stream_forward(e, P, nb * S, b0 * S, vel_dev, wav_ptr(e), wav_stride(e), traces, true, st))) return rc;
            if ((rc = stream_adjoint(e, P, nb * S, b0 * S, vel_dev, resid_dev, (float*)e->gunit.p, dwav_dev, st)))
                return rc;
        }
        if (grad_dev) LAUNCH(wg::launch_fold(g, (const float*)e->gunit.p, vel_dev, nullptr, B, grad_dev, e->slay, st));
        return WG_OK;
    }
    const int mc = models_per_chunk(e, B);
    if ((rc = ensure(e, e->snaps, snap_bytes_per_unit(e) * S * mc))) return rc;
    for (int b0 = 0; b0 < B; b0 += mc) {
        const int nb = std::min(mc, B - b0);
        if ((rc = run_chunk_forward(e, b0, nb, vel_dev, traces, true, st))) return rc;
        if ((rc = run_chunk_adjoint(e, b0, nb, vel_dev, resid_dev, (float*)e->gunit.p, dwav_dev, st))) return rc;
    }
    if (grad_dev) LAUNCH(wg::launch_fold(g, (const float*)e->gunit.p, vel_dev, nullptr, B, grad_dev, st));
    return WG_OK;
}

int32_t wg_born_dev(wg_engine* e, int32_t B, const float* vel_dev, const float* dvel_dev, float* dtraces_dev,
                    void* stream) {
    if (!e || !vel_dev || !dvel_dev || !dtraces_dev || B < 0) return fail(WG_E_INVALID, "invalid argument");
    Guard G(e);
    HIPCK(hipSetDevice(e->device));
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = fence(e, st))) return rc;
    if (B == 0) return WG_OK;
    const wg::Grid& g = e->g;
    const int S = e->cfg.shots;
    const size_t per = (size_t)S * g.nt * g.nrec;
    if ((rc = ensure(e, e->traces, B * per * sizeof(float)))) return rc;
    if (e->streamed) {
        const StreamPlan P = plan_stream(e, B, true);
        if (P.nseg > 1) return fail(WG_E_TOO_LARGE, "the Born operator needs every snapshot of a model in memory");
        if ((rc = stream_alloc(e, P, P.models * S))) return rc;
        const size_t cells = e->slay.ustride;
        for (int b0 = 0; b0 < B; b0 += P.models) {
            const int nb = std::min(P.models, B - b0);
            if ((rc = stream_forward(e, P, nb * S, b0 * S, vel_dev, wav_ptr(e), wav_stride(e), (float*)e->traces.p, true, st)))
                return rc;
            LAUNCH(wg::launch_zero((float*)e->astate.p, 8 * (size_t)nb * S * cells, st));
            wg::StepParams p = step_params(e, vel_dev, nb * S, b0 * S);
            p.traces = dtraces_dev;
            p.seg_len = P.seg_len;
            p.state = (float*)e->astate.p + e->slay.off;
            p.dvel = dvel_dev;
            p.born_snaps = (const float*)e->snaps.p + e->s_soff;
            for (int t = 0; t < g.nt; ++t) {  // single steps: wavefield buffers 0 and 1, zeta half 0
                p.t = t;
                p.u_cur = t & 1;
                p.u_prev = 1 - (t & 1);
                p.ps_half = t & 1;
                p.ze_half = 0;
                LAUNCH(wg::launch_fwd_step(p, nb * S, false, true, st));
            }
        }
        return WG_OK;
    }
    const int mc = models_per_chunk(e, B);
    if ((rc = ensure(e, e->snaps, snap_bytes_per_unit(e) * S * mc))) return rc;
    for (int b0 = 0; b0 < B; b0 += mc) {
        const int nb = std::min(mc, B - b0);
        if ((rc = run_chunk_forward(e, b0, nb, vel_dev, (float*)e->traces.p, true, st))) return rc;
        wg::FwdParams p = fwd_params(e, vel_dev, dtraces_dev);
        p.unit0 = b0 * S;
        p.dvel = dvel_dev;
        p.born_snaps = (const float*)e->snaps.p;
        LAUNCH(wg::launch_forward(p, nb * S, false, true, st));
    }
    return WG_OK;
}

static int32_t simulate_wavelets_locked(wg_engine* e, int32_t B, const float* vel_dev, const float* wav_dev,
                                        float* traces_dev, void* stream) {
    if (!vel_dev || !wav_dev || !traces_dev || B < 0) return fail(WG_E_INVALID, "invalid argument");
    HIPCK(hipSetDevice(e->device));
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = fence(e, st))) return rc;
    if (B == 0) return WG_OK;
    if (use_bands(e)) return band_simulate(e, 0, B, vel_dev, wav_dev, e->g.nt, traces_dev, st);
    if (e->streamed) return stream_simulate(e, 0, B, vel_dev, wav_dev, e->g.nt, traces_dev, st);
    wg::FwdParams p = fwd_params(e, vel_dev, traces_dev);
    p.wavelet = wav_dev;
    p.wav_stride = e->g.nt;
    LAUNCH(wg::launch_forward(p, B * e->cfg.shots, false, false, st));
    return WG_OK;
}

int32_t wg_simulate_wavelets_dev(wg_engine* e, int32_t B, const float* vel_dev, const float* wav_dev,
                                 float* traces_dev, void* stream) {
    if (!e) return fail(WG_E_INVALID, "invalid argument");
    Guard G(e);
    return simulate_wavelets_locked(e, B, vel_dev, wav_dev, traces_dev, stream);
}

int32_t wg_simulate_units_dev(wg_engine* e, int32_t B, const int32_t* src_cell_dev, const float* src_w_dev,
                              const uint8_t* active_dev, const float* vel_dev, const float* wav_dev,
                              float* traces_dev, void* stream) {
    if (!e || !src_cell_dev || !src_w_dev) return fail(WG_E_INVALID, "invalid argument");
    Guard G(e);
    UnitScope U(e, src_cell_dev, src_w_dev, active_dev);
    return simulate_wavelets_locked(e, B, vel_dev, wav_dev, traces_dev, stream);
}

int32_t wg_set_snapshot_budget(wg_engine* e, uint64_t bytes) {
    if (!e || bytes == 0) return fail(WG_E_INVALID, "budget must be positive");
    Guard G(e);
    e->snap_budget = (size_t)bytes;
    return WG_OK;
}

int32_t wg_engine_streamed(const wg_engine* e) { return e ? (e->streamed ? 1 : 0) : -1; }

int32_t wg_causal_filter_dev(const float* x_dev, int64_t nseries, int32_t nt, int32_t nrec, const float* taps_dev,
                             int32_t ntaps, float* y_dev, void* stream) {
    if (!x_dev || !y_dev || !taps_dev || nseries < 0 || nt < 1 || nrec < 1 || ntaps < 1 || x_dev == y_dev)
        return fail(WG_E_INVALID, "invalid argument");
    const hipError_t err = wg::launch_causal_fir(x_dev, nseries, nt, nrec, taps_dev, ntaps, y_dev, (hipStream_t)stream);
    if (err != hipSuccess) return fail(WG_E_DEVICE, std::string("causal filter launch: ") + hipGetErrorString(err));
    return WG_OK;
}

int32_t wg_illumination_dev(wg_engine* e, int32_t B, const float* vel_dev, float* illum_dev, void* stream) {
    if (!e || !vel_dev || !illum_dev || B < 0) return fail(WG_E_INVALID, "invalid argument");
    Guard G(e);
    HIPCK(hipSetDevice(e->device));
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = fence(e, st))) return rc;
    if (B == 0) return WG_OK;
    if (e->streamed) return fail(WG_E_TOO_LARGE, "illumination needs the resident kernels (grid fits one workgroup)");
    const wg::Grid& g = e->g;
    const int S = e->cfg.shots;
    const size_t per = (size_t)S * g.nt * g.nrec;
    if ((rc = ensure(e, e->traces, B * per * sizeof(float)))) return rc;
    const int mc = models_per_chunk(e, B);
    if ((rc = ensure(e, e->snaps, snap_bytes_per_unit(e) * S * mc))) return rc;
    for (int b0 = 0; b0 < B; b0 += mc) {
        const int nb = std::min(mc, B - b0);
        if ((rc = run_chunk_forward(e, b0, nb, vel_dev, (float*)e->traces.p, true, st))) return rc;
        LAUNCH(wg::launch_illum(g, (const float*)e->snaps.p, nb, illum_dev + (size_t)b0 * g.nvx * g.nvz, st));
    }
    return WG_OK;
}

int32_t wg_stream_tiles(void) { return wg::kStreamTiles; }

int32_t wg_engine_bands(const wg_engine* e) { return e ? e->bands : -1; }

int32_t wg_set_band_split(wg_engine* e, int32_t on) {
    if (!e) return fail(WG_E_INVALID, "null engine");
    Guard G(e);
    if (on < 0 || on > 2) return fail(WG_E_INVALID, "band split must be 0, 1 or 2");
    e->band_on = on;
    return WG_OK;
}

int32_t wg_set_overlap(wg_engine* e, int32_t on) {
    if (!e) return fail(WG_E_INVALID, "null engine");
    Guard G(e);
    e->overlap = on != 0;
    return WG_OK;
}

int32_t wg_set_time_split(wg_engine* e, int32_t parts) {
    if (!e || parts < 0 || parts > 16) return fail(WG_E_INVALID, "parts must be in [0, 16]");
    Guard G(e);
    e->qparts = parts;
    return WG_OK;
}

int32_t wg_set_queue_wait_limit(wg_engine* e, int32_t polls) {
    if (!e) return fail(WG_E_INVALID, "null engine");
    Guard G(e);
    e->qspin = polls < 0 ? -1 : polls;
    return WG_OK;
}

int32_t wg_set_time_block(wg_engine* e, int32_t steps) {
    if (!e || steps < 1 || steps > 2) return fail(WG_E_INVALID, "steps must be 1 or 2");
    Guard G(e);
    e->time_block = steps;
    return WG_OK;
}

int32_t wg_set_forward_time_block(wg_engine* e, int32_t steps) {
    if (!e || steps < 1 || steps > 2) return fail(WG_E_INVALID, "steps must be 1 or 2");
    Guard G(e);
    e->fwd_block = steps;
    return WG_OK;
}

int32_t wg_set_stream_tile(wg_engine* e, int32_t tile) {
    if (!e || tile < 0 || tile >= wg::kStreamTiles) return fail(WG_E_INVALID, "tile must be in [0, wg_stream_tiles())");
    Guard G(e);
    e->tile = tile;
    return WG_OK;
}

int32_t wg_timing_enable(wg_engine* e, int32_t on) {
    if (!e) return fail(WG_E_INVALID, "null engine");
    Guard G(e);
    e->timing = on != 0;
    return WG_OK;
}

int32_t wg_timing_read(wg_engine* e, double* fwd_ms, int64_t* fwd_n, double* adj_ms, int64_t* adj_n) {
    if (!e) return fail(WG_E_INVALID, "null engine");
    Guard G(e);
    HIPCK(hipSetDevice(e->device));
    double sums[2] = {0.0, 0.0};
    int64_t counts[2] = {0, 0};
    int which = 0;
    for (auto* l : {&e->ev_fwd, &e->ev_adj}) {
        for (auto& pr : *l) {
            HIPCK(hipEventSynchronize(pr.b));
            float ms = 0.0f;
            HIPCK(hipEventElapsedTime(&ms, pr.a, pr.b));
            sums[which] += ms;
            counts[which] += pr.n;
            (void)hipEventDestroy(pr.a);
            (void)hipEventDestroy(pr.b);
        }
        l->clear();
        ++which;
    }
    if (fwd_ms) *fwd_ms = sums[0];
    if (fwd_n) *fwd_n = counts[0];
    if (adj_ms) *adj_ms = sums[1];
    if (adj_n) *adj_n = counts[1];
    return WG_OK;
}

}  // extern "C"
