// Internal interface between the host runtime (engine.hip) and the gfx950
// kernels (wave_kernels.hip).  Not part of the public C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wg {

// Padded-grid description shared by every kernel.  nx = nvx + 2 pad,
// pad = 2 + pml; the inner (updated) cells are x, z in [2, n-2).
struct Grid {
    int nx, nz;           // padded grid
    int ni, nzi;          // inner grid (nx-4, nz-4)
    int inner;            // ni * nzi
    int pad;              // FD_PAD + pml
    int region;           // forward C-PML region, pml + 4 (wave_wgpu.rs:291)
    int region_adj;       // adjoint C-PML region, pml + 6 (wave_wgpu.rs:295)
    int swf, swa;         // LDS strip widths: region + 2, region_adj + 2
    int nvx, nvz;         // visible grid
    int nt, nrec, shots;
    float rh, rh2, dt2;   // 1/h, rh*rh, dt*dt (float32, lib.rs:95-97)
    float vmin, vmax;
};

// Per-engine device tables.
struct Geometry {
    const float* prof;       // ax bx dbx (nx each) az bz dbz (nz each)
    const int* src_cell;     // [shots*4] corner cells (padded index)
    const float* src_w;      // [shots*4] bilinear weights
    const int* rec_cell;     // [nrec*4]
    const float* rec_w;      // [nrec*4]
    // Receiver CSR restricted to the rows [band_z0, band_z0 + band_rows):
    // band_off[(z - band_z0) * nx + x] .. band_off[... + 1] index ent_r/ent_w.
    const int* band_off;
    const int* ent_r;
    const float* ent_w;
    int band_z0, band_rows, n_ent;
    // Uneven batches: per-unit source slots (src_cell / src_w hold one entry
    // of 4 corners per unit) and an activity mask; inactive units are padding
    // and produce zero traces and gradients.  Defaults: slot = unit % shots,
    // every unit active.
    int src_by_unit;
    const unsigned char* active;
};

__device__ __forceinline__ int src_slot(const Geometry& geo, int unit, int shot) {
    return geo.src_by_unit ? unit : shot;
}
__device__ __forceinline__ bool unit_active(const Geometry& geo, int unit) { return !geo.active || geo.active[unit]; }

// Outputs of a padding unit: zero `n` floats at a (and `m` at b), block-strided.
// Kept out of line so that it costs the resident kernels no registers.
static __device__ __noinline__ void zero_unit_outputs(float* a, int n, float* b, int m) {
    for (int j = threadIdx.x; j < n; j += blockDim.x) a[j] = 0.0f;
    for (int j = threadIdx.x; j < m; j += blockDim.x) b[j] = 0.0f;
}

struct FwdParams {
    Grid g;
    Geometry geo;
    const float* vel;        // [B, nvz, nvx] m/s
    const float* wavelet;    // [units or 1, nt]
    long long wav_stride;    // 0 = one wavelet for every unit
    float* traces;           // [units, nt, nrec]
    float* snaps;            // [launch units, nt, inner] or null (indexed by
                             // the unit's position in this launch)
    const float* dvel;       // Born: [B, nvz, nvx] perturbation
    const float* born_snaps; // Born: base snapshots, same indexing as snaps
    int unit0;               // first unit of this launch (model b = unit / shots)
};

struct AdjParams {
    Grid g;
    Geometry geo;
    const float* vel;        // [B, nvz, nvx]
    const float* wavelet;    // [nt] (source term of the gradient)
    long long wav_stride;
    const float* resid;      // [units, nt, nrec] scaled residual
    const float* snaps;      // [launch units, nt, inner] (position in launch)
    float* grad;             // [units, cells] padded-grid gradient per unit
    float* dwav;             // [units, nt] or null
    int unit0;
};

// One time step of the HBM-streamed kernels (stream_kernels.hip) for the units
// [unit0, unit0 + units) of a chunk.  state holds the chunk's arrays, each
// [units, nx * nz]: forward u0 u1 psix0 psix1 psiz0 psiz1 zetax zetaz; adjoint
// lambda0 lambda1 Px0 Px1 Zx0 Zx1 Pz0 Pz1 Zz0 Zz1.
struct StepParams {
    Grid g;
    Geometry geo;
    const float* vel;        // [B, nvz, nvx]
    const float* wavelet;    // [units or 1, nt]
    long long wav_stride;
    float* traces;           // forward: [all units, nt, nrec] or null (no recording)
    float* snaps;            // [seg_len, units] of sinner floats: written (forward STORE) or read (adjoint)
    int snap_t0, seg_len;    // snapshot of step t is entry t - snap_t0
    float* state;
    int units, unit0, t;
    const float* dvel;       // Born
    const float* born_snaps; // Born: base snapshots, same layout as snaps
    const float* resid;      // adjoint: [all units, nt, nrec]
    float* grad;             // adjoint: [all units] in the storage layout below, accumulated
    float* dwav;             // adjoint: [all units, nt] or null
    int tile;                // tile shape: 0 64x16, 1 128x8, 2 32x32, 3 64x32
    // adjoint state buffers: lambda buffers 0..3 (lambda^t, lambda^{t+1}, and
    // for two-step launches the two written ones) and the half (0 / 1) the
    // C-PML memory variables are read from (the other half is written)
    int lam_cur, lam_prev, lam_out0, lam_out1, half;
    // forward state buffers: wavefield buffers 0..3 (u^t, u^{t-1}, and for
    // two-step launches the two written ones, u^{t+1} and u^{t+2}), the half
    // psi is read from (the other half is written) and the half holding zeta
    // (a single step updates it in place, a two-step launch writes the other
    // half); trace_prev: a two-step launch also records the gathers of step
    // t - 1 (u^{t-1}, which the launch before it computed but could not record)
    int u_cur, u_prev, u_out0, u_out1, ps_half, ze_half, trace_prev;
    int xcd_order;           // set by the launchers: tiles renumbered per XCD (large grids)
    // The clamped velocity on the padded grid, [models, nz, nx], entry 0 for
    // model vpad_b0 (launch_pad_velocity), or null.  With it, tiles with no
    // C-PML cell, receiver or source take a lean path with no per-cell
    // clamping, bounds or C-PML tests.
    const float* vpad;
    int vpad_b0;
    // Set by the launchers: ceil(2^32 / d) for d = the tile-grid width, the
    // tiles of a unit, the shots and nx (0: divide); n / d = mulhi(n, m)
    // exactly for n < 2^32 / d, which the launchers check.
    unsigned m_gx, m_gxy, m_shots, m_nx;
    // Framed launches (set by the launchers): a lean kernel covers the
    // rectangle of nlx x nlz tiles whose first cell is inner cell (lx0, lz0),
    // whose cells and 2-cell ring lie outside the adjoint C-PML region and
    // the receiver rows; with bnd_n > 0 a general kernel covers the rest of
    // the grid, bnd_n tiles a unit in four bands: the rows above the
    // rectangle (gxa x gzt tiles), the rows below it (gxa x gzb), then the
    // columns left (gxl x gzs) and right (gxr x gzs) of it, each tile
    // storing only its band's cells.  m_*: multipliers of gxa, gxl, gxr.
    // The band kernel's tiles may differ from the lean kernel's: [lx0, lx1)
    // x [lz0, lz1) is the rectangle, the side bands are gzs tile rows.
    int lx0, lz0, nlx, nlz, lx1, lz1;
    int bnd_n, gxa, gzt, gzb, gxl, gxr, gzs;
    unsigned m_gxa, m_gxl, m_gxr;
    // Storage layout of the per-unit arrays (state, gradient, vpad): padded
    // cell (x, z) of unit u at u * sc + z * sp + x from the pointers above,
    // which include the offset of cell (0, 0) within a unit's block.  On
    // wide grids the row pitch sp is a multiple of 32 floats and that offset
    // puts the first inner column at a 128-byte boundary, so that the tiles'
    // rows are whole cache lines.  Snapshots: inner cell (x, z) of unit u
    // and step t at (t * units + u) * sinner + (z - 2) * sni + (x - 2).
    int sp;
    size_t sc;
    int sni;
    size_t sinner;
};
// Layout of per-unit padded-grid arrays: cell (x, z) of unit u at
// u * ustride + off + z * pitch + x.  plain_layout(g) is (nx, 0, nx * nz).
struct CellLayout {
    int pitch, off;
    size_t ustride;
};
inline CellLayout plain_layout(const Grid& g) { return CellLayout{g.nx, 0, (size_t)g.nx * g.nz}; }
// vpad[m, z, x] = clamp(vel[b0 + m] at the clamped model cell of padded (x, z)),
// in layout `lay`
hipError_t launch_pad_velocity(const Grid& g, const float* vel, int b0, int nb, float* vpad, const CellLayout& lay,
                               hipStream_t st);
constexpr int kStreamTiles = 4;

// Arrays of the streamed forward state, each [units] in the storage layout: wavefield
// buffers 0..3, psi_x halves (2, 3), psi_z halves (4, 5), zeta_x and zeta_z
// halves.  Single steps use wavefield buffers 0 and 1, the psi halves by
// turns and one zeta half: arrays 0..7.
constexpr int kFwdArrays = 12;
__host__ __device__ inline int fwd_u_arr(int k) { return k < 2 ? k : k + 6; }
__host__ __device__ inline int fwd_zx_arr(int h) { return h ? 10 : 6; }
__host__ __device__ inline int fwd_zz_arr(int h) { return h ? 11 : 7; }

// Band launches of the resident kernels: a unit whose grid does not fit one
// workgroup is cut into P row bands, one workgroup each, which exchange their
// edge rows after every step (wave_kernels.hip).  A workgroup takes its
// (unit, band) from a device counter when it starts, so a unit's bands start
// one after another and bands that wait for each other run together.
constexpr int kMaxBands = 4;
struct BandLaunch {
    int P;
    int bounds[kMaxBands + 1];  // band p holds the inner rows [bounds[p], bounds[p + 1])
    unsigned long long* xbuf;   // band_xbuf_words(g, units, P) words, zero at launch
    int* next;                  // task counter, zero at launch
    int* err;                   // set to 1 when a wait gives up
    int spin;                   // polls before a wait gives up (< 0: at once)
    int fixed;                  // 1: the compiled-grid band kernels where the grid and split have them
};
// The split of a grid too large for one workgroup into bands that fit (P in
// [2, kMaxBands], boundaries in the interior rows at least one adjoint strip
// width from either edge, the receiver rows in the top band), or false.
bool band_plan(const Grid& g, const Geometry& geo, int lds_limit, int* P, int bounds[kMaxBands + 1]);
size_t band_xbuf_words(const Grid& g, int units, int P);
hipError_t launch_forward_band(const FwdParams& p, int units, bool store, const BandLaunch& b, hipStream_t st);
hipError_t launch_adjoint_band(const AdjParams& p, int units, const BandLaunch& b, hipStream_t st);

hipError_t launch_zero(float* p, size_t n, hipStream_t st);
hipError_t launch_illum(const Grid& g, const float* snaps, int nb, float* out, hipStream_t st);
hipError_t launch_causal_fir(const float* x, long long nseries, int nt, int nrec, const float* h, int L, float* y,
                             hipStream_t st);
hipError_t launch_fwd_step(const StepParams& p, int units, bool store, bool born, hipStream_t st);
// Two forward steps (t, t + 1) per launch; no Born variant.
hipError_t launch_fwd_step2(const StepParams& p, int units, bool store, hipStream_t st);
// Gathers of step p.t from wavefield buffer p.u_cur (after a two-step launch,
// whose second step's input no launch of its own records).
hipError_t launch_fwd_record(const StepParams& p, int units, hipStream_t st);
hipError_t launch_adj_step(const StepParams& p, int units, hipStream_t st);
// Two adjoint steps (t, t - 1) per launch; p.dwav must be null.
hipError_t launch_adj_step2(const StepParams& p, int units, hipStream_t st);

// Returns hipSuccess or the launch error.  lds_bytes_* report the dynamic
// LDS a launch needs (0 when the grid does not fit one workgroup).
size_t forward_lds_bytes(const Grid& g);
size_t adjoint_lds_bytes(const Grid& g, const Geometry& geo);
// Register slots per thread for the resident kernels (a compiled K >=
// ceil(inner / 512)), or -1 when the grid has too many inner cells.
int slots_needed(const Grid& g);

hipError_t launch_forward(const FwdParams& p, int units, bool store, bool born,
                          hipStream_t st);

// Kernels specialised for the reference grid (70 x 70, pml 20): return false
// when the grid is another one (then the runtime-geometry kernels run).
bool launch_forward_fx(const FwdParams& p, int units, bool store, bool born, hipStream_t st,
                       hipError_t* err);
bool launch_adjoint_fx(const AdjParams& p, int units, hipStream_t st, hipError_t* err);

// Time-split work queues of the compiled-grid kernels: the forward (storing
// snapshots, f.unit0 = a.unit0 = 0) or the adjoint of `units` units, each cut
// into `parts` time segments, in one launch of `workgroups` workgroups.
// next_f, next_a, err and the done flags (parts * units ints each) start at
// zero.  They return false when the grid is not the compiled one.
struct QueueParams {
    FwdParams f;
    AdjParams a;
    float* fstate;  // per unit: queue_state_fx's forward floats
    float* astate;  // per unit: its adjoint floats
    int *next_f, *next_a, *err;
    int *done_f, *done_a;
    int units, parts;
    int spin_limit;  // polls before a wait gives up and sets *err (< 0: at once)
};
bool queue_state_fx(const Grid& g, size_t* fwd_floats, size_t* adj_floats);
bool launch_queue_fx(const QueueParams& q, bool adjoint, int workgroups, hipStream_t st, hipError_t* err);
bool launch_queue_fwd_fx(const QueueParams& q, int workgroups, hipStream_t st, hipError_t* err);
// After a queued loss and gradient: NaN losses and gradients [B, nvis] (grad
// may be null) when a queue wait timed out (*err != 0).
hipError_t launch_flag_nan(const int* err, float* loss, int B, float* grad, long long nvis, hipStream_t st);
size_t forward_lds_bytes_fx(const Grid& g);
size_t adjoint_lds_bytes_fx(const Grid& g, const Geometry& geo);
// Row-band kernels of the compiled grids (the 100 x 100 and 141 x 141
// models, pml 20, and small batches of the 70 x 70 one): band_split_fx gives
// the compiled split of a grid too large for one workgroup
// (false for another grid, or receivers outside its top band); the launchers
// return false when the grid or the split b is not a compiled one.
bool band_split_fx(const Grid& g, const Geometry& geo, int* P, int bounds[kMaxBands + 1]);
// The compiled split of one call over `units` units: band_split_fx's, or a
// split into more bands where that call ends sooner on `cus` compute units.
bool band_split_call_fx(const Grid& g, const Geometry& geo, int units, int cus, int* P,
                        int bounds[kMaxBands + 1]);
// The 70 x 70 model (pml 20) as 4 or 2 bands when `units` units of that many
// bands fit `cus` compute units (one band workgroup per compute unit): the
// split, or false (another grid, or too many units).
bool band_split_small_fx(const Grid& g, const Geometry& geo, int units, int cus, int* P,
                         int bounds[kMaxBands + 1]);
bool launch_forward_band_fx(const FwdParams& p, int units, bool store, const BandLaunch& b, hipStream_t st,
                            hipError_t* err);
bool launch_adjoint_band_fx(const AdjParams& p, int units, const BandLaunch& b, hipStream_t st, hipError_t* err);

// An uneven batch (one group per (grid, nt)) in one launch: every group's
// units (70 x 70: one task per unit) or bands (100 x 100 and 141 x 141 in
// their compiled splits: one task per band) are tasks of ONE launch; a
// workgroup takes a task from a device counter and runs that unit's or
// band's forward, the unit's residual and its adjoint, with the snapshots in
// a slot it holds for the task's lifetime.  The table lists the groups in
// task order.
struct MultiGroup {
    FwdParams f;          // the group's forward (f.snaps is set per task)
    AdjParams a;          // and adjoint (a.resid: written by the task)
    BandLaunch b;         // its exchange words and error flag (b.P: bands per unit, 1 for whole units)
    const float* obs;     // [units, nt, nrec]
    const float* scale;   // [models]
    float* resid;         // [units, nt, nrec] (= a.resid)
    int kind;             // multi_kind
    int task0;            // first task of the group
};
constexpr int kMultiMaxGroups = 64;
// The kind of a grid in the one-launch batch (0: 70 x 70 as whole units, 1:
// 100 x 100 in 2 bands, 2: 141 x 141 in 4 bands; bounds must be the compiled
// split), or -1 when the grid has no compiled kernel for it.
int multi_kind(const Grid& g, int P, const int* bounds);
// Tasks per unit and the dynamic LDS a group of that kind needs.
int multi_tasks_per_unit(int kind);
size_t multi_lds_bytes(int kind, const Grid& g, const Geometry& geo, bool adjoint);
// Writes group entry i of the device table (a kernel: capturable).
hipError_t launch_multi_put(MultiGroup* table, int i, const MultiGroup& m, hipStream_t st);
// The one launch: next (task counter) zero and slot_flags (ns of them) zero
// at launch; ns snapshot slots of slot_floats floats; err set when a task
// found no free slot.
hipError_t launch_fused(const MultiGroup* table, int n, int tasks, int* next, float* slots,
                        unsigned long long slot_floats, int* slot_flags, int ns, int* err, size_t lds, hipStream_t st);
hipError_t launch_adjoint(const AdjParams& p, int units, hipStream_t st);

// Misfit: energy (f64), residual (p-d)*scale and the reference's 256-chunk
// float32 tree partials summed in f64.
hipError_t launch_energy(const float* obs, long long per_model, int B,
                         double* energy, float* scale, hipStream_t st);
hipError_t launch_residual(const float* pred, const float* obs,
                           const float* scale, long long per_model, int B,
                           float* resid, float* partials, hipStream_t st);
hipError_t launch_loss_sum(const float* partials, long long chunks, int B,
                           const double* energy, float* loss, hipStream_t st);
// grad_vis[b] = fold(mask * sum_s grad_unit[b*S+s]); NaN where energy == 0.
// grad_unit in layout `lay` (the plain layout unless given).
hipError_t launch_fold(const Grid& g, const float* grad_unit, const float* vel,
                       const double* energy, int B, float* grad_vis,
                       hipStream_t st);
hipError_t launch_fold(const Grid& g, const float* grad_unit, const float* vel,
                       const double* energy, int B, float* grad_vis,
                       const CellLayout& lay, hipStream_t st);

}  // namespace wg
