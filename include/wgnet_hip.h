/*
 * wgnet_hip: C-ABI of the MI355X wave engine (forward, exact discrete adjoint,
 * misfit and gradient) that replaces the PyO3 module `wave_gauss_rs` of the
 * WGNet reference (crates/wave-gauss-rs/src/lib.rs).
 *
 * Conventions
 *  - Every entry point returns 0 on success and a negative WG_E_* code on
 *    failure; wg_last_error() then returns a message (thread-local).
 *  - Host-pointer calls (wg_simulate, wg_evaluate_adjoint) are synchronous,
 *    as the reference's `simulate` / `evaluate_adjoint` are (lib.rs:108-137).
 *  - *_dev calls take device pointers and a hipStream_t (passed as void*;
 *    NULL = the null stream).  They enqueue work and return without waiting.
 *  - Layouts follow the reference: velocity [nvz, nvx] row-major in m/s,
 *    gathers [model, shot, t, receiver] (lib.rs:115), gradients [nvz, nvx]
 *    in d(loss)/d(v in m/s) (lib.rs:118-119).
 *  - One engine must not be used by two threads at once; calls on one engine
 *    are serialised by an internal mutex.
 */
#ifndef WGNET_HIP_H
#define WGNET_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WG_OK 0
#define WG_E_INVALID (-1)   /* bad argument: maps to ValueError */
#define WG_E_DEVICE (-2)    /* HIP runtime failure: maps to RuntimeError */
#define WG_E_TOO_LARGE (-3) /* grid does not fit the resident kernel */
#define WG_E_ZERO_ENERGY (-4) /* observed traces have zero energy (ValueError) */

typedef struct wg_engine wg_engine;

/* Engine parameters.  Defaults (wg_config_default) are those of
 * PyWaveEngine(shots=5, nt=2001, dt=0.0005, peak_hz=15.0, pml=20,
 * vmin=1000.0, vmax=5000.0) (lib.rs:52) on the 70 x 70 grid at h = 10 m with
 * one receiver per surface cell (wave_wgpu.rs:175-215). */
typedef struct {
    int32_t shots;
    int32_t nt;
    float dt;
    float peak_hz;
    int32_t pml;
    float vmin;
    float vmax;
    int32_t nvx; /* visible grid columns (reference: 70) */
    int32_t nvz; /* visible grid rows (reference: 70) */
    float h;     /* spacing in metres (reference: 10) */
} wg_config;

/* Fills the reference defaults (lib.rs:52, geology-core lib.rs:5-9). */
void wg_config_default(wg_config *cfg);

/* Thread-local message of the last failing call ("" if none). */
const char *wg_last_error(void);

/* Replaces PyWaveEngine::new / NativeWaveEngine::new_async
 * (lib.rs:51-57, wave_wgpu.rs:156-712): geometry (default sources at
 * x_k = xmax*k/(S-1), z = h, wave_wgpu.rs:204-214), C-PML profiles,
 * Ricker wavelet, device buffers.  Uses the current HIP device. */
int32_t wg_engine_create(const wg_config *cfg, wg_engine **out);
void wg_engine_destroy(wg_engine *e);

/* Getters of lib.rs:59-72 plus the padded-grid sizes. */
int32_t wg_engine_shots(const wg_engine *e);
int32_t wg_engine_nt(const wg_engine *e);
int32_t wg_engine_nrec(const wg_engine *e);
int32_t wg_engine_dims(const wg_engine *e, int32_t *nx, int32_t *nz,
                       int32_t *pad);

/* Source positions [shots, 2] (x, z) in metres: lib.rs:74-90,
 * wave_wgpu.rs:761-785 (bilinear corners recomputed). */
int32_t wg_get_sources(const wg_engine *e, float *xz, int32_t n);
int32_t wg_set_sources(wg_engine *e, const float *xz, int32_t n);

/* Source wavelet [nt]: lib.rs:92-106, wave_wgpu.rs:788-795. */
int32_t wg_get_wavelet(const wg_engine *e, float *w, int32_t n);
int32_t wg_set_wavelet(wg_engine *e, const float *w, int32_t n);

/* PyWaveEngine.simulate (lib.rs:108-116, wave_wgpu.rs:799-844): host
 * velocity [nvz*nvx] m/s -> host gathers [shots, nt, nrec].  Synchronous. */
int32_t wg_simulate(wg_engine *e, const float *vel_host, float *traces_host);

/* PyWaveEngine.evaluate_adjoint (lib.rs:118-137, wave_wgpu.rs:848-988):
 * loss = 0.5 ||p - d||^2 / ||d||^2 and grad [nvz*nvx] = dloss/dv (m/s).
 * Returns WG_E_ZERO_ENERGY when sum(obs^2) == 0 (the reference panics with
 * "Observed traces have zero energy", wave_wgpu.rs:853).  Synchronous. */
int32_t wg_evaluate_adjoint(wg_engine *e, const float *vel_host,
                            const float *obs_host, float *loss_host,
                            float *grad_host);

/* Batched device path of WavePhysics.simulate (physics.py:59-64): B models
 * vel_dev [B, nvz, nvx] -> traces_dev [B, shots, nt, nrec]. */
int32_t wg_simulate_dev(wg_engine *e, int32_t B, const float *vel_dev,
                        float *traces_dev, void *stream);

/* Batched device path of WavePhysics.loss_and_grad (physics.py:66-76) in
 * m/s units: per model b, loss_dev[b] and grad_dev [B, nvz, nvx].
 * traces_dev [B, shots, nt, nrec] receives the predicted gathers (may be
 * NULL).  Models whose observed energy is zero get loss = NaN and a NaN
 * gradient; energy_dev [B] (may be NULL) receives sum(obs^2) per model in
 * double so the caller can raise the reference's error. */
int32_t wg_loss_grad_dev(wg_engine *e, int32_t B, const float *vel_dev,
                         const float *obs_dev, float *loss_dev,
                         float *grad_dev, float *traces_dev,
                         double *energy_dev, void *stream);

/* The phys stage's two calls in one (train.py:262-263: physics.simulate of
 * the true models, physics.py:59-64, then physics.loss_and_grad of the
 * predicted ones, physics.py:66-76): observed gathers of true_dev
 * [B, nvz, nvx] m/s into obs_dev [B, shots, nt, nrec], then wg_loss_grad_dev
 * of vel_dev against them.  When the batch is one resident round of units
 * the observed forward runs beside the storing forward; results equal
 * wg_simulate_dev followed by wg_loss_grad_dev bit for bit. */
int32_t wg_loss_grad_true_dev(wg_engine *e, int32_t B, const float *vel_dev,
                              const float *true_dev, float *obs_dev,
                              float *loss_dev, float *grad_dev,
                              float *traces_dev, double *energy_dev,
                              void *stream);

/* wg_loss_grad_dev with per-(model, shot) wavelets wav_dev [B, shots, nt]
 * in place of the engine wavelet, so that models of one batch can sit in
 * different passbands (the reference changes the engine wavelet per band,
 * physics.py:37-46; here each model of a batched call carries its own).
 * The source term of the gradient uses the same wavelets. */
int32_t wg_loss_grad_wavelets_dev(wg_engine *e, int32_t B, const float *vel_dev,
                                  const float *wav_dev, const float *obs_dev,
                                  float *loss_dev, float *grad_dev,
                                  float *traces_dev, double *energy_dev,
                                  void *stream);

/* The loss alone, for evaluations that need no gradient (a misfit taken
 * without autograd; the reference runs the adjoint inside every
 * evaluate_adjoint, wave_wgpu.rs:848-988): the forward without snapshots,
 * the residual and the loss of
 * wg_loss_grad_dev, whose losses it returns bit for bit.  wav_dev [B, shots,
 * nt] per-(model, shot) wavelets, or NULL for the engine wavelet; traces_dev
 * and energy_dev may be NULL. */
int32_t wg_loss_dev(wg_engine *e, int32_t B, const float *vel_dev,
                    const float *wav_dev, const float *obs_dev, float *loss_dev,
                    float *traces_dev, double *energy_dev, void *stream);

/* wg_loss_dev that allocates nothing (no reference counterpart: it is for
 * callers that cannot allocate device memory, such as code under HIP graph
 * capture, tight serving loops and harnesses that forbid hipMalloc).  Its
 * kernels keep their working data (per-model loss scale, observed energy
 * when energy_dev is NULL, the residual tree's partial sums) in one fixed
 * device array of 4 MiB per GPU, held until the process ends: 4 + 8 bytes per
 * model plus 4 bytes per 256 samples of its gathers, about 760 models of 5
 * shots x nt 1000 x 70 receivers per call; larger batches return
 * WG_E_TOO_LARGE with the limit in the message.  traces_dev [B, shots, nt, nrec] is required;
 * resident grids only (WG_E_TOO_LARGE on an HBM-streamed engine).  Losses
 * equal wg_loss_dev's bit for bit.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other engines, other
 * threads) race. */
int32_t wg_loss_noalloc_dev(wg_engine *e, int32_t B, const float *vel_dev,
                            const float *wav_dev, const float *obs_dev,
                            float *loss_dev, float *traces_dev,
                            double *energy_dev, void *stream);

/* Uneven batches (no reference counterpart; the reference solves one model
 * with the engine's own source list per call, wave_wgpu.rs:761-785).  Each
 * of the B models has `shots` unit slots; unit u = b * shots + k carries its
 * own source position and wavelet, and units whose active flag is 0 are
 * padding (a model with fewer shots): they cost no time steps and give zero
 * traces and zero gradient, so the misfit and gradient of each model are
 * those of its active units alone.
 *   wg_source_corners: host helper, bilinear corners (padded cell index) and
 *     weights of n positions xz [n, 2] in metres, as the kernels use them.
 *   src_cell_dev [B*shots*4], src_w_dev [B*shots*4]: per-unit corners;
 *   active_dev [B*shots] (NULL: all active); wav_dev [B, shots, nt]. */
int32_t wg_source_corners(const wg_engine *e, int32_t n, const float *xz,
                          int32_t *cell, float *w);
int32_t wg_loss_grad_units_dev(wg_engine *e, int32_t B,
                               const int32_t *src_cell_dev,
                               const float *src_w_dev,
                               const uint8_t *active_dev, const float *vel_dev,
                               const float *wav_dev, const float *obs_dev,
                               float *loss_dev, float *grad_dev,
                               float *traces_dev, double *energy_dev,
                               void *stream);
/* The loss and gradient of several engines' batches in one call (no
 * reference counterpart: an uneven batch, one engine per (grid, nt) group;
 * the reference solves one model per engine call, physics.py:62-75).  Each
 * group carries wg_loss_grad_units_dev's arguments for its engine, and the
 * results are the same bits as n such calls.  When every group's grid has
 * compiled kernels (the 70 x 70, 100 x 100 and 141 x 141 models, pml 20, on
 * the default band split), all groups run in ONE launch: a workgroup takes a
 * task (a unit, or one row band of a unit) from a device counter, longest
 * tasks first, and runs its forward, residual and adjoint, with the
 * snapshots in a slot it holds while the task runs (about one slot per
 * compute unit, each the largest group's nt x inner floats, held by the
 * first engine; the engines' snapshot budgets do not apply).  Otherwise the
 * groups run one after another on `stream`.  Each engine may appear once; at
 * most 64 groups. */
typedef struct wg_group_args {
    wg_engine *engine;
    int32_t B;
    const int32_t *src_cell_dev;
    const float *src_w_dev;
    const uint8_t *active_dev;
    const float *vel_dev;
    const float *wav_dev;
    const float *obs_dev;
    float *loss_dev;
    float *grad_dev;
    float *traces_dev;
    double *energy_dev;
} wg_group_args;
int32_t wg_loss_grad_groups_dev(int32_t n, const wg_group_args *groups, void *stream);
int32_t wg_simulate_units_dev(wg_engine *e, int32_t B,
                              const int32_t *src_cell_dev,
                              const float *src_w_dev, const uint8_t *active_dev,
                              const float *vel_dev, const float *wav_dev,
                              float *traces_dev, void *stream);

/* Transpose of the linearised engine at vel_dev, for the dot-product test
 * and gradient checks (no reference counterpart: the reference exposes the
 * adjoint only through evaluate_adjoint, splatAdjoint.wgsl:58-131).
 *   resid_dev [B, shots, nt, nrec]: trace-space vector y (already scaled)
 *   grad_dev  [B, nvz, nvx]: fold(mask * sum_s J_s^T y)    (may be NULL)
 *   dwav_dev  [B, shots, nt]: d<traces, y>/d(wavelet)       (may be NULL)
 *   traces_dev [B, shots, nt, nrec]: forward gathers       (may be NULL) */
int32_t wg_adjoint_apply_dev(wg_engine *e, int32_t B, const float *vel_dev,
                             const float *resid_dev, float *grad_dev,
                             float *dwav_dev, float *traces_dev, void *stream);

/* Tangent-linear (Born) forward: dtraces_dev [B, shots, nt, nrec] =
 * J(vel) dvel, dvel_dev [B, nvz, nvx] (no reference counterpart; the other
 * half of the dot-product test). */
int32_t wg_born_dev(wg_engine *e, int32_t B, const float *vel_dev,
                    const float *dvel_dev, float *dtraces_dev, void *stream);

/* Wavelet-only forward used with dwav of wg_adjoint_apply_dev: gathers for
 * per-(model, shot) wavelets wav_dev [B, shots, nt] (no reference
 * counterpart). */
int32_t wg_simulate_wavelets_dev(wg_engine *e, int32_t B, const float *vel_dev,
                                 const float *wav_dev, float *traces_dev,
                                 void *stream);

/* Upper bound on device memory for wavefield snapshots and checkpoints (no
 * reference counterpart: the reference keeps every snapshot,
 * wave_wgpu.rs:180-189).  Batches are processed in model chunks that fit;
 * grids too large for one workgroup switch to a checkpoint every 100 steps
 * with segment recompute when a model's snapshots do not fit.  Default:
 * half the free device memory at engine creation, at most 96 GiB. */
int32_t wg_set_snapshot_budget(wg_engine *e, uint64_t bytes);
/* 1 when the engine runs the HBM-streamed kernels (grid too large for the
 * resident ones), 0 otherwise. */
int32_t wg_engine_streamed(const wg_engine *e);
/* Row bands (no reference counterpart; the reference's grid is fixed at
 * 70 x 70, wave_wgpu.rs:714-734): a grid too large for one workgroup whose
 * units split into 2 to 4 row bands that each fit (the 100 x 100 and
 * 141 x 141 models) runs forward and adjoint on the band kernels: one
 * workgroup per band holds its rows in LDS and registers for all nt steps,
 * with every snapshot stored, and exchanges two edge rows with each
 * neighbour band per step through global memory.  wg_engine_bands returns
 * the bands per unit (0: the grid does not split, or fits one workgroup).
 * The 100 x 100 and 141 x 141 grids (pml 20) have band kernels compiled for
 * their geometry and split (2 and 3 bands; a 141 x 141 call runs 4 where its
 * unit count fills the compute units in fewer rounds that way, and
 * wg_engine_bands returns the default 3); other grids run runtime-geometry
 * band kernels.  Results equal the HBM-streamed kernels' bit for bit either
 * way.  The band kernels serve simulate and the loss and gradient when a
 * model's snapshots fit the snapshot budget; the Born operator, the wavelet
 * gradient and checkpointed adjoints stay streamed.
 * The reference's 70 x 70 grid fits one workgroup, but a call with at most a
 * quarter (half) as many (model, shot) units as the device has compute units
 * runs it as 4 (2) compiled bands, so that a small batch is not bound by one
 * workgroup's step latency; results equal the whole-unit kernels' bit for
 * bit.  The Born forward and the wavelet-gradient adjoint stay whole-unit.
 * wg_set_band_split(e, on): 1 (default) the band kernels, compiled where the
 * grid has them; 2 the runtime-geometry band kernels only; 0 no bands (the
 * HBM-streamed kernels, or whole-unit workgroups on the 70 x 70 grid). */
int32_t wg_engine_bands(const wg_engine *e);
int32_t wg_set_band_split(wg_engine *e, int32_t on);
/* Tile shapes of the HBM-streamed kernels (no reference counterpart): a
 * 256-thread workgroup computes 64x16 (0, default; the two-step forward's
 * lean interior tiles are 64x32 in it), 128x8 (1), 32x32 (2) or 64x32 (3)
 * cells.  Results do not depend on the shape; an autotuner times them per
 * grid and sets the fastest. */
int32_t wg_stream_tiles(void);
/* Overlapped chunk pipeline of the resident loss/gradient (no reference
 * counterpart; default on): a batch larger than one unit per compute unit is
 * cut into device-sized model chunks, and each chunk's forward runs on a
 * second stream beside the previous chunk's adjoint.  Results are identical
 * either way; off runs the chunks one after another on the caller's stream. */
int32_t wg_set_overlap(wg_engine *e, int32_t on);
int32_t wg_set_stream_tile(wg_engine *e, int32_t tile);
/* Temporal blocking of the HBM-streamed adjoint (no reference counterpart):
 * time steps per launch, 1 or 2 (default).  A two-step launch recomputes a
 * 2-cell ring around each tile and reads and writes the wavefield, memory
 * variables and gradient once per two steps.  Results are identical either
 * way; which is faster depends on the grid (an autotuner times both); the
 * wavelet-gradient adjoint (wg_adjoint_apply_dev with dwav) always runs
 * single steps. */
int32_t wg_set_time_block(wg_engine *e, int32_t steps);
/* Temporal blocking of the HBM-streamed forward (no reference counterpart):
 * time steps per launch, 1 or 2 (default).  A two-step launch recomputes a
 * 2-cell ring around each tile and reads the wavefield, velocity and memory
 * variables and writes them once per two steps (snapshots every step).
 * Gathers, snapshots, checkpoints and everything computed from them are
 * identical either way; the Born forward always runs single steps. */
int32_t wg_set_forward_time_block(wg_engine *e, int32_t steps);
/* Resident work queues (no reference counterpart; grids of the compiled
 * 70 x 70 kernels): the loss and gradient of a whole batch run as one
 * forward launch and one adjoint launch, in which every (model, shot) unit's
 * forward or adjoint is cut into `parts` time segments (default 4) taken from
 * a device-side queue by one workgroup per compute unit; a segment restores
 * its unit's state from HBM.  Results are identical to whole-unit launches;
 * 0 turns the queues off (the chunk pipeline runs instead).  The queues also
 * stay off when a batch has no more units than the device has compute units
 * (nothing to balance) or its snapshots exceed the snapshot budget.
 * Timing (below) counts the forward and adjoint queue launches. */
int32_t wg_set_time_split(wg_engine *e, int32_t parts);

/* Polls a work-queue task makes for its unit's previous part before it gives
 * up (no reference counterpart; default 1 << 24, seconds, where a part takes
 * milliseconds).  A wait that gives up makes every loss and gradient of that
 * call NaN.  A negative value makes every wait give up at once: it drives
 * that failure path in tests, with every kernel still ending normally. */
int32_t wg_set_queue_wait_limit(wg_engine *e, int32_t polls);

/* Kernel timing (no reference counterpart; the reference records no
 * timestamps, wave_wgpu.rs:829).  When enabled, HIP events are recorded on the
 * launch stream around forward and adjoint kernel launches (around each launch
 * of the resident kernels; around each run of per-step launches of the
 * HBM-streamed kernels, a checkpoint recompute counting as forward);
 * wg_timing_read waits for them and returns the summed durations (ms) and
 * kernel launch counts since the last read, then clears them. */
int32_t wg_timing_enable(wg_engine *e, int32_t on);
int32_t wg_timing_read(wg_engine *e, double *fwd_ms, int64_t *fwd_launches,
                       double *adj_ms, int64_t *adj_launches);

/* Source illumination illum_dev [B, nvz, nvx]: for each visible cell, the
 * sum over the model's shots and all steps of the squared snapshot 2 v dt^2 w
 * at that cell (the pseudo-Hessian diagonal that splatAdjoint.wgsl:113-117
 * accumulates and the reference never reads back).  Runs one storing forward;
 * resident grids only (WG_E_TOO_LARGE otherwise). */
int32_t wg_illumination_dev(wg_engine *e, int32_t B, const float *vel_dev,
                            float *illum_dev, void *stream);

/* Causal low-pass of gathers along time, the data half of frequency
 * continuation (python/wgnet/physics.py:48-57, a conv1d with the taps
 * left-padded by ntaps - 1 zeros): for each of the nseries blocks of x_dev
 * [nseries, nt, nrec], y[t][r] = sum_k taps[k] x[t + k - (ntaps - 1)][r],
 * k ascending, float32.  y_dev must not alias x_dev.  Enqueued on stream. */
int32_t wg_causal_filter_dev(const float *x_dev, int64_t nseries, int32_t nt,
                             int32_t nrec, const float *taps_dev,
                             int32_t ntaps, float *y_dev, void *stream);

/* Reference Ricker (geology-core lib.rs:12-21), float32 with expf. */
void wg_ricker_wavelet(int32_t samples, float dt, float peak_hz, float *out);

/* Reference C-PML profiles (geology-core lib.rs:25-68), f64 inside. */
void wg_cpml_profiles(int32_t n, float h, float dt, int32_t width,
                      float max_vel, float freq, float *a, float *b,
                      float *db);

/* ------------------------------------------------------------------------
 * Learned gradient preconditioner (no reference counterpart: the reference
 * refines with plain projected L-BFGS, python/wgnet/refine.py:92-169; this
 * network is applied at its gradient hook, refine.py:105-106).
 *
 * A network of dilated 3x3 convolutions, 2 -> 32 -> 32 -> 32 -> 1 channels
 * with dilations 4, 16, 8, 1, GELU after the first three layers and zero
 * padding equal to the dilation (wg_pc_layer_shape reports each layer), mapping
 * x [N, 2, H, W] (gradient, normalised velocity) to y [N, H, W].  bfloat16
 * weights and activations, float32 MFMA accumulation, float32 master weights
 * and AdamW moments.  Parameters in PyTorch Conv2d order (per layer
 * W[co][ci][ky][kx], then b[co]).  Status codes as above; messages from
 * wg_pc_last_error().  One network must not be driven by two threads at once
 * (calls are serialised by a mutex).
 * ------------------------------------------------------------------------ */
typedef struct wg_pc wg_pc;

const char *wg_pc_last_error(void);
/* max_batch images of height x width; deterministic U(+-1/sqrt(fan_in)) init */
int32_t wg_pc_create(int32_t max_batch, int32_t height, int32_t width,
                     uint64_t seed, wg_pc **out);
void wg_pc_destroy(wg_pc *pc);
int32_t wg_pc_num_params(void);
int32_t wg_pc_num_layers(void);
/* input and output channels and dilation of layer 0 .. num_layers-1 */
int32_t wg_pc_layer_shape(int32_t layer, int32_t *cin, int32_t *cout,
                          int32_t *dilation);
int32_t wg_pc_get_params(wg_pc *pc, float *host, int32_t n);
/* also resets the AdamW moments and step count */
int32_t wg_pc_set_params(wg_pc *pc, const float *host, int32_t n);
/* y_dev [N, H, W] = net(x_dev [N, 2, H, W]), enqueued on stream */
int32_t wg_pc_forward_dev(wg_pc *pc, int32_t N, const float *x_dev,
                          float *y_dev, void *stream);
/* loss = mean((net(x) - target)^2) and its parameter gradient (host,
 * PyTorch order), without updating; waits for the stream */
int32_t wg_pc_loss_grad(wg_pc *pc, int32_t N, const float *x_dev,
                        const float *target_dev, float *loss_host,
                        float *grad_host, void *stream);
/* one AdamW step (beta 0.9 / 0.999, eps 1e-8, decoupled weight decay) on the
 * mean squared error; loss_dev [1] (may be NULL) receives the loss before
 * the step */
int32_t wg_pc_train_step_dev(wg_pc *pc, int32_t N, const float *x_dev,
                             const float *target_dev, float lr,
                             float weight_decay, float *loss_dev,
                             void *stream);

#ifdef __cplusplus
}
#endif
#endif
