/*
 * CPU oracle for the WGNet wave engine: TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker (or the timed CPU baseline);
 * the product path (wgnet_amd) never calls it.
 *
 * What it restates, function by function (paths relative to the WGNet
 * reference repository):
 *   or_ricker_*        crates/geology-core/src/lib.rs:12-21
 *   or_cpml_*          crates/geology-core/src/lib.rs:25-68 (f64 inside)
 *   or_corners         crates/wave-gauss-rs/src/wave_wgpu.rs:217-230, 761-785
 *   or_pad_velocity    crates/wave-gauss-rs/src/wave_wgpu.rs:714-734
 *   or_fold            crates/wave-gauss-rs/src/wave_wgpu.rs:993-1005
 *   or_forward_*       crates/geology-core/src/lib.rs:74-221 (simulate_shot_cpu),
 *                      generalised to an nvx x nvz grid and bilinear stations
 *                      with splatForward.wgsl:61-67 (receiver sum) and
 *                      splatForward.wgsl:108-110 (snapshot 2 v dt^2 w)
 *   or_adjoint_*       crates/wave-gauss-rs/src/shaders/splatAdjoint.wgsl:58-131
 *                      (no CPU adjoint exists in the reference; this is a
 *                      restatement of the GPU kernel's recurrence)
 *   or_born_*          tangent-linear of or_forward (no reference counterpart;
 *                      used for the dot-product test <J dv, y> = <dv, J^T y>)
 *   or_misfit_f32      RESIDUAL_WGSL, wave_wgpu.rs:13-54, 852-854, 960-965
 *   or_evaluate_adjoint_f32  NativeWaveEngine::evaluate_adjoint,
 *                      wave_wgpu.rs:848-988 (shot sum x clamp mask + fold)
 *   (the low-pass taps and causal band-limit, physics.py:18-57, are
 *    restated in numpy in oracle/bandlimit.py)
 *
 * Pinning (what checks each function against something other than itself)
 * is listed in tests/test_oracle.py.
 *
 * Precision: the *_f32 variants follow the reference's float32 operation
 * order exactly (build with -ffp-contract=off so no FMA is formed).  The
 * *_f64 variants run the same algorithm with every operation in double and
 * the derived constants (profiles, 1/h, wavelet) kept in double.
 */
#ifndef WGNET_WAVE_ORACLE_H
#define WGNET_WAVE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Problem geometry for one (model, shot).  Receivers default to one per
 * surface column at z = h (wave_wgpu.rs:215) when rec_xz is NULL. */
typedef struct {
    int32_t nvx, nvz;   /* visible grid (70 x 70 in the reference) */
    int32_t pml;        /* C-PML width in cells (default 20) */
    int32_t nt;         /* time steps */
    float h;            /* grid spacing, m */
    float dt;           /* time step, s */
    float vmin, vmax;   /* velocity clamp, m/s */
    float peak_hz;      /* C-PML frequency and Ricker peak */
    float src_x, src_z; /* source position, m (visible-grid coordinates) */
    int32_t nrec;       /* number of receivers */
    const float *rec_xz;/* [nrec, 2] receiver (x, z) in m, or NULL = default */
} or_problem;

/* Derived padded-grid sizes. */
typedef struct {
    int32_t pad, nx, nz, cells, region, region_adj;
} or_dims;

void or_dims_of(const or_problem *p, or_dims *d);

/* a1: delayed Ricker wavelet. */
void or_ricker_f32(int32_t samples, float dt, float peak_hz, float *out);
void or_ricker_f64(int32_t samples, double dt, double peak_hz, double *out);

/* a2: C-PML profiles (computed in double; *_f32 casts as the reference does). */
void or_cpml_f64(int32_t n, float h, float dt, int32_t width, float max_vel,
                 float freq, double *a, double *b, double *db);
void or_cpml_f32(int32_t n, float h, float dt, int32_t width, float max_vel,
                 float freq, float *a, float *b, float *db);

/* a4: bilinear corners of a station at (x, z) metres. */
void or_corners(float x, float z, float h, int32_t pad, int32_t nx, int32_t nz,
                int32_t cells_out[4], float w_out[4]);

/* a6: edge-replicating pad with clamp, and the strict clamp mask. */
void or_pad_velocity(const or_problem *p, const float *vel_vis, float *padded,
                     float *mask);
/* a10: transpose of the pad (f64 sums, f32 result). */
void or_fold(const or_problem *p, const float *padded, float *vis);

/* a3/a7: forward of one shot.  wavelet has nt samples.  traces [nt, nrec].
 * snaps, when non-NULL, receives [nt, cells] snapshots 2 v dt^2 w
 * (inner cells; the FD_PAD ring is left untouched). */
void or_forward_f32(const or_problem *p, const float *vel_vis,
                    const float *wavelet, float *traces, float *snaps);
void or_forward_f64(const or_problem *p, const float *vel_vis,
                    const double *wavelet, double *traces, double *snaps);

/* a9: adjoint of one shot given snapshots [nt, cells] and the trace residual
 * resid [nt, nrec] (already scaled).  grad [cells] is OVERWRITTEN with the
 * padded-grid gradient of this shot (before clamp mask and fold).  dwav,
 * when non-NULL, receives [nt] = dJ/d(wavelet[t]). */
void or_adjoint_f32(const or_problem *p, const float *vel_vis,
                    const float *wavelet, const float *snaps,
                    const float *resid, float *grad, float *dwav);
void or_adjoint_f64(const or_problem *p, const float *vel_vis,
                    const double *wavelet, const double *snaps,
                    const double *resid, double *grad, double *dwav);

/* Tangent-linear (Born) forward: dtraces [nt, nrec] = J dv where
 * dv_vis is a perturbation of the visible model (m/s). */
void or_born_f64(const or_problem *p, const float *vel_vis,
                 const double *dvel_vis, const double *wavelet,
                 double *dtraces);

void or_born_f32(const or_problem *p, const float *vel_vis,
                 const float *dvel_vis, const float *wavelet, float *dtraces);

/* a8: residual in place semantics.  count elements; returns the loss and
 * writes resid = (pred - obs) * scale.  scale = f32(1 / sum obs^2 [f64]). */
float or_misfit_f32(int64_t count, const float *pred, const float *obs,
                    float scale, float *resid);
/* energy sum in f64 (sequential) and the f32 scale, as wave_wgpu.rs:852-854 */
double or_energy(int64_t count, const float *obs);

/* Full reference evaluate_adjoint for `shots` sources (src_xz [shots,2]).
 * obs [shots, nt, nrec].  grad_vis [nvz*nvx].  Returns 0, or -1 on zero
 * observed energy. */
int32_t or_evaluate_adjoint_f32(const or_problem *p, int32_t shots,
                                const float *src_xz, const float *vel_vis,
                                const float *wavelet, const float *obs,
                                float *loss, float *grad_vis);

/* Same contract in double arithmetic (loss and gradient in double). */
int32_t or_evaluate_adjoint_f64(const or_problem *p, int32_t shots,
                                const float *src_xz, const float *vel_vis,
                                const double *wavelet, const double *obs,
                                double *loss, double *grad_vis);

/* The loss of or_evaluate_adjoint_f64 alone (forward of every shot, no
 * snapshots, no adjoint), with the same additions in the same order, so the
 * two return the same double.  Used for central finite differences of the
 * float64 misfit.  Returns 0, or -1 on zero observed energy. */
int32_t or_loss_f64(const or_problem *p, int32_t shots, const float *src_xz,
                    const float *vel_vis, const double *wavelet,
                    const double *obs, double *loss);

/* CPU baseline: `units` independent (model, shot) forward+adjoint solves on
 * `threads` POSIX threads.  vel_vis [units, nvz*nvx], obs [units, nt, nrec],
 * src_xz [units, 2].  grads [units, cells] (padded), losses [units].
 * Each unit's residual is scaled by its own 1/sum(obs^2). */
void or_batch_fwd_adj_f32(const or_problem *p, int32_t units,
                          const float *src_xz, const float *vel_vis,
                          const float *wavelet, const float *obs,
                          float *losses, float *grads, int32_t threads);

/* CPU forward-only batch (simulate), traces [units, nt, nrec]. */
void or_batch_forward_f32(const or_problem *p, int32_t units,
                          const float *src_xz, const float *vel_vis,
                          const float *wavelet, float *traces,
                          int32_t threads);

#ifdef __cplusplus
}
#endif
#endif
