/*
 * CPU oracle for the WGNet wave engine: TEST INFRASTRUCTURE ONLY (see
 * wave_oracle.h for what each function restates and who may call it).
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off, no -march=native).
 */
#include "wave_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define FD_PAD 2

void or_dims_of(const or_problem *p, or_dims *d) {
    d->pad = FD_PAD + p->pml;
    d->nx = p->nvx + 2 * d->pad;
    d->nz = p->nvz + 2 * d->pad;
    d->cells = d->nx * d->nz;
    d->region = p->pml + 2 * FD_PAD;     /* wave_wgpu.rs:291 */
    d->region_adj = p->pml + 3 * FD_PAD; /* wave_wgpu.rs:295 */
}

/* lib.rs:12-21: all float32, expf. */
void or_ricker_f32(int32_t samples, float dt, float peak_hz, float *out) {
    const float delay = 1.5f / peak_hz;
    for (int32_t i = 0; i < samples; ++i) {
        const float t = (float)i * dt - delay;
        const float u = 3.14159265358979323846f * peak_hz * t;
        const float a = u * u;
        out[i] = (1.0f - 2.0f * a) * expf(-a);
    }
}

void or_ricker_f64(int32_t samples, double dt, double peak_hz, double *out) {
    const double delay = 1.5 / peak_hz;
    for (int32_t i = 0; i < samples; ++i) {
        const double t = (double)i * dt - delay;
        const double u = 3.14159265358979323846 * peak_hz * t;
        const double a = u * u;
        out[i] = (1.0 - 2.0 * a) * exp(-a);
    }
}

/* lib.rs:25-68, in double. */
void or_cpml_f64(int32_t n, float h, float dt, int32_t width, float max_vel,
                 float freq, double *a, double *b, double *db) {
    const double start0 = (double)(FD_PAD + width);
    const double start1 = (double)(n - 1 - FD_PAD - width);
    const double width_f = (double)width;
    const double h_f = (double)h;
    const double dt_f = (double)dt;
    const double sigma0 = -(3.0 * (double)max_vel * log(0.001)) / (2.0 * width_f * h_f);
    for (int32_t x = 0; x < n; ++x) {
        const double xf = (double)x;
        const double f0 = (start0 - xf) / width_f;
        const double f1 = (xf - start1) / width_f;
        double frac = f0 > f1 ? f0 : f1;
        if (frac < 0.0) frac = 0.0;
        if (frac > 1.0) frac = 1.0;
        const double sigma = sigma0 * frac * frac;
        const double alpha = 3.14159265358979323846 * (double)freq * (1.0 - frac);
        const double axv = exp(-(sigma + alpha) * dt_f);
        b[x] = (sigma / (sigma + alpha)) * (axv - 1.0);
        a[x] = frac == 0.0 ? 0.0 : axv;
    }
    for (int32_t x = 0; x < n; ++x) db[x] = 0.0;
    for (int32_t x = 2; x < n - 2; ++x)
        db[x] = ((8.0 / 12.0) * (b[x + 1] - b[x - 1]) - (1.0 / 12.0) * (b[x + 2] - b[x - 2])) / h_f;
}

void or_cpml_f32(int32_t n, float h, float dt, int32_t width, float max_vel,
                 float freq, float *a, float *b, float *db) {
    double *tmp = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    or_cpml_f64(n, h, dt, width, max_vel, freq, tmp, tmp + n, tmp + 2 * n);
    for (int32_t i = 0; i < n; ++i) {
        a[i] = (float)tmp[i];
        b[i] = (float)tmp[n + i];
        db[i] = (float)tmp[2 * n + i];
    }
    free(tmp);
}

/* Rust's saturating `f32 as usize`: NaN and negatives -> 0. */
static int64_t sat_floor(float g) {
    const float f = floorf(g);
    if (!(f > 0.0f)) return 0;
    if (f > 9.0e18f) return INT64_MAX;
    return (int64_t)f;
}

/* wave_wgpu.rs:217-230 (and the identical set_sources, :761-785). */
void or_corners(float x, float z, float h, int32_t pad, int32_t nx, int32_t nz,
                int32_t cells_out[4], float w_out[4]) {
    const float gx = x / h + (float)pad;
    const float gz = z / h + (float)pad;
    int64_t x0 = sat_floor(gx), z0 = sat_floor(gz);
    if (x0 > nx - 2) x0 = nx - 2;
    if (z0 > nz - 2) z0 = nz - 2;
    const float fx = gx - (float)x0;
    const float fz = gz - (float)z0;
    cells_out[0] = (int32_t)(z0 * nx + x0);
    cells_out[1] = (int32_t)(z0 * nx + x0 + 1);
    cells_out[2] = (int32_t)((z0 + 1) * nx + x0);
    cells_out[3] = (int32_t)((z0 + 1) * nx + x0 + 1);
    w_out[0] = (1.0f - fx) * (1.0f - fz);
    w_out[1] = fx * (1.0f - fz);
    w_out[2] = (1.0f - fx) * fz;
    w_out[3] = fx * fz;
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* wave_wgpu.rs:714-734.  Rust f32::clamp keeps NaN. */
void or_pad_velocity(const or_problem *p, const float *vel_vis, float *padded,
                     float *mask) {
    or_dims d;
    or_dims_of(p, &d);
    for (int z = 0; z < d.nz; ++z) {
        const int vz = clampi(z - d.pad, 0, p->nvz - 1);
        for (int x = 0; x < d.nx; ++x) {
            const int vx = clampi(x - d.pad, 0, p->nvx - 1);
            const float v = vel_vis[vz * p->nvx + vx];
            const int idx = z * d.nx + x;
            if (mask) mask[idx] = (v > p->vmin && v < p->vmax) ? 1.0f : 0.0f;
            padded[idx] = v < p->vmin ? p->vmin : (v > p->vmax ? p->vmax : v);
        }
    }
}

/* wave_wgpu.rs:993-1005. */
void or_fold(const or_problem *p, const float *padded, float *vis) {
    or_dims d;
    or_dims_of(p, &d);
    double *sum = (double *)calloc((size_t)p->nvx * p->nvz, sizeof(double));
    for (int z = 0; z < d.nz; ++z) {
        const int vz = clampi(z - d.pad, 0, p->nvz - 1);
        for (int x = 0; x < d.nx; ++x) {
            const int vx = clampi(x - d.pad, 0, p->nvx - 1);
            sum[vz * p->nvx + vx] += (double)padded[z * d.nx + x];
        }
    }
    for (int i = 0; i < p->nvx * p->nvz; ++i) vis[i] = (float)sum[i];
    free(sum);
}

static void fold_f64(const or_problem *p, const double *padded, double *vis) {
    or_dims d;
    or_dims_of(p, &d);
    memset(vis, 0, sizeof(double) * (size_t)p->nvx * p->nvz);
    for (int z = 0; z < d.nz; ++z) {
        const int vz = clampi(z - d.pad, 0, p->nvz - 1);
        for (int x = 0; x < d.nx; ++x) {
            const int vx = clampi(x - d.pad, 0, p->nvx - 1);
            vis[vz * p->nvx + vx] += padded[z * d.nx + x];
        }
    }
}

#define T float
#define FN(x) x##_f32
#define LIT(x) x##f
#include "wave_oracle_impl.inc"
#undef T
#undef FN
#undef LIT

#define T double
#define FN(x) x##_f64
#define LIT(x) x
#include "wave_oracle_impl.inc"
#undef T
#undef FN
#undef LIT

/* Tangent-linear forward in double: base forward for the snapshots
 * S^t = 2 v dt^2 w, then the same recurrence on the tangent state with the
 * distributed source dv_pad * S^t.  dv_pad = pad(dv) x clamp mask (the
 * derivative of the clamp). */
void or_born_f64(const or_problem *p, const float *vel_vis,
                 const double *dvel_vis, const double *wavelet,
                 double *dtraces) {
    or_dims d;
    or_dims_of(p, &d);
    double *snaps = (double *)calloc((size_t)p->nt * d.cells, sizeof(double));
    double *traces = (double *)malloc(sizeof(double) * (size_t)p->nt * (p->nrec > 0 ? p->nrec : 1));
    or_forward_f64(p, vel_vis, wavelet, traces, snaps);
    float *velp = (float *)malloc(sizeof(float) * 2 * (size_t)d.cells);
    or_pad_velocity(p, vel_vis, velp, velp + d.cells);
    double *dvp = (double *)malloc(sizeof(double) * (size_t)d.cells);
    for (int z = 0; z < d.nz; ++z) {
        const int vz = clampi(z - d.pad, 0, p->nvz - 1);
        for (int x = 0; x < d.nx; ++x) {
            const int vx = clampi(x - d.pad, 0, p->nvx - 1);
            const int i = z * d.nx + x;
            dvp[i] = dvel_vis[vz * p->nvx + vx] * (double)velp[d.cells + i];
        }
    }
    forward_core_f64(p, vel_vis, wavelet, dtraces, NULL, dvp, snaps);
    free(dvp);
    free(velp);
    free(traces);
    free(snaps);
}

/* The same tangent-linear model in float32 arithmetic (the GPU Born kernel's
 * operation order), for the roundoff level of the FP32 dot-product test. */
void or_born_f32(const or_problem *p, const float *vel_vis, const float *dvel_vis,
                 const float *wavelet, float *dtraces) {
    or_dims d;
    or_dims_of(p, &d);
    float *snaps = (float *)calloc((size_t)p->nt * d.cells, sizeof(float));
    float *traces = (float *)malloc(sizeof(float) * (size_t)p->nt * (p->nrec > 0 ? p->nrec : 1));
    or_forward_f32(p, vel_vis, wavelet, traces, snaps);
    float *velp = (float *)malloc(sizeof(float) * 2 * (size_t)d.cells);
    or_pad_velocity(p, vel_vis, velp, velp + d.cells);
    float *dvp = (float *)malloc(sizeof(float) * (size_t)d.cells);
    for (int z = 0; z < d.nz; ++z) {
        const int vz = clampi(z - d.pad, 0, p->nvz - 1);
        for (int x = 0; x < d.nx; ++x) {
            const int vx = clampi(x - d.pad, 0, p->nvx - 1);
            const int i = z * d.nx + x;
            dvp[i] = dvel_vis[vz * p->nvx + vx] * velp[d.cells + i];
        }
    }
    forward_core_f32(p, vel_vis, wavelet, dtraces, NULL, dvp, snaps);
    free(dvp);
    free(velp);
    free(traces);
    free(snaps);
}

double or_energy(int64_t count, const float *obs) {
    double e = 0.0;
    for (int64_t i = 0; i < count; ++i) e += (double)obs[i] * (double)obs[i];
    return e;
}

/* RESIDUAL_WGSL (wave_wgpu.rs:28-53): per 256-element chunk, a float32 tree
 * sum of 0.5 d^2 scale; the host sums the chunk partials in double
 * (wave_wgpu.rs:960-965). */
float or_misfit_f32(int64_t count, const float *pred, const float *obs,
                    float scale, float *resid) {
    float wg[256];
    double loss = 0.0;
    for (int64_t base = 0; base < count; base += 256) {
        for (int li = 0; li < 256; ++li) {
            const int64_t i = base + li;
            float e = 0.0f;
            if (i < count) {
                const float dd = pred[i] - obs[i];
                if (resid) resid[i] = dd * scale;
                e = 0.5f * dd * dd * scale;
            }
            wg[li] = e;
        }
        for (int s = 128; s > 0; s /= 2)
            for (int li = 0; li < s; ++li) wg[li] += wg[li + s];
        loss += (double)wg[0];
    }
    return (float)loss;
}

int32_t or_evaluate_adjoint_f32(const or_problem *p0, int32_t shots,
                                const float *src_xz, const float *vel_vis,
                                const float *wavelet, const float *obs,
                                float *loss, float *grad_vis) {
    or_problem p = *p0;
    or_dims d;
    or_dims_of(&p, &d);
    const int64_t per = (int64_t)p.nt * p.nrec;
    const int64_t count = per * shots;
    const double energy = or_energy(count, obs);
    if (!(energy > 0.0)) return -1;
    const float scale = (float)(1.0 / energy);
    float *pred = (float *)malloc(sizeof(float) * (size_t)count);
    float *res = (float *)malloc(sizeof(float) * (size_t)count);
    float *snaps = (float *)calloc((size_t)shots * p.nt * d.cells, sizeof(float));
    for (int s = 0; s < shots; ++s) {
        p.src_x = src_xz[2 * s];
        p.src_z = src_xz[2 * s + 1];
        or_forward_f32(&p, vel_vis, wavelet, pred + s * per, snaps + (size_t)s * p.nt * d.cells);
    }
    *loss = or_misfit_f32(count, pred, obs, scale, res);
    float *g = (float *)malloc(sizeof(float) * (size_t)shots * d.cells);
    for (int s = 0; s < shots; ++s) {
        p.src_x = src_xz[2 * s];
        p.src_z = src_xz[2 * s + 1];
        or_adjoint_f32(&p, vel_vis, wavelet, snaps + (size_t)s * p.nt * d.cells,
                       res + s * per, g + (size_t)s * d.cells, NULL);
    }
    float *velp = (float *)malloc(sizeof(float) * 2 * (size_t)d.cells);
    or_pad_velocity(&p, vel_vis, velp, velp + d.cells);
    float *padded = (float *)malloc(sizeof(float) * (size_t)d.cells);
    for (int c = 0; c < d.cells; ++c) {
        float sum_s = 0.0f;
        for (int s = 0; s < shots; ++s) sum_s += g[(size_t)s * d.cells + c];
        padded[c] = sum_s * velp[d.cells + c];
    }
    or_fold(&p, padded, grad_vis);
    free(padded);
    free(velp);
    free(g);
    free(snaps);
    free(res);
    free(pred);
    return 0;
}

int32_t or_evaluate_adjoint_f64(const or_problem *p0, int32_t shots,
                                const float *src_xz, const float *vel_vis,
                                const double *wavelet, const double *obs,
                                double *loss, double *grad_vis) {
    or_problem p = *p0;
    or_dims d;
    or_dims_of(&p, &d);
    const int64_t per = (int64_t)p.nt * p.nrec;
    const int64_t count = per * shots;
    double energy = 0.0;
    for (int64_t i = 0; i < count; ++i) energy += obs[i] * obs[i];
    if (!(energy > 0.0)) return -1;
    const double scale = 1.0 / energy;
    double *pred = (double *)malloc(sizeof(double) * (size_t)count);
    double *res = (double *)malloc(sizeof(double) * (size_t)count);
    double *snaps = (double *)calloc((size_t)p.nt * d.cells, sizeof(double));
    double *g = (double *)malloc(sizeof(double) * (size_t)d.cells);
    double *gsum = (double *)calloc((size_t)d.cells, sizeof(double));
    double l = 0.0;
    for (int s = 0; s < shots; ++s) {
        p.src_x = src_xz[2 * s];
        p.src_z = src_xz[2 * s + 1];
        or_forward_f64(&p, vel_vis, wavelet, pred + s * per, snaps);
        for (int64_t i = 0; i < per; ++i) {
            const double dd = pred[s * per + i] - obs[s * per + i];
            res[s * per + i] = dd * scale;
            l += 0.5 * dd * dd * scale;
        }
        or_adjoint_f64(&p, vel_vis, wavelet, snaps, res + s * per, g, NULL);
        for (int c = 0; c < d.cells; ++c) gsum[c] += g[c];
    }
    *loss = l;
    float *velp = (float *)malloc(sizeof(float) * 2 * (size_t)d.cells);
    or_pad_velocity(&p, vel_vis, velp, velp + d.cells);
    for (int c = 0; c < d.cells; ++c) gsum[c] *= (double)velp[d.cells + c];
    fold_f64(&p, gsum, grad_vis);
    free(velp);
    free(gsum);
    free(g);
    free(snaps);
    free(res);
    free(pred);
    return 0;
}

int32_t or_loss_f64(const or_problem *p0, int32_t shots, const float *src_xz,
                    const float *vel_vis, const double *wavelet,
                    const double *obs, double *loss) {
    or_problem p = *p0;
    const int64_t per = (int64_t)p.nt * p.nrec;
    const int64_t count = per * shots;
    double energy = 0.0;
    for (int64_t i = 0; i < count; ++i) energy += obs[i] * obs[i];
    if (!(energy > 0.0)) return -1;
    const double scale = 1.0 / energy;
    double *pred = (double *)malloc(sizeof(double) * (size_t)per);
    double l = 0.0;
    for (int s = 0; s < shots; ++s) {
        p.src_x = src_xz[2 * s];
        p.src_z = src_xz[2 * s + 1];
        or_forward_f64(&p, vel_vis, wavelet, pred, NULL);
        for (int64_t i = 0; i < per; ++i) {
            const double dd = pred[i] - obs[s * per + i];
            l += 0.5 * dd * dd * scale;
        }
    }
    *loss = l;
    free(pred);
    return 0;
}

/* ---- multi-threaded CPU baseline over independent (model, shot) units ---- */

typedef struct {
    const or_problem *p;
    int32_t units, next_unit;
    const float *src_xz, *vel_vis, *wavelet, *obs;
    float *losses, *grads, *traces;
    int forward_only;
    pthread_mutex_t lock;
} batch_job;

static void *batch_worker(void *arg) {
    batch_job *job = (batch_job *)arg;
    or_problem p = *job->p;
    or_dims d;
    or_dims_of(&p, &d);
    const int64_t per = (int64_t)p.nt * p.nrec;
    const int nvis = p.nvx * p.nvz;
    float *snaps = job->forward_only ? NULL : (float *)calloc((size_t)p.nt * d.cells, sizeof(float));
    float *pred = (float *)malloc(sizeof(float) * (size_t)per);
    float *res = (float *)malloc(sizeof(float) * (size_t)per);
    for (;;) {
        pthread_mutex_lock(&job->lock);
        const int u = job->next_unit++;
        pthread_mutex_unlock(&job->lock);
        if (u >= job->units) break;
        p.src_x = job->src_xz[2 * u];
        p.src_z = job->src_xz[2 * u + 1];
        if (job->forward_only) {
            or_forward_f32(&p, job->vel_vis + (size_t)u * nvis, job->wavelet,
                           job->traces + (size_t)u * per, NULL);
            continue;
        }
        or_forward_f32(&p, job->vel_vis + (size_t)u * nvis, job->wavelet, pred, snaps);
        const float *obs = job->obs + (size_t)u * per;
        const double energy = or_energy(per, obs);
        const float scale = (float)(1.0 / (energy > 0.0 ? energy : 1.0));
        job->losses[u] = or_misfit_f32(per, pred, obs, scale, res);
        or_adjoint_f32(&p, job->vel_vis + (size_t)u * nvis, job->wavelet, snaps, res,
                       job->grads + (size_t)u * d.cells, NULL);
    }
    free(res);
    free(pred);
    free(snaps);
    return NULL;
}

static void run_batch(batch_job *job, int32_t threads) {
    if (threads < 1) threads = 1;
    pthread_mutex_init(&job->lock, NULL);
    pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int i = 0; i < threads; ++i) pthread_create(&tid[i], NULL, batch_worker, job);
    for (int i = 0; i < threads; ++i) pthread_join(tid[i], NULL);
    free(tid);
    pthread_mutex_destroy(&job->lock);
}

void or_batch_fwd_adj_f32(const or_problem *p, int32_t units,
                          const float *src_xz, const float *vel_vis,
                          const float *wavelet, const float *obs,
                          float *losses, float *grads, int32_t threads) {
    batch_job job;
    memset(&job, 0, sizeof(job));
    job.p = p; job.units = units; job.src_xz = src_xz; job.vel_vis = vel_vis;
    job.wavelet = wavelet; job.obs = obs; job.losses = losses; job.grads = grads;
    run_batch(&job, threads);
}

void or_batch_forward_f32(const or_problem *p, int32_t units,
                          const float *src_xz, const float *vel_vis,
                          const float *wavelet, float *traces,
                          int32_t threads) {
    batch_job job;
    memset(&job, 0, sizeof(job));
    job.p = p; job.units = units; job.src_xz = src_xz; job.vel_vis = vel_vis;
    job.wavelet = wavelet; job.traces = traces; job.forward_only = 1;
    run_batch(&job, threads);
}
