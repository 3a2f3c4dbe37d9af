/*
 * refprec.c - one train step restated on the host in the reference's own
 * arithmetic types: the CPU baseline of the benchmark.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by bench.py's cpu_baseline leg and by
 * tests (tests/test_refprec.py pins it against oracle.c).  The product path
 * (libpxl.so) never uses it.
 *
 * oracle.c computes the step in FP64 throughout (it is the parity judge);
 * this file follows the types of the reference's kernels expression by
 * expression, so that its speed is that of the reference's maths on a CPU:
 *   - FP32 parameter planes; interpolation weights in FP64 (trilerpRGB,
 *     RenderUtils.cu:826-1004), accumulated into float sums;
 *   - float activations: softplus log(1 + exp(x - 13)) and its derivative in
 *     float, the colour sigmoid through a double exp (RenderUtils.cu:
 *     479-515);
 *   - FP64 position t, transmittance, weights and opacity, float ray
 *     planes (volumeRender_RGB_train, RenderUtils.cu:1582-1718);
 *   - FP64 residuals and loss, float per-sample gradients and float atomic
 *     adds into the gradient planes (renderRays_withGradients_RGBGPU and
 *     trilerpGradientsRGB, RenderUtils.cu:2051-2256, 1074-1364);
 *   - addVarianceLossGPU (RenderUtils.cu:2469-2547) with float sums;
 *   - Adam with updateRGBSceneGPU_ADAM's float/double mix (Optimizer.cu:
 *     359-415).
 * Parallel with OpenMP: rays (dynamic schedule), the TV's (x, y) columns in
 * four parity passes (columns of one pass share no vertex, where the
 * reference's threads race), Adam over the parameters.  Compiled -O3 for
 * x86-64-v3 without fast-math (the Makefile).  The reference's per-step
 * jitter offsets (cuRAND) are zero here, as in the GPU path.
 */
#include "refprec.h"

#include <math.h>
#include <omp.h>

/* Step cap of a march, the oracle's and the GPU path's. */
#define RP_MAX_STEPS (1 << 20)

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* getSceneIndex (RenderUtils.cu:3-13). */
static uint32_t scene_index(int x, int y, int z, int W, int H, int D) {
    x = clampi(x, 0, W - 1);
    y = clampi(y, 0, H - 1);
    z = clampi(z, 0, D - 1);
    return (uint32_t)z * (uint32_t)W * (uint32_t)H + (uint32_t)y * (uint32_t)W + (uint32_t)x;
}

/* The cell of a coordinate (the GPU path's convention for p <= 0). */
static int base_of(double p) {
    if (!(p > 0.0)) return 0;
    if (p >= 1073741824.0) return 1073741824;
    return (int)p;
}

/* RenderUtils.cu:479-515. */
static float sigmoid_ref(float x) { return (float)(1.f / (1.f + exp(-0.1 * x))); }
static float sigmoid_derivative_ref(float x) {
    const float s = sigmoid_ref(x);
    return (float)(0.1 * s * (1.f - s));
}
static float density_activation_ref(float x) { return logf(1.f + expf(x - 13.f)); }
static float density_derivation_ref(float x) { return 1.f / (1.f + expf(13.f - x)); }

/* The unnormalised SH basis in float (RenderUtils.cu:1639-1670). */
static void shs_ref(const float d[3], float *shs) {
    shs[0] = 1.f;
    shs[1] = d[1];
    shs[2] = d[2];
    shs[3] = d[0];
    shs[4] = d[0] * d[1];
    shs[5] = d[1] * d[2];
    shs[6] = 3.f * d[2] * d[2] - 1.f;
    shs[7] = d[0] * d[2];
    shs[8] = d[0] * d[0] - d[1] * d[1];
    shs[9] = d[1] * (3.f * d[0] * d[0] - d[1] * d[1]);
    shs[10] = d[0] * d[1] * d[2];
    shs[11] = d[1] * (5.f * d[2] * d[2] - 1.f);
    shs[12] = 5.f * d[2] * d[2] * d[2] - 3.f * d[2];
    shs[13] = d[0] * (5.f * d[2] * d[2] - 1.f);
    shs[14] = (d[0] * d[0] - d[1] * d[1]) * d[2];
    shs[15] = d[0] * (d[0] * d[0] - 3.f * d[1] * d[1]);
}

/* A sample's cell, its 8 corner records and FP64 weights; 0 when the cell
 * is empty (caged grids). */
typedef struct corners {
    uint32_t v[8];
    double w[8];
} corners;

static int locate(const orc_grid *g, const double p[3], corners *c) {
    const int bx = base_of(p[0]), by = base_of(p[1]), bz = base_of(p[2]);
    if (g->sparse && !g->occupancy[scene_index(bx, by, bz, g->W - 1, g->H - 1, g->D - 1)]) return 0;
    const double x = p[0] - bx, y = p[1] - by, z = p[2] - bz;
    for (int j = 0; j < 8; j++) {
        const int dx = j & 1, dy = (j >> 1) & 1, dz = j >> 2;
        uint32_t v = scene_index(bx + dx, by + dy, bz + dz, g->W, g->H, g->D);
        if (g->sparse) v = g->index_offsets[v];
        c->v[j] = v;
        c->w[j] = (dx ? x : 1. - x) * (dy ? y : 1. - y) * (dz ? z : 1. - z);
    }
    return 1;
}

/* trilerpRGB: float raw density and colours from double weights. */
static void trilerp_ref(const orc_grid *g, const corners *c, const float *shs, float tri[4]) {
    const int S = g->S;
    const int64_t n = g->n_vox;
    float sh_tri[48];
    for (int i = 0; i < 3 * S; i++) sh_tri[i] = 0.f;
    tri[0] = 0.f;
    for (int j = 0; j < 8; j++) {
        const uint32_t v = c->v[j];
        tri[0] += c->w[j] * g->planes[v];
        for (int ch = 0; ch < 3; ch++)
            for (int k = 0; k < S; k++)
                sh_tri[ch * S + k] += c->w[j] * g->planes[(int64_t)(1 + ch * S + k) * n + v];
    }
    for (int ch = 0; ch < 3; ch++) {
        tri[ch + 1] = 0.f;
        for (int k = 0; k < S; k++) tri[ch + 1] += sh_tri[ch * S + k] * shs[k];
    }
}

/* Marching position of step k: the fixed rule (t_near + k step) or the
 * reference's adaptive one (t += |step - w| after an evaluated sample). */
static double next_t(const orc_grid *g, double t, double tn, int k, double w, int evaluated) {
    if (!g->adaptive) return fma((double)k, (double)g->step, tn);
    return evaluated ? t + fabs(g->step - w) : t + g->step;
}

/* volumeRender_RGB_train for one ray: the float ray planes. */
static void forward_ray(const orc_grid *g, const orc_rays *r, int ray, float *out) {
    float acc_r = 0.f, acc_g = 0.f, acc_b = 0.f, acc_d = 0.f;
    double acc = 0., lam = 0.;
    const double tn = r->near_t[ray], tf = r->far_t[ray];
    const float d[3] = {r->dx[ray], r->dy[ray], r->dz[ray]};
    float shs[16];
    shs_ref(d, shs);
    double t = tn;
    int k = 0;
    while (t < tf && k < RP_MAX_STEPS) {
        const double p[3] = {r->ox[ray] + t * d[0], r->oy[ray] + t * d[1], r->oz[ray] + t * d[2]};
        corners c;
        if (!locate(g, p, &c)) {
            ++k;
            t = next_t(g, t, tn, k, 0., 0);
            continue;
        }
        float tri[4];
        trilerp_ref(g, &c, shs, tri);
        const float sigma = density_activation_ref(tri[0]);
        const double Ti = exp(-lam);
        const double w = Ti * (1. - expf(-sigma * g->step));
        acc += w;
        acc_d += sigma;
        acc_r += w * sigmoid_ref(tri[1]);
        acc_g += w * sigmoid_ref(tri[2]);
        acc_b += w * sigmoid_ref(tri[3]);
        lam += sigma * g->step;
        ++k;
        t = next_t(g, t, tn, k, w, 1);
    }
    float accw = (float)acc;
    if (r->white_bg) {
        acc_r += 1. - acc;
        acc_g += 1. - acc;
        acc_b += 1. - acc;
    }
    out[0] = acc_r;
    out[1] = acc_g;
    out[2] = acc_b;
    out[3] = accw;
    out[4] = acc_d;
}

static void atomic_addf(float *p, float v) {
#pragma omp atomic
    *p += v;
}

/* renderRays_withGradients_RGBGPU for one ray given its forward planes;
 * returns the ray's loss. */
static double backward_ray(const orc_grid *g, const orc_rays *r, int ray, const float *fw,
                           double mask_coef, float huber, float *grad) {
    const int S = g->S;
    const int64_t n = g->n_vox;
    const float d[3] = {r->dx[ray], r->dy[ray], r->dz[ray]};
    float shs[16];
    shs_ref(d, shs);
    double res[3] = {fw[0] - r->gt_r[ray] / 255., fw[1] - r->gt_g[ray] / 255.,
                     fw[2] - r->gt_b[ray] / 255.};
    const double loss = (fabs(res[0]) + fabs(res[1]) + fabs(res[2])) * 255. / 3.;
    if (isnan(loss)) return 255.;
    for (int i = 0; i < 3; i++) {
        if (res[i] > huber) res[i] = huber - 0.5 * huber * huber;
        else if (res[i] < -huber) res[i] = 0.5 * huber * huber - huber;
    }
    float mask = 0.f;
    if (r->gt_mask[ray] && fw[3] < 1.) mask = (r->gt_mask[ray] / 255.0 - fw[3]) * mask_coef;
    float wpre = fw[3] > 1e-6 ? fw[3] : (float)1e-6;
    wpre = wpre < 1 - 1e-6 ? wpre : (float)(1 - 1e-6);
    mask += (logf(wpre) - log(1. - wpre)) * 1e-3;
    const double tn = r->near_t[ray], tf = r->far_t[ray];
    double t = tn, lam = 0., accw = 0., acc_rgb[3] = {0., 0., 0.};
    int k = 0;
    while (t < tf && k < RP_MAX_STEPS) {
        const double p[3] = {r->ox[ray] + t * d[0], r->oy[ray] + t * d[1], r->oz[ray] + t * d[2]};
        corners c;
        if (!locate(g, p, &c)) {
            ++k;
            t = next_t(g, t, tn, k, 0., 0);
            continue;
        }
        float tri[4], act[4], grd[4];
        trilerp_ref(g, &c, shs, tri);
        act[0] = density_activation_ref(tri[0]);
        for (int ch = 1; ch < 4; ch++) act[ch] = sigmoid_ref(tri[ch]);
        double Ti = exp(-lam);
        const double w = Ti * (1. - expf(-act[0] * g->step));
        accw += w;
        for (int ch = 0; ch < 3; ch++) {
            acc_rgb[ch] += w * act[ch + 1];
            grd[ch + 1] = 0.6667f * w * res[ch] * sigmoid_derivative_ref(tri[ch + 1]);
        }
        lam += act[0] * g->step;
        Ti = exp(-lam);
        grd[0] = res[0] * (Ti * act[1] - (fw[0] - acc_rgb[0]));
        grd[0] += res[1] * (Ti * act[2] - (fw[1] - acc_rgb[1]));
        grd[0] += res[2] * (Ti * act[3] - (fw[2] - acc_rgb[2]));
        grd[0] *= 0.6667 * g->step * density_derivation_ref(tri[0]);
        grd[0] += g->step * density_derivation_ref(tri[0]) * (fw[3] - accw - Ti) * mask;
        /* trilerpGradientsRGB: float atomics per corner and parameter. */
        for (int j = 0; j < 8; j++) {
            const uint32_t v = c.v[j];
            atomic_addf(&grad[v], (float)(grd[0] * c.w[j]));
            for (int ch = 0; ch < 3; ch++)
                for (int kk = 0; kk < S; kk++)
                    atomic_addf(&grad[(int64_t)(1 + ch * S + kk) * n + v],
                                (float)(grd[1 + ch] * c.w[j] * shs[kk]));
        }
        ++k;
        t = next_t(g, t, tn, k, w, 1);
    }
    return loss;
}

void rp_render_train(const orc_grid *g, const orc_rays *r, const int32_t *sample_index,
                     int64_t len, double mask_coef, float huber_delta, float *fwd, double *losses,
                     float *grad, int32_t threads) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
    for (int64_t b = 0; b < len; b++) {
        const int ray = sample_index[b];
        float *fw = fwd + 5 * b;
        if (r->near_t[ray] > r->far_t[ray]) {
            for (int i = 0; i < 5; i++) fw[i] = 0.f;
            losses[b] = 0.;
            continue;
        }
        forward_ray(g, r, ray, fw);
        losses[b] = backward_ray(g, r, ray, fw, mask_coef, huber_delta, grad);
    }
}

/* A TV addition: record 0, where every corner outside the cage of a caged
 * grid lands (SceneUtils.cu:1106-1113), is shared by columns of one pass and
 * takes an atomic add; every other record belongs to one column. */
static void tv_add(float *p, double v, int shared) {
    if (shared) {
#pragma omp atomic
        *p += v;
    } else {
        *p += v;
    }
}

/* addVarianceLossGPU: per (x, y) column a walk over z; the cells of a column
 * pass through one thread, columns of one parity class (x mod 2, y mod 2)
 * run in parallel. */
void rp_tv_disparity(const orc_grid *g, double lambda_density, double lambda_sh,
                     int32_t optimize_step, float *grad, int32_t threads) {
    const int W = g->W, H = g->H, D = g->D, S = g->S;
    const int64_t n = g->n_vox;
    for (int pass = 0; pass < 4; pass++) {
        const int px = pass & 1, py = pass >> 1;
        const int nx = (W - 1 - px + 1) / 2, ny = (H - 1 - py + 1) / 2;
#pragma omp parallel for schedule(dynamic, 8) num_threads(threads)
        for (int64_t col = 0; col < (int64_t)nx * ny; col++) {
            const int x = px + 2 * (int)(col % nx), y = py + 2 * (int)(col / nx);
            float sum_ch[3][16];
            uint32_t v8[8];
            for (int z = 0; z < D - 1; z++) {
                if (g->sparse &&
                    !g->occupancy[scene_index(x, y, z, W - 1, H - 1, D - 1)])
                    continue;
                for (int j = 0; j < 8; j++) {
                    uint32_t v = scene_index(x + (j & 1), y + ((j >> 1) & 1), z + (j >> 2), W, H, D);
                    if (g->sparse) v = g->index_offsets[v];
                    v8[j] = v;
                }
                if (optimize_step <= 1) {
                    for (int ch = 0; ch < 3; ch++)
                        for (int k = 0; k < S; k++) sum_ch[ch][k] = 0.f;
                    for (int j = 0; j < 8; j++)
                        for (int ch = 0; ch < 3; ch++)
                            for (int k = 0; k < S; k++)
                                sum_ch[ch][k] += g->planes[(int64_t)(1 + ch * S + k) * n + v8[j]] / 8.f;
                    for (int j = 0; j < 8; j++)
                        for (int ch = 0; ch < 3; ch++)
                            for (int k = 0; k < S; k++) {
                                float *gp = &grad[(int64_t)(1 + ch * S + k) * n + v8[j]];
                                const double add =
                                    lambda_sh * (g->planes[(int64_t)(1 + ch * S + k) * n + v8[j]] -
                                                 sum_ch[ch][k]);
                                tv_add(gp, add, v8[j] == 0);
                            }
                }
                float sum_d = 0.f;
                for (int j = 0; j < 8; j++) sum_d += g->planes[v8[j]];
                sum_d /= 8.f;
                for (int j = 0; j < 8; j++)
                    tv_add(&grad[v8[j]], lambda_density * (g->planes[v8[j]] - sum_d), v8[j] == 0);
            }
        }
    }
}

/* updateRGBSceneGPU_ADAM over P planes of n values. */
void rp_adam(float *params, float *grad, float *m, float *v, int32_t P, int64_t n, double lr,
             double sh_lr, int32_t t, int32_t threads) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1.e-6f;
    const double bc1 = 1.0 - (double)powf(b1, (float)t);
    const double bc2 = 1.0 - (double)powf(b2, (float)t);
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int64_t i = 0; i < (int64_t)P * n; i++) {
        const int p = (int)(i / n);
        const float gg = grad[i];
        const float mm = (float)((double)(b1 * m[i]) + (1.0 - (double)b1) * (double)gg);
        const float vv = (float)((double)(b2 * v[i]) + (1.0 - (double)b2) * (double)gg * (double)gg);
        const float cm = (float)((double)mm / bc1);
        const float cv = (float)((double)vv / bc2);
        const double nt = (double)params[i] - (p == 0 ? lr : sh_lr) * (double)cm /
                                                  (double)(sqrtf(cv) + eps);
        params[i] = p == 0 ? (nt < 0.0 ? 0.0f : (float)nt) : (float)nt;
        m[i] = mm;
        v[i] = vv;
        grad[i] = 0.0f;
    }
}

double rp_now(void) { return omp_get_wtime(); }
