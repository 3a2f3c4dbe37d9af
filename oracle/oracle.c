/*
 * oracle.c - CPU restatement of the reference's train-step maths.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled with gcc
 * -ffp-contract=off so every operation happens as written: the ray
 * generator and the marching positions are bit-replicas of the GPU's
 * operation sequence; everything else is FP64.
 *
 * Every function cites the reference lines it restates
 * (SparseVolumeRadianceField/...).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* Threads used by the loops over rays, cells and parameters (default 1:
 * every sum in a fixed order, results reproducible bit for bit).  With more
 * threads the gradient additions become atomic and their order varies. */
static int g_threads = 1;

void orc_set_threads(int32_t n) { g_threads = n > 0 ? n : 1; }

static inline void add_d(double *p, double v) {
    if (g_threads > 1) {
#pragma omp atomic
        *p += v;
    } else {
        *p += v;
    }
}

#define ORC_PI 3.141592653589793

/* ------------------------------------------------------------------ */
/* Small dense linear algebra (Utils.cu:132-299).                       */
/* ------------------------------------------------------------------ */

static void mat_mul(const double *a, const double *b, double *out, int n, int k, int m) {
    /* out[n x m] = a[n x k] * b[k x m], accumulation in index order
     * (tinyMatMul, Utils.cu:154-180). */
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m; j++) {
            double s = 0.0;
            for (int t = 0; t < k; t++) s += a[i * k + t] * b[t * m + j];
            out[i * m + j] = s;
        }
}

/* Gauss-Jordan inverse with the reference's pivot rule: a pivot below 1e-8
 * in magnitude is replaced by swapping in the next rows (getMatInvert,
 * Utils.cu:233-299).  Returns 0 on a singular matrix. */
static int mat_inv(const double *src, double *dst, int n) {
    double aug[4][8];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < 2 * n; j++)
            aug[i][j] = j < n ? src[i * n + j] : (j - n == i ? 1.0 : 0.0);
    for (int r = 0; r < n; r++) {
        int swaps = 0;
        double piv = aug[r][r];
        while (fabs(piv) < 1e-8) {
            ++swaps;
            if (swaps > n - r - 1) return 0;
            for (int j = 0; j < 2 * n; j++) {
                const double t = aug[r][j];
                aug[r][j] = aug[r + swaps][j];
                aug[r + swaps][j] = t;
            }
            piv = aug[r][r];
        }
        for (int j = r; j < 2 * n; j++) aug[r][j] /= piv;
        for (int s = r + 1; s < n; s++) {
            const double f = aug[s][r];
            for (int j = 0; j < 2 * n; j++) aug[s][j] -= f * aug[r][j];
        }
    }
    for (int r = n - 2; r >= 0; r--)
        for (int s = n - 1; s > r; s--) {
            const double f = aug[r][s];
            for (int j = r + 1; j < 2 * n; j++) aug[r][j] -= f * aug[s][j];
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) dst[i * n + j] = aug[i][n + j];
    return 1;
}

/* eulerAngle2RotationMat (Utils.cu:182-231): degrees, R = inv(Rx Ry Rz). */
static void euler_rot(const double deg[3], double *R) {
    const double a = deg[0] * ORC_PI / 180.0, b = deg[1] * ORC_PI / 180.0,
                 c = deg[2] * ORC_PI / 180.0;
    const double rx[9] = {1, 0, 0, 0, cos(a), -sin(a), 0, sin(a), cos(a)};
    const double ry[9] = {cos(b), 0, sin(b), 0, 1, 0, -sin(b), 0, cos(b)};
    const double rz[9] = {cos(c), -sin(c), 0, sin(c), cos(c), 0, 0, 0, 1};
    double rxy[9], c2w[9];
    mat_mul(rx, ry, rxy, 3, 3, 3);
    mat_mul(rxy, rz, c2w, 3, 3, 3);
    mat_inv(c2w, R, 3);
}

static void normalize3(double *v) {
    double s = 0.0;
    for (int i = 0; i < 3; i++) s += v[i] * v[i];
    s = sqrt(s);
    for (int i = 0; i < 3; i++) v[i] /= s;
}

static void cross3(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

/* setCameraLookAtXOY (Utils.cu:404-446): camera at the target's height. */
static void look_at_xoy(const double *cam, const double *obj, double *R) {
    double z[3] = {-(obj[0] - cam[0]), -(obj[1] - cam[1]), -(obj[2] - cam[2])};
    if (fabs(z[0]) < 1e-8 && fabs(z[2]) < 1e-8) {
        const double ang[3] = {z[1] < 0.0 ? 90.0 : -90.0, 0.0, z[1] < 0.0 ? 0.0 : 180.0};
        euler_rot(ang, R);
        return;
    }
    normalize3(z);
    const double up[3] = {0.0, 1.0, 0.0};
    double x[3], y[3];
    cross3(up, z, x);
    normalize3(x);
    cross3(z, x, y);
    double w2c[9] = {x[0], x[1], x[2], y[0], y[1], y[2], z[0], z[1], z[2]};
    double c2w[9], tmp[9], flip[9];
    mat_inv(w2c, c2w, 3);
    memcpy(tmp, c2w, sizeof(tmp));
    const double a90[3] = {0.0, 0.0, 90.0};
    euler_rot(a90, flip);
    mat_mul(tmp, flip, c2w, 3, 3, 3);
    const double ey[3] = {0.0, 1.0, 0.0};
    double yw[3];
    mat_mul(c2w, ey, yw, 3, 3, 1);
    if (yw[2] < 0.0) {
        memcpy(tmp, c2w, sizeof(tmp));
        const double a180[3] = {0.0, 0.0, 180.0};
        euler_rot(a180, flip);
        mat_mul(tmp, flip, c2w, 3, 3, 3);
    }
    mat_inv(c2w, R, 3);
}

/* setCameraLookAt (Utils.cu:448-486): world-to-camera rotation whose rows
 * are the camera axes; z axis = camera - target. */
static void look_at(const double *cam, const double *obj, double *R) {
    const double A = cam[0] - obj[0], B = cam[1] - obj[1], C = cam[2] - obj[2];
    if (fabs(C) < 1e-8) {
        look_at_xoy(cam, obj, R);
        return;
    }
    const double a = 1.0;
    const double sc = -a * B / sqrt(B * B + C * C);
    const double cc = -a * A * C / sqrt(A * A + B * B + C * C) / sqrt(B * B + C * C);
    const double phi = atan2(cc, sc);
    const double t = ORC_PI / 2 - phi;
    double y[3], z[3] = {A, B, C}, x[3];
    y[0] = a * sqrt((B * B + C * C) / (A * A + B * B + C * C)) * cos(t);
    y[1] = a / sqrt(B * B + C * C) * (C * sin(t) - A * B / sqrt(A * A + B * B + C * C) * cos(t));
    y[2] = sc * sin(t) + cc * cos(t);
    normalize3(y);
    normalize3(z);
    cross3(y, z, x);
    for (int i = 0; i < 3; i++) {
        R[i] = x[i];
        R[3 + i] = y[i];
        R[6 + i] = z[i];
    }
}

void orc_camera_from_pose(const double translation[3], const double object_centre[3],
                          double camera_distance, int32_t need_rescale, int32_t z_reversed,
                          double fx, double fy, double cx, double cy, double c2w[16],
                          double w2c[16], double centre[4]) {
    /* Rescale of the pose translation into voxel units
     * (CameraUtils.cu:362-374). */
    double c[3] = {translation[0], translation[1], translation[2]};
    const double rr = (double)sqrtf((float)(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]));
    if (need_rescale && rr < 10.0)
        for (int i = 0; i < 3; i++) c[i] = c[i] * (object_centre[i] / camera_distance) + object_centre[i];

    /* formulateProjectionMat (CameraUtils.cu:45-106). */
    double R[9];
    look_at(c, object_centre, R);
    if (z_reversed) {
        const double a180[3] = {180.0, 0.0, 0.0};
        double flip[9], fr[9];
        euler_rot(a180, flip);
        mat_mul(flip, R, fr, 3, 3, 3);
        memcpy(R, fr, sizeof(R));
    }
    double r4[16] = {R[0], R[1], R[2], 0, R[3], R[4], R[5], 0, R[6], R[7], R[8], 0, 0, 0, 0, 1};
    double tr[16] = {1, 0, 0, -c[0], 0, 1, 0, -c[1], 0, 0, 1, -c[2], 0, 0, 0, 1};
    mat_mul(r4, tr, w2c, 4, 4, 4);
    mat_inv(w2c, c2w, 4);
    const double kI0[12] = {fx, 0.0, cx, 0.0, 0.0, fy, cy, 0.0, 0.0, 0.0, 1.0, 0.0};
    double P[12];
    mat_mul(kI0, w2c, P, 3, 4, 4);

    /* Camera centre = -M^-1 p4 (CameraUtils.cu:589-605). */
    const double M[9] = {P[0], P[1], P[2], P[4], P[5], P[6], P[8], P[9], P[10]};
    const double p4[3] = {P[3], P[7], P[11]};
    double Mi[9], cc[3];
    mat_inv(M, Mi, 3);
    mat_mul(Mi, p4, cc, 3, 3, 1);
    centre[0] = -cc[0];
    centre[1] = -cc[1];
    centre[2] = -cc[2];
    centre[3] = 1.0;
}

/* ------------------------------------------------------------------ */
/* Rays (Rays.cu:13-144), FP32 bit-replica.                             */
/* ------------------------------------------------------------------ */

static double dminr(double a, double b) { return a > b ? b : a; }
static double dmaxr(double a, double b) { return a > b ? a : b; }

void orc_rays_init(const double c2w[16], const double centre[4], double fx, double fy,
                   double cx, double cy, const int32_t bbox[6], int32_t img_w, int32_t img_h,
                   const uint8_t *img_r, const uint8_t *img_g, const uint8_t *img_b,
                   const uint8_t *img_mask, int32_t white_bg, float *ox, float *oy, float *oz,
                   float *dx, float *dy, float *dz, float *near_t, float *far_t, uint8_t *gt_r,
                   uint8_t *gt_g, uint8_t *gt_b, uint8_t *gt_mask) {
    for (int64_t pix = 0; pix < (int64_t)img_w * img_h; pix++) {
        const int row = (int)(pix / img_w), col = (int)(pix % img_w);
        float xc[4];
        xc[0] = (float)((((double)col + 0.5) - cx) / fx);
        xc[1] = (float)((((double)row + 0.5) - cy) / fy);
        xc[2] = 1.0f;
        xc[3] = 1.0f;
        float sp[4];
        for (int r = 0; r < 4; r++) {
            sp[r] = 0.0f;
            for (int c = 0; c < 4; c++) {
                const double prod = c2w[r * 4 + c] * (double)xc[c];
                sp[r] = (float)((double)sp[r] + prod);
            }
            sp[r] = (float)((double)sp[r] - centre[r]);
        }
        const float s0 = sp[0] * sp[0];
        const float s1 = sp[1] * sp[1];
        const float s2 = sp[2] * sp[2];
        const float ss = (s0 + s1) + s2;
        const double norm = sqrt((double)ss);
        sp[0] = (float)((double)sp[0] / norm);
        sp[1] = (float)((double)sp[1] / norm);
        sp[2] = (float)((double)sp[2] / norm);
        const double t0 = ((double)bbox[1] - centre[0]) / (double)sp[0];
        const double t1 = ((double)bbox[0] - centre[0]) / (double)sp[0];
        const double t2 = ((double)bbox[3] - centre[1]) / (double)sp[1];
        const double t3 = ((double)bbox[2] - centre[1]) / (double)sp[1];
        const double t4 = ((double)bbox[5] - centre[2]) / (double)sp[2];
        const double t5 = ((double)bbox[4] - centre[2]) / (double)sp[2];
        double tmin = dminr(t0, t1), tmax = dmaxr(t0, t1);
        tmin = dmaxr(tmin, dminr(t2, t3));
        tmax = dminr(tmax, dmaxr(t2, t3));
        tmin = dmaxr(tmin, dminr(t4, t5));
        tmax = dminr(tmax, dmaxr(t4, t5));
        ox[pix] = (float)centre[0];
        oy[pix] = (float)centre[1];
        oz[pix] = (float)centre[2];
        near_t[pix] = (float)tmin;
        far_t[pix] = (float)tmax;
        dx[pix] = sp[0];
        dy[pix] = sp[1];
        dz[pix] = sp[2];
        uint8_t gr = 0, gg = 0, gb = 0, gm = 0;
        if (img_r) {
            gm = img_mask[pix];
            if (white_bg) {
                const int m = gm;
                gr = (uint8_t)((float)(img_r[pix] * m) / 255.0f + 255.0f - (float)m);
                gg = (uint8_t)((float)(img_g[pix] * m) / 255.0f + 255.0f - (float)m);
                gb = (uint8_t)((float)(img_b[pix] * m) / 255.0f + 255.0f - (float)m);
            } else {
                gr = img_r[pix];
                gg = img_g[pix];
                gb = img_b[pix];
            }
        }
        gt_r[pix] = gr;
        gt_g[pix] = gg;
        gt_b[pix] = gb;
        gt_mask[pix] = gm;
    }
}

/* ------------------------------------------------------------------ */
/* Occupancy and caged index (SceneUtils.cu:800-913, 1081-1120).       */
/* ------------------------------------------------------------------ */

int64_t orc_build_occupancy(const float *density, int32_t W, int32_t H, int32_t D,
                            float threshold, uint8_t *occupancy, uint8_t *in_cage) {
    const int64_t WH = (int64_t)W * H;
    memset(in_cage, 0, (size_t)(WH * D));
    int64_t n = 0;
    for (int z = 0; z < D - 1; z++)
        for (int y = 0; y < H - 1; y++)
            for (int x = 0; x < W - 1; x++) {
                int occ = 0;
                for (int dz = 0; dz <= 1 && !occ; dz++)
                    for (int dy = 0; dy <= 1 && !occ; dy++)
                        for (int dx = 0; dx <= 1 && !occ; dx++) {
                            /* getIndex clamps only the upper end (SceneUtils.cu:250-260). */
                            int xx = x + dx, yy = y + dy, zz = z + dz;
                            if (xx > W - 1) xx = W - 1;
                            if (yy > H - 1) yy = H - 1;
                            if (zz > D - 1) zz = D - 1;
                            if (density[(int64_t)zz * WH + (int64_t)yy * W + xx] > threshold) occ = 1;
                        }
                const int64_t cell = (int64_t)z * (W - 1) * (H - 1) + (int64_t)y * (W - 1) + x;
                occupancy[cell] = (uint8_t)occ;
                if (occ) {
                    n++;
                    in_cage[(int64_t)z * WH + (int64_t)y * W + x] = 1;
                    if (z == D - 2) in_cage[(int64_t)(D - 1) * WH + (int64_t)y * W + x] = 1;
                    if (y == H - 2) in_cage[(int64_t)z * WH + (int64_t)(H - 1) * W + x] = 1;
                    if (x == W - 2) in_cage[(int64_t)z * WH + (int64_t)H * W + x + 1] = 1;
                }
            }
    return n;
}

int64_t orc_build_index(const uint8_t *in_cage, int32_t W, int32_t H, int32_t D,
                        uint32_t *index_offsets) {
    uint32_t k = 0;
    const int64_t n = (int64_t)W * H * D;
    for (int64_t i = 0; i < n; i++) index_offsets[i] = in_cage[i] ? k++ : 0u;
    return (int64_t)k;
}

/* ------------------------------------------------------------------ */
/* Marching (RenderUtils.cu:3-13, 480-515, 826-1004, 1582-1718).        */
/* ------------------------------------------------------------------ */

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

static uint32_t scene_index(int x, int y, int z, int W, int H, int D) {
    x = clampi(x, 0, W - 1);
    y = clampi(y, 0, H - 1);
    z = clampi(z, 0, D - 1);
    return (uint32_t)z * (uint32_t)W * (uint32_t)H + (uint32_t)y * (uint32_t)W + (uint32_t)x;
}

/* Iteration cap per ray, the same as the GPU's: it only bounds rays whose
 * far distance is infinite or whose adaptive step collapses. */
#define ORC_MAX_STEPS (1 << 20)

/* (unsigned int) truncation of a position; (-1, 0) maps to 0. */
static int base_of(double p) {
    if (!(p > 0.0)) return 0;
    if (p >= 1073741824.0) return 1073741824;
    return (int)p;
}

static double softplus13(double raw) {
    const double x = raw - 13.0;
    return x > 30.0 ? x + log1p(exp(-x)) : log1p(exp(x));
}
static double dsoftplus13(double raw) { return 1.0 / (1.0 + exp(13.0 - raw)); }
static double sigmoid01(double raw) { return 1.0 / (1.0 + exp(-0.1 * raw)); }

/* Unnormalised SH basis (RenderUtils.cu:1639-1670) from the FP32 direction. */
static void sh_basis(double x, double y, double z, double *Y) {
    Y[0] = 1.0;
    Y[1] = y;
    Y[2] = z;
    Y[3] = x;
    Y[4] = x * y;
    Y[5] = y * z;
    Y[6] = 3.0 * z * z - 1.0;
    Y[7] = x * z;
    Y[8] = x * x - y * y;
    Y[9] = y * (3.0 * x * x - y * y);
    Y[10] = x * y * z;
    Y[11] = y * (5.0 * z * z - 1.0);
    Y[12] = 5.0 * z * z * z - 3.0 * z;
    Y[13] = x * (5.0 * z * z - 1.0);
    Y[14] = (x * x - y * y) * z;
    Y[15] = x * (x * x - 3.0 * y * y);
}

typedef struct sample {
    int bx, by, bz;
    uint32_t v[8];
    double w[8];
} sample;

/* Sample location; 0 if the sample is skipped (empty cell of a sparse grid,
 * RenderUtils.cu:835-842). */
static int locate(const orc_grid *g, double px, double py, double pz, sample *s) {
    s->bx = base_of(px);
    s->by = base_of(py);
    s->bz = base_of(pz);
    if (g->sparse) {
        const int cx = clampi(s->bx, 0, g->W - 2), cy = clampi(s->by, 0, g->H - 2),
                  cz = clampi(s->bz, 0, g->D - 2);
        const int64_t cell = (int64_t)cz * (g->W - 1) * (g->H - 1) + (int64_t)cy * (g->W - 1) + cx;
        if (!g->occupancy[cell]) return 0;
    }
    const double fx = px - (double)s->bx, fy = py - (double)s->by, fz = pz - (double)s->bz;
    for (int j = 0; j < 8; j++) {
        const int dx = j & 1, dy = (j >> 1) & 1, dz = j >> 2;
        uint32_t v = scene_index(s->bx + dx, s->by + dy, s->bz + dz, g->W, g->H, g->D);
        if (g->sparse) v = g->index_offsets[v];
        s->v[j] = v;
        s->w[j] = (dx ? fx : 1.0 - fx) * (dy ? fy : 1.0 - fy) * (dz ? fz : 1.0 - fz);
    }
    return 1;
}

/* trilerpRGB (RenderUtils.cu:826-1004): raw density and raw colours. */
static int sh_active_of(const orc_grid *g) { return g->sh_active > 0 ? g->sh_active : g->S; }

static void interp(const orc_grid *g, const sample *s, const double *Y, double *sraw,
                   double craw[3]) {
    const int S = g->S;
    const int act = sh_active_of(g);
    const float *pl = g->planes;
    const int64_t n = g->n_vox;
    double sd = 0.0;
    for (int j = 0; j < 8; j++) sd += s->w[j] * (double)pl[s->v[j]];
    *sraw = sd;
    for (int ch = 0; ch < 3; ch++) {
        double c = 0.0;
        for (int k = 0; k < act; k++) {
            double shk = 0.0;
            const float *plane = pl + (int64_t)(1 + ch * S + k) * n;
            for (int j = 0; j < 8; j++) shk += s->w[j] * (double)plane[s->v[j]];
            c += shk * Y[k];
        }
        craw[ch] = c;
    }
}

static uint32_t fnv_mix(uint32_t h, uint32_t x) {
    for (int i = 0; i < 4; i++) {
        h ^= (x >> (8 * i)) & 0xFFu;
        h *= 16777619u;
    }
    return h;
}

/* Proposal bin of step k of a ray: 64 equal bins of [t_near, t_far). */
static int bin_of(int k, double step, double tn, double tf) {
    const double span = tf - tn;
    if (!(span > 0.0)) return 0;
    const double q = floor((double)k * step * 64.0 / span);
    return q > 63.0 ? 63 : (int)q;
}

/* Whether batch entry b may evaluate step k (no mask: always). */
static int bin_allowed(const uint64_t *bin_mask, int64_t b, int k, double step, double tn,
                       double tf) {
    if (!bin_mask) return 1;
    return (int)((bin_mask[b] >> bin_of(k, step, tn, tf)) & 1u);
}

static void render_fwd(const orc_grid *g, const orc_rays *r, const int32_t *sample_index,
                       int64_t start, int64_t len, double *out_r, double *out_g, double *out_b,
                       double *out_w, double *out_density, uint32_t *trace,
                       const uint64_t *bin_mask, double target_wmin, uint64_t *targets) {
    const double step = (double)g->step;
#pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads) if (g_threads > 1)
    for (int64_t b = 0; b < len; b++) {
        const int ray = sample_index[start + b];
        const double ox = r->ox[ray], oy = r->oy[ray], oz = r->oz[ray];
        const double dx = r->dx[ray], dy = r->dy[ray], dz = r->dz[ray];
        const double tn = r->near_t[ray], tf = r->far_t[ray];
        double C[3] = {0.0, 0.0, 0.0}, A = 0.0, sig = 0.0;
        uint32_t cnt = 0, h = 2166136261u;
        if (!(tn > tf)) {
            double Y[16];
            sh_basis(dx, dy, dz, Y);
            double t = tn, lam = 0.0;
            int k = 0;
            uint64_t tgt = 0;
            while (t < tf && k < ORC_MAX_STEPS) {
                const double px = fma(t, dx, ox), py = fma(t, dy, oy), pz = fma(t, dz, oz);
                sample s;
                if (!locate(g, px, py, pz, &s) || !bin_allowed(bin_mask, b, k, step, tn, tf)) {
                    if (g->adaptive) t += step;
                    ++k;
                    if (!g->adaptive) t = fma((double)k, step, tn);
                    continue;
                }
                cnt++;
                h = fnv_mix(h, (uint32_t)k);
                h = fnv_mix(h, (uint32_t)s.bx);
                h = fnv_mix(h, (uint32_t)s.by);
                h = fnv_mix(h, (uint32_t)s.bz);
                double sraw, craw[3];
                interp(g, &s, Y, &sraw, craw);
                const double sigma = softplus13(sraw);
                const double T = exp(-lam);
                const double w = T * (-expm1(-sigma * step));
                if (w > target_wmin) tgt |= (uint64_t)1 << bin_of(k, step, tn, tf);
                A += w;
                for (int ch = 0; ch < 3; ch++) C[ch] += w * sigmoid01(craw[ch]);
                sig += sigma;
                lam += sigma * step;
                ++k;
                if (g->adaptive) t += fabs(step - w);
                else t = fma((double)k, step, tn);
            }
            if (r->white_bg)
                for (int ch = 0; ch < 3; ch++) C[ch] += 1.0 - A;
            if (targets) targets[b] = tgt;
        } else if (targets) {
            targets[b] = 0;
        }
        out_r[ray] = C[0];
        out_g[ray] = C[1];
        out_b[ray] = C[2];
        out_w[ray] = A;
        out_density[ray] = sig;
        if (trace) {
            trace[2 * b] = cnt;
            trace[2 * b + 1] = h;
        }
    }
}

void orc_render_fwd(const orc_grid *g, const orc_rays *r, const int32_t *sample_index,
                    int64_t start, int64_t len, double *out_r, double *out_g, double *out_b,
                    double *out_w, double *out_density, uint32_t *trace) {
    render_fwd(g, r, sample_index, start, len, out_r, out_g, out_b, out_w, out_density, trace,
               NULL, 0.0, NULL);
}

/* The forward with a proposal mask (steps in bins whose bit of
 * bin_mask[b] is clear are skipped like unoccupied cells; bin_mask may be
 * NULL) and, when targets is not NULL, the bins holding a sample of weight
 * T alpha above target_wmin as a 64-bit mask per batch entry. */
void orc_render_fwd_masked(const orc_grid *g, const orc_rays *r, const int32_t *sample_index,
                           int64_t start, int64_t len, const uint64_t *bin_mask,
                           double target_wmin, uint64_t *targets, double *out_r, double *out_g,
                           double *out_b, double *out_w, double *out_density, uint32_t *trace) {
    render_fwd(g, r, sample_index, start, len, out_r, out_g, out_b, out_w, out_density, trace,
               bin_mask, target_wmin, targets);
}

/* renderImageGPU + getBound + volumeRender_RGB (RenderUtils.cu:1720-1810,
 * 34-97, 1443-1580): eval image of one camera, adaptive marching rule.
 * Outputs planar [3][h*w] (any may be NULL). */
void orc_render_image(const orc_grid *g, const double c2w[16], const double centre[4],
                      double fx, double fy, double cx, double cy, int32_t width, int32_t height,
                      int32_t white_bg, double *rgb, uint8_t *rgb8, uint8_t *mask,
                      double *disparity, double *rays_out) {
    const double step = (double)g->step;
    const size_t n = (size_t)width * (size_t)height;
    for (size_t px = 0; px < n; px++) {
        const int row = (int)(px / (size_t)width), col = (int)(px % (size_t)height);
        const double x[4] = {(col + 0.5 - cx) / fx, (row + 0.5 - cy) / fy, 1.0, 1.0};
        double sp[3];
        for (int r = 0; r < 3; r++) {
            double v = 0.0;
            for (int c = 0; c < 4; c++) v += c2w[r * 4 + c] * x[c];
            sp[r] = v - centre[r];
        }
        const double nn = sqrt(sp[0] * sp[0] + sp[1] * sp[1] + sp[2] * sp[2]);
        const double d[3] = {sp[0] / nn, sp[1] / nn, sp[2] / nn};
        const double lo[3] = {g->bbox[0], g->bbox[2], g->bbox[4]};
        const double hi[3] = {g->bbox[1], g->bbox[3], g->bbox[5]};
        double tmin = 0.0, tmax = 0.0;
        for (int i = 0; i < 3; i++) {
            const double t1 = (hi[i] - centre[i]) / d[i], t2 = (lo[i] - centre[i]) / d[i];
            const double mn = t1 < t2 ? t1 : t2, mx = t1 > t2 ? t1 : t2;
            if (i == 0) {
                tmin = mn;
                tmax = mx;
            } else {
                tmin = tmin > mn ? tmin : mn;
                tmax = tmax < mx ? tmax : mx;
            }
        }
        if (rays_out) {
            rays_out[5 * px + 0] = tmin;
            rays_out[5 * px + 1] = tmax;
            for (int i = 0; i < 3; i++) rays_out[5 * px + 2 + i] = d[i];
        }
        if (tmin > tmax) {
            for (int c = 0; c < 3; c++) {
                if (rgb) rgb[c * n + px] = white_bg ? 1.0 : 0.0;
                if (rgb8) rgb8[c * n + px] = 255;
            }
            if (mask) mask[px] = 0;
            if (disparity) disparity[px] = 0.0;
            continue;
        }
        double Y[16];
        sh_basis((float)d[0], (float)d[1], (float)d[2], Y);
        double C[3] = {0.0, 0.0, 0.0}, A = 0.0, disp = 0.0, lam = 0.0, t = tmin;
        int k = 0;
        while (t < tmax && k < ORC_MAX_STEPS) {
            const double p0 = fma(t, d[0], centre[0]), p1 = fma(t, d[1], centre[1]),
                         p2 = fma(t, d[2], centre[2]);
            sample s;
            if (!locate(g, p0, p1, p2, &s)) {
                t += step;
                ++k;
                continue;
            }
            double sraw, craw[3];
            interp(g, &s, Y, &sraw, craw);
            const double sigma = softplus13(sraw);
            const double T = exp(-lam);
            const double w = T * (-expm1(-sigma * step));
            A += w;
            disp += w * t;
            for (int ch = 0; ch < 3; ch++) C[ch] += w * sigmoid01(craw[ch]);
            lam += sigma * step;
            ++k;
            t += fabs(step - w);
        }
        for (int c = 0; c < 3; c++) {
            if (rgb8) rgb8[c * n + px] = (uint8_t)(int)((C[c] + 1.0 - A) * 255.0 + 0.5);
            if (rgb) rgb[c * n + px] = C[c] + (white_bg ? 1.0 - A : 0.0);
        }
        if (mask) mask[px] = (uint8_t)(int)(A * 255.0);
        if (disparity) disparity[px] = disp;
    }
}

/* renderRays_withGradients_RGBGPU + trilerpGradientsRGB
 * (RenderUtils.cu:2149-2254, 1074-1364); bin_mask as in the forward. */
static void render_bwd(const orc_grid *g, const orc_rays *r, const int32_t *sample_index,
                       int64_t start, int64_t len, const double *fwd_r, const double *fwd_g,
                       const double *fwd_b, const double *fwd_w, double mask_coef,
                       float huber_delta, double *losses, double *grad,
                       const uint64_t *bin_mask, double *grad_abs) {
    const double step = (double)g->step;
    const int S = g->S;
    const int64_t n = g->n_vox;
    const double kc = (double)0.6667f, kd = 0.6667;
#pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads) if (g_threads > 1)
    for (int64_t b = 0; b < len; b++) {
        const int ray = sample_index[start + b];
        const double tn = r->near_t[ray], tf = r->far_t[ray];
        if (tn > tf) {
            losses[b] = 0.0;
            continue;
        }
        const double Cf[3] = {fwd_r[ray], fwd_g[ray], fwd_b[ray]};
        const double Af = fwd_w[ray];
        double res[3] = {Cf[0] - r->gt_r[ray] / 255.0, Cf[1] - r->gt_g[ray] / 255.0,
                         Cf[2] - r->gt_b[ray] / 255.0};
        const double loss = (fabs(res[0]) + fabs(res[1]) + fabs(res[2])) * 255.0 / 3.0;
        if (isnan(loss)) {
            losses[b] = 255.0;
            continue;
        }
        losses[b] = loss;
        /* Clip (RenderUtils.cu:2164-2174): r > d -> d - d^2/2, r < -d -> d^2/2 - d. */
        const double hd = (double)huber_delta;
        for (int i = 0; i < 3; i++) {
            if (res[i] > hd) res[i] = hd - 0.5 * hd * hd;
            else if (res[i] < -hd) res[i] = 0.5 * hd * hd - hd;
        }
        /* Mask and opacity-logit term (RenderUtils.cu:2184-2193). */
        double m = 0.0;
        if (r->gt_mask[ray] && Af < 1.0) m = (r->gt_mask[ray] / 255.0 - Af) * mask_coef;
        double wp = Af > 1e-6 ? Af : 1e-6;
        wp = wp < 1.0 - 1e-6 ? wp : 1.0 - 1e-6;
        m += (log(wp) - log(1.0 - wp)) * 1e-3;

        const double ox = r->ox[ray], oy = r->oy[ray], oz = r->oz[ray];
        const double dx = r->dx[ray], dy = r->dy[ray], dz = r->dz[ray];
        double Y[16];
        sh_basis(dx, dy, dz, Y);
        double t = tn, lam = 0.0, Ai = 0.0, Ch[3] = {0.0, 0.0, 0.0};
        int k = 0;
        while (t < tf && k < ORC_MAX_STEPS) {
            const double px = fma(t, dx, ox), py = fma(t, dy, oy), pz = fma(t, dz, oz);
            sample s;
            if (!locate(g, px, py, pz, &s) || !bin_allowed(bin_mask, b, k, step, tn, tf)) {
                if (g->adaptive) t += step;
                ++k;
                if (!g->adaptive) t = fma((double)k, step, tn);
                continue;
            }
            double sraw, craw[3], col[3];
            interp(g, &s, Y, &sraw, craw);
            const double sigma = softplus13(sraw);
            double T = exp(-lam);
            const double w = T * (-expm1(-sigma * step));
            Ai += w;
            double gc[3];
            for (int ch = 0; ch < 3; ch++) {
                col[ch] = sigmoid01(craw[ch]);
                Ch[ch] += w * col[ch];
                gc[ch] = kc * w * res[ch] * (0.1 * col[ch] * (1.0 - col[ch]));
            }
            lam += sigma * step;
            T = exp(-lam);
            const double ds = dsoftplus13(sraw);
            double gs = 0.0, gs_abs = 0.0;
            for (int ch = 0; ch < 3; ch++) {
                gs += res[ch] * (T * col[ch] - (Cf[ch] - Ch[ch]));
                gs_abs += fabs(res[ch]) * (fabs(T * col[ch]) + fabs(Cf[ch] - Ch[ch]));
            }
            gs *= kd * step * ds;
            gs += step * ds * (Af - Ai - T) * m;
            gs_abs = kd * step * ds * gs_abs + step * ds * (fabs(Af - Ai) + T) * fabs(m);
            for (int j = 0; j < 8; j++) {
                add_d(&grad[s.v[j]], gs * s.w[j]);
                if (grad_abs) add_d(&grad_abs[s.v[j]], gs_abs * fabs(s.w[j]));
                for (int ch = 0; ch < 3; ch++)
                    for (int kk = g->sh_first; kk < sh_active_of(g); kk++) {
                        const double c = gc[ch] * s.w[j] * Y[kk];
                        add_d(&grad[(int64_t)(1 + ch * S + kk) * n + s.v[j]], c);
                        if (grad_abs) add_d(&grad_abs[(int64_t)(1 + ch * S + kk) * n + s.v[j]], fabs(c));
                    }
            }
            ++k;
            if (g->adaptive) t += fabs(step - w);
            else t = fma((double)k, step, tn);
        }
    }
}

void orc_render_bwd(const orc_grid *g, const orc_rays *r, const int32_t *sample_index,
                    int64_t start, int64_t len, const double *fwd_r, const double *fwd_g,
                    const double *fwd_b, const double *fwd_w, double mask_coef,
                    float huber_delta, double *losses, double *grad) {
    render_bwd(g, r, sample_index, start, len, fwd_r, fwd_g, fwd_b, fwd_w, mask_coef, huber_delta,
               losses, grad, NULL, NULL);
}

void orc_render_bwd_cond(const orc_grid *g, const orc_rays *r, const int32_t *sample_index,
                         int64_t start, int64_t len, const double *fwd_r, const double *fwd_g,
                         const double *fwd_b, const double *fwd_w, double mask_coef,
                         float huber_delta, double *losses, double *grad, double *grad_abs) {
    render_bwd(g, r, sample_index, start, len, fwd_r, fwd_g, fwd_b, fwd_w, mask_coef, huber_delta,
               losses, grad, NULL, grad_abs);
}

void orc_render_bwd_masked(const orc_grid *g, const orc_rays *r, const int32_t *sample_index,
                           int64_t start, int64_t len, const uint64_t *bin_mask,
                           const double *fwd_r, const double *fwd_g, const double *fwd_b,
                           const double *fwd_w, double mask_coef, float huber_delta,
                           double *losses, double *grad) {
    render_bwd(g, r, sample_index, start, len, fwd_r, fwd_g, fwd_b, fwd_w, mask_coef, huber_delta,
               losses, grad, bin_mask, NULL);
}

/* ------------------------------------------------------------------ */
/* Variance losses (RenderUtils.cu:2442-2763).                          */
/* ------------------------------------------------------------------ */

void orc_tv_disparity(const orc_grid *g, double lambda_density, double lambda_sh,
                      int32_t optimize_step, double *grad) {
    const int W = g->W, H = g->H, D = g->D, S = g->S;
    const int64_t n = g->n_vox;
    const int P = 1 + 3 * S;
#pragma omp parallel for collapse(2) schedule(dynamic, 4) num_threads(g_threads) if (g_threads > 1)
    for (int z = 0; z < D - 1; z++)
        for (int y = 0; y < H - 1; y++)
            for (int x = 0; x < W - 1; x++) {
                const int64_t cell = (int64_t)z * (W - 1) * (H - 1) + (int64_t)y * (W - 1) + x;
                if (g->sparse && !g->occupancy[cell]) continue;
                uint32_t v[8];
                for (int j = 0; j < 8; j++) {
                    const int dx = j & 1, dy = (j >> 1) & 1, dz = j >> 2;
                    v[j] = scene_index(x + dx, y + dy, z + dz, W, H, D);
                    if (g->sparse) v[j] = g->index_offsets[v[j]];
                }
                for (int p = 0; p < P; p++) {
                    double lam;
                    if (p == 0) lam = lambda_density;
                    else if (optimize_step <= 1 && (p - 1) % S >= g->sh_first &&
                             (p - 1) % S < sh_active_of(g))
                        lam = lambda_sh;
                    else continue;
                    const float *pl = g->planes + (int64_t)p * n;
                    double mean = 0.0;
                    for (int j = 0; j < 8; j++) mean += pl[v[j]];
                    mean /= 8.0;
                    for (int j = 0; j < 8; j++)
                        add_d(&grad[(int64_t)p * n + v[j]], lam * (pl[v[j]] - mean));
                }
            }
}

void orc_tv_neighbour6_cols(const orc_grid *g, double lambda_density, double lambda_sh,
                            int64_t col_limit, double *grad) {
    const int W = g->W, H = g->H, D = g->D, S = g->S;
    const int64_t n = g->n_vox;
    const int P = 1 + 3 * S;
    /* Neighbour order of the 27-index stencil entries 4, 10, 12, 14, 16, 22. */
    static const int off[6][3] = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0},
                                  {1, 0, 0},  {0, 1, 0},  {0, 0, 1}};
    for (int y = 1; y < H - 1; y++)
        for (int x = 1; x < W - 1; x++) {
            if ((int64_t)y * W + x >= col_limit) continue;
            for (int z = 1; z < D - 1; z++) {
                uint32_t self = scene_index(x, y, z, W, H, D);
                if (g->sparse) {
                    self = g->index_offsets[self];
                    if (!g->in_cage[self]) continue;
                }
                uint32_t nb[6];
                for (int i = 0; i < 6; i++) {
                    nb[i] = scene_index(x + off[i][0], y + off[i][1], z + off[i][2], W, H, D);
                    if (g->sparse) nb[i] = g->index_offsets[nb[i]];
                }
                for (int p = 0; p < P; p++) {
                    double lam;
                    if (p == 0) lam = lambda_density;
                    else if ((p - 1) % S >= g->sh_first && (p - 1) % S < sh_active_of(g))
                        lam = lambda_sh;
                    else continue;
                    const float *pl = g->planes + (int64_t)p * n;
                    const double u = pl[self];
                    double sq = 0.0, dis = 0.0;
                    for (int i = 0; i < 6; i++) {
                        const double un = nb[i] != 0u ? (double)pl[nb[i]] : u;
                        sq += (u - un) * (u - un);
                        dis += u - un;
                    }
                    const double sd = sqrt(sq) / 6.0;
                    grad[(int64_t)p * n + self] += lam / (sd + 1e-6) * dis;
                }
            }
        }
}

void orc_tv_neighbour6(const orc_grid *g, double lambda_density, double lambda_sh,
                       double *grad) {
    orc_tv_neighbour6_cols(g, lambda_density, lambda_sh, (int64_t)g->W * g->H, grad);
}

/* ------------------------------------------------------------------ */
/* Adam (Optimizer.cu:359-415) and the loss reduction.                  */
/* ------------------------------------------------------------------ */

void orc_adam(float *params, float *grad, float *m, float *v, const float *adaptive_lr,
              int32_t P, int64_t n, double lr, double sh_lr, int32_t t) {
    orc_adam_band(params, grad, m, v, adaptive_lr, P, 0, n, lr, sh_lr, t, 0, 0);
}

void orc_adam_band(float *params, float *grad, float *m, float *v, const float *adaptive_lr,
                   int32_t P, int32_t S, int64_t n, double lr, double sh_lr, int32_t t,
                   int32_t sh_lo, int32_t sh_hi) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1.e-6f;
    const double bc1 = 1.0 - (double)powf(b1, (float)t);
    const double bc2 = 1.0 - (double)powf(b2, (float)t);
#pragma omp parallel for collapse(2) schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int p = 0; p < P; p++)
        for (int64_t i = 0; i < n; i++) {
            if (S > 0 && p > 0 && ((p - 1) % S < sh_lo || (p - 1) % S >= sh_hi)) continue;
            const int64_t k = (int64_t)p * n + i;
            const float g = grad[k];
            const float mm = (float)((double)(b1 * m[k]) + (1.0 - (double)b1) * (double)g);
            const float vv = (float)((double)(b2 * v[k]) + (1.0 - (double)b2) * (double)g * (double)g);
            const float cm = (float)((double)mm / bc1);
            const float cv = (float)((double)vv / bc2);
            const double lrp = p == 0 ? (adaptive_lr ? lr * (double)adaptive_lr[i] : lr) : sh_lr;
            const double nt = (double)params[k] - lrp * (double)cm / (double)(sqrtf(cv) + eps);
            params[k] = p == 0 ? (nt < 0.0 ? 0.0f : (float)nt) : (float)nt;
            m[k] = mm;
            v[k] = vv;
            grad[k] = 0.0f;
        }
}

/* ------------------------------------------------------------------ */
/* Upsampling (SceneUtils.cu:250-485).                                  */
/* ------------------------------------------------------------------ */

/* getIndex (SceneUtils.cu:250-260): upper clamp only (unsigned inputs). */
static uint32_t up_index(uint32_t x, uint32_t y, uint32_t z, int W, int H, int D) {
    if (x > (uint32_t)(W - 1)) x = (uint32_t)(W - 1);
    if (y > (uint32_t)(H - 1)) y = (uint32_t)(H - 1);
    if (z > (uint32_t)(D - 1)) z = (uint32_t)(D - 1);
    return z * (uint32_t)W * (uint32_t)H + y * (uint32_t)W + x;
}

/* upsample_trilinear (SceneUtils.cu:262-416) at destination vertex
 * (x, y, z) of a W2 x H2 x D2 grid, sampling the source at
 * x * 1.f * W / W2 (float operations, upsampleScene :428-433, :448).  Corners
 * in the reference's order, each added to a float accumulator as
 * (float)((double)acc + weight * value).  Fills tri[0 .. P_src) (bands
 * beyond the first when low_band) and returns 0 when the source is caged and
 * the cell holding the sample is unoccupied. */
static int up_trilinear(const orc_grid *s, int x, int y, int z, int W2, int H2, int D2,
                        int low_band, float *tri) {
    const int S = s->S, P = 1 + 3 * S;
    float bc[3];
    bc[0] = (float)x * 1.f * (float)s->W / (float)W2;
    bc[1] = (float)y * 1.f * (float)s->H / (float)H2;
    bc[2] = (float)z * 1.f * (float)s->D / (float)D2;
    const uint32_t b0 = (uint32_t)bc[0], b1 = (uint32_t)bc[1], b2 = (uint32_t)bc[2];
    if (s->sparse) {
        const uint32_t cell = up_index(b0, b1, b2, s->W - 1, s->H - 1, s->D - 1);
        if (!s->occupancy[cell]) return 0;
    }
    const double fx = (double)(bc[0] - (float)b0);
    const double fy = (double)(bc[1] - (float)b1);
    const double fz = (double)(bc[2] - (float)b2);
    for (int i = 0; i < P; i++) tri[i] = 0.f;
    const int nsh = low_band ? 1 : S;
    /* (dx, dy, dz) in the order the reference visits the corners */
    static const int ord[8][3] = {{0, 0, 0}, {0, 1, 0}, {1, 0, 0}, {1, 1, 0},
                                  {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
    for (int c = 0; c < 8; c++) {
        const int dx = ord[c][0], dy = ord[c][1], dz = ord[c][2];
        uint32_t idx = up_index(b0 + (uint32_t)dx, b1 + (uint32_t)dy, b2 + (uint32_t)dz, s->W,
                                s->H, s->D);
        if (s->sparse) idx = s->index_offsets[idx];
        const double w = (dx ? fx : 1. - fx) * (dy ? fy : 1. - fy) * (dz ? fz : 1. - fz);
        const float *pl = s->planes;
        const int64_t n = s->n_vox;
        tri[0] = (float)((double)tri[0] + w * (double)pl[idx]);
        for (int i = 0; i < nsh; i++)
            for (int ch = 0; ch < 3; ch++) {
                const int q = 1 + ch * S + i;
                tri[q] = (float)((double)tri[q] + w * (double)pl[(int64_t)q * n + idx]);
            }
    }
    return 1;
}

void orc_upsample_density(const orc_grid *src, int32_t W2, int32_t H2, int32_t D2,
                          float threshold, float *density) {
    float *tri = (float *)malloc(sizeof(float) * (size_t)(1 + 3 * src->S));
    for (int z = 0; z < D2; z++)
        for (int y = 0; y < H2; y++)
            for (int x = 0; x < W2; x++) {
                const size_t d = ((size_t)z * H2 + y) * W2 + x;
                if (up_trilinear(src, x, y, z, W2, H2, D2, 1, tri))
                    density[d] = tri[0] < threshold ? 0.f : tri[0];
                else
                    density[d] = 0.f;
            }
    free(tri);
}

void orc_upsample(const orc_grid *src, int32_t W2, int32_t H2, int32_t D2, int32_t S2,
                  const uint8_t *dst_in_cage, const uint32_t *dst_index, int64_t dst_n,
                  float threshold, int32_t low_band, float *dst) {
    const int S = src->S, ncopy = low_band ? 1 : S;
    float *tri = (float *)malloc(sizeof(float) * (size_t)(1 + 3 * S));
    for (int z = 0; z < D2; z++)
        for (int y = 0; y < H2; y++)
            for (int x = 0; x < W2; x++) {
                size_t r = ((size_t)z * H2 + y) * W2 + x;
                if (dst_in_cage) {
                    if (!dst_in_cage[r]) continue;
                    r = dst_index[r];
                }
                const int val = up_trilinear(src, x, y, z, W2, H2, D2, low_band, tri);
                dst[r] = val ? (tri[0] < threshold ? 0.f : tri[0]) : 0.f;
                for (int i = 0; i < ncopy; i++)
                    for (int ch = 0; ch < 3; ch++)
                        dst[(size_t)(1 + ch * S2 + i) * (size_t)dst_n + r] =
                            val ? tri[1 + ch * S + i] : 0.f;
            }
    free(tri);
}

double orc_mse(const double *losses, int64_t len) {
    double s = 0.0;
    for (int64_t i = 0; i < len; i++) s += losses[i] * losses[i];
    return s / (double)len;
}

double orc_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
