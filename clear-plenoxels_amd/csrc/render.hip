// Ray marching: forward compositing, analytic backward, their fusion, and
// two bookkeeping variants (index trace, touched-record bitmap).
//
// Work mapping (CDNA4, wave64): one wavefront marches one ray.  The ray's
// samples are visited in order; per sample the 64 lanes split the record's
// parameters: DPP row 0/1/2 hold the red/green/blue SH coefficient k at lane
// k (k < S), row 3 lane 0 holds the density.  Each lane gathers its
// parameter from the 8 cell corners (8 loads per lane, all lanes of a load
// hitting one 128-byte record), interpolates, multiplies by its SH basis
// value, and a 4-step DPP rotate-sum per row gives sigma_raw and the three
// raw colours.  The compositing scalars are wave-uniform; transmittance,
// opacity and colour accumulate in FP64 as in volumeRender_RGB_train
// (RenderUtils.cu:1582-1718).  The backward re-marches the ray and each
// lane owns the gradient of its parameter: per-corner sums are merged in
// registers while consecutive samples stay in the same cell and flushed
// with one coalesced global_atomic_add_f32 per corner (instead of
// 8 x (1+3S) scattered atomicAdds per sample, RenderUtils.cu:1074-1364).
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <type_traits>

#include "common.hpp"

namespace pxl {
namespace {

// kLog: forward plus the sample log of the two-kernel fused step (group
// marcher only).
// kImage: eval render of one camera image (renderImageGPU), adaptive rule.
enum Mode { kFwd = 0, kBwd = 1, kFused = 2, kTrace = 3, kTouch = 4, kLog = 5, kImage = 6 };

// Hard cap on marching iterations per ray.  Legitimate rays need at most a
// few thousand (diagonal / step); the cap only bounds rays whose far
// distance is infinite or whose adaptive step |step - w| collapses.
constexpr int kMaxSteps = 1 << 20;

struct RenderArgs {
    const float *params;
    int64_t n_vox;
    const uint32_t *occ;
    const uint32_t *bricks;
    const uint32_t *dbricks;  // the brick map dilated by one brick (after the map)
    int nbx, nby;
    const uint32_t *io;
    int W, H, D, stride;
    int gstride;  // floats per gradient record
    // SH coefficients k < sh_active enter the colour; gradients go to the
    // density and to k in [sh_first, sh_active).
    int sh_active, sh_first;
    float step;
    int white_bg;
    pxl_rays rays;
    const int32_t *sample_index;
    int64_t start, len;
    float *grad;
    double *losses;
    double mask_coef;
    float huber;
    uint32_t *trace;
    uint32_t *bitmap;
    // Sample log of the two-kernel fused path (all null when unused):
    // FP64 forward totals C0, C1, C2, A per batch entry, logged sample count
    // (-1: backward done by the marcher itself), and per logged sample the 8
    // corner records and (sraw, col0, col1, col2, k) padded to 8 words.
    double *log_tot;
    int32_t *log_count;
    uint32_t *log_rec;
    float *log_val;
    int log_cap;
    int log_redo;  // kFused: only the batch entries whose count is -1
    // Chunk-parallel gradient: per batch entry and 64-sample chunk of its log
    // the FP64 (T, A, C0, C1, C2) in front of the chunk; log_chunks =
    // ceil(log_cap / 64) chunks per entry.
    double *log_pre;
    int log_chunks;
    // Batch entries longer than the log, appended by k_march8<kLog>; their
    // count is work[kRedoCounter].
    int32_t *redo_list;
    // Logged step: per parameter p and batch entry b the entry's FP64 sum
    // into record 0 ([32][len], zeroed by the forward), folded into the
    // gradient in a fixed order by k_rec0_fold.  Null: record 0 takes one
    // float atomic per ray.
    double *rec0;
    double *rec0_part;  // [32][kFoldParts] partial sums of the fold
    // Brick-binned gradient (schedule gradient PXL_GRAD_BINNED): per bin key
    // (brick of 8^3 vertices x 8 classes) the sample count and the fill
    // pointer into gb_list (the log rows in key order), the bricks with
    // work (count in work[kTargetCounter]), the scan's block sums, per batch
    // entry the ray's SH basis; vertex bricks per axis.
    uint32_t *gb_count, *gb_fill, *gb_bricks, *gb_list, *gb_blk;
    float *gb_y;
    int vbx, vby, vbz;
    // Persistent launches: 8 claim counters per kernel (zeroed before launch).
    unsigned long long *work;
    // Learned proposal: per batch entry, the bins a ray may evaluate (null:
    // all), and the bins where the logged samples carry weight > target_wmin
    // (written by k_log_grad; target_valid 0 for rays it does not see).
    const uint64_t *bin_mask;
    uint64_t *bin_targets;
    uint8_t *target_valid;
    double target_wmin;
    // Eval image (kImage): camera, size, grid box and the output planes
    // (any may be null): RGB float [3][h*w], RGB bytes [3][h*w], mask and
    // disparity [h*w].
    pxl_camera cam;
    pxl_bbox box;
    int img_w, img_h;
    float *img_rgb;
    uint8_t *img_rgb8, *img_mask;
    float *img_disp;
};

// The dilated half of a brick map written by pxl_build_brick_occupancy.
inline const uint32_t *dilated_bricks(const uint32_t *bricks, const pxl_grid *g) {
    if (bricks == nullptr) return nullptr;
    const int64_t nb = (int64_t)((g->W - 1 + 7) / 8) * ((g->H - 1 + 7) / 8) * ((g->D - 1 + 7) / 8);
    return bricks + (nb + 31) / 32;
}

// Eval ray of pixel (row, col) (getBound, RenderUtils.cu:34-97): the pixel
// centre through the camera, c2w x - centre, normalised, all FP64 with the
// reference's summation order; near / far from the slabs of the integer box
// (no clamp at 0).
__device__ inline void eval_ray(const RenderArgs &a, int row, int col, double d[3],
                                double &tmin, double &tmax) {
    const double x[4] = {((double)col + 0.5 - a.cam.cx) / a.cam.fx,
                         ((double)row + 0.5 - a.cam.cy) / a.cam.fy, 1.0, 1.0};
    double sp[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        double v = 0.0;
#pragma unroll
        for (int c = 0; c < 4; c++) v += a.cam.c2w[r * 4 + c] * x[c];
        sp[r] = v - a.cam.centre[r];
    }
    const double n = sqrt(sp[0] * sp[0] + sp[1] * sp[1] + sp[2] * sp[2]);
    d[0] = sp[0] / n;
    d[1] = sp[1] / n;
    d[2] = sp[2] / n;
    const double lo[3] = {(double)a.box.x_min, (double)a.box.y_min, (double)a.box.z_min};
    const double hi[3] = {(double)a.box.x_max, (double)a.box.y_max, (double)a.box.z_max};
    tmin = -1.0 / 0.0;
    tmax = 1.0 / 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double t1 = (hi[i] - a.cam.centre[i]) / d[i], t2 = (lo[i] - a.cam.centre[i]) / d[i];
        const double mn = t1 < t2 ? t1 : t2, mx = t1 > t2 ? t1 : t2;
        if (i == 0) {
            tmin = mn;
            tmax = mx;
        } else {
            tmin = tmin > mn ? tmin : mn;
            tmax = tmax < mx ? tmax : mx;
        }
    }
}

// Proposal bin of step k: 64 equal bins of [t_near, t_far).
__device__ __forceinline__ int bin_of(int k, double step, double tn, double tf) {
    const double span = tf - tn;
    if (!(span > 0.0)) return 0;
    const double q = floor((double)k * step * 64.0 / span);
    return q > 63.0 ? 63 : (int)q;
}

// Float building blocks on the hardware exp2 / log2 / reciprocal (about one
// ulp each).  exp(z) = 2^h (1 + l ln 2) with h + l = z log2 e split by an
// FMA (h the rounded product), so large |z| keeps a few-ulp result.  log1p
// and expm1 use Kahan's identities log1p(y) = y log(u) / (u - 1), u = 1 + y,
// and expm1(x) = (u - 1) x / log(u), u = e^x, which keep the relative
// accuracy of small arguments to a few ulps (W. Kahan, HP-15C Advanced
// Functions Handbook, 1982).
__device__ __forceinline__ float fexp(float z) {
    constexpr float kL2e = 1.44269504088896341f, kL2eLo = 1.925963e-08f;
    const float h = z * kL2e;
    const float l = __builtin_fmaf(z, kL2e, -h) + z * kL2eLo;
    return __builtin_amdgcn_exp2f(h) * (1.0f + l * 0.693147180559945309f);
}
__device__ __forceinline__ float flog(float u) {
    return __builtin_amdgcn_logf(u) * 0.693147180559945309f;
}
__device__ __forceinline__ float frcp(float u) { return __builtin_amdgcn_rcpf(u); }
__device__ __forceinline__ float flog1p(float y) {  // y >= 0
    const float u = 1.0f + y;
    return u == 1.0f ? y : flog(u) * (y / (u - 1.0f));
}
__device__ __forceinline__ float fexpm1(float x) {  // |x| moderate
    const float u = fexp(x);
    if (u == 1.0f) return x;
    const float um1 = u - 1.0f;
    return um1 == -1.0f ? -1.0f : um1 * (x / flog(u));
}

// Activations (RenderUtils.cu:480-515) in float or double.
template <typename R> __device__ __forceinline__ R softplus_shift(R raw);
template <> __device__ __forceinline__ float softplus_shift<float>(float raw) {
    // log(1 + e^x) = max(x, 0) + log1p(e^-|x|)
    const float x = raw - 13.0f;
    return fmaxf(x, 0.0f) + flog1p(fexp(-fabsf(x)));
}
template <> __device__ __forceinline__ double softplus_shift<double>(double raw) {
    const double x = raw - 13.0;
    return x > 30.0 ? x + log1p(exp(-x)) : log1p(exp(x));
}
template <typename R> __device__ __forceinline__ R dsoftplus_shift(R raw);
template <> __device__ __forceinline__ float dsoftplus_shift<float>(float raw) {
    return frcp(1.0f + fexp(13.0f - raw));
}
template <> __device__ __forceinline__ double dsoftplus_shift<double>(double raw) {
    return 1.0 / (1.0 + exp(13.0 - raw));
}
template <typename R> __device__ __forceinline__ R sigmoid01(R raw);
template <> __device__ __forceinline__ float sigmoid01<float>(float raw) {
    return frcp(1.0f + fexp(-0.1f * raw));
}
template <> __device__ __forceinline__ double sigmoid01<double>(double raw) {
    return 1.0 / (1.0 + exp(-0.1 * raw));
}
// 1 - exp(-x) without cancellation for small x.
template <typename R> __device__ __forceinline__ R one_minus_exp_neg(R x);
template <> __device__ __forceinline__ float one_minus_exp_neg<float>(float x) {
    return -fexpm1(-x);
}
template <> __device__ __forceinline__ double one_minus_exp_neg<double>(double x) {
    return -expm1(-x);
}

template <typename R> __device__ __forceinline__ R row_sum(R v);
template <> __device__ __forceinline__ float row_sum<float>(float v) { return row_sum_f(v); }
template <> __device__ __forceinline__ double row_sum<double>(double v) { return row_sum_d(v); }
template <typename R> __device__ __forceinline__ R readlane(R v, int l);
template <> __device__ __forceinline__ float readlane<float>(float v, int l) { return readlane_f(v, l); }
template <> __device__ __forceinline__ double readlane<double>(double v, int l) { return readlane_d(v, l); }

// Unnormalised real SH basis (RenderUtils.cu:1639-1670) of the FP32
// direction; entry k of the first 16.  Evaluated in FP64 (once per ray):
// entries such as 3 z^2 - 1 cancel, and their FP32 evaluation (the
// reference's) carries an absolute error of ~1e-7 that the gradient sums
// of many rays then pile up on entries that nearly cancel.
__device__ __forceinline__ double sh_basis_d(int k, double x, double y, double z) {
    switch (k) {
        case 0: return 1.0;
        case 1: return y;
        case 2: return z;
        case 3: return x;
        case 4: return x * y;
        case 5: return y * z;
        case 6: return 3.0 * z * z - 1.0;
        case 7: return x * z;
        case 8: return x * x - y * y;
        case 9: return y * (3.0 * x * x - y * y);
        case 10: return x * y * z;
        case 11: return y * (5.0 * z * z - 1.0);
        case 12: return 5.0 * z * z * z - 3.0 * z;
        case 13: return x * (5.0 * z * z - 1.0);
        case 14: return (x * x - y * y) * z;
        case 15: return x * (x * x - 3.0 * y * y);
        default: return 0.0;
    }
}
__device__ __forceinline__ float sh_basis(int k, float x, float y, float z) {
    return (float)sh_basis_d(k, (double)x, (double)y, (double)z);
}

// One sample position: base cell, corner records and trilinear weights.
// The cell part (bx, by, bz, occ, v) is kept across samples: consecutive
// samples that stay in one cell reuse its occupancy test and the 8
// corner-record lookups.
template <typename R>
struct Sample {
    int bx, by, bz;
    bool occ;
    uint32_t v[8];
    R w[8];
};

template <typename R>
__device__ __forceinline__ void sample_reset(Sample<R> &s) {
    s.bx = s.by = s.bz = -2;
    s.occ = false;
}

// Base cell of a position: the reference truncates with (unsigned int),
// which maps (-1, 0) to 0 (RenderUtils.cu:828-832).
__device__ __forceinline__ int base_of(double p) {
    if (!(p > 0.0)) return 0;
    if (p >= 1073741824.0) return 1073741824;
    return (int)p;
}

// Returns false if the sample is skipped (sparse grid, empty cell:
// RenderUtils.cu:835-842 and 1681-1686).  A set bit of the optional
// 8x8x8-cell brick map is necessary for any occupied cell inside the
// brick, so a clear brick bit skips the sample without reading the
// cell bitfield; the set of evaluated samples is unchanged.
template <typename R, bool SPARSE>
__device__ __forceinline__ bool locate(const RenderArgs &a, double px, double py,
                                       double pz, Sample<R> &s) {
    const int bx = base_of(px), by = base_of(py), bz = base_of(pz);
    if (bx != s.bx || by != s.by || bz != s.bz) {
        s.bx = bx;
        s.by = by;
        s.bz = bz;
        bool occ = true;
        if (SPARSE) {
            const int cx = clampi(bx, 0, a.W - 2);
            const int cy = clampi(by, 0, a.H - 2);
            const int cz = clampi(bz, 0, a.D - 2);
            if (a.bricks != nullptr) {
                const uint32_t br = (uint32_t)(cz >> 3) * (uint32_t)a.nbx * (uint32_t)a.nby +
                                    (uint32_t)(cy >> 3) * (uint32_t)a.nbx + (uint32_t)(cx >> 3);
                occ = (a.bricks[br >> 5] >> (br & 31)) & 1u;
            }
            if (occ) {
                const uint32_t cell = (uint32_t)cz * (uint32_t)(a.W - 1) * (uint32_t)(a.H - 1) +
                                      (uint32_t)cy * (uint32_t)(a.W - 1) + (uint32_t)cx;
                occ = (a.occ[cell >> 5] >> (cell & 31)) & 1u;
            }
        }
        s.occ = occ;
        if (occ) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int dx = j & 1, dy = (j >> 1) & 1, dz = j >> 2;
                uint32_t v = scene_index(bx + dx, by + dy, bz + dz, a.W, a.H, a.D);
                if (SPARSE) v = min(a.io[v], (uint32_t)(a.n_vox - 1));
                s.v[j] = v;
            }
        }
    }
    if (!s.occ) return false;
    const R fx = (R)(px - (double)bx);
    const R fy = (R)(py - (double)by);
    const R fz = (R)(pz - (double)bz);
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int dx = j & 1, dy = (j >> 1) & 1, dz = j >> 2;
        const R wx = dx ? fx : (R)1 - fx;
        const R wy = dy ? fy : (R)1 - fy;
        const R wz = dz ? fz : (R)1 - fz;
        s.w[j] = wx * wy * wz;
    }
    return true;
}

// One stored parameter as float: FP32, bfloat16 (upper half of an FP32)
// or IEEE half.
template <int STORE>
__device__ __forceinline__ float load_param(const float *base, size_t idx) {
    if (STORE == PXL_STORE_F32) return base[idx];
    const uint16_t h = reinterpret_cast<const uint16_t *>(base)[idx];
    if (STORE == PXL_STORE_BF16) return __uint_as_float((uint32_t)h << 16);
    return __half2float(__ushort_as_half(h));
}

// Interpolated raw density and raw colours of a located sample.
template <typename R, int STORE>
__device__ __forceinline__ void interpolate(const RenderArgs &a, const Sample<R> &s,
                                            int p, R ylane, R &sraw, R &c0, R &c1,
                                            R &c2) {
    R acc = 0;
    if (p >= 0) {
#pragma unroll
        for (int j = 0; j < 8; j++)
            acc = fma(s.w[j], (R)load_param<STORE>(a.params, (size_t)s.v[j] * (size_t)a.stride + p),
                      acc);
    }
    R val = acc * ylane;
    val = row_sum<R>(val);
    c0 = readlane<R>(val, 0);
    c1 = readlane<R>(val, 16);
    c2 = readlane<R>(val, 32);
    sraw = readlane<R>(val, 48);
}

__device__ __forceinline__ uint32_t fnv_mix(uint32_t h, uint32_t x) {
    for (int i = 0; i < 4; i++) {
        h ^= (x >> (8 * i)) & 0xFFu;
        h *= 16777619u;
    }
    return h;
}

template <int S, bool SPARSE, bool ADAPT, int MODE, int STORE>
__global__ __launch_bounds__(256) void k_render(const RenderArgs a) {
    using R = typename std::conditional<ADAPT, double, float>::type;
    const int lane = threadIdx.x & 63;
    const int64_t b =
        (int64_t)xcd_slice_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
    if (b >= a.len) return;
    // Second launch of the adaptive rule's logged step: only the rays the
    // log could not hold.
    if (MODE == kFused && a.log_redo && uniform_i(a.log_count[b]) != -1) return;
    const int ray = MODE == kImage ? (int)b : uniform_i(a.sample_index[a.start + b]);
    if (MODE != kImage && (ray < 0 || (int64_t)ray >= a.rays.n)) {
        // Out-of-range sample index: no outputs, zero loss.
        if ((MODE == kBwd || MODE == kFused) && lane == 0) a.losses[b] = 0.0;
        if (MODE == kTrace && lane == 0) {
            a.trace[2 * b] = 0u;
            a.trace[2 * b + 1] = 2166136261u;
        }
        return;
    }

    double ox, oy, oz, dx, dy, dz, tnear, tfar;
    float fdx, fdy, fdz;
    if (MODE == kImage) {
        // Pixel b: row b / w, column b % h (the reference's pixelCol,
        // RenderUtils.cu:1727; the same as b % w for square images).
        double d[3];
        eval_ray(a, (int)(b / a.img_w), (int)(b % a.img_h), d, tnear, tfar);
        ox = a.cam.centre[0];
        oy = a.cam.centre[1];
        oz = a.cam.centre[2];
        dx = d[0];
        dy = d[1];
        dz = d[2];
        fdx = (float)d[0];
        fdy = (float)d[1];
        fdz = (float)d[2];
    } else {
        ox = a.rays.ox[ray];
        oy = a.rays.oy[ray];
        oz = a.rays.oz[ray];
        fdx = a.rays.dx[ray];
        fdy = a.rays.dy[ray];
        fdz = a.rays.dz[ray];
        dx = fdx;
        dy = fdy;
        dz = fdz;
        tnear = a.rays.near_t[ray];
        tfar = a.rays.far_t[ray];
    }
    const double step = (double)a.step;
    const R rstep = (R)a.step;

    // Lane role: rows 0..2 = colour channel, lane k = SH coefficient k;
    // row 3 lane 0 = density.
    const int row = lane >> 4, kk = lane & 15;
    int p = -1;
    if (row < 3 && kk < S) p = 1 + row * S + kk;
    else if (row == 3 && kk == 0) p = 0;
    // Coefficients from sh_active on do not enter the colour (the band limit
    // of trilerpRGB, RenderUtils.cu:871).
    const R ylane = row < 3 ? (kk < a.sh_active ? (R)sh_basis_d(kk, fdx, fdy, fdz) : (R)0) : (R)1;

    const bool do_fwd =
        MODE == kFwd || MODE == kFused || MODE == kTrace || MODE == kTouch || MODE == kImage;
    const bool do_bwd = MODE == kBwd || MODE == kFused;

    // Forward totals (FP64).
    double C0 = 0.0, C1 = 0.0, C2 = 0.0, A = 0.0, sigsum = 0.0, disp = 0.0;

    if (MODE == kImage && tnear > tfar) {
        // Missed the box: white bytes, float colour 1 on a white background
        // (else 0), no mask or disparity (RenderUtils.cu:1753-1791).
        if (lane == 0) {
            const size_t n = (size_t)a.img_w * (size_t)a.img_h;
            for (int c = 0; c < 3; c++) {
                if (a.img_rgb) a.img_rgb[c * n + b] = a.white_bg ? 1.0f : 0.0f;
                if (a.img_rgb8) a.img_rgb8[c * n + b] = 255;
            }
            if (a.img_mask) a.img_mask[b] = 0;
            if (a.img_disp) a.img_disp[b] = 0.0f;
        }
        return;
    }
    if (tnear > tfar) {
        if (MODE == kFwd || MODE == kFused || MODE == kTrace) {
            if (lane == 0) {
                a.rays.acc_r[ray] = 0.0f;
                a.rays.acc_g[ray] = 0.0f;
                a.rays.acc_b[ray] = 0.0f;
                a.rays.acc_w[ray] = 0.0f;
                a.rays.acc_density[ray] = 0.0f;
            }
        }
        if (MODE == kTrace && lane == 0) {
            a.trace[2 * b] = 0u;
            a.trace[2 * b + 1] = 2166136261u;
        }
        if (do_bwd && lane == 0) a.losses[b] = 0.0;
        return;
    }

    if (do_fwd) {
        double t = tnear, T = 1.0;
        int k = 0;
        uint32_t hcount = 0, hash = 2166136261u;
        int last_bx = -1, last_by = -1, last_bz = -1;
        Sample<R> s;
        sample_reset(s);
        while (t < tfar && k < kMaxSteps) {
            const double px = fma(t, dx, ox), py = fma(t, dy, oy), pz = fma(t, dz, oz);
            if (!locate<R, SPARSE>(a, px, py, pz, s)) {
                if (ADAPT) t += step;
                ++k;
                if (!ADAPT) t = fma((double)k, step, tnear);
                continue;
            }
            if (MODE == kTrace) {
                hcount++;
                hash = fnv_mix(hash, (uint32_t)k);
                hash = fnv_mix(hash, (uint32_t)s.bx);
                hash = fnv_mix(hash, (uint32_t)s.by);
                hash = fnv_mix(hash, (uint32_t)s.bz);
            }
            if (MODE == kTouch) {
                if (s.bx != last_bx || s.by != last_by || s.bz != last_bz) {
                    if (lane < 8) {
                        uint32_t r8 = s.v[0];
#pragma unroll
                        for (int j = 1; j < 8; j++)
                            if (lane == j) r8 = s.v[j];
                        atomicOr(a.bitmap + (r8 >> 5), 1u << (r8 & 31));
                    }
                    last_bx = s.bx;
                    last_by = s.by;
                    last_bz = s.bz;
                }
                // The touched set does not depend on the marching weights
                // under the fixed rule; the adaptive rule needs them.
                if (!ADAPT) {
                    ++k;
                    t = fma((double)k, step, tnear);
                    continue;
                }
            }
            R sraw, cr0, cr1, cr2;
            interpolate<R, STORE>(a, s, p, ylane, sraw, cr0, cr1, cr2);
            const R sigma = softplus_shift<R>(sraw);
            const R col0 = sigmoid01<R>(cr0), col1 = sigmoid01<R>(cr1), col2 = sigmoid01<R>(cr2);
            const R alpha = one_minus_exp_neg<R>(sigma * rstep);
            const double w = T * (double)alpha;
            A += w;
            if (MODE == kImage) disp += w * t;
            C0 += w * (double)col0;
            C1 += w * (double)col1;
            C2 += w * (double)col2;
            sigsum += (double)sigma;
            T = T * (1.0 - (double)alpha);
            ++k;
            if (ADAPT) t += fabs(step - w);
            else t = fma((double)k, step, tnear);
        }
        if (MODE == kImage) {
            // Bytes always over white, floats over white when white_bg, the
            // mask A * 255 truncated (RenderUtils.cu:1565-1579).
            if (lane == 0) {
                const size_t n = (size_t)a.img_w * (size_t)a.img_h;
                const double cc[3] = {C0, C1, C2};
                for (int c = 0; c < 3; c++) {
                    if (a.img_rgb8)
                        a.img_rgb8[c * n + b] = (uint8_t)(int)((cc[c] + 1.0 - A) * 255.0 + 0.5);
                    if (a.img_rgb) a.img_rgb[c * n + b] = (float)(cc[c] + (a.white_bg ? 1.0 - A : 0.0));
                }
                if (a.img_mask) a.img_mask[b] = (uint8_t)(int)(A * 255.0);
                if (a.img_disp) a.img_disp[b] = (float)disp;
            }
            return;
        }
        if (a.white_bg) {
            C0 += 1.0 - A;
            C1 += 1.0 - A;
            C2 += 1.0 - A;
        }
        if (MODE == kFwd || MODE == kFused || MODE == kTrace) {
            if (lane == 0) {
                a.rays.acc_r[ray] = (float)C0;
                a.rays.acc_g[ray] = (float)C1;
                a.rays.acc_b[ray] = (float)C2;
                a.rays.acc_w[ray] = (float)A;
                a.rays.acc_density[ray] = (float)sigsum;
            }
        }
        if (MODE == kTrace && lane == 0) {
            a.trace[2 * b] = hcount;
            a.trace[2 * b + 1] = hash;
        }
    }

    if (!do_bwd) return;

    // Backward (renderRays_withGradients_RGBGPU, RenderUtils.cu:2149-2254).
    if (MODE == kBwd) {
        C0 = a.rays.acc_r[ray];
        C1 = a.rays.acc_g[ray];
        C2 = a.rays.acc_b[ray];
        A = a.rays.acc_w[ray];
    }
    double r0 = C0 - (double)a.rays.gt_r[ray] / 255.0;
    double r1 = C1 - (double)a.rays.gt_g[ray] / 255.0;
    double r2 = C2 - (double)a.rays.gt_b[ray] / 255.0;
    double loss = (fabs(r0) + fabs(r1) + fabs(r2)) * 255.0 / 3.0;
    if (isnan(loss)) {
        if (lane == 0) a.losses[b] = 255.0;
        return;
    }
    if (lane == 0) a.losses[b] = loss;
    const double hd = (double)a.huber;
    const double hi = hd - 0.5 * hd * hd, lo = 0.5 * hd * hd - hd;
    r0 = r0 > hd ? hi : (r0 < -hd ? lo : r0);
    r1 = r1 > hd ? hi : (r1 < -hd ? lo : r1);
    r2 = r2 > hd ? hi : (r2 < -hd ? lo : r2);

    // Mask and opacity-logit term (RenderUtils.cu:2188-2193).
    double mterm = 0.0;
    const int gm = a.rays.gt_mask[ray];
    if (gm != 0 && A < 1.0) mterm = ((double)gm / 255.0 - A) * a.mask_coef;
    double wpre = A > 1e-6 ? A : 1e-6;
    wpre = wpre < 1.0 - 1e-6 ? wpre : 1.0 - 1e-6;
    mterm += (log(wpre) - log(1.0 - wpre)) * 1e-3;

    double t = tnear, T = 1.0, Ai = 0.0, H0 = 0.0, H1 = 0.0, H2 = 0.0;
    int k = 0;
    // Lanes whose parameter takes a gradient: the density and the SH
    // coefficients of the optimised band (RenderUtils.cu:1117-1132).
    const bool active = p >= 0 && (row == 3 || (kk >= a.sh_first && kk < a.sh_active));
    float acc[8];
    double acc0 = 0.0;
    uint32_t cv[8];
    int cbx = -1, cby = -1, cbz = -1;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        acc[j] = 0.0f;
        cv[j] = 0u;
    }
    float *gbase = a.grad + p;
    const size_t stride = (size_t)a.gstride;

    Sample<R> s;
    sample_reset(s);
    while (t < tfar && k < kMaxSteps) {
        const double px = fma(t, dx, ox), py = fma(t, dy, oy), pz = fma(t, dz, oz);
        if (!locate<R, SPARSE>(a, px, py, pz, s)) {
            if (ADAPT) t += step;
            ++k;
            if (!ADAPT) t = fma((double)k, step, tnear);
            continue;
        }
        R sraw, cr0, cr1, cr2;
        interpolate<R, STORE>(a, s, p, ylane, sraw, cr0, cr1, cr2);
        const R sigma = softplus_shift<R>(sraw);
        const R col0 = sigmoid01<R>(cr0), col1 = sigmoid01<R>(cr1), col2 = sigmoid01<R>(cr2);
        const R alpha = one_minus_exp_neg<R>(sigma * rstep);
        const double w = T * (double)alpha;
        Ai += w;
        H0 += w * (double)col0;
        H1 += w * (double)col1;
        H2 += w * (double)col2;
        // Colour gradients: 0.6667 * w * r * sigmoid'(raw).
        // (0.6667f is a float literal in the reference, RenderUtils.cu:2220-2222.)
        const double kc = (double)0.6667f;
        const double g0 = kc * w * r0 * (double)((R)0.1 * col0 * ((R)1 - col0));
        const double g1 = kc * w * r1 * (double)((R)0.1 * col1 * ((R)1 - col1));
        const double g2 = kc * w * r2 * (double)((R)0.1 * col2 * ((R)1 - col2));
        const double Tn = T * (1.0 - (double)alpha);
        const double ds = (double)dsoftplus_shift<R>(sraw);
        double gs = r0 * (Tn * (double)col0 - (C0 - H0)) + r1 * (Tn * (double)col1 - (C1 - H1)) +
                    r2 * (Tn * (double)col2 - (C2 - H2));
        gs *= 0.6667 * step * ds;
        gs += step * ds * (A - Ai - Tn) * mterm;
        T = Tn;

        // This lane's coefficient of d(param)/d(corner value).
        const double gch = row == 0 ? g0 : (row == 1 ? g1 : g2);
        const float coef = (float)(row == 3 ? gs : gch * (double)ylane);

        if (s.bx != cbx || s.by != cby || s.bz != cbz) {
            // New cell.  Corners it shares with the previous cell (same
            // vertex coordinates) keep accumulating; the others are flushed
            // with one coalesced atomic per corner.  The shift (db) is
            // wave-uniform, so every branch here is uniform.
            const int dbx = s.bx - cbx, dby = s.by - cby, dbz = s.bz - cbz;
            float nacc[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int ix = (j & 1) + dbx, iy = ((j >> 1) & 1) + dby, iz = (j >> 2) + dbz;
                const bool ok = cbx >= 0 && (unsigned)ix < 2u && (unsigned)iy < 2u &&
                                (unsigned)iz < 2u;
                const int src = iz * 4 + iy * 2 + ix;
                float v = 0.0f;
#pragma unroll
                for (int i = 0; i < 8; i++) v = (ok && src == i) ? acc[i] : v;
                nacc[j] = v;
            }
            if (cbx >= 0) {
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const int jx = (i & 1) - dbx, jy = ((i >> 1) & 1) - dby, jz = (i >> 2) - dbz;
                    const bool kept = (unsigned)jx < 2u && (unsigned)jy < 2u && (unsigned)jz < 2u;
                    // Record 0 (also every corner outside the cage) is summed
                    // in acc0 and added once per ray.
                    if (!kept && cv[i] == 0u) acc0 += (double)acc[i];
                    else if (!kept && active)
                        unsafeAtomicAdd(gbase + (size_t)cv[i] * stride, acc[i]);
                }
            }
            cbx = s.bx;
            cby = s.by;
            cbz = s.bz;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                cv[j] = s.v[j];
                acc[j] = nacc[j];
            }
        }
#pragma unroll
        for (int j = 0; j < 8; j++) acc[j] += coef * (float)s.w[j];

        ++k;
        if (ADAPT) t += fabs(step - w);
        else t = fma((double)k, step, tnear);
    }
    if (cbx >= 0) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (cv[j] == 0u) acc0 += (double)acc[j];
            else if (active) unsafeAtomicAdd(gbase + (size_t)cv[j] * stride, acc[j]);
        }
    }
    // Record 0: into the entry's FP64 slot of the logged step, else one
    // float atomic per ray.
    if (active && acc0 != 0.0) {
        if (a.rec0 != nullptr) unsafeAtomicAdd(a.rec0 + (size_t)p * (size_t)a.len + (size_t)b, acc0);
        else unsafeAtomicAdd(gbase, (float)acc0);
    }
}

#include "march8.hpp"
#include "adapt8.hpp"
#include "binned.hpp"

// Marcher selection: the eight-samples-per-wave kernel serves the fixed rule
// for S <= 9 in the forward, backward, fused and trace modes; the one-sample
// kernel serves the adaptive rule, S = 16 and the touched-record count.
// The PXL_OPT_MARCHER option forces the one-sample kernel (both are tested).
bool use_march8() { return option_group_marcher(); }

template <int S, bool SPARSE, bool ADAPT, int MODE, int STORE>
int32_t launch3(const RenderArgs &a, hipStream_t st) {
    const unsigned blocks = (unsigned)((a.len + 3) / 4);
    // The adaptive rule's forward: 64 / kAdaptLanes rays per wave, four
    // waves per block.
    constexpr bool kAdapt8 = ADAPT && S <= 9 && (MODE == kFwd || MODE == kTrace);
    if (kAdapt8 && use_march8()) {
        constexpr int64_t per_block = 4 * (64 / kAdaptLanes);
        hipLaunchKernelGGL((k_adapt8<(S <= 9 ? S : 9), SPARSE, MODE>),
                           dim3((unsigned)((a.len + per_block - 1) / per_block)), dim3(256), 0,
                           st, a);
        return PXL_OK;
    }
    constexpr bool kGroup = !ADAPT && S <= 9 && MODE != kTouch;
    if (kGroup && use_march8()) {
        hipLaunchKernelGGL((k_march8<(S <= 9 ? S : 9), SPARSE, (MODE == kTouch ? kFwd : MODE),
                                     STORE>),
                           dim3(blocks), dim3(256), 0, st, a);
        return PXL_OK;
    }
    hipLaunchKernelGGL((k_render<S, SPARSE, ADAPT, MODE, STORE>), dim3(blocks), dim3(256), 0, st,
                       a);
    return PXL_OK;
}

template <int S, bool SPARSE, int MODE>
int32_t launch_store(int store, bool adapt, const RenderArgs &a, hipStream_t st) {
    if (adapt) {
        if (store != PXL_STORE_F32) {
            set_error("render: the adaptive marching rule needs FP32 storage");
            return PXL_EINVAL;
        }
        return launch3<S, SPARSE, true, MODE, PXL_STORE_F32>(a, st);
    }
    switch (store) {
        case PXL_STORE_F32: return launch3<S, SPARSE, false, MODE, PXL_STORE_F32>(a, st);
        case PXL_STORE_BF16: return launch3<S, SPARSE, false, MODE, PXL_STORE_BF16>(a, st);
        case PXL_STORE_F16: return launch3<S, SPARSE, false, MODE, PXL_STORE_F16>(a, st);
        default:
            set_error("render: unknown storage %d", store);
            return PXL_EINVAL;
    }
}

template <int MODE>
int32_t dispatch(int S, bool sparse, bool adapt, int store, const RenderArgs &a,
                 hipStream_t st) {
#define PXL_CASE(SV)                                                              \
    case SV:                                                                      \
        return sparse ? launch_store<SV, true, MODE>(store, adapt, a, st)         \
                      : launch_store<SV, false, MODE>(store, adapt, a, st);
    switch (S) {
        PXL_CASE(1)
        PXL_CASE(4)
        PXL_CASE(9)
        PXL_CASE(16)
        default:
            set_error("render: unsupported sh_dim %d (1, 4, 9 or 16)", S);
            return PXL_EINVAL;
    }
#undef PXL_CASE
}

int32_t check_grid(const pxl_grid *g, const char *who) {
    PXL_REQUIRE(g != nullptr, "%s: null grid", who);
    PXL_REQUIRE(g->W >= 2 && g->H >= 2 && g->D >= 2, "%s: grid %dx%dx%d too small", who, g->W,
                g->H, g->D);
    PXL_REQUIRE((int64_t)g->W * g->H * g->D < (int64_t)1 << 32,
                "%s: grid of %lld vertices exceeds 32-bit indexing", who,
                (long long)g->W * g->H * g->D);
    PXL_REQUIRE(g->sh_dim == 1 || g->sh_dim == 4 || g->sh_dim == 9 || g->sh_dim == 16,
                "%s: unsupported sh_dim %d (1, 4, 9 or 16)", who, g->sh_dim);
    PXL_REQUIRE(g->storage == PXL_STORE_F32 || g->storage == PXL_STORE_BF16 ||
                    g->storage == PXL_STORE_F16,
                "%s: unknown storage %d", who, g->storage);
    const int align = g->storage == PXL_STORE_F32 ? 4 : 8;
    PXL_REQUIRE(g->stride >= 1 + 3 * g->sh_dim && g->stride % align == 0,
                "%s: stride %d must be a multiple of %d and >= %d", who, g->stride, align,
                1 + 3 * g->sh_dim);
    PXL_REQUIRE(g->params != nullptr && g->n_vox > 0, "%s: empty grid", who);
    PXL_REQUIRE(((uintptr_t)g->params & 15) == 0, "%s: params must be 16-byte aligned", who);
    PXL_REQUIRE(g->sparse || g->n_vox == (int64_t)g->W * g->H * g->D,
                "%s: dense grid needs n_vox = W*H*D", who);
    PXL_REQUIRE(!g->sparse || (g->occupancy && g->index_offsets),
                "%s: sparse grid needs occupancy and index_offsets", who);
    PXL_REQUIRE(g->step > 0.0f, "%s: step must be positive", who);
    PXL_REQUIRE(g->march_mode == PXL_MARCH_FIXED || g->march_mode == PXL_MARCH_REF_ADAPTIVE,
                "%s: bad march_mode %d", who, g->march_mode);
    PXL_REQUIRE(g->sh_active >= 0 && g->sh_active <= g->sh_dim,
                "%s: sh_active %d outside [0, %d]", who, g->sh_active, g->sh_dim);
    return PXL_OK;
}

int32_t prepare(const pxl_grid *g, const pxl_rays *r, const int32_t *sample_index,
                int64_t start, int64_t len, const char *who, RenderArgs &a) {
    int32_t st = check_grid(g, who);
    if (st != PXL_OK) return st;
    PXL_REQUIRE(r != nullptr && (sample_index != nullptr || len == 0),
                "%s: null rays or sample_index", who);
    PXL_REQUIRE(start >= 0 && len >= 0, "%s: negative start/len", who);
    PXL_REQUIRE(r->ox && r->oy && r->oz && r->dx && r->dy && r->dz && r->near_t && r->far_t,
                "%s: ray geometry planes must be non-null", who);
    memset(&a, 0, sizeof(a));
    a.params = g->params;
    a.n_vox = g->n_vox;
    a.occ = g->occupancy;
    a.bricks = g->sparse ? g->brick_occupancy : nullptr;
    a.nbx = (g->W - 1 + 7) / 8;
    a.nby = (g->H - 1 + 7) / 8;
    a.dbricks = dilated_bricks(a.bricks, g);
    a.io = g->index_offsets;
    a.W = g->W;
    a.H = g->H;
    a.D = g->D;
    a.stride = g->stride;
    a.gstride = g->stride;
    a.sh_active = g->sh_active > 0 ? g->sh_active : g->sh_dim;
    a.sh_first = 0;
    a.step = g->step;
    a.white_bg = r->white_bg;
    a.rays = *r;
    a.sample_index = sample_index;
    a.start = start;
    a.len = len;
    return PXL_OK;
}

}  // namespace
}  // namespace pxl

using namespace pxl;

extern "C" int32_t pxl_render_fwd(const pxl_grid *grid, const pxl_rays *rays,
                                  const int32_t *sample_index, int64_t start, int64_t len,
                                  void *stream) {
    clear_error();
    RenderArgs a;
    int32_t st = prepare(grid, rays, sample_index, start, len, "pxl_render_fwd", a);
    if (st != PXL_OK) return st;
    PXL_REQUIRE(rays->acc_r && rays->acc_g && rays->acc_b && rays->acc_w && rays->acc_density,
                "pxl_render_fwd: output planes must be non-null");
    if (len == 0) return PXL_OK;
    st = dispatch<kFwd>(grid->sh_dim, grid->sparse, grid->march_mode == PXL_MARCH_REF_ADAPTIVE,
                        grid->storage, a, as_stream(stream));
    return st != PXL_OK ? st : check_launch("pxl_render_fwd");
}

extern "C" int32_t pxl_render_fwd_masked(const pxl_grid *grid, const pxl_rays *rays,
                                         const int32_t *sample_index, int64_t start, int64_t len,
                                         const uint64_t *bin_mask, uint32_t *trace,
                                         void *stream) {
    clear_error();
    RenderArgs a;
    int32_t st = prepare(grid, rays, sample_index, start, len, "pxl_render_fwd_masked", a);
    if (st != PXL_OK) return st;
    PXL_REQUIRE(rays->acc_r && rays->acc_g && rays->acc_b && rays->acc_w && rays->acc_density,
                "pxl_render_fwd_masked: output planes must be non-null");
    PXL_REQUIRE(grid->march_mode == PXL_MARCH_FIXED && grid->sh_dim <= 9 && use_march8(),
                "pxl_render_fwd_masked: needs the eight-sample marcher (fixed rule, sh_dim <= 9)");
    if (len == 0) return PXL_OK;
    a.bin_mask = bin_mask;
    a.trace = trace;
    st = trace ? dispatch<kTrace>(grid->sh_dim, grid->sparse, false, grid->storage, a,
                                  as_stream(stream))
               : dispatch<kFwd>(grid->sh_dim, grid->sparse, false, grid->storage, a,
                                as_stream(stream));
    return st != PXL_OK ? st : check_launch("pxl_render_fwd_masked");
}

extern "C" int32_t pxl_render_image(const pxl_grid *grid, const pxl_camera *cam, int32_t width,
                                    int32_t height, int32_t white_bg, float *rgb, uint8_t *rgb8,
                                    uint8_t *mask, float *disparity, void *stream) {
    clear_error();
    int32_t st = check_grid(grid, "pxl_render_image");
    if (st != PXL_OK) return st;
    PXL_REQUIRE(cam != nullptr && width > 0 && height > 0, "pxl_render_image: bad camera / size");
    PXL_REQUIRE(grid->storage == PXL_STORE_F32,
                "pxl_render_image: the adaptive marching rule needs FP32 storage");
    RenderArgs a;
    memset(&a, 0, sizeof(a));
    a.params = grid->params;
    a.n_vox = grid->n_vox;
    a.occ = grid->occupancy;
    a.bricks = grid->sparse ? grid->brick_occupancy : nullptr;
    a.nbx = (grid->W - 1 + 7) / 8;
    a.nby = (grid->H - 1 + 7) / 8;
    a.dbricks = dilated_bricks(a.bricks, grid);
    a.io = grid->index_offsets;
    a.W = grid->W;
    a.H = grid->H;
    a.D = grid->D;
    a.stride = grid->stride;
    a.gstride = grid->stride;
    a.sh_active = grid->sh_active > 0 ? grid->sh_active : grid->sh_dim;
    a.step = grid->step;
    a.white_bg = white_bg;
    a.len = (int64_t)width * (int64_t)height;
    a.cam = *cam;
    a.box = grid->bbox;
    a.img_w = width;
    a.img_h = height;
    a.img_rgb = rgb;
    a.img_rgb8 = rgb8;
    a.img_mask = mask;
    a.img_disp = disparity;
    const unsigned blocks = (unsigned)((a.len + 3) / 4);
    const hipStream_t s = as_stream(stream);
#define PXL_IMG(SV)                                                                           \
    case SV:                                                                                  \
        if (grid->sparse)                                                                     \
            hipLaunchKernelGGL((k_render<SV, true, true, kImage, PXL_STORE_F32>), dim3(blocks),  \
                               dim3(256), 0, s, a);                                           \
        else                                                                                  \
            hipLaunchKernelGGL((k_render<SV, false, true, kImage, PXL_STORE_F32>), dim3(blocks), \
                               dim3(256), 0, s, a);                                           \
        break;
    switch (grid->sh_dim) {
        PXL_IMG(1)
        PXL_IMG(4)
        PXL_IMG(9)
        PXL_IMG(16)
        default:
            set_error("pxl_render_image: unsupported sh_dim %d", grid->sh_dim);
            return PXL_EINVAL;
    }
#undef PXL_IMG
    return check_launch("pxl_render_image");
}

extern "C" int32_t pxl_render_fwd_trace(const pxl_grid *grid, const pxl_rays *rays,
                                        const int32_t *sample_index, int64_t start,
                                        int64_t len, uint32_t *trace, void *stream) {
    clear_error();
    RenderArgs a;
    int32_t st = prepare(grid, rays, sample_index, start, len, "pxl_render_fwd_trace", a);
    if (st != PXL_OK) return st;
    PXL_REQUIRE(trace != nullptr, "pxl_render_fwd_trace: null trace");
    PXL_REQUIRE(rays->acc_r && rays->acc_g && rays->acc_b && rays->acc_w && rays->acc_density,
                "pxl_render_fwd_trace: output planes must be non-null");
    if (len == 0) return PXL_OK;
    a.trace = trace;
    st = dispatch<kTrace>(grid->sh_dim, grid->sparse,
                          grid->march_mode == PXL_MARCH_REF_ADAPTIVE, grid->storage, a,
                          as_stream(stream));
    return st != PXL_OK ? st : check_launch("pxl_render_fwd_trace");
}

namespace pxl {
namespace {
// Workspace of the two-kernel fused step: per batch entry 4 FP64 totals, a
// count and 32 FP64 record-0 slots, per logged sample 8 records and 16 value
// words, per 64-sample chunk of an entry's log 5 FP64 prefixes, and the
// record-0 fold's 32 x kFoldParts partial sums; 256-byte aligned.
struct LogLayout {
    size_t work, tot, count, rec, val, pre, redo, rec0, part, total;
    // Brick-binned gradient (only with nkeys > 0): counts and fill pointers
    // per key, scan block sums, target bricks, per-entry SH basis.
    size_t gb_count, gb_fill, gb_blk, gb_bricks, gb_y;
};
// Slices of the batch the record-0 fold's first pass sums separately.
constexpr int kFoldParts = 64;
inline size_t align256(size_t x) { return (x + 255) / 256 * 256; }
LogLayout log_layout(int64_t len, int32_t cap, int64_t nkeys = 0) {
    LogLayout L;
    size_t o = 0;
    L.work = o;  // 16 claim counters of the persistent launches
    o += 256;
    L.tot = o;
    o += align256((size_t)len * 4 * sizeof(double));
    L.count = o;
    o += align256((size_t)len * sizeof(int32_t));
    L.rec = o;
    o += align256((size_t)len * (size_t)cap * 8 * sizeof(uint32_t));
    L.val = o;  // 8 words per sample (fixed rule) or 8 FP64 words (adaptive rule)
    o += align256((size_t)len * (size_t)cap * 16 * sizeof(float));
    L.pre = o;
    o += align256((size_t)len * (size_t)((cap + 63) / 64) * 5 * sizeof(double));
    L.redo = o;
    o += align256((size_t)len * sizeof(int32_t));
    L.rec0 = o;
    o += align256((size_t)len * 32 * sizeof(double));
    L.part = o;
    o += align256((size_t)32 * kFoldParts * sizeof(double));
    L.gb_count = L.gb_fill = L.gb_blk = L.gb_bricks = L.gb_y = 0;
    if (nkeys > 0) {
        const int64_t nblk = (nkeys + kBinScanBlockKeys - 1) / kBinScanBlockKeys;
        L.gb_count = o;
        o += align256((size_t)nkeys * sizeof(uint32_t));
        L.gb_fill = o;
        o += align256((size_t)nkeys * sizeof(uint32_t));
        L.gb_blk = o;
        o += align256((size_t)nblk * sizeof(uint32_t));
        L.gb_bricks = o;
        o += align256((size_t)(nkeys / 8) * sizeof(uint32_t));
        L.gb_y = o;
        o += align256((size_t)len * kBinYStride * sizeof(float));
    }
    L.total = o;
    return L;
}
// Bin keys of the brick-binned gradient for a grid: 8 per brick of 8^3
// vertices (0 for grids it does not serve: each side 2 .. 1025 vertices,
// so a cell coordinate fits in 10 bits).
int64_t bin_keys(const pxl_grid *g) {
    if (!g || g->W < 2 || g->H < 2 || g->D < 2 || g->W > 1025 || g->H > 1025 || g->D > 1025)
        return 0;
    return (int64_t)((g->W + 7) / 8) * ((g->H + 7) / 8) * ((g->D + 7) / 8) * 8;
}
constexpr int32_t kMaxLogSamples = 1 << 16;

__global__ void k_zero_u64(unsigned long long *p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// Record 0 of the logged step, summed in a fixed order in two passes:
// block (c, p) of k_rec0_part sums parameter p's FP64 slots of the c-th of
// kFoldParts equal slices of the batch (thread t: entries t, t + 256, ... of
// the slice, then a fixed tree over the threads); k_rec0_fold (one wave per
// parameter) sums the slices with a fixed butterfly and adds the total to
// the gradient.  Every corner outside the cage of a caged grid lands on
// record 0 (SceneUtils.cu:1106-1113), so this sum has ~10^5 largely
// cancelling terms.
__global__ __launch_bounds__(256) void k_rec0_part(const double *__restrict__ rec0, int64_t len,
                                                   double *__restrict__ part) {
    __shared__ double red[256];
    const int c = blockIdx.x, p = blockIdx.y, t = threadIdx.x;
    const int64_t per = (len + kFoldParts - 1) / kFoldParts;
    const int64_t lo = (int64_t)c * per, hi = lo + per < len ? lo + per : len;
    const double *src = rec0 + (size_t)p * (size_t)len;
    double s = 0.0;
    for (int64_t i = lo + t; i < hi; i += 256) s += src[i];
    red[t] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    if (t == 0) part[p * kFoldParts + c] = red[0];
}

__global__ __launch_bounds__(64) void k_rec0_fold(const double *__restrict__ part, float *grad) {
    static_assert(kFoldParts == 64, "one slice per lane");
    const int p = blockIdx.x;
    const double v = wave_sum_d(part[p * kFoldParts + threadIdx.x]);
    if (threadIdx.x == 0 && v != 0.0) grad[p] += (float)v;
}

// The two passes over the P parameters' slots.
inline void launch_rec0_fold(const RenderArgs &a, int P, hipStream_t st) {
    hipLaunchKernelGGL(k_rec0_part, dim3(kFoldParts, (unsigned)P), dim3(256), 0, st, a.rec0,
                       a.len, a.rec0_part);
    hipLaunchKernelGGL(k_rec0_fold, dim3((unsigned)P), dim3(64), 0, st, a.rec0_part, a.grad);
}

// Blocks of `threads` threads of `kernel` that fit on the whole device at
// once (the grid of a persistent launch), or PXL_OPT_CLAIM_GRID if smaller.
template <typename K>
unsigned resident_blocks(K kernel, int threads, size_t lds = 0) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds) != hipSuccess ||
        cus <= 0 || per_cu <= 0)
        cus = 1024, per_cu = 1;
    const unsigned resident = (unsigned)cus * (unsigned)per_cu;
    const int64_t cap = option_claim_grid();
    return cap > 0 && cap < (int64_t)resident ? (unsigned)cap : resident;
}

// The brick-binned gradient after the forward + log kernel (and after
// grad_wait): per-sample scalars and bin counts, the scan and the target
// bricks, the scatter into key order, the LDS accumulation per brick.
// gwpb: waves per workgroup of the per-ray kernels.  The bin counts are
// cleared here; the target counter with the claim counters.
template <int S, bool SPARSE, bool TGT, bool ADAPT>
void launch_binned(const RenderArgs &a, int gwpb, hipStream_t st) {
    const int64_t nkeys = (int64_t)a.vbx * a.vby * a.vbz * 8;
    const unsigned nblk = (unsigned)((nkeys + kBinScanBlockKeys - 1) / kBinScanBlockKeys);
    const int gthreads = 64 * gwpb;
    const unsigned gblocks = (unsigned)((a.len + gwpb - 1) / gwpb);
    hipLaunchKernelGGL(k_bin_zero, dim3((unsigned)((nkeys + 255) / 256)), dim3(256), 0, st,
                       a.gb_count, nkeys);
    hipLaunchKernelGGL((k_bin_samples<S, TGT, ADAPT>), dim3(gblocks), dim3(gthreads), 0, st, a);
    hipLaunchKernelGGL(k_bin_scan_sum, dim3(nblk), dim3(kBinScanThreads), 0, st, a.gb_count,
                       nkeys, a.gb_blk);
    hipLaunchKernelGGL(k_bin_scan_blocks, dim3(1), dim3(kBinScanThreads), 0, st, a.gb_blk,
                       (int64_t)nblk);
    hipLaunchKernelGGL(k_bin_scan_final, dim3(nblk), dim3(kBinScanThreads), 0, st, a, nkeys);
    hipLaunchKernelGGL(k_bin_scatter, dim3(gblocks), dim3(gthreads), 0, st, a);
    constexpr size_t lds = bin_accum_lds<S>();
    static std::once_flag attr_once;
    std::call_once(attr_once, [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bin_accum<S, SPARSE>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    });
    const unsigned nb = (unsigned)(nkeys / 8);
    const unsigned ga = std::min(nb, resident_blocks(k_bin_accum<S, SPARSE>, 256, lds));
    hipLaunchKernelGGL((k_bin_accum<S, SPARSE>), dim3(ga), dim3(256), lds, st, a);
}

// The logged step's kernels in two halves.
//
// Forward half: the claim / redo / target counters cleared, then the
// forward + log kernel (wpb waves per block; persistent launches size the
// grid to what is resident and let the waves claim batch entries).
template <int S, bool SPARSE, int STORE>
void fwd_logged(const RenderArgs &a, int wpb, bool persist, hipStream_t st) {
    const int threads = 64 * wpb;
    const unsigned blocks = (unsigned)((a.len + wpb - 1) / wpb);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, a.work, kTargetCounter + 1);
    if (persist) {
        const unsigned ga =
            std::min(blocks, resident_blocks(k_march8<S, SPARSE, kLog, STORE, true>, threads));
        hipLaunchKernelGGL((k_march8<S, SPARSE, kLog, STORE, true>), dim3(ga), dim3(threads), 0,
                           st, a);
    } else {
        hipLaunchKernelGGL((k_march8<S, SPARSE, kLog, STORE>), dim3(blocks), dim3(threads), 0, st,
                           a);
    }
}
// Gradient half: the gradient of the logged rays under the schedule's
// gradient mode, the fused marcher again for the rays longer than the log
// (the list the forward appended to), the record-0 fold.
template <int S, bool SPARSE, int STORE>
void grad_logged(const RenderArgs &a, int wpb, int gwpb, bool persist, int gmode,
                 hipStream_t st) {
    const int threads = 64 * wpb;
    const unsigned blocks = (unsigned)((a.len + wpb - 1) / wpb);
    const int gthreads = 64 * gwpb;
    const size_t glds = (size_t)gwpb * kWalkLdsBytes;  // the gradient walk's rows
    const unsigned gblocks = (unsigned)((a.len + gwpb - 1) / gwpb);
    const bool tgt = a.bin_targets != nullptr;
    if (gmode == PXL_GRAD_BINNED) {
        if (tgt) launch_binned<S, SPARSE, true, false>(a, gwpb, st);
        else launch_binned<S, SPARSE, false, false>(a, gwpb, st);
    } else if (persist) {
        if (tgt) {
            const unsigned gb =
                std::min(gblocks, resident_blocks(k_log_grad<S, true, true>, gthreads, glds));
            hipLaunchKernelGGL((k_log_grad<S, true, true>), dim3(gb), dim3(gthreads), glds, st, a);
        } else {
            const unsigned gb =
                std::min(gblocks, resident_blocks(k_log_grad<S, true>, gthreads, glds));
            hipLaunchKernelGGL((k_log_grad<S, true>), dim3(gb), dim3(gthreads), glds, st, a);
        }
    } else if (gmode == PXL_GRAD_CHUNKED && !tgt) {
        // Chunk-parallel gradient (the proposal's target words need the
        // whole ray in one wave, so learning batches keep the per-ray kernel).
        hipLaunchKernelGGL((k_log_prefix<S>), dim3(gblocks), dim3(gthreads), 0, st, a);
        const int64_t items = a.len * (int64_t)a.log_chunks;
        hipLaunchKernelGGL((k_log_grad_chunked<S>), dim3((unsigned)((items + gwpb - 1) / gwpb)),
                           dim3(gthreads), glds, st, a, items);
    } else if (tgt) {
        hipLaunchKernelGGL((k_log_grad<S, false, true>), dim3(gblocks), dim3(gthreads), glds, st,
                           a);
    } else {
        hipLaunchKernelGGL((k_log_grad<S>), dim3(gblocks), dim3(gthreads), glds, st, a);
    }
    RenderArgs redo = a;
    redo.log_redo = 1;
    hipLaunchKernelGGL((k_march8_redo<S, SPARSE, STORE>), dim3(std::min(blocks, 1024u)),
                       dim3(threads), 0, st, redo);
    launch_rec0_fold(a, 1 + 3 * S, st);
}
// The adaptive rule's forward half: k_adapt8 (forward + log, 64 /
// kAdaptLanes rays per wave, wpb waves per block, slots claiming batch
// entries from a resident grid).
template <int S, bool SPARSE>
void fwd_logged_adapt(const RenderArgs &a, int wpb, hipStream_t st) {
    const int threads = 64 * wpb;
    const int64_t per_block = (int64_t)wpb * (64 / kAdaptLanes);
    const unsigned blocks =
        std::min((unsigned)((a.len + per_block - 1) / per_block),
                 resident_blocks(k_adapt8<S, SPARSE, kLog>, threads));
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, a.work, kTargetCounter + 1);
    hipLaunchKernelGGL((k_adapt8<S, SPARSE, kLog>), dim3(blocks), dim3(threads), 0, st, a);
}
// The adaptive rule's gradient half: the gradient kernels over the FP64
// log (a persistent schedule applies to the gradient kernel only), the
// one-sample fused kernel for the rays longer than the log, the record-0
// fold.
template <int S, bool SPARSE>
void grad_logged_adapt(const RenderArgs &a, int gwpb, bool persist, int gmode, hipStream_t st) {
    const int gthreads = 64 * gwpb;
    const size_t glds = (size_t)gwpb * kWalkLdsBytes;
    const unsigned gblocks = (unsigned)((a.len + gwpb - 1) / gwpb);
    if (gmode == PXL_GRAD_BINNED) {
        launch_binned<S, SPARSE, false, true>(a, gwpb, st);
    } else if (persist) {
        const unsigned gb =
            std::min(gblocks, resident_blocks(k_log_grad<S, true, false, true>, gthreads, glds));
        hipLaunchKernelGGL((k_log_grad<S, true, false, true>), dim3(gb), dim3(gthreads), glds, st,
                           a);
    } else if (gmode == PXL_GRAD_CHUNKED) {
        hipLaunchKernelGGL((k_log_prefix<S, true>), dim3(gblocks), dim3(gthreads), 0, st, a);
        const int64_t items = a.len * (int64_t)a.log_chunks;
        hipLaunchKernelGGL((k_log_grad_chunked<S, true>),
                           dim3((unsigned)((items + gwpb - 1) / gwpb)), dim3(gthreads), glds, st, a,
                           items);
    } else {
        hipLaunchKernelGGL((k_log_grad<S, false, false, true>), dim3(gblocks), dim3(gthreads),
                           glds, st, a);
    }
    RenderArgs redo = a;
    redo.log_redo = 1;
    hipLaunchKernelGGL((k_render<S, SPARSE, true, kFused, PXL_STORE_F32>),
                       dim3((unsigned)((a.len + 3) / 4)), dim3(256), 0, st, redo);
    launch_rec0_fold(a, 1 + 3 * S, st);
}

// The workspace pointers of a batch's sample log (layout L at w).
void set_log_args(RenderArgs &r, char *w, const LogLayout &L, int64_t nkeys) {
    r.work = reinterpret_cast<unsigned long long *>(w + L.work);
    r.log_tot = reinterpret_cast<double *>(w + L.tot);
    r.log_count = reinterpret_cast<int32_t *>(w + L.count);
    r.log_rec = reinterpret_cast<uint32_t *>(w + L.rec);
    r.log_val = reinterpret_cast<float *>(w + L.val);
    r.log_pre = reinterpret_cast<double *>(w + L.pre);
    r.redo_list = reinterpret_cast<int32_t *>(w + L.redo);
    r.rec0 = reinterpret_cast<double *>(w + L.rec0);
    r.rec0_part = reinterpret_cast<double *>(w + L.part);
    if (nkeys > 0) {
        r.gb_count = reinterpret_cast<uint32_t *>(w + L.gb_count);
        r.gb_fill = reinterpret_cast<uint32_t *>(w + L.gb_fill);
        r.gb_blk = reinterpret_cast<uint32_t *>(w + L.gb_blk);
        r.gb_bricks = reinterpret_cast<uint32_t *>(w + L.gb_bricks);
        r.gb_y = reinterpret_cast<float *>(w + L.gb_y);
        // The log rows in key order take the value words the per-sample
        // pass has read (one u32 per logged sample).
        r.gb_list = reinterpret_cast<uint32_t *>(w + L.val);
    }
}

// Forward, then (after grad_wait) the gradient half, on one stream.
template <typename FwdF, typename GradF>
int32_t run_logged(const RenderArgs &a, hipEvent_t grad_wait, hipStream_t st, FwdF fwd,
                   GradF grad) {
    fwd(a, st);
    if (grad_wait && hipStreamWaitEvent(st, grad_wait, 0) != hipSuccess) {
        set_error("render: hipStreamWaitEvent before the gradient kernels failed");
        return PXL_EHIP;
    }
    grad(a, st);
    return PXL_OK;
}

template <int S, bool SPARSE, int STORE>
int32_t launch_logged(const RenderArgs &a, int wpb, int gwpb, bool persist, int gmode,
                      hipEvent_t grad_wait, hipStream_t st) {
    return run_logged(
        a, grad_wait, st,
        [&](const RenderArgs &ai, hipStream_t s) { fwd_logged<S, SPARSE, STORE>(ai, wpb, persist, s); },
        [&](const RenderArgs &ai, hipStream_t s) {
            grad_logged<S, SPARSE, STORE>(ai, wpb, gwpb, persist, gmode, s);
        });
}
template <int S, bool SPARSE>
int32_t launch_logged_adapt(const RenderArgs &a, int wpb, int gwpb, bool persist, int gmode,
                            hipEvent_t grad_wait, hipStream_t st) {
    return run_logged(
        a, grad_wait, st,
        [&](const RenderArgs &ai, hipStream_t s) { fwd_logged_adapt<S, SPARSE>(ai, wpb, s); },
        [&](const RenderArgs &ai, hipStream_t s) {
            grad_logged_adapt<S, SPARSE>(ai, gwpb, persist, gmode, s);
        });
}

template <int S>
int32_t dispatch_logged_s(bool sparse, int store, const RenderArgs &a, int wpb, int gwpb,
                          bool persist, int gmode, hipEvent_t grad_wait, hipStream_t st) {
#define PXL_LOGGED(SP, ST)                                                        \
    if (sparse == SP && store == ST)                                              \
        return launch_logged<S, SP, ST>(a, wpb, gwpb, persist, gmode, grad_wait, st);
    PXL_LOGGED(true, PXL_STORE_F32)
    PXL_LOGGED(true, PXL_STORE_BF16)
    PXL_LOGGED(true, PXL_STORE_F16)
    PXL_LOGGED(false, PXL_STORE_F32)
    PXL_LOGGED(false, PXL_STORE_BF16)
    PXL_LOGGED(false, PXL_STORE_F16)
#undef PXL_LOGGED
    set_error("render: unknown storage %d", store);
    return PXL_EINVAL;
}
int32_t dispatch_logged(int S, bool sparse, int store, const RenderArgs &a, int wpb, int gwpb,
                        bool persist, int gmode, hipEvent_t gw, hipStream_t st,
                        bool adapt = false) {
    if (adapt) {
        if (store != PXL_STORE_F32) {
            set_error("render: the adaptive marching rule needs FP32 storage");
            return PXL_EINVAL;
        }
#define PXL_ADAPT(SV)                                                                         \
    case SV:                                                                                  \
        return sparse ? launch_logged_adapt<SV, true>(a, wpb, gwpb, persist, gmode, gw, st) \
                      : launch_logged_adapt<SV, false>(a, wpb, gwpb, persist, gmode, gw, st);
        switch (S) {
            PXL_ADAPT(1)
            PXL_ADAPT(4)
            PXL_ADAPT(9)
            default:
                set_error("render: the logged step supports sh_dim 1, 4 and 9");
                return PXL_EINVAL;
        }
#undef PXL_ADAPT
    }
    switch (S) {
        case 1: return dispatch_logged_s<1>(sparse, store, a, wpb, gwpb, persist, gmode, gw, st);
        case 4: return dispatch_logged_s<4>(sparse, store, a, wpb, gwpb, persist, gmode, gw, st);
        case 9: return dispatch_logged_s<9>(sparse, store, a, wpb, gwpb, persist, gmode, gw, st);
        default:
            set_error("render: the logged fused step supports sh_dim 1, 4 and 9");
            return PXL_EINVAL;
    }
}
}  // namespace
}  // namespace pxl

namespace pxl {
// Grids whose train step runs as the two-kernel logged step (forward + sample
// log, then the gradient kernels): the fixed rule with the eight-sample
// marcher, sh_dim 1, 4 or 9.
bool logged_step_serves(const pxl_grid *g) {
    if (!g || g->sh_dim > 9 || !use_march8()) return false;
    return g->march_mode == PXL_MARCH_FIXED ||
           (g->march_mode == PXL_MARCH_REF_ADAPTIVE && g->storage == PXL_STORE_F32);
}
}  // namespace pxl

static int32_t bwd_common(const pxl_grid *grid, const pxl_optim *optim, const pxl_rays *rays,
                          double *losses, const int32_t *sample_index, int64_t start,
                          int64_t len, double mask_coef, float huber_delta, void *stream,
                          bool fused, const TrainExtra &x = TrainExtra()) {
    const char *who = x.workspace ? "pxl_train_rays_ws" : fused ? "pxl_train_rays" : "pxl_render_bwd_rgb";
    clear_error();
    RenderArgs a;
    int32_t st = prepare(grid, rays, sample_index, start, len, who, a);
    if (st != PXL_OK) return st;
    PXL_REQUIRE(optim && optim->grad, "%s: null gradient", who);
    PXL_REQUIRE(losses != nullptr, "%s: null losses", who);
    PXL_REQUIRE(rays->gt_r && rays->gt_g && rays->gt_b && rays->gt_mask,
                "%s: ground-truth planes must be non-null", who);
    PXL_REQUIRE(rays->acc_r && rays->acc_g && rays->acc_b && rays->acc_w && rays->acc_density,
                "%s: forward planes must be non-null", who);
    PXL_REQUIRE(optim->sh_first >= 0 && optim->sh_first < a.sh_active,
                "%s: sh_first %d outside [0, %d)", who, optim->sh_first, a.sh_active);
    if (len == 0) return PXL_OK;
    a.grad = optim->grad;
    a.gstride = grad_stride(grid, optim);
    a.sh_first = optim->sh_first;
    PXL_REQUIRE(a.gstride >= 1 + 3 * grid->sh_dim && a.gstride % 4 == 0,
                "%s: gradient stride %d must be a multiple of 4 and >= %d", who, a.gstride,
                1 + 3 * grid->sh_dim);
    PXL_REQUIRE(grid->storage == PXL_STORE_F32 || optim->stride > 0,
                "%s: 16-bit grids need an explicit optimiser stride", who);
    a.losses = losses;
    a.mask_coef = mask_coef;
    a.huber = huber_delta;
    const bool adapt = grid->march_mode == PXL_MARCH_REF_ADAPTIVE;
    const bool group = logged_step_serves(grid);
    PXL_REQUIRE(!x.bin_mask || (group && !adapt),
                "%s: proposal masks need the eight-sample marcher (fixed rule, sh_dim <= 9)", who);
    PXL_REQUIRE(!x.bin_targets || (group && !adapt && x.workspace && x.target_valid),
                "%s: proposal targets need the logged step and a validity array", who);
    a.bin_mask = x.bin_mask;
    a.bin_targets = x.bin_targets;
    a.target_valid = x.target_valid;
    a.target_wmin = x.target_wmin;
    const pxl_schedule *sched = x.sched;
    if (x.workspace && group) {
        const int32_t max_samples = x.max_samples;
        PXL_REQUIRE(max_samples > 0 && max_samples <= kMaxLogSamples,
                    "%s: max_samples %d outside [1, %d]", who, max_samples, kMaxLogSamples);
        const int gmode = sched ? sched->gradient : PXL_GRAD_PER_RAY;
        PXL_REQUIRE(gmode == PXL_GRAD_PER_RAY || gmode == PXL_GRAD_CHUNKED ||
                        gmode == PXL_GRAD_BINNED,
                    "%s: unknown gradient schedule %d", who, gmode);
        const int64_t nkeys = gmode == PXL_GRAD_BINNED ? bin_keys(grid) : 0;
        PXL_REQUIRE(gmode != PXL_GRAD_BINNED || nkeys > 0,
                    "%s: the brick-binned gradient needs 2 to 1025 vertices per side (grid "
                    "%d x %d x %d)",
                    who, grid->W, grid->H, grid->D);
        const LogLayout L = log_layout(len, max_samples, nkeys);
        PXL_REQUIRE(x.workspace_bytes >= L.total,
                    "%s: workspace of %zu bytes, %zu needed for %lld rays x %d samples%s", who,
                    x.workspace_bytes, L.total, (long long)len, max_samples,
                    nkeys > 0 ? " with the brick-binned gradient "
                                "(pxl_train_rays_workspace_bytes_grid)"
                              : "");
        PXL_REQUIRE(((uintptr_t)x.workspace & 255) == 0,
                    "%s: workspace must be 256-byte aligned", who);
        set_log_args(a, static_cast<char *>(x.workspace), L, nkeys);
        a.log_cap = max_samples;
        a.log_chunks = (max_samples + 63) / 64;
        if (nkeys > 0) {
            a.vbx = (grid->W + 7) / 8;
            a.vby = (grid->H + 7) / 8;
            a.vbz = (grid->D + 7) / 8;
        }
        const int wpb = sched ? sched->waves_per_block : 4;
        PXL_REQUIRE(wpb == 1 || wpb == 2 || wpb == 4, "%s: waves_per_block %d (1, 2 or 4)", who,
                    wpb);
        const int gwpb = sched && sched->grad_waves_per_block ? sched->grad_waves_per_block : wpb;
        PXL_REQUIRE(gwpb == 1 || gwpb == 2 || gwpb == 4,
                    "%s: grad_waves_per_block %d (0, 1, 2 or 4)", who, gwpb);
        PXL_REQUIRE(!sched || !(sched->persistent && gmode == PXL_GRAD_CHUNKED),
                    "%s: the chunk-parallel gradient has no persistent form", who);
        st = dispatch_logged(grid->sh_dim, grid->sparse, grid->storage, a, wpb, gwpb,
                             sched && sched->persistent, gmode, x.grad_wait,
                             as_stream(stream), adapt);
        return st != PXL_OK ? st : check_launch(who);
    }
    // One kernel writes the gradient: it starts after grad_wait.
    if (x.grad_wait && hipStreamWaitEvent(as_stream(stream), x.grad_wait, 0) != hipSuccess) {
        set_error("%s: hipStreamWaitEvent failed", who);
        return PXL_EHIP;
    }
    st = fused ? dispatch<kFused>(grid->sh_dim, grid->sparse, adapt, grid->storage, a,
                                  as_stream(stream))
               : dispatch<kBwd>(grid->sh_dim, grid->sparse, adapt, grid->storage, a,
                                as_stream(stream));
    return st != PXL_OK ? st : check_launch(who);
}

namespace pxl {
int32_t train_rays_ex(const pxl_grid *grid, const pxl_optim *optim, const pxl_rays *rays,
                      double *losses, const int32_t *sample_index, int64_t start, int64_t len,
                      double mask_coef, float huber_delta, void *stream, const TrainExtra &x) {
    return bwd_common(grid, optim, rays, losses, sample_index, start, len, mask_coef,
                      huber_delta, stream, true, x);
}
}  // namespace pxl

extern "C" int32_t pxl_render_bwd_rgb(const pxl_grid *grid, const pxl_optim *optim,
                                      const pxl_rays *rays, double *losses,
                                      const int32_t *sample_index, int64_t start, int64_t len,
                                      double mask_coef, float huber_delta, void *stream) {
    return bwd_common(grid, optim, rays, losses, sample_index, start, len, mask_coef,
                      huber_delta, stream, false);
}

extern "C" int32_t pxl_train_rays(const pxl_grid *grid, const pxl_optim *optim,
                                  const pxl_rays *rays, double *losses,
                                  const int32_t *sample_index, int64_t start, int64_t len,
                                  double mask_coef, float huber_delta, void *stream) {
    return bwd_common(grid, optim, rays, losses, sample_index, start, len, mask_coef,
                      huber_delta, stream, true);
}

extern "C" size_t pxl_train_rays_workspace_bytes(int64_t len, int32_t max_samples) {
    if (len <= 0 || max_samples <= 0 || max_samples > kMaxLogSamples) return 0;
    return log_layout(len, max_samples).total;
}

extern "C" size_t pxl_train_rays_workspace_bytes_grid(const pxl_grid *grid, int64_t len,
                                                      int32_t max_samples) {
    if (len <= 0 || max_samples <= 0 || max_samples > kMaxLogSamples) return 0;
    return log_layout(len, max_samples, bin_keys(grid)).total;
}

extern "C" int32_t pxl_train_rays_ws(const pxl_grid *grid, const pxl_optim *optim,
                                     const pxl_rays *rays, double *losses,
                                     const int32_t *sample_index, int64_t start, int64_t len,
                                     double mask_coef, float huber_delta, void *workspace,
                                     size_t workspace_bytes, int32_t max_samples, void *stream) {
    if (!workspace) {
        clear_error();
        set_error("pxl_train_rays_ws: null workspace");
        return PXL_EINVAL;
    }
    TrainExtra x;
    x.workspace = workspace;
    x.workspace_bytes = workspace_bytes;
    x.max_samples = max_samples;
    return bwd_common(grid, optim, rays, losses, sample_index, start, len, mask_coef,
                      huber_delta, stream, true, x);
}

extern "C" int32_t pxl_train_rays_prop(const pxl_grid *grid, const pxl_optim *optim,
                                       const pxl_rays *rays, double *losses,
                                       const int32_t *sample_index, int64_t start, int64_t len,
                                       double mask_coef, float huber_delta, void *workspace,
                                       size_t workspace_bytes, int32_t max_samples,
                                       const pxl_schedule *sched, const uint64_t *bin_mask,
                                       uint64_t *bin_targets, uint8_t *target_valid,
                                       double target_wmin, void *stream) {
    if (!workspace) {
        clear_error();
        set_error("pxl_train_rays_prop: null workspace");
        return PXL_EINVAL;
    }
    const pxl_schedule dflt = {4, 0, 0, 0};
    TrainExtra x;
    x.workspace = workspace;
    x.workspace_bytes = workspace_bytes;
    x.max_samples = max_samples;
    x.sched = sched ? sched : &dflt;
    x.bin_mask = bin_mask;
    x.bin_targets = bin_targets;
    x.target_valid = target_valid;
    x.target_wmin = target_wmin;
    return bwd_common(grid, optim, rays, losses, sample_index, start, len, mask_coef,
                      huber_delta, stream, true, x);
}

extern "C" int32_t pxl_train_rays_sched(const pxl_grid *grid, const pxl_optim *optim,
                                        const pxl_rays *rays, double *losses,
                                        const int32_t *sample_index, int64_t start, int64_t len,
                                        double mask_coef, float huber_delta, void *workspace,
                                        size_t workspace_bytes, int32_t max_samples,
                                        const pxl_schedule *sched, void *stream) {
    if (!workspace || !sched) {
        clear_error();
        set_error("pxl_train_rays_sched: null workspace or schedule");
        return PXL_EINVAL;
    }
    TrainExtra x;
    x.workspace = workspace;
    x.workspace_bytes = workspace_bytes;
    x.max_samples = max_samples;
    x.sched = sched;
    return bwd_common(grid, optim, rays, losses, sample_index, start, len, mask_coef,
                      huber_delta, stream, true, x);
}

// Sample log of pxl_train_rays_noalloc: one array per GPU, held by the
// library until the process ends.  Its name is a host-side stand-in on the
// host: kernels get the device address hipGetSymbolAddress returns for the
// current device (scratch_of_device).
static __device__ __align__(256) unsigned char scratch[PXL_NOALLOC_SCRATCH_BYTES];

namespace pxl {
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call on that device.
int32_t scratch_of_device(void **out, const char *who) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_error("%s: no current device (or device %d beyond %d)", who, dev, kMaxDevices);
        return PXL_EHIP;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            set_error("%s: hipGetSymbolAddress of the sample log failed on device %d: %s", who,
                      dev, hipGetErrorString(e));
            return PXL_EHIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PXL_OK;
}
}  // namespace
}  // namespace pxl

extern "C" int32_t pxl_train_rays_noalloc(const pxl_grid *grid, const pxl_optim *optim,
                                          const pxl_rays *rays, double *losses,
                                          const int32_t *sample_index, int64_t start,
                                          int64_t len, double mask_coef, float huber_delta,
                                          int32_t max_samples, const pxl_schedule *sched,
                                          void *stream) {
    clear_error();
    const char *who = "pxl_train_rays_noalloc";
    PXL_REQUIRE(len >= 0 && max_samples > 0 && max_samples <= kMaxLogSamples,
                "%s: len %lld / max_samples %d outside [1, %d]", who, (long long)len, max_samples,
                kMaxLogSamples);
    const size_t need = log_layout(len, max_samples).total;
    PXL_REQUIRE(need <= (size_t)PXL_NOALLOC_SCRATCH_BYTES,
                "%s: %lld rays x %d samples need a %zu-byte sample log, beyond the library's "
                "%zu bytes (PXL_NOALLOC_SCRATCH_BYTES); use pxl_train_rays_sched with a workspace",
                who, (long long)len, max_samples, need, (size_t)PXL_NOALLOC_SCRATCH_BYTES);
    void *ws = nullptr;
    const int32_t st = scratch_of_device(&ws, who);
    if (st != PXL_OK) return st;
    const pxl_schedule dflt = {4, 0, 0, 0};
    TrainExtra x;
    x.workspace = ws;
    x.workspace_bytes = (size_t)PXL_NOALLOC_SCRATCH_BYTES;
    x.max_samples = max_samples;
    x.sched = sched ? sched : &dflt;
    return bwd_common(grid, optim, rays, losses, sample_index, start, len, mask_coef,
                      huber_delta, stream, true, x);
}

namespace pxl {
namespace {
__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_popcount(const uint32_t *bm, int64_t nwords, unsigned long long *count) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t c = 0;
    for (; i < nwords; i += (int64_t)gridDim.x * blockDim.x) c += __popc(bm[i]);
    c = wave_sum_u32(c);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, (unsigned long long)c);
}
}  // namespace
}  // namespace pxl

extern "C" int32_t pxl_count_touched(const pxl_grid *grid, const pxl_rays *rays,
                                     const int32_t *sample_index, int64_t start, int64_t len,
                                     uint32_t *bitmap, uint64_t *count, void *stream) {
    clear_error();
    RenderArgs a;
    int32_t st = prepare(grid, rays, sample_index, start, len, "pxl_count_touched", a);
    if (st != PXL_OK) return st;
    PXL_REQUIRE(bitmap && count, "pxl_count_touched: null bitmap/count");
    const int64_t nwords = (grid->n_vox + 31) / 32;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, s,
                       bitmap, nwords, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, reinterpret_cast<uint32_t *>(count),
                       (int64_t)2, 0u);
    if (len > 0) {
        a.bitmap = bitmap;
        st = dispatch<kTouch>(grid->sh_dim, grid->sparse,
                              grid->march_mode == PXL_MARCH_REF_ADAPTIVE, grid->storage, a, s);
        if (st != PXL_OK) return st;
    }
    hipLaunchKernelGGL(k_popcount, dim3(256), dim3(256), 0, s, bitmap, nwords,
                       reinterpret_cast<unsigned long long *>(count));
    return check_launch("pxl_count_touched");
}

extern "C" const char *pxl_render_kernel_name(const pxl_grid *grid, int32_t mode) {
    clear_error();
    if (!grid || mode < 0 || mode > 3) {
        set_error("pxl_render_kernel_name: bad argument");
        return nullptr;
    }
    static thread_local char name[96];
    const bool adapt = grid->march_mode == PXL_MARCH_REF_ADAPTIVE;
    const bool group = logged_step_serves(grid);
    const char *modes[4] = {"fwd", "bwd", "fused", "fused"};
    const char *store = grid->storage == PXL_STORE_BF16  ? "bf16"
                        : grid->storage == PXL_STORE_F16 ? "f16"
                                                         : "f32";
    if (group && adapt && mode == 3)
        snprintf(name, sizeof(name), "k_adapt8<%d,%s,log>+k_log_grad<%d,adaptive>",
                 grid->sh_dim, grid->sparse ? "sparse" : "dense", grid->sh_dim);
    else if (group && adapt && mode == 0)
        snprintf(name, sizeof(name), "k_adapt8<%d,%s,fwd>", grid->sh_dim,
                 grid->sparse ? "sparse" : "dense");
    else if (group && mode == 3)
        snprintf(name, sizeof(name), "k_march8<%d,%s,log,%s>+k_log_grad<%d>", grid->sh_dim,
                 grid->sparse ? "sparse" : "dense", store, grid->sh_dim);
    else if (group && !adapt)
        snprintf(name, sizeof(name), "k_march8<%d,%s,%s,%s>", grid->sh_dim,
                 grid->sparse ? "sparse" : "dense", modes[mode], store);
    else
        snprintf(name, sizeof(name), "k_render<%d,%s,%s,%s,%s>", grid->sh_dim,
                 grid->sparse ? "sparse" : "dense", adapt ? "adaptive" : "fixed", modes[mode],
                 store);
    return name;
}
