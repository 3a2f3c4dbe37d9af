// Ray generation: one thread per pixel of one view.
//
// Follows getBound + initialRays (Rays.cu:22-144): camera-space point of the
// pixel centre in FP32, c2w product accumulated into FP32 with FP64
// products, FP32 normalisation through an FP64 square root, FP64 slab test
// against the integer bounding box.  This file is compiled with
// -ffp-contract=off so the operation sequence is exactly the one written
// here; the oracle's ray generator uses the same sequence, which makes the
// two bit-identical.
#include "common.hpp"

namespace pxl {
namespace {

struct CamArgs {
    double c2w[16];
    double centre[4];
    double fx, fy, cx, cy;
};

__device__ __forceinline__ double dmin(double a, double b) { return a > b ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }

// Pixels of the box rows [r0, r0 + bh) x cols [c0, c0 + bw) in row-major
// order; box entry i becomes ray start + i (the whole image is the box
// (0, 0, img_h, img_w)).
__global__ __launch_bounds__(256) void k_rays_init(
    CamArgs cam, pxl_bbox bb, int img_w, int img_h, int r0, int c0, int bw, int bh,
    const uint8_t *__restrict__ ir, const uint8_t *__restrict__ ig,
    const uint8_t *__restrict__ ib, const uint8_t *__restrict__ im, int64_t start,
    pxl_rays rays) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)bw * bh) return;
    const int row = r0 + (int)(i / bw);
    const int col = c0 + (int)(i % bw);
    const int64_t pix = (int64_t)row * img_w + col;

    // Camera-space pixel centre (Rays.cu:29-35).
    float xc[4];
    xc[0] = (float)((((double)col + 0.5) - cam.cx) / cam.fx);
    xc[1] = (float)((((double)row + 0.5) - cam.cy) / cam.fy);
    xc[2] = 1.0f;
    xc[3] = 1.0f;

    // World-space direction before normalisation (Rays.cu:38-46).
    float sp[4];
    for (int r = 0; r < 4; r++) {
        sp[r] = 0.0f;
        for (int c = 0; c < 4; c++) {
            const double prod = cam.c2w[r * 4 + c] * (double)xc[c];
            sp[r] = (float)((double)sp[r] + prod);
        }
        sp[r] = (float)((double)sp[r] - cam.centre[r]);
    }
    // normalizePoint (Rays.cu:13-20): FP32 squares, FP64 root and divide.
    const float s0 = sp[0] * sp[0];
    const float s1 = sp[1] * sp[1];
    const float s2 = sp[2] * sp[2];
    const float ssum = (s0 + s1) + s2;
    const double norm = sqrt((double)ssum);
    sp[0] = (float)((double)sp[0] / norm);
    sp[1] = (float)((double)sp[1] / norm);
    sp[2] = (float)((double)sp[2] / norm);

    // Slab test (Rays.cu:48-66).
    const double t0 = ((double)bb.x_max - cam.centre[0]) / (double)sp[0];
    const double t1 = ((double)bb.x_min - cam.centre[0]) / (double)sp[0];
    const double t2 = ((double)bb.y_max - cam.centre[1]) / (double)sp[1];
    const double t3 = ((double)bb.y_min - cam.centre[1]) / (double)sp[1];
    const double t4 = ((double)bb.z_max - cam.centre[2]) / (double)sp[2];
    const double t5 = ((double)bb.z_min - cam.centre[2]) / (double)sp[2];
    double tmin = dmin(t0, t1);
    double tmax = dmax(t0, t1);
    tmin = dmax(tmin, dmin(t2, t3));
    tmax = dmin(tmax, dmax(t2, t3));
    tmin = dmax(tmin, dmin(t4, t5));
    tmax = dmin(tmax, dmax(t4, t5));

    const int64_t ray = start + i;
    rays.ox[ray] = (float)cam.centre[0];
    rays.oy[ray] = (float)cam.centre[1];
    rays.oz[ray] = (float)cam.centre[2];
    rays.near_t[ray] = (float)tmin;
    rays.far_t[ray] = (float)tmax;
    rays.dx[ray] = sp[0];
    rays.dy[ray] = sp[1];
    rays.dz[ray] = sp[2];

    // Ground truth (Rays.cu:118-130); white background composites the
    // colour over white with the mask as alpha and truncates to a byte.
    uint8_t gr = 0, gg = 0, gb = 0, gm = 0;
    if (ir != nullptr) {
        gm = im[pix];
        if (rays.white_bg) {
            const int m = gm;
            gr = (uint8_t)((float)(ir[pix] * m) / 255.0f + 255.0f - (float)m);
            gg = (uint8_t)((float)(ig[pix] * m) / 255.0f + 255.0f - (float)m);
            gb = (uint8_t)((float)(ib[pix] * m) / 255.0f + 255.0f - (float)m);
        } else {
            gr = ir[pix];
            gg = ig[pix];
            gb = ib[pix];
        }
    }
    rays.gt_mask[ray] = gm;
    rays.gt_r[ray] = gr;
    rays.gt_g[ray] = gg;
    rays.gt_b[ray] = gb;
}

// Image::bbox before readMask (ImageUtils.cu:12-15).
__global__ void k_mask_bounds_init(int img_w, int img_h, int32_t *bounds) {
    if (threadIdx.x == 0) {
        bounds[0] = img_h - 1;
        bounds[1] = 0;
        bounds[2] = img_w - 1;
        bounds[3] = 0;
    }
}

// readMask's bounds (Utils.cu:508-548): rows and columns of the nonzero
// mask pixels, reduced per wave before one atomic per bound.
__global__ __launch_bounds__(256) void k_mask_bounds(const uint8_t *__restrict__ mask, int img_w,
                                                     int img_h, int32_t *bounds) {
    const int64_t pix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = pix < (int64_t)img_w * img_h && mask[pix] != 0;
    int rmin = INT32_MAX, rmax = INT32_MIN, cmin = INT32_MAX, cmax = INT32_MIN;
    if (on) {
        rmin = rmax = (int)(pix / img_w);
        cmin = cmax = (int)(pix % img_w);
    }
    for (int off = 32; off > 0; off >>= 1) {
        rmin = min(rmin, __shfl_xor(rmin, off, 64));
        rmax = max(rmax, __shfl_xor(rmax, off, 64));
        cmin = min(cmin, __shfl_xor(cmin, off, 64));
        cmax = max(cmax, __shfl_xor(cmax, off, 64));
    }
    if ((threadIdx.x & 63) == 0 && rmin != INT32_MAX) {
        atomicMin(bounds + 0, rmin);
        atomicMax(bounds + 1, rmax);
        atomicMin(bounds + 2, cmin);
        atomicMax(bounds + 3, cmax);
    }
}

// shuffleToNewGroup (Rays.cu:242-269): dst ray j = src ray perm[j]; an
// index outside the source set gives a zero ray.
__global__ __launch_bounds__(256) void k_rays_permute(const pxl_rays src, const pxl_rays dst,
                                                      const int32_t *__restrict__ perm,
                                                      int64_t n) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int32_t k = perm[j];
    const bool ok = k >= 0 && (int64_t)k < src.n;
    dst.ox[j] = ok ? src.ox[k] : 0.0f;
    dst.oy[j] = ok ? src.oy[k] : 0.0f;
    dst.oz[j] = ok ? src.oz[k] : 0.0f;
    dst.dx[j] = ok ? src.dx[k] : 0.0f;
    dst.dy[j] = ok ? src.dy[k] : 0.0f;
    dst.dz[j] = ok ? src.dz[k] : 0.0f;
    dst.near_t[j] = ok ? src.near_t[k] : 0.0f;
    dst.far_t[j] = ok ? src.far_t[k] : 0.0f;
    dst.gt_r[j] = ok ? src.gt_r[k] : (uint8_t)0;
    dst.gt_g[j] = ok ? src.gt_g[k] : (uint8_t)0;
    dst.gt_b[j] = ok ? src.gt_b[k] : (uint8_t)0;
    dst.gt_mask[j] = ok ? src.gt_mask[k] : (uint8_t)0;
}

int32_t check_ray_planes(const pxl_rays *r, const char *what) {
    PXL_REQUIRE(r->ox && r->oy && r->oz && r->dx && r->dy && r->dz && r->near_t && r->far_t &&
                    r->gt_r && r->gt_g && r->gt_b && r->gt_mask,
                "%s: ray planes must be non-null", what);
    return PXL_OK;
}

int32_t launch_rays(const pxl_camera *cam, const pxl_bbox *bbox, int32_t img_w, int32_t img_h,
                    int r0, int c0, int bw, int bh, const uint8_t *img_r, const uint8_t *img_g,
                    const uint8_t *img_b, const uint8_t *img_mask, int64_t start_index,
                    const pxl_rays *rays, void *stream, const char *what) {
    PXL_REQUIRE(start_index >= 0 && start_index + (int64_t)bw * bh <= rays->n,
                "%s: rays [%lld, %lld) exceed the ray set of %lld", what, (long long)start_index,
                (long long)(start_index + (int64_t)bw * bh), (long long)rays->n);
    if (check_ray_planes(rays, what) != PXL_OK) return PXL_EINVAL;
    const bool any_img = img_r || img_g || img_b || img_mask;
    PXL_REQUIRE(!any_img || (img_r && img_g && img_b && img_mask),
                "%s: image planes must be all set or all null", what);
    CamArgs a;
    memcpy(a.c2w, cam->c2w, sizeof(a.c2w));
    memcpy(a.centre, cam->centre, sizeof(a.centre));
    a.fx = cam->fx;
    a.fy = cam->fy;
    a.cx = cam->cx;
    a.cy = cam->cy;
    const int64_t n = (int64_t)bw * bh;
    if (n == 0) return PXL_OK;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_rays_init, dim3(blocks), dim3(256), 0, as_stream(stream), a, *bbox,
                       img_w, img_h, r0, c0, bw, bh, img_r, img_g, img_b, img_mask, start_index,
                       *rays);
    return check_launch(what);
}

}  // namespace
}  // namespace pxl

extern "C" int32_t pxl_rays_init(const pxl_camera *cam, const pxl_bbox *bbox,
                                 int32_t img_w, int32_t img_h,
                                 const uint8_t *img_r, const uint8_t *img_g,
                                 const uint8_t *img_b, const uint8_t *img_mask,
                                 int64_t start_index, const pxl_rays *rays,
                                 void *stream) {
    using namespace pxl;
    clear_error();
    PXL_REQUIRE(cam && bbox && rays, "pxl_rays_init: null argument");
    PXL_REQUIRE(img_w > 0 && img_h > 0, "pxl_rays_init: bad image size %dx%d", img_w, img_h);
    return launch_rays(cam, bbox, img_w, img_h, 0, 0, img_w, img_h, img_r, img_g, img_b,
                       img_mask, start_index, rays, stream, "pxl_rays_init");
}

extern "C" int32_t pxl_rays_init_box(const pxl_camera *cam, const pxl_bbox *bbox,
                                     int32_t img_w, int32_t img_h, const int32_t pixel_box[4],
                                     const uint8_t *img_r, const uint8_t *img_g,
                                     const uint8_t *img_b, const uint8_t *img_mask,
                                     int64_t start_index, const pxl_rays *rays, void *stream) {
    using namespace pxl;
    clear_error();
    PXL_REQUIRE(cam && bbox && rays && pixel_box, "pxl_rays_init_box: null argument");
    PXL_REQUIRE(img_w > 0 && img_h > 0, "pxl_rays_init_box: bad image size %dx%d", img_w, img_h);
    const int r0 = pixel_box[0], r1 = pixel_box[1], c0 = pixel_box[2], c1 = pixel_box[3];
    PXL_REQUIRE(0 <= r0 && r0 <= r1 && r1 < img_h && 0 <= c0 && c0 <= c1 && c1 < img_w,
                "pxl_rays_init_box: pixel box rows [%d, %d] cols [%d, %d] outside %dx%d or "
                "empty",
                r0, r1, c0, c1, img_w, img_h);
    return launch_rays(cam, bbox, img_w, img_h, r0, c0, c1 - c0 + 1, r1 - r0 + 1, img_r, img_g,
                       img_b, img_mask, start_index, rays, stream, "pxl_rays_init_box");
}

extern "C" int32_t pxl_mask_bounds(const uint8_t *mask, int32_t img_w, int32_t img_h,
                                   int32_t *bounds, void *stream) {
    using namespace pxl;
    clear_error();
    PXL_REQUIRE(mask && bounds, "pxl_mask_bounds: null argument");
    PXL_REQUIRE(img_w > 0 && img_h > 0, "pxl_mask_bounds: bad image size %dx%d", img_w, img_h);
    const hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_mask_bounds_init, dim3(1), dim3(64), 0, st, img_w, img_h, bounds);
    const int64_t n = (int64_t)img_w * img_h;
    hipLaunchKernelGGL(k_mask_bounds, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, mask,
                       img_w, img_h, bounds);
    return check_launch("pxl_mask_bounds");
}

extern "C" int32_t pxl_rays_permute(const pxl_rays *src, const pxl_rays *dst,
                                    const int32_t *perm, int64_t n, void *stream) {
    using namespace pxl;
    clear_error();
    PXL_REQUIRE(src && dst && perm, "pxl_rays_permute: null argument");
    PXL_REQUIRE(n >= 0 && n <= dst->n, "pxl_rays_permute: %lld rays into a set of %lld",
                (long long)n, (long long)dst->n);
    if (check_ray_planes(src, "pxl_rays_permute") != PXL_OK ||
        check_ray_planes(dst, "pxl_rays_permute") != PXL_OK)
        return PXL_EINVAL;
    if (n == 0) return PXL_OK;
    hipLaunchKernelGGL(k_rays_permute, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       as_stream(stream), *src, *dst, perm, n);
    return check_launch("pxl_rays_permute");
}
