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

__global__ __launch_bounds__(256) void k_rays_init(
    CamArgs cam, pxl_bbox bb, int img_w, int img_h, const uint8_t *__restrict__ ir,
    const uint8_t *__restrict__ ig, const uint8_t *__restrict__ ib,
    const uint8_t *__restrict__ im, int64_t start, pxl_rays rays) {
    const int64_t pix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= (int64_t)img_w * img_h) return;
    const int row = (int)(pix / img_w);
    const int col = (int)(pix % img_w);

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

    const int64_t ray = start + pix;
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
    PXL_REQUIRE(start_index >= 0 && start_index + (int64_t)img_w * img_h <= rays->n,
                "pxl_rays_init: rays [%lld, %lld) exceed the ray set of %lld",
                (long long)start_index, (long long)(start_index + (int64_t)img_w * img_h),
                (long long)rays->n);
    PXL_REQUIRE(rays->ox && rays->oy && rays->oz && rays->dx && rays->dy && rays->dz &&
                    rays->near_t && rays->far_t && rays->gt_r && rays->gt_g &&
                    rays->gt_b && rays->gt_mask,
                "pxl_rays_init: ray planes must be non-null");
    const bool any_img = img_r || img_g || img_b || img_mask;
    PXL_REQUIRE(!any_img || (img_r && img_g && img_b && img_mask),
                "pxl_rays_init: image planes must be all set or all null");
    CamArgs a;
    memcpy(a.c2w, cam->c2w, sizeof(a.c2w));
    memcpy(a.centre, cam->centre, sizeof(a.centre));
    a.fx = cam->fx;
    a.fy = cam->fy;
    a.cx = cam->cx;
    a.cy = cam->cy;
    const int64_t npix = (int64_t)img_w * img_h;
    const unsigned blocks = (unsigned)((npix + 255) / 256);
    hipLaunchKernelGGL(k_rays_init, dim3(blocks), dim3(256), 0, as_stream(stream), a,
                       *bbox, img_w, img_h, img_r, img_g, img_b, img_mask, start_index,
                       *rays);
    return check_launch("pxl_rays_init");
}
