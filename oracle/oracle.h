/*
 * oracle.h - CPU restatement of the reference's train-step maths.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker:
 * the product (libpxl.so) never calls it.
 *
 * Layout here is the reference's: P = 1 + 3S parameter planes of n_vox
 * values each, plane-major (plane 0 density, 1..S red SH, S+1..2S green,
 * 2S+1..3S blue; SceneUtils.cu:148-171).  Arithmetic is FP64 except where a
 * function says it replicates FP32 operations bit for bit.
 */
#ifndef PXL_ORACLE_H
#define PXL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid {
    int32_t W, H, D, S;
    int32_t sparse;
    int64_t n_vox;
    const float *planes;           /* [P][n_vox] */
    const uint8_t *occupancy;      /* sparse: one byte per cell */
    const uint32_t *index_offsets; /* sparse: [W*H*D] */
    int32_t bbox[6];               /* x_min, x_max, y_min, y_max, z_min, z_max */
    float step;
    int32_t adaptive;              /* 0 fixed, 1 reference adaptive step */
    /* SH band (RenderUtils.cu:859-1000, 1117-1132): coefficients k <
     * sh_active enter the colour (0 = all S); gradients and the SH terms of
     * the variance losses cover k in [sh_first, sh_active). */
    int32_t sh_active;
    int32_t sh_first;
    const uint8_t *in_cage;        /* sparse: one byte per dense vertex (TV type 0) */
} orc_grid;

typedef struct orc_rays {
    const float *ox, *oy, *oz, *dx, *dy, *dz, *near_t, *far_t;
    const uint8_t *gt_r, *gt_g, *gt_b, *gt_mask;
    int32_t white_bg;
} orc_rays;

/* Camera of one pose file entry (readFromMyBlender, CameraUtils.cu:271-386):
 * look-at from `translation` (after the optional rescale) towards
 * object_centre, OpenGL->OpenCV flip when z_reversed, w2c = R [I | -c],
 * c2w = inverse, centre = -M^-1 p4 (CameraUtils.cu:577-608). */
void orc_camera_from_pose(const double translation[3], const double object_centre[3],
                          double camera_distance, int32_t need_rescale, int32_t z_reversed,
                          double fx, double fy, double cx, double cy, double c2w[16],
                          double w2c[16], double centre[4]);

/* Ray generation, FP32 bit-replica of the operation sequence of getBound +
 * initialRays (Rays.cu:22-144) without FP contraction. */
void orc_rays_init(const double c2w[16], const double centre[4], double fx, double fy,
                   double cx, double cy, const int32_t bbox[6], int32_t img_w, int32_t img_h,
                   const uint8_t *img_r, const uint8_t *img_g, const uint8_t *img_b,
                   const uint8_t *img_mask, int32_t white_bg, float *ox, float *oy, float *oz,
                   float *dx, float *dy, float *dz, float *near_t, float *far_t, uint8_t *gt_r,
                   uint8_t *gt_g, uint8_t *gt_b, uint8_t *gt_mask);

/* pruneScene + occupiedVoxel (SceneUtils.cu:800-881): occupancy byte per
 * cell, in-cage byte per vertex (cleared first).  Returns occupied cells. */
int64_t orc_build_occupancy(const float *density, int32_t W, int32_t H, int32_t D,
                            float threshold, uint8_t *occupancy, uint8_t *in_cage);

/* getCagedNums (SceneUtils.cu:1081-1120).  Returns inCageNums. */
int64_t orc_build_index(const uint8_t *in_cage, int32_t W, int32_t H, int32_t D,
                        uint32_t *index_offsets);

/* Forward march of batch rays sample_index[start .. start+len)
 * (volumeRender_RGB_train + renderRaysGPU, RenderUtils.cu:1582-1718,
 * 1919-1959).  out_* are indexed by ray.  trace (optional) gets
 * [count, hash] per batch position, same hash as pxl_render_fwd_trace. */
void orc_render_fwd(const orc_grid *g, const orc_rays *r, const int32_t *sample_index,
                    int64_t start, int64_t len, double *out_r, double *out_g, double *out_b,
                    double *out_w, double *out_density, uint32_t *trace);

/* Backward (renderRays_withGradients_RGBGPU, RenderUtils.cu:2051-2256)
 * given the forward colour and opacity per ray (fwd_*).  Adds into
 * grad[P][n_vox]; losses per batch position. */
void orc_render_bwd(const orc_grid *g, const orc_rays *r, const int32_t *sample_index,
                    int64_t start, int64_t len, const double *fwd_r, const double *fwd_g,
                    const double *fwd_b, const double *fwd_w, double mask_coef,
                    float huber_delta, double *losses, double *grad);
/* orc_render_bwd that also adds into grad_abs[P][n_vox] the magnitude of
 * every term of every contribution: |colour contribution|, and for the
 * density 0.6667 step sigma' sum_ch |r_ch| (|T c_ch| + |C_ch - H_ch|) +
 * step sigma' (|A - A_i| + T) |m| times the corner weight.  A gradient entry
 * computed in FP32 arithmetic carries an error of a few FP32 rounding units
 * of this sum, which can exceed 1e-4 of the entry where its terms cancel. */
void orc_render_bwd_cond(const orc_grid *g, const orc_rays *r, const int32_t *sample_index,
                         int64_t start, int64_t len, const double *fwd_r, const double *fwd_g,
                         const double *fwd_b, const double *fwd_w, double mask_coef,
                         float huber_delta, double *losses, double *grad, double *grad_abs);
/* Proposal variants: steps in bins whose bit of bin_mask[b] is clear are
 * skipped (bin k of a ray: floor(k step 64 / (t_far - t_near)), at most 63);
 * the forward also returns, per batch entry, the bins holding a sample of
 * weight above target_wmin.  Fixed marching rule only. */
/* Eval image of one camera (renderImageGPU): adaptive rule, planar
 * outputs [3][h*w] / [h*w], any may be NULL; rays_out (optional) gets
 * (t_min, t_max, dx, dy, dz) per pixel. */
void orc_render_image(const orc_grid *g, const double c2w[16], const double centre[4],
                      double fx, double fy, double cx, double cy, int32_t width, int32_t height,
                      int32_t white_bg, double *rgb, uint8_t *rgb8, uint8_t *mask,
                      double *disparity, double *rays_out);
void orc_render_fwd_masked(const orc_grid *g, const orc_rays *r, const int32_t *sample_index,
                           int64_t start, int64_t len, const uint64_t *bin_mask,
                           double target_wmin, uint64_t *targets, double *out_r, double *out_g,
                           double *out_b, double *out_w, double *out_density, uint32_t *trace);
void orc_render_bwd_masked(const orc_grid *g, const orc_rays *r, const int32_t *sample_index,
                           int64_t start, int64_t len, const uint64_t *bin_mask,
                           const double *fwd_r, const double *fwd_g, const double *fwd_b,
                           const double *fwd_w, double mask_coef, float huber_delta,
                           double *losses, double *grad);

/* addVarianceLossGPU (RenderUtils.cu:2469-2547), cells visited serially so
 * the reference's unsynchronised += becomes an exact sum.  Adds into grad. */
void orc_tv_disparity(const orc_grid *g, double lambda_density, double lambda_sh,
                      int32_t optimize_step, double *grad);

/* totalVarianceLoss, 6-neighbour branch (RenderUtils.cu:2660-2763): the
 * reference's loop over interior vertices whose column y W + x is below
 * col_limit (W H for type 0; the launch of type 2, RenderUtils.cu:2840-2843,
 * covers ceil((W-1)/16) ceil((H-1)/16) 256), serially.  Caged grids: the
 * vertex's record r = index_offsets[v] is skipped unless in_cage[r] (the
 * flag array is per dense vertex; the reference reads it at the record
 * number, :2669); a neighbour whose index (record) is 0 reads the vertex
 * itself. */
void orc_tv_neighbour6(const orc_grid *g, double lambda_density, double lambda_sh,
                       double *grad);
void orc_tv_neighbour6_cols(const orc_grid *g, double lambda_density, double lambda_sh,
                            int64_t col_limit, double *grad);

/* updateRGBSceneGPU_ADAM (Optimizer.cu:359-415) in the reference's own
 * float/double mix.  Planes [P][n]; grad is zeroed. */
void orc_adam(float *params, float *grad, float *m, float *v, const float *adaptive_lr,
              int32_t P, int64_t n, double lr, double sh_lr, int32_t t);
/* The same restricted, when S > 0 (P = 1 + 3 S), to the density and the SH
 * coefficients [sh_lo, sh_hi) of each channel (hierarchyStrategy,
 * Optimizer.cu:393-412); other planes keep their values. */
void orc_adam_band(float *params, float *grad, float *m, float *v, const float *adaptive_lr,
                   int32_t P, int32_t S, int64_t n, double lr, double sh_lr, int32_t t,
                   int32_t sh_lo, int32_t sh_hi);

/* getMSEError (Utils.cu:311-320). */
/* upsampleScene + upsample_trilinear (SceneUtils.cu:262-485).
 * orc_upsample_density: the thresholded density of every vertex of the
 * W2 x H2 x D2 grid (the dense temporary initSelfByUpsample prunes,
 * SceneUtils.cu:540-546).  orc_upsample: dst planes [1 + 3 S2][dst_n]
 * (zero-initialised by the caller, as initSelf does); caged when
 * dst_in_cage is given (records via dst_index). */
void orc_upsample_density(const orc_grid *src, int32_t W2, int32_t H2, int32_t D2,
                          float threshold, float *density);
void orc_upsample(const orc_grid *src, int32_t W2, int32_t H2, int32_t D2, int32_t S2,
                  const uint8_t *dst_in_cage, const uint32_t *dst_index, int64_t dst_n,
                  float threshold, int32_t low_band, float *dst);
/* Threads for the ray, cell and parameter loops (default 1, bit-reproducible;
 * more threads make the gradient additions atomic, in varying order). */
void orc_set_threads(int32_t n);
double orc_mse(const double *losses, int64_t len);

/* Wall-clock seconds (CLOCK_MONOTONIC), for the CPU baseline. */
double orc_now(void);

#ifdef __cplusplus
}
#endif

#endif
