/*
 * refprec.h - the train step in the reference's arithmetic types, on the
 * host: the benchmark's CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by bench.py's cpu_baseline leg and the
 * tests.  Grids and rays use the oracle's descriptors (oracle.h): FP32
 * parameter planes [P][n_vox], byte occupancy, index offsets.
 */
#ifndef PXL_REFPREC_H
#define PXL_REFPREC_H

#include "oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Forward then backward of the batch entries sample_index[0 .. len)
 * (volumeRender_RGB_train, renderRays_withGradients_RGBGPU,
 * RenderUtils.cu:1582-1718, 2051-2256): fwd[5 b .. 5 b + 4] = the ray's
 * float colour, opacity and density sum, losses[b] its loss, gradients added
 * into grad[P][n_vox] with float atomics.  OpenMP over the entries. */
void rp_render_train(const orc_grid *g, const orc_rays *r, const int32_t *sample_index,
                     int64_t len, double mask_coef, float huber_delta, float *fwd, double *losses,
                     float *grad, int32_t threads);

/* addVarianceLossGPU (RenderUtils.cu:2469-2547) into grad[P][n_vox]. */
void rp_tv_disparity(const orc_grid *g, double lambda_density, double lambda_sh,
                     int32_t optimize_step, float *grad, int32_t threads);

/* updateRGBSceneGPU_ADAM (Optimizer.cu:359-415) over P planes of n values;
 * gradient := 0. */
void rp_adam(float *params, float *grad, float *m, float *v, int32_t P, int64_t n, double lr,
             double sh_lr, int32_t t, int32_t threads);

/* Wall clock in seconds (omp_get_wtime). */
double rp_now(void);

#ifdef __cplusplus
}
#endif

#endif
