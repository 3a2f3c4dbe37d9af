// C++ host layer of libpxl: camera maths, the reference's operator API as
// static methods over the C descriptors, and the batch loop of the trainer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pxl.h"

namespace pxl {
namespace host {

bool invert(const double *src, double *dst, int n);
void euler_to_rotation(const double deg[3], double *R);
void look_at(const double *cam, const double *obj, double *R);

// RenderUtils (RenderUtils.cuh:7-85): the reference passes Rays,
// ObjectScene and Optimizer by value; here the same roles are the C
// descriptors, plus an explicit stream.
struct RenderUtils {
    static int32_t renderRays(const pxl_rays &rays, const pxl_grid &scene,
                              const int32_t *sampleIndex, int64_t startSampleIndex, int64_t len,
                              hipStream_t stream);
    static int32_t renderRays_withGradients_RGB(const pxl_optim &optimizer, const pxl_rays &rays,
                                                const pxl_grid &scene, double *losses,
                                                const int32_t *sampleIndex,
                                                int64_t startSampleIndex, int64_t len,
                                                double maskRayLossCoefficient, float huberDelta,
                                                hipStream_t stream);
    static int32_t addVarianceLoss(const pxl_grid &scene, const pxl_optim &optimizer,
                                   double densityCoefficient, double shCoefficient,
                                   int32_t varianceType, int32_t neighbourNums,
                                   int32_t optimizeStep, void *scratch, size_t scratchBytes,
                                   hipStream_t stream);
};

// Optimizer::updateRGBScene (Optimizer.cuh:35-37).
struct Optimizer {
    static int32_t updateRGBScene(const pxl_optim &optimizer, const pxl_grid &scene,
                                  double learningRate, double shLearningRate, int32_t timeStep,
                                  hipStream_t stream);
};

}  // namespace host
}  // namespace pxl
