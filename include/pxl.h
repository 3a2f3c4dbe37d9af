/*
 * pxl.h - C ABI of the MI355X-native Plenoxel train-step library (libpxl.so).
 *
 * One train step of the reference trainer is four calls in
 * Render::objectSceneLearner (SparseVolumeRadianceField/Render.cu:407, 430,
 * 471, 492).  Each has an entry point here with the same argument meaning;
 * the reference's structs-of-device-pointers (Rays, ObjectScene, Optimizer)
 * become the POD descriptors below.  Every buffer is caller-owned device
 * memory; no entry point allocates.  Every entry point is asynchronous on the
 * HIP stream passed as `stream` (NULL = default stream) and returns a status
 * code instead of exiting (the reference prints and calls exit(-1) in
 * checkGPUStatus, Utils.cu:700-717).
 *
 * Grid layout in HBM: one record of `stride` floats per stored vertex,
 * record[0] = raw density, record[1 + ch*S + k] = SH coefficient k of colour
 * channel ch (R, G, B).  This is the reference's plane order
 * (SceneUtils.cu:148-171) packed per vertex; pxl_grid_import_planes /
 * pxl_grid_export_planes convert from / to the reference's planes.
 */
#ifndef PXL_H
#define PXL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
enum {
    PXL_OK = 0,
    PXL_EINVAL = 1,   /* bad argument (null pointer, size, unsupported S) */
    PXL_ENOMEM = 2,   /* caller-provided scratch too small                */
    PXL_EHIP = 3,     /* HIP runtime error (launch failure)              */
    PXL_ENAN = 4      /* non-finite batch loss (Render.cu:450-457)       */
};

/* Marching rule. */
enum {
    /* t_k = t_near + k * step; the rule the benchmark configurations name. */
    PXL_MARCH_FIXED = 0,
    /* t += |step - w| after every evaluated sample, t += step after a
     * skipped one (RenderUtils.cu:1681-1705); evaluated in FP64. */
    PXL_MARCH_REF_ADAPTIVE = 1
};

/* Storage of the grid parameters.  Interpolation, compositing, gradients
 * and Adam moments are FP32/FP64 in every case. */
enum {
    PXL_STORE_F32 = 0,
    PXL_STORE_BF16 = 1,   /* bfloat16, round to nearest even   */
    PXL_STORE_F16 = 2     /* IEEE binary16, round to nearest even */
};

/* Variance-loss (TV) types of RenderUtils::addVarianceLoss
 * (RenderUtils.cu:2815-2851). */
enum {
    PXL_TV_NEIGHBOUR = 0,   /* totalVarianceLoss, 6-neighbour branch      */
    PXL_TV_DISPARITY = 1,   /* addVarianceLossGPU, per-cell mean disparity */
    PXL_TV_BOTH = 2         /* totalVarianceLoss over the launch of type 1,
                               then addVarianceLossGPU (RenderUtils.cu:2840-2845) */
};

/* Process-wide library options (pxl_set_option / pxl_get_option).  Each
 * launch reads the current value; nothing is read from the environment. */
enum {
    /* Marcher of the fixed rule: PXL_MARCHER_GROUP (default, eight samples
     * per wave) or PXL_MARCHER_WAVE (one sample per wave, as the adaptive
     * rule's kernel).  Both give the reference's results. */
    PXL_OPT_MARCHER = 1,
    /* 1: the TV gathers always take their 64-bit address form (otherwise
     * only for grids beyond 2^24 records or 4 GiB); default 0. */
    PXL_OPT_TV_ADDR64 = 2,
    /* Largest grid (blocks) of the launches whose waves claim batch entries
     * (the persistent schedules and the adaptive rule's logged forward);
     * 0 (default): what is resident on the device.  A small cap makes every
     * claim path run at small batch sizes; results do not depend on it. */
    PXL_OPT_CLAIM_GRID = 3
};
enum { PXL_MARCHER_GROUP = 0, PXL_MARCHER_WAVE = 1 };

/* Pinhole camera, host memory.  Mirrors the fields of Camera that
 * initialRays reads (CameraUtils.cuh:4-84): c2wExternalMat, cameraCenter,
 * alpha and principle. */
typedef struct pxl_camera {
    double c2w[16];     /* row-major camera-to-world 4x4 */
    double centre[4];   /* camera centre; centre[3] = 1  */
    double fx, fy;      /* focal lengths in pixels       */
    double cx, cy;      /* principal point in pixels     */
} pxl_camera;

/* Integer bounding box in voxel units (objectScene::bbox_*,
 * SceneUtils.cuh:13-15). */
typedef struct pxl_bbox {
    int32_t x_min, x_max, y_min, y_max, z_min, z_max;
} pxl_bbox;

/* Ray set, structure of arrays as in Rays (Rays.cuh:7-56).  Geometry and
 * ground truth are written by pxl_rays_init; the five acc_* planes are the
 * forward outputs. */
typedef struct pxl_rays {
    float *ox, *oy, *oz;          /* origin (cameraCenter*_Device)      */
    float *dx, *dy, *dz;          /* unit direction                     */
    float *near_t, *far_t;        /* slab-test interval                 */
    uint8_t *gt_r, *gt_g, *gt_b;  /* ground-truth colour, 0..255        */
    uint8_t *gt_mask;             /* ground-truth mask, 0..255          */
    float *acc_r, *acc_g, *acc_b; /* rendered colour                    */
    float *acc_w;                 /* accumulated weights (opacity)      */
    float *acc_density;           /* sum of activated density samples   */
    int64_t n;                    /* number of rays in every plane      */
    int32_t white_bg;             /* composite onto white (WhiteBG:1)   */
    int32_t reserved;
} pxl_rays;

/* Voxel grid.  Dense: n_vox = W*H*D records in z-y-x order.  Sparse (the
 * reference's caged layout, SceneUtils.cu:516-640): n_vox = inCageNums
 * records, `index_offsets` maps a dense vertex to its record,
 * `occupancy` is one bit per cell, `in_cage` one byte per dense vertex and
 * `compact_to_dense` maps a record back to its dense vertex.
 * `brick_occupancy` (optional, sparse only, as written by
 * pxl_build_brick_occupancy) has one bit per brick of 8x8x8 cells, set when
 * any cell in it is occupied, followed by its dilation; marching then skips
 * empty bricks without reading `occupancy`. */
typedef struct pxl_grid {
    int32_t W, H, D;              /* vertices per axis                 */
    int32_t sh_dim;               /* S = (band+1)^2: 1, 4, 9 or 16     */
    int32_t stride;               /* elements per record, >= 1+3S; a multiple of 4
                                     (FP32) or 8 (16-bit storage) */
    int32_t sparse;               /* 0 dense, 1 caged                  */
    int64_t n_vox;                /* number of records                 */
    float *params;                /* [n_vox][stride]                   */
    const uint32_t *occupancy;    /* bit c = cell z(W-1)(H-1)+y(W-1)+x */
    const uint32_t *index_offsets;/* [W*H*D]                           */
    const uint8_t *in_cage;       /* [W*H*D]                           */
    const uint32_t *compact_to_dense; /* [n_vox]                       */
    const uint32_t *brick_occupancy;  /* optional: bit per 8^3-cell brick */
    pxl_bbox bbox;
    float step;                   /* marching step in voxels (StepSize) */
    int32_t march_mode;           /* PXL_MARCH_*                       */
    int32_t storage;              /* PXL_STORE_*: element type of params */
    /* SH coefficients k < sh_active of each channel enter the colour; the
     * reference's (optimizingBand + 1)^2 (trilerpRGB, RenderUtils.cu:
     * 859-1000).  0 = all sh_dim. */
    int32_t sh_active;
} pxl_grid;

/* Optimiser state (Optimizer.cuh:5-37): FP32 records of `stride` floats
 * (0 = the grid's stride, FP32 grids only).  With 16-bit grid storage,
 * `master` (optional) holds the FP32 parameters Adam updates; the grid
 * then receives their rounded copy.  The gradient may use its own record
 * stride (`grad_stride`, 0 = `stride`): 32 floats per record put each row
 * the gradient kernels add to the planes in one 128-byte cache line, while
 * the moments Adam streams stay unpadded. */
typedef struct pxl_optim {
    float *grad;                  /* [n_vox][grad_stride] gradient     */
    float *m;                     /* [n_vox][stride] Adam first moment */
    float *v;                     /* [n_vox][stride] Adam second moment*/
    const float *adaptive_lr;     /* dense only: per-vertex LR factor or NULL */
    float *master;                /* 16-bit grids: FP32 parameters [n_vox][stride] or NULL */
    int32_t stride;               /* floats per record, multiple of 4  */
    /* Gradients, variance losses and Adam touch the SH coefficients k in
     * [sh_first, sh_active) of each channel, and the density: the
     * reference's band of the per-band strategy (hierarchyStrategy with
     * optimizingBand^2, RenderUtils.cu:1117-1132, Optimizer.cu:393-412);
     * 0 = from the first. */
    int32_t sh_first;
    int32_t grad_stride;          /* floats per gradient record, multiple of 4; 0 = stride */
} pxl_optim;

/* Last error message of the calling thread ("" if none). */
const char *pxl_last_error(void);

/* Library options (PXL_OPT_*): PXL_EINVAL for an unknown key or value.
 * Not tied to a stream: set them between calls, not during one.  No
 * counterpart in the reference. */
int32_t pxl_set_option(int32_t key, int64_t value);
int32_t pxl_get_option(int32_t key, int64_t *value);

/* Library version, (major << 16) | minor. */
int32_t pxl_version(void);

/* sizeof of pxl_camera, pxl_bbox, pxl_rays, pxl_grid, pxl_optim,
 * pxl_train_config, pxl_level_config, pxl_level and pxl_schedule, for
 * bindings to check their mirrors; returns 9. */
int32_t pxl_abi_sizes(int64_t *out, int32_t n);

/* Host-side camera of one pose (readFromMyBlender per view,
 * CameraUtils.cu:339-377; formulateProjectionMat :45-106; centre as in
 * transferSelf2GPU :577-608).  The pose translation is rescaled into voxel
 * units (t * object_centre / camera_distance + object_centre) when
 * need_rescale is set and |t| < 10; the rotation looks at object_centre
 * with +z up (setCameraLookAt, Utils.cu:448-486) and is flipped 180 degrees
 * about x when z_reversed.  w2c_out (optional) receives the 4x4
 * world-to-camera matrix.  No GPU work. */
int32_t pxl_camera_from_pose(const double translation[3], const double object_centre[3],
                             double camera_distance, int32_t need_rescale,
                             int32_t z_reversed, double fx, double fy, double cx, double cy,
                             pxl_camera *out, double w2c_out[16]);

/* Rays of one view (replaces the per-view initialRays launch in
 * _rays::initialSelf, Rays.cu:77-144 and 187-240; getBound Rays.cu:22-75).
 * Pixel (row, col) of the img_w x img_h image becomes ray
 * start_index + row*img_w + col.  img_* are device planes of img_w*img_h
 * bytes; all four may be NULL (ground truth then 0). */
int32_t pxl_rays_init(const pxl_camera *cam, const pxl_bbox *bbox,
                      int32_t img_w, int32_t img_h,
                      const uint8_t *img_r, const uint8_t *img_g,
                      const uint8_t *img_b, const uint8_t *img_mask,
                      int64_t start_index, const pxl_rays *rays, void *stream);

/* Mask-area rays of one view (_rays::initialSelf with MaskAreaOnly,
 * Rays.cu:187-240; initialRays :86-92; Image::processMasksInfo,
 * ImageUtils.cu:130-141): the pixels of the box rows pixel_box[0] ..
 * pixel_box[1], columns pixel_box[2] .. pixel_box[3] (inclusive, Image::bbox
 * order) in row-major order become rays start_index, start_index + 1, ...
 * Otherwise as pxl_rays_init. */
int32_t pxl_rays_init_box(const pxl_camera *cam, const pxl_bbox *bbox,
                          int32_t img_w, int32_t img_h, const int32_t pixel_box[4],
                          const uint8_t *img_r, const uint8_t *img_g,
                          const uint8_t *img_b, const uint8_t *img_mask,
                          int64_t start_index, const pxl_rays *rays, void *stream);

/* Bounding box of the nonzero pixels of a device mask (readMask,
 * Utils.cu:508-548, from the start values of Image::initSelf,
 * ImageUtils.cu:12-15): bounds (device int32[4]) = row min, row max, column
 * min, column max.  An all-zero mask leaves (img_h - 1, 0, img_w - 1, 0),
 * which processMasksInfo rejects (ImageUtils.cu:132). */
int32_t pxl_mask_bounds(const uint8_t *mask, int32_t img_w, int32_t img_h,
                        int32_t *bounds, void *stream);

/* _rays::shuffleTo (Rays.cu:242-287): ray j of dst = ray perm[j] of src for
 * j < n, geometry and ground truth (the acc_* planes are not copied).  An
 * index outside [0, src->n) gives a zero ray. */
int32_t pxl_rays_permute(const pxl_rays *src, const pxl_rays *dst, const int32_t *perm,
                         int64_t n, void *stream);

/* ShuffleRank (Utils.cu:488-506), host memory: with regenerate, dest[i] = i
 * on [min_index, max_index); with shuffle, for i from max_index - 1 down to
 * min_index + 1 swaps dest[i] and dest[rand() % i + min_index].  Draws from
 * the C library's rand(): after the same srand() and rand() calls it gives
 * the reference's order on the same C library.  The draw can reach
 * min_index + max_index - 2, past the range, unless min_index <= 1: larger
 * values are rejected when shuffling (the reference uses 0).  No GPU work. */
int32_t pxl_shuffle_rank(int32_t *dest, int32_t max_index, int32_t regenerate,
                         int32_t min_index, int32_t shuffle);

/* Forward render of rays sample_index[start .. start+len)
 * (RenderUtils::renderRays, RenderUtils.cu:1919-1971).  In the render and
 * train entry points (pxl_render_*, pxl_train_rays*) and pxl_batch_order,
 * len == 0 is a no-op and sample_index (and out_index) may then be null. */
int32_t pxl_render_fwd(const pxl_grid *grid, const pxl_rays *rays,
                       const int32_t *sample_index, int64_t start,
                       int64_t len, void *stream);

/* Same as pxl_render_fwd and also writes, per batch position i, trace[2i] =
 * number of evaluated samples and trace[2i+1] = a 32-bit FNV-1a hash of the
 * (sample number, base cell x, y, z) sequence.  For index-exactness tests. */
int32_t pxl_render_fwd_trace(const pxl_grid *grid, const pxl_rays *rays,
                             const int32_t *sample_index, int64_t start,
                             int64_t len, uint32_t *trace, void *stream);

/* Loss, Huber clip, mask/logit term and gradient scatter of a batch whose
 * forward outputs are in rays->acc_* (RenderUtils::renderRays_withGradients_RGB,
 * RenderUtils.cu:2051-2278).  losses[i] is the per-ray L1 loss x 255/3 of
 * batch position i. Gradients are added into optim->grad. */
int32_t pxl_render_bwd_rgb(const pxl_grid *grid, const pxl_optim *optim,
                           const pxl_rays *rays, double *losses,
                           const int32_t *sample_index, int64_t start,
                           int64_t len, double mask_coef, float huber_delta,
                           void *stream);

/* Fused forward + backward of one batch: one launch, forward results kept
 * in FP64 registers between the two marches.  Writes the same outputs as
 * pxl_render_fwd followed by pxl_render_bwd_rgb. */
int32_t pxl_train_rays(const pxl_grid *grid, const pxl_optim *optim,
                       const pxl_rays *rays, double *losses,
                       const int32_t *sample_index, int64_t start,
                       int64_t len, double mask_coef, float huber_delta,
                       void *stream);

/* Variance (TV) loss gradient added into optim->grad
 * (RenderUtils::addVarianceLoss, RenderUtils.cu:2815-2851).
 * type PXL_TV_DISPARITY: addVarianceLossGPU (RenderUtils.cu:2469-2547),
 *   computed race-free as a per-vertex gather; SH terms only when
 *   optimize_step <= 1 (RenderUtils.cu:2493).
 * type PXL_TV_NEIGHBOUR: totalVarianceLoss 6-neighbour branch
 *   (RenderUtils.cu:2660-2763); neighbours must be 6.  Interior vertices
 *   (no face of the grid); a neighbour whose index (dense, or record when
 *   caged) is 0 counts as the vertex itself.  Caged grids keep the
 *   reference's test of in_cage at the vertex's record number
 *   (RenderUtils.cu:2669: the flag array is indexed by dense vertex); a
 *   vertex whose record is 0 adds into record 0 atomically.
 * type PXL_TV_BOTH: VarianceType 2 (RenderUtils.cu:2840-2845): the
 *   6-neighbour term over the columns of a (W-1) x (H-1) launch rounded up
 *   to 16 x 16 (columns y W + x below ceil((W-1)/16) ceil((H-1)/16) 256),
 *   then the disparity term.
 * SH terms cover the coefficients [optim->sh_first, grid->sh_active) of
 * each channel (startSHIdx .. (optimizingBand+1)^2, RenderUtils.cu:2480,
 * 2594).
 * scratch: sparse grids only - pxl_tv_scratch_bytes() bytes of device
 *   memory, 256-byte aligned, filled once by pxl_tv_prepare for the grid's
 *   cage (it holds the per-record weights of the corners that alias record
 *   0, the list of records whose weight is not 0, and each record's 27
 *   neighbour records, occupied-cell mask and multiplicities, at most about
 *   140 B per record); NULL / 0 for dense grids.  Prepare again after the
 *   cage changes.  Calls that share one scratch on one stream are safe;
 *   calls on different streams must not share it. */
size_t pxl_tv_scratch_bytes(const pxl_grid *grid);
int32_t pxl_tv_prepare(const pxl_grid *grid, void *scratch, size_t scratch_bytes,
                       void *stream);
int32_t pxl_add_variance_loss(const pxl_grid *grid, const pxl_optim *optim,
                              double lambda_density, double lambda_sh,
                              int32_t type, int32_t neighbours,
                              int32_t optimize_step, void *scratch,
                              size_t scratch_bytes, void *stream);

/* Adam step over every record, then gradient := 0
 * (Optimizer::updateRGBScene / updateRGBSceneGPU_ADAM, Optimizer.cu:359-479).
 * t is the 1-based time step; density is clamped at 0. */
int32_t pxl_adam_step(const pxl_grid *grid, const pxl_optim *optim,
                      double lr, double sh_lr, int32_t t, void *stream);

/* pxl_add_variance_loss of type PXL_TV_DISPARITY followed by pxl_adam_step,
 * as one pass over the records: the TV sums of the records in
 * grid->params enter each record's gradient and its Adam step, and the
 * stepped records go to params_out, so that no lane reads a neighbour that
 * has already been stepped.  Fuses addVarianceLossGPU (RenderUtils.cu:
 * 2469-2547) and updateRGBSceneGPU_ADAM (Optimizer.cu:359-415), which the
 * reference runs one after the other (Render.cu:471, 492).
 * params_out: n_vox x stride elements of the grid's storage type, not
 *   overlapping grid->params; on return (stream-ordered) it holds what
 *   grid->params would hold after the two separate calls, bit for bit,
 *   and grid->params is unchanged.  The caller then uses params_out as the
 *   grid's records (the trainer swaps two buffers).
 * grad := 0 and m, v (and the FP32 master of 16-bit grids) are updated in
 * place as by pxl_adam_step.  Records of at most 32 elements (64 when
 * sh_dim > 10); scratch as for pxl_add_variance_loss. */
int32_t pxl_tv_adam_step(const pxl_grid *grid, const pxl_optim *optim,
                         double lambda_density, double lambda_sh,
                         int32_t optimize_step, void *scratch,
                         size_t scratch_bytes, double lr, double sh_lr,
                         int32_t t, void *params_out, void *stream);

/* Cell occupancy and in-cage flags from a dense raw-density plane
 * (objectScene::pruneSelf / pruneScene / occupiedVoxel,
 * SceneUtils.cu:800-913).  occupancy: ceil((W-1)(H-1)(D-1)/32) words.
 * in_cage: W*H*D bytes, cleared by this call.  n_occupied: optional device
 * uint64 receiving the number of occupied cells. */
int32_t pxl_build_occupancy(const float *density, int32_t W, int32_t H,
                            int32_t D, float threshold, uint32_t *occupancy,
                            uint8_t *in_cage, uint64_t *n_occupied,
                            void *stream);

/* Brick map of a cell bitfield: bit b of brick (bx, by, bz) =
 * bz*nbx*nby + by*nbx + bx, nbx = ceil((W-1)/8) (likewise y, z), set when
 * any of its cells is occupied.  bricks: pxl_brick_words(W, H, D) words =
 * 2 * ceil(nbx*nby*nbz / 32): the map, then the map dilated by one brick
 * (bit b set when brick b or one of its 26 neighbours has an occupied
 * cell), from which the fixed-rule marcher skips empty space 8 cells per
 * position probed. */
int64_t pxl_brick_words(int32_t W, int32_t H, int32_t D);
int32_t pxl_build_brick_occupancy(const uint32_t *occupancy, int32_t W, int32_t H,
                                  int32_t D, uint32_t *bricks, void *stream);

/* Sparse index: index_offsets[v] = number of in-cage vertices before v if
 * v is in cage, else 0; *in_cage_nums = total (objectScene::getCagedNums,
 * SceneUtils.cu:1081-1120).  compact_to_dense (optional) receives the
 * inverse map.  in_cage must be 16-byte aligned. */
size_t pxl_sparse_index_scratch_bytes(int64_t n_vertices);
int32_t pxl_build_sparse_index(const uint8_t *in_cage, int32_t W, int32_t H,
                               int32_t D, uint32_t *index_offsets,
                               uint32_t *compact_to_dense,
                               uint32_t *in_cage_nums, void *scratch,
                               size_t scratch_bytes, void *stream);

/* Compaction of dense records into caged order (the host loop in
 * objectScene::initSelfByUpsample, SceneUtils.cu:583-635). */
int32_t pxl_grid_compact(const float *dense_records, int32_t stride,
                         const uint32_t *compact_to_dense, int64_t n_vox,
                         float *sparse_records, void *stream);

/* Coarse-to-fine upsampling (objectScene::initSelfByUpsample,
 * SceneUtils.cu:487-657, through upsampleScene + upsample_trilinear,
 * SceneUtils.cu:250-485).  Destination vertex (x, y, z) samples the source
 * at (x W / W2, y H / H2, z D / D2) in float, trilinearly, in the
 * reference's corner order and float accumulation; when the source is
 * caged, a sample in an unoccupied cell gives zeros and corners outside the
 * cage read record 0.  Density below `threshold` becomes 0.
 *
 * pxl_upsample_density writes that density for every vertex of the
 * W2 x H2 x D2 grid (the dense temporary the reference prunes, :540-546);
 * pxl_build_occupancy(density, ..., threshold) and pxl_build_sparse_index
 * then give the caged layout of the finer grid.
 *
 * pxl_upsample fills dst->params (FP32 records, stride <= 64), dense or
 * caged (record r is vertex compact_to_dense[r]).  low_band_only carries
 * only SH coefficient 0 of each channel (the reference's lowBoundOnly flag);
 * dst->sh_dim must hold the carried coefficients; every other element of a
 * written record, padding included, becomes 0 (initSelf's zero start). */
int32_t pxl_upsample_density(const pxl_grid *src, int32_t W2, int32_t H2, int32_t D2,
                             float threshold, float *density, void *stream);
int32_t pxl_upsample(const pxl_grid *src, const pxl_grid *dst, float threshold,
                     int32_t low_band_only, void *stream);

/* Text formats (host code; grid functions copy to or from the device and
 * synchronise `stream`).
 *
 * pxl_read_cameras: Camera::readFromMyBlender (CameraUtils.cu:271-386) -
 * "total N", the 3x3 intrinsics (rescaled when the principal point is not
 * the centre of target_w x target_h), then per view "i:", the translation
 * and a 4x4 matrix the reference reads and ignores; each view becomes
 * pxl_camera_from_pose(translation, object_centre, camera_distance,
 * need_rescale, z_reversed, fx, fy, cx, cy).  camera_nums != 0 overrides N.
 * With out == NULL only *count is set.
 *
 * pxl_read_image_text: Image::initSelfFromFile + readMask
 * (ImageUtils.cu:84-128, Utils.cu:508-572): host planes from
 * prefix_mask.txt and prefix_red/_green/_blue.txt (prefix.txt into r when
 * gray_only), integers per pixel in row-major order; bounds = row min, row
 * max, column min, column max of the nonzero mask pixels, from
 * (height - 1, 0, width - 1, 0).
 *
 * pxl_grid_save_text: objectScene::saveSelf (SceneUtils.cu:1246-1299) into
 * an existing directory: occupancy.txt, opacity_band<b>.txt or the caged
 * sparse_opacity_band<b>.txt, sparse_inCageInfo_band<b>.txt and
 * sparse_indexOffset.txt, and the SH files <colour>_<i>[_sparse].txt for
 * i < (band + 1)^2 (from band^2 with hierarchy_mode; none with
 * density_only); "%.6lf " per float, "%d " per integer, a row per depth
 * slice (dense) or per record (caged).  The caged record count goes to
 * InCageNumsInfo.txt; the reference writes it over the in-cage file
 * instead (SceneUtils.cu:1279-1283).
 *
 * pxl_grid_text_records: records of a saved grid (W*H*D dense; caged: the
 * count of InCageNumsInfo.txt, else the first number of the in-cage file).
 *
 * pxl_grid_load_text: objectScene::initFromFile (SceneUtils.cu:703-790)
 * into caller-allocated device buffers: FP32 records [n][stride] (elements
 * without a file set to 0), occupancy bits (have_occupancy), and for a
 * caged grid in_cage, index_offsets and compact_to_dense.  In-cage flags
 * come from their file, or, when it holds only the count (a directory the
 * reference saved), from occupancy.txt by the pruning rule; the index file
 * must agree with them. */
int32_t pxl_read_cameras(const char *path, const double object_centre[3], int32_t target_w,
                         int32_t target_h, int32_t camera_nums, double camera_distance,
                         int32_t need_rescale, int32_t z_reversed, pxl_camera *out,
                         int32_t capacity, int32_t *count);
int32_t pxl_read_image_text(const char *prefix, int32_t width, int32_t height, int32_t gray_only,
                            uint8_t *r, uint8_t *g, uint8_t *b, uint8_t *mask,
                            int32_t bounds[4]);
int32_t pxl_grid_save_text(const pxl_grid *g, const char *dir, int32_t band,
                           int32_t density_only, int32_t hierarchy_mode, void *stream);
int32_t pxl_grid_text_records(const char *dir, int32_t band, int32_t sparse, int32_t W,
                              int32_t H, int32_t D, int64_t *n_records);
int32_t pxl_grid_load_text(const char *dir, int32_t band, int32_t sparse, int32_t have_occupancy,
                           int32_t W, int32_t H, int32_t D, int32_t sh_dim, int32_t stride,
                           int64_t n_records, float *params, uint32_t *occupancy,
                           uint8_t *in_cage, uint32_t *index_offsets,
                           uint32_t *compact_to_dense, void *stream);

/* Reference plane layout <-> packed records.  `planes` is a host array of
 * n_planes device pointers to n_vox floats each (objectScene plane table,
 * SceneUtils.cu:148-171).  Import zero-fills the padding floats. */
int32_t pxl_grid_import_planes(const float *const *planes, int32_t n_planes,
                               int64_t n_vox, float *records, int32_t stride,
                               void *stream);
int32_t pxl_grid_export_planes(const float *records, int32_t stride,
                               int64_t n_vox, float *const *planes,
                               int32_t n_planes, void *stream);

/* FP32 records (stride32 floats) -> 16-bit records (stride16 elements,
 * storage PXL_STORE_BF16 or PXL_STORE_F16), round to nearest even; the
 * first min(stride32, stride16) elements are converted, the rest of each
 * 16-bit record is zeroed. */
int32_t pxl_records_to_16bit(const float *rec32, int32_t stride32, int64_t n_vox,
                             void *rec16, int32_t stride16, int32_t storage, void *stream);

/* FP64 batch loss on the GPU: out[0] = mean(losses^2) (getMSEError,
 * Utils.cu:311-320), out[1] = 10 log10(65025 / out[0]) (Render.cu:466). */
int32_t pxl_loss_reduce_f64(const double *losses, int64_t len, double *out,
                            void *stream);

/* Number of distinct records touched by the gathers of a batch (untimed
 * bookkeeping for the roofline's algorithmic bytes).  bitmap: caller
 * scratch of ceil(n_vox/32) words, cleared by this call; *count: device
 * uint64. */
int32_t pxl_count_touched(const pxl_grid *grid, const pxl_rays *rays,
                          const int32_t *sample_index, int64_t start,
                          int64_t len, uint32_t *bitmap, uint64_t *count,
                          void *stream);

/* The fused step as two kernels with a sample log in between (same
 * arguments and results as pxl_train_rays).  The marcher renders each ray
 * and writes, per sample, its 8 corner records and raw density / colours
 * into `workspace`; a second kernel then scatters the gradient of every
 * logged ray with no global load between its atomics.  A ray with more than
 * max_samples evaluated samples is differentiated by the marcher itself.
 * Grids the eight-sample marcher does not serve (adaptive rule, S = 16) run
 * as pxl_train_rays.  workspace: 256-byte aligned device memory of at least
 * pxl_train_rays_workspace_bytes(len, max_samples) bytes (every schedule
 * below except the brick-binned gradient); calls sharing a workspace must
 * not overlap.  No counterpart in the reference (an
 * implementation of renderRays_withGradients_RGB, RenderUtils.cu:2257-2278,
 * fused with renderRays). */
size_t pxl_train_rays_workspace_bytes(int64_t len, int32_t max_samples);
/* Workspace bytes of the logged step on `grid` under every gradient
 * schedule, the brick-binned one (PXL_GRAD_BINNED) included: its bin arrays
 * (8 bytes per bin key, 8 keys per brick of 8^3 vertices) and one SH basis
 * of 9 floats per batch entry come on top of the sample log.  0 for bad
 * sizes.  No counterpart in the reference. */
size_t pxl_train_rays_workspace_bytes_grid(const pxl_grid *grid, int64_t len,
                                           int32_t max_samples);

/* Launch schedule of the two-kernel step: waves (rays) per workgroup, 1, 2
 * or 4, of the forward + log kernel and (grad_waves_per_block, 0 = the same)
 * of the per-ray gradient kernels; persistent != 0 launches as many
 * workgroups as fit on the device at once and lets each wave claim batch
 * entries (XCD-sliced claim counters in the workspace) until the batch is
 * done, for batches of very uneven ray lengths.  gradient selects how the
 * gradient of the logged samples reaches the planes:
 *   PXL_GRAD_PER_RAY  one wave per ray walks its samples and adds the
 *                     corners each cell change leaves with float atomics;
 *   PXL_GRAD_CHUNKED  (not with persistent) the same walk as one wave per
 *                     64-sample chunk of each ray's log after a pass that
 *                     stores each chunk's FP64 prefix, for batches too small
 *                     to fill the device (batches with proposal targets keep
 *                     one wave per ray);
 *   PXL_GRAD_BINNED   the samples' gradient scalars are binned by brick of
 *                     8^3 vertices and one workgroup per brick sums them in
 *                     LDS and updates each touched record once, without
 *                     float atomics (workspace:
 *                     pxl_train_rays_workspace_bytes_grid; grids of 2 to
 *                     1025 vertices per side).
 * The results do not depend on the schedule beyond the order of float
 * additions. */
enum {
    PXL_GRAD_PER_RAY = 0,
    PXL_GRAD_CHUNKED = 1,
    PXL_GRAD_BINNED = 2
};
typedef struct pxl_schedule {
    int32_t waves_per_block;
    int32_t persistent;
    int32_t gradient;            /* PXL_GRAD_* */
    int32_t grad_waves_per_block;
} pxl_schedule;
int32_t pxl_train_rays_ws(const pxl_grid *grid, const pxl_optim *optim,
                          const pxl_rays *rays, double *losses,
                          const int32_t *sample_index, int64_t start, int64_t len,
                          double mask_coef, float huber_delta, void *workspace,
                          size_t workspace_bytes, int32_t max_samples, void *stream);
/* pxl_train_rays_ws with an explicit schedule (pxl_train_rays_ws uses 4
 * waves per block, not persistent).  No counterpart in the reference. */
int32_t pxl_train_rays_sched(const pxl_grid *grid, const pxl_optim *optim,
                             const pxl_rays *rays, double *losses,
                             const int32_t *sample_index, int64_t start, int64_t len,
                             double mask_coef, float huber_delta, void *workspace,
                             size_t workspace_bytes, int32_t max_samples,
                             const pxl_schedule *sched, void *stream);

/* pxl_train_rays_sched without a workspace argument and without any
 * allocation: the sample log lives in an array of PXL_NOALLOC_SCRATCH_BYTES
 * (64 MiB) of device memory that the library holds on every GPU it runs on
 * until the process ends (its address is looked up on the first call on a
 * device).  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops).  Rejects (PXL_EINVAL) a batch whose log would exceed the array:
 * pxl_train_rays_workspace_bytes(len, max_samples) must be at most
 * PXL_NOALLOC_SCRATCH_BYTES (for example 4096 rays x 240 samples).  Every
 * call on one GPU uses the same array, which keeps the last call's
 * contents: calls that overlap on one GPU (other streams, other threads)
 * race and must be serialised by the caller.  sched may be NULL (4 waves per
 * block).  Calls no hipMalloc / hipFree. */
#define PXL_NOALLOC_SCRATCH_BYTES (64u << 20)
int32_t pxl_train_rays_noalloc(const pxl_grid *grid, const pxl_optim *optim,
                               const pxl_rays *rays, double *losses,
                               const int32_t *sample_index, int64_t start, int64_t len,
                               double mask_coef, float huber_delta, int32_t max_samples,
                               const pxl_schedule *sched, void *stream);

/* Batch orders (pxl_batch_order). */
enum {
    PXL_ORDER_MORTON = 1,        /* Morton code of the grid entry point */
    PXL_ORDER_LENGTH_MORTON = 2  /* longest-first length buckets, Morton inside */
};

/* The ray ids sample_index[start .. start+len) in the given order, written
 * to out_index[0 .. len) (use it as the sample index of the train call with
 * start 0).  The same rays as the reference's batch (Render.cu:399-406), in
 * an order that keeps neighbouring rays on one XCD and, for uneven batches,
 * starts the longest rays first; rays that miss the grid box go last.
 * scratch: 256-byte aligned, pxl_batch_order_scratch_bytes(len) bytes.  No
 * counterpart in the reference. */
size_t pxl_batch_order_scratch_bytes(int64_t len);
int32_t pxl_batch_order(const pxl_grid *grid, const pxl_rays *rays,
                        const int32_t *sample_index, int64_t start, int64_t len,
                        int32_t mode, int32_t *out_index, void *scratch,
                        size_t scratch_bytes, void *stream);

/* ------------------------------------------------------------------ */
/* Learned proposal (no counterpart in the reference): a per-ray MLP that */
/* predicts which of PXL_PROP_BINS equal bins of [t_near, t_far) hold     */
/* samples of visible weight.                                            */
/* ------------------------------------------------------------------ */

#define PXL_PROP_IN 128     /* 126 positional-encoding features + 2 zero */
#define PXL_PROP_WIDTH 64
#define PXL_PROP_BINS 64
/* Weights W1 [128][64], W2 [64][64], W3 [64][64] (row-major, input-major),
 * then biases b1, b2, b3 [64] each. */
#define PXL_PROP_WEIGHTS (PXL_PROP_IN * PXL_PROP_WIDTH + 2 * PXL_PROP_WIDTH * PXL_PROP_WIDTH)
#define PXL_PROP_PARAMS (PXL_PROP_WEIGHTS + 3 * PXL_PROP_WIDTH)

typedef struct pxl_proposal {
    uint16_t *weights;   /* PXL_PROP_WEIGHTS bfloat16 (the copy the kernels use) */
    float *master;       /* PXL_PROP_PARAMS FP32: weights then biases */
    float *m, *v;        /* Adam moments, PXL_PROP_PARAMS each */
    float *grad;         /* PXL_PROP_PARAMS: gradient of the last train step */
} pxl_proposal;

/* Workspace of the proposal calls for a batch of len rays (256-byte
 * aligned).  pxl_proposal_features fills its feature rows; forward and
 * train_step read them. */
size_t pxl_proposal_workspace_bytes(int64_t len);
/* He-uniform weights from `seed`, zero biases and moments. */
int32_t pxl_proposal_init(const pxl_proposal *p, uint64_t seed, void *stream);
/* Features of batch entries sample_index[start .. start+len): positional
 * encoding (10 frequencies) of the grid entry point o + max(t_near, 0) d,
 * mapped to [-1, 1]^3, and of the direction. */
int32_t pxl_proposal_features(const pxl_grid *grid, const pxl_rays *rays,
                              const int32_t *sample_index, int64_t start, int64_t len,
                              void *workspace, size_t workspace_bytes, void *stream);
/* Bin logits [len][64] (FP32) of the featured batch. */
int32_t pxl_proposal_forward(const pxl_proposal *p, int64_t len, float *logits,
                             void *workspace, size_t workspace_bytes, void *stream);
/* One training step on the featured batch: binary cross entropy of every
 * bin against bit j of targets[b] (rows with valid[b] == 0 ignored; valid
 * may be NULL), positives weighted pos_weight, mean over n_valid rows x 64
 * bins; gradient into p->grad; then Adam (lr, step t >= 1) when lr > 0.
 * loss_out (device, optional): the weighted BCE sum (FP64). */
int32_t pxl_proposal_train_step(const pxl_proposal *p, int64_t len, const uint64_t *targets,
                                const uint8_t *valid, int64_t n_valid, float pos_weight,
                                double lr, int32_t t, void *workspace, size_t workspace_bytes,
                                double *loss_out, void *stream);
/* Allowed-bin masks from logits: bit j set when sigmoid(logit j) >
 * threshold, dilated by `dilate` bins on each side. */
int32_t pxl_proposal_masks(const float *logits, int64_t len, float threshold, int32_t dilate,
                           uint64_t *masks, void *stream);

/* Eval render of one width x height image from `cam`
 * (RenderUtils::renderImage, RenderUtils.cu:1720-1818, RGB path): per pixel
 * an FP64 camera ray (getBound, RenderUtils.cu:34-97; pixel row i / width,
 * column i % height as the reference computes it), the adaptive marching
 * rule t += |step - w|, and the outputs of volumeRender_RGB
 * (RenderUtils.cu:1443-1580): rgb [3][h*w] float (white-composited when
 * white_bg), rgb8 [3][h*w] bytes (always white-composited, rounded), mask
 * [h*w] = A * 255 truncated, disparity [h*w] = sum w t.  Pixels whose ray
 * misses the grid box are white.  Any output may be NULL.  FP32 grids. */
int32_t pxl_render_image(const pxl_grid *grid, const pxl_camera *cam, int32_t width,
                         int32_t height, int32_t white_bg, float *rgb, uint8_t *rgb8,
                         uint8_t *mask, float *disparity, void *stream);

/* pxl_render_fwd with a proposal mask: batch entry b evaluates step k only
 * when bit floor(k step 64 / (t_far - t_near)) (at most 63) of
 * bin_mask[b] is set (bin_mask NULL: every step, as pxl_render_fwd).  With
 * trace != NULL, also the per-entry sample count and hash as
 * pxl_render_fwd_trace.  Eight-sample marcher only (fixed rule, sh_dim <= 9). */
int32_t pxl_render_fwd_masked(const pxl_grid *grid, const pxl_rays *rays,
                              const int32_t *sample_index, int64_t start, int64_t len,
                              const uint64_t *bin_mask, uint32_t *trace, void *stream);
/* pxl_train_rays_sched with a proposal mask (as above, may be NULL) and, when
 * bin_targets is not NULL, the proposal's training targets: per batch entry
 * the bins whose logged samples carry weight T alpha > target_wmin, and
 * target_valid[b] = 0 for rays longer than the log.  sched may be NULL. */
int32_t pxl_train_rays_prop(const pxl_grid *grid, const pxl_optim *optim,
                            const pxl_rays *rays, double *losses,
                            const int32_t *sample_index, int64_t start, int64_t len,
                            double mask_coef, float huber_delta, void *workspace,
                            size_t workspace_bytes, int32_t max_samples,
                            const pxl_schedule *sched, const uint64_t *bin_mask,
                            uint64_t *bin_targets, uint8_t *target_valid, double target_wmin,
                            void *stream);

/* Name of the kernel the render entry points launch for this grid:
 * mode 0 = pxl_render_fwd, 1 = pxl_render_bwd_rgb, 2 = pxl_train_rays,
 * 3 = pxl_train_rays_ws (two names joined by '+' when it runs as two kernels).
 * Returns a thread-local string such as "k_march8<9,sparse,fused,f32>" (or
 * NULL for a bad argument); for profilers and benchmark reports.  No
 * counterpart in the reference. */
const char *pxl_render_kernel_name(const pxl_grid *grid, int32_t mode);

/* ------------------------------------------------------------------ */
/* Trainer: the batch loop of Render::objectSceneLearner                 */
/* (Render.cu:352-503) as a host object driving the entry points above.  */
/* ------------------------------------------------------------------ */

/* The config_linux.txt keys the batch loop reads (BatchSize, LearningRate,
 * SHLearningRate, VarianceCoefficient, ColorDisparityCoefficient,
 * VarianceType, NeighboorNumsForTVLoss, VarianceLossInterval,
 * MaskRayCoefficient, HuberDelta; Utils.cu:765-932). */
typedef struct pxl_train_config {
    int64_t batch_size;
    double learning_rate;
    double sh_learning_rate;
    double variance_coefficient;
    double sh_coefficient;
    int32_t variance_type;
    int32_t neighbour_nums;
    int32_t variance_interval;
    int32_t optimize_step;   /* current hierarchy level (Render.cu:264) */
    double mask_coefficient;
    float huber_delta;
    int32_t fused;           /* 1: fused forward+backward launch per batch */
    int32_t log_samples;     /* fused only: > 0 runs the two-kernel step
                                (pxl_train_rays_sched) with a trainer-owned
                                workspace logging up to this many samples per
                                ray; 0 runs pxl_train_rays */
    int32_t autotune;        /* log_samples > 0: time every batch order x
                                schedule on the first batch, keep the fastest */
    int32_t batch_order;     /* 0 as given, PXL_ORDER_* (when not autotuned) */
    int32_t waves_per_block; /* 1, 2 or 4 (when not autotuned; 0 = 4) */
    int32_t persistent;      /* when not autotuned */
    int32_t gradient;        /* when not autotuned: PXL_GRAD_* (pxl_schedule);
                                PXL_GRAD_CHUNKED not with persistent */
    int32_t tv_inline;       /* 1: the variance loss runs on the trainer's
                                stream after the render (no overlap) */
    int32_t tv_separate;     /* 1: the disparity variance loss and Adam run
                                as two passes; 0: one fused pass
                                (pxl_tv_adam_step) on the batches that take
                                the variance loss, the trainer keeping a
                                second record buffer (pxl_trainer_sync) */
} pxl_train_config;

typedef struct pxl_trainer pxl_trainer;

/* Trainer over caller-owned buffers: rays order[0 .. n_rays) are visited in
 * batches of batch_size; losses holds batch_size doubles, loss_out 2;
 * tv_scratch as for pxl_add_variance_loss (already prepared by
 * pxl_tv_prepare for sparse grids).  Returns NULL on a bad argument (see
 * pxl_last_error). */
pxl_trainer *pxl_trainer_create(const pxl_train_config *cfg, const pxl_grid *grid,
                                const pxl_optim *optim, const pxl_rays *rays,
                                const int32_t *order, int64_t n_rays, double *losses,
                                double *loss_out, void *tv_scratch, size_t tv_scratch_bytes,
                                void *stream);

/* One batch: render + backward (fused or the two reference launches), FP64
 * batch loss, variance loss (every batch when dense, every
 * variance_interval-th when sparse, Render.cu:469) fused with Adam or
 * before it (tv_separate), Adam with time step
 * epoch*batchCount + batch + 1, LR decay lr /= 2^(1/(batchCount (epoch+1)))
 * (Render.cu:499-500).  Asynchronous.  A variance loss not fused with Adam:
 * with the logged step, once its schedule is chosen and without the
 * proposal, it runs on a second stream beside the forward kernel and the
 * gradient kernels wait for it (the same sum into the gradient planes in
 * another order); tv_inline
 * keeps it after the render on the trainer's stream.  The overlap is used
 * only where the render runs as the two-kernel logged step (the fixed rule
 * or the reference's adaptive rule, sh_dim 1, 4 or 9, the eight-sample
 * marcher).  The batch loss is reduced on that second stream too; the
 * trainer's stream waits for the reduction before the call returns its
 * work, so synchronising that stream makes losses and loss_out readable. */
int32_t pxl_trainer_step(pxl_trainer *tr);

/* With the fused TV + Adam pass (tv_separate 0, variance type
 * PXL_TV_DISPARITY, records of at most 32 elements, 64 when sh_dim > 10)
 * the trainer allocates a second record buffer at creation and the current
 * records alternate between it and the caller's grid->params after each
 * regulariser batch.  pxl_trainer_sync copies the current records into the
 * caller's buffer when they are in the other one (on the trainer's stream)
 * and makes the caller's current again; pxl_trainer_grid writes the
 * trainer's grid descriptor, current records included, to *out.  The
 * trainer syncs at the end of pxl_trainer_learn_level and in
 * pxl_trainer_destroy.  No counterpart in the reference, whose Adam steps
 * its planes in place. */
int32_t pxl_trainer_sync(pxl_trainer *tr);
int32_t pxl_trainer_grid(const pxl_trainer *tr, pxl_grid *out);

/* Record HIP events around the render launch of the next max_steps
 * batches; pxl_trainer_render_ms returns the average duration and count of
 * every timed batch and, when the pointers are not NULL, of the batches
 * whose render ran without the variance loss beside it (alone_ms,
 * alone_count; on dense grids every batch carries the variance loss). */
int32_t pxl_trainer_time_render(pxl_trainer *tr, int32_t max_steps);
int32_t pxl_trainer_render_ms(pxl_trainer *tr, double *avg_ms, int64_t *count);
int32_t pxl_trainer_render_ms_split(pxl_trainer *tr, double *avg_ms, int64_t *count,
                                    double *alone_ms, int64_t *alone_count);

/* Last batch's FP64 MSE and PSNR (synchronises); PXL_ENAN if not finite. */
int32_t pxl_trainer_loss(pxl_trainer *tr, double *mse, double *psnr);

/* Steps taken and current learning rates. */
int32_t pxl_trainer_state(const pxl_trainer *tr, int64_t *t, double *lr, double *sh_lr);
/* Turns on the learned proposal for a trainer running the logged step.
 * Every interval-th batch (every batch when prune == 0) runs unpruned,
 * takes the bins whose samples carry weight > target_wmin as targets and
 * makes one proposal training step (Adam lr, positives weighted
 * pos_weight); the other batches evaluate only the bins the proposal allows
 * (sigmoid > threshold, dilated by `dilate` bins).  The network in *p stays
 * owned by the caller; the trainer allocates its own buffers.  No
 * counterpart in the reference. */
int32_t pxl_trainer_set_proposal(pxl_trainer *tr, const pxl_proposal *p, int32_t interval,
                                 int32_t prune, float threshold, int32_t dilate, double lr,
                                 float pos_weight, double target_wmin);
/* The weighted BCE sum of the last proposal training batch (synchronises). */
int32_t pxl_trainer_proposal_loss(pxl_trainer *tr, double *bce_sum);

/* The batch order and launch schedule the trainer runs (after autotuning,
 * the chosen one); any pointer may be NULL. */
int32_t pxl_trainer_schedule(const pxl_trainer *tr, int32_t *batch_order,
                             pxl_schedule *schedule);

void pxl_trainer_destroy(pxl_trainer *tr);

/* ------------------------------------------------------------------ */
/* Per-band SH strategy and the coarse-to-fine level loop               */
/* ------------------------------------------------------------------ */

/* Optimiser state at the start of a band (optimizer::initialBand /
 * bandMemoryManage and optimizer::initialAll, Optimizer.cu:11-64,
 * 139-187), on the gradient and both Adam moments: band_mode 0 clears every
 * element; band_mode 1 clears SH coefficients [band^2, (band+1)^2) of each
 * channel, and the density too when band == 0 or !density_initialized (the
 * other elements keep their values, which the band's updates never read). */
int32_t pxl_optim_init_band(const pxl_grid *grid, const pxl_optim *optim, int32_t band,
                            int32_t band_mode, int32_t density_initialized, void *stream);

/* Switches a trainer to band `band` of its grid (Render.cu:282-315): the
 * colour uses coefficients k < (band+1)^2 (grid sh_active), gradients and
 * Adam touch [band^2, (band+1)^2) in band mode, else every coefficient
 * (optim sh_first 0); pxl_optim_init_band; learning rates back to their
 * configured values and the time step back to 1. */
int32_t pxl_trainer_set_band(pxl_trainer *tr, int32_t band, int32_t band_mode,
                             int32_t density_initialized);

/* The band settings of one level of Render::objectSceneLearner
 * (Render.cu:264-300; config keys SphericalHarmonicsBand,
 * HierarchyBandStrategy, SHsForTargetResolutionOnly, InitialFromFile,
 * AvailableSHBands, Utils.cu:765-932). */
typedef struct pxl_level_config {
    int32_t epochs;
    int32_t sh_band;                   /* the grid's band: sh_dim = (sh_band+1)^2 */
    int32_t hierarchy_band_strategy;
    int32_t sh_for_target_only;
    int32_t current_hierarchy;
    int32_t minimum_hierarchy;
    int32_t initial_from_file;
    int32_t upsampled;                 /* the grid came from the level before */
    int32_t available_sh_band;
    int32_t check_nan;                 /* 1: stop with PXL_ENAN after a NaN batch loss */
} pxl_level_config;

/* One level of the coarse-to-fine loop on a trainer (the band loop, its
 * epochs and batches of Render::objectSceneLearner, Render.cu:264-503):
 * bands from the start band (0, the file's next band, or sh_band; Render.cu:
 * 267-276) to sh_band; per band pxl_trainer_set_band with band mode when
 * hierarchy_band_strategy, or when sh_for_target_only below the minimum
 * hierarchy, else the whole optimiser cleared; epochs x batch_count
 * pxl_trainer_step calls.  With check_nan the batch loss is read after
 * every batch (one synchronisation each; the reference stops on NaN,
 * Render.cu:450-457), otherwise once at the end.  The grid's sh_dim must be
 * (sh_band+1)^2. */
int32_t pxl_trainer_learn_level(pxl_trainer *tr, const pxl_level_config *lc);

/* One level of the coarse-to-fine plan (trainSynthesisBlender,
 * VolumeFieldRender.cu:456-577). */
typedef struct pxl_level {
    int32_t level;       /* hierarchy index, coarsest first            */
    int32_t epochs;      /* epochs of the level (:459, 476-489)         */
    int64_t batch_size;  /* batch_size << (3 - level) unless fixed (:491-495) */
    int32_t upsample;    /* after the level: 0 none (last), 1 dense, 2 caged (:549-557) */
    int32_t reserved;
} pxl_level;

/* Levels hierarchy .. minimum_hierarchy of the plan: epochs start at
 * epochs + 3 - hierarchy and grow by one per level with shuffled rays (10
 * without); the batch doubles per level below 3 unless fixed (levels above
 * 3 with doubling are rejected: the reference's shift count would be
 * negative); the upsampling after level i > minimum is caged when
 * i <= minimum_hierarchy - to_sparse_hierarchy.  out == NULL: only *count. */
int32_t pxl_hierarchy_plan(int32_t hierarchy, int32_t minimum_hierarchy, int32_t epochs,
                           int64_t batch_size, int32_t shuffle_rays, int32_t fix_batch_size,
                           int32_t to_sparse_hierarchy, pxl_level *out, int32_t capacity,
                           int32_t *count);

#ifdef __cplusplus
}
#endif

#endif /* PXL_H */
