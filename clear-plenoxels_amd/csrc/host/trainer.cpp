// The reference's operator API over the C ABI, and the batch loop of
// Render::objectSceneLearner (Render.cu:352-503) as a host object.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <new>
#include <vector>

#include "common.hpp"
#include "pxl_host.hpp"

namespace pxl {
namespace host {

int32_t RenderUtils::renderRays(const pxl_rays &rays, const pxl_grid &scene,
                                const int32_t *sampleIndex, int64_t startSampleIndex,
                                int64_t len, hipStream_t stream) {
    return pxl_render_fwd(&scene, &rays, sampleIndex, startSampleIndex, len, stream);
}

int32_t RenderUtils::renderRays_withGradients_RGB(const pxl_optim &optimizer,
                                                  const pxl_rays &rays, const pxl_grid &scene,
                                                  double *losses, const int32_t *sampleIndex,
                                                  int64_t startSampleIndex, int64_t len,
                                                  double maskRayLossCoefficient,
                                                  float huberDelta, hipStream_t stream) {
    return pxl_render_bwd_rgb(&scene, &optimizer, &rays, losses, sampleIndex, startSampleIndex,
                              len, maskRayLossCoefficient, huberDelta, stream);
}

int32_t RenderUtils::addVarianceLoss(const pxl_grid &scene, const pxl_optim &optimizer,
                                     double densityCoefficient, double shCoefficient,
                                     int32_t varianceType, int32_t neighbourNums,
                                     int32_t optimizeStep, void *scratch, size_t scratchBytes,
                                     hipStream_t stream) {
    return pxl_add_variance_loss(&scene, &optimizer, densityCoefficient, shCoefficient,
                                 varianceType, neighbourNums, optimizeStep, scratch,
                                 scratchBytes, stream);
}

int32_t Optimizer::updateRGBScene(const pxl_optim &optimizer, const pxl_grid &scene,
                                  double learningRate, double shLearningRate, int32_t timeStep,
                                  hipStream_t stream) {
    return pxl_adam_step(&scene, &optimizer, learningRate, shLearningRate, timeStep, stream);
}

}  // namespace host
}  // namespace pxl

struct pxl_trainer {
    pxl_train_config cfg;
    pxl_grid grid;
    pxl_optim optim;
    pxl_rays rays;
    const int32_t *order;
    int64_t n_rays;
    double *losses;
    double *loss_out;
    void *tv_scratch;
    size_t tv_scratch_bytes;
    hipStream_t stream;
    // The batch-loss reduction runs on a second stream (default priority),
    // after the render and beside TV and Adam; the next render waits for it
    // before it rewrites the losses.  side == nullptr: all on `stream`.
    hipStream_t side;
    hipEvent_t ev_rendered, ev_reduced;
    bool side_used;
    // Regulariser batches of the logged step: TV runs on the side stream
    // beside the forward kernel, after the previous batch's Adam (ev_step);
    // the gradient kernels wait for it (ev_tv).  cfg.tv_inline keeps TV on
    // `stream` after the render.
    hipEvent_t ev_step, ev_tv;
    bool tv_overlap;
    int64_t batch_count;
    int64_t t;
    double lr, sh_lr;
    // Optional HIP events around each batch's render launch, and whether
    // the variance loss ran beside that render.
    std::vector<hipEvent_t> ev;
    std::vector<char> ev_beside;
    int64_t ev_used;
    // Fused TV + Adam: the caller's records (params_home) and a second,
    // owned buffer of params_bytes; grid.params names the current one.
    bool tv_fused;
    float *params_home;
    void *params_alt;
    size_t params_bytes;
    // Sample-log workspace of the two-kernel fused step (owned).
    void *march_ws;
    size_t march_ws_bytes;
    // Batch order buffer and sort scratch (owned), and the schedule in use.
    int32_t *order_buf;
    void *order_ws;
    size_t order_ws_bytes;
    int32_t order_mode;
    pxl_schedule sched;
    bool tuned;
    bool binned_ok;  // the workspace holds the brick-binned gradient's arrays
    // Learned proposal (caller-owned network; trainer-owned buffers).
    bool prop_on;
    pxl_proposal prop;
    int32_t prop_interval, prop_dilate, prop_prune;
    float prop_threshold, prop_pos_weight;
    double prop_lr, prop_wmin;
    int32_t prop_t;
    void *prop_ws;
    size_t prop_ws_bytes;
    float *prop_logits;
    uint64_t *prop_masks, *prop_targets;
    uint8_t *prop_valid;
    double *prop_loss;
};

using namespace pxl;
using namespace pxl::host;

// Releases a trainer whose creation failed and keeps the error text that
// says why (the release itself clears the error state).
static pxl_trainer *create_failed(pxl_trainer *tr) {
    const std::string why = pxl_last_error();
    pxl_trainer_destroy(tr);
    set_error("%s", why.c_str());
    return nullptr;
}

extern "C" pxl_trainer *pxl_trainer_create(const pxl_train_config *cfg, const pxl_grid *grid,
                                           const pxl_optim *optim, const pxl_rays *rays,
                                           const int32_t *order, int64_t n_rays,
                                           double *losses, double *loss_out, void *tv_scratch,
                                           size_t tv_scratch_bytes, void *stream) {
    clear_error();
    if (!cfg || !grid || !optim || !rays || !order || !losses || !loss_out) {
        set_error("pxl_trainer_create: null argument");
        return nullptr;
    }
    if (cfg->batch_size <= 0 || n_rays <= 0 || n_rays > rays->n || cfg->variance_interval <= 0) {
        set_error("pxl_trainer_create: bad batch size %lld / ray count %lld / interval %d",
                  (long long)cfg->batch_size, (long long)n_rays, cfg->variance_interval);
        return nullptr;
    }
    pxl_trainer *tr = new (std::nothrow) pxl_trainer();
    if (!tr) {
        set_error("pxl_trainer_create: out of host memory");
        return nullptr;
    }
    tr->cfg = *cfg;
    tr->grid = *grid;
    tr->optim = *optim;
    tr->rays = *rays;
    tr->order = order;
    tr->n_rays = n_rays;
    tr->losses = losses;
    tr->loss_out = loss_out;
    tr->stream = as_stream(stream);
    tr->side = nullptr;
    tr->ev_rendered = tr->ev_reduced = tr->ev_step = tr->ev_tv = nullptr;
    tr->side_used = false;
    if (hipStreamCreateWithFlags(&tr->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&tr->ev_rendered, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&tr->ev_reduced, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&tr->ev_step, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&tr->ev_tv, hipEventDisableTiming) != hipSuccess) {
        for (hipEvent_t e : {tr->ev_rendered, tr->ev_reduced, tr->ev_step, tr->ev_tv})
            if (e) (void)hipEventDestroy(e);
        if (tr->side) (void)hipStreamDestroy(tr->side);
        tr->side = nullptr;
        tr->ev_rendered = tr->ev_reduced = tr->ev_step = tr->ev_tv = nullptr;
    }
    tr->params_home = grid->params;
    tr->params_alt = nullptr;
    tr->params_bytes = (size_t)grid->n_vox * (size_t)grid->stride *
                       (grid->storage == PXL_STORE_F32 ? 4u : 2u);
    const int32_t P = 1 + 3 * grid->sh_dim;
    tr->tv_fused = !cfg->tv_separate && cfg->variance_type == PXL_TV_DISPARITY &&
                   grid->stride <= (P > 32 ? 64 : 32) && tr->params_bytes > 0;
    if (tr->tv_fused && hipMalloc(&tr->params_alt, tr->params_bytes) != hipSuccess) {
        tr->params_alt = nullptr;
        set_error("pxl_trainer_create: cannot allocate a second %zu-byte record buffer for the "
                  "fused TV + Adam pass (tv_separate = 1 runs without it)",
                  tr->params_bytes);
        return create_failed(tr);
    }
    tr->tv_overlap = tr->side != nullptr && !cfg->tv_inline && !tr->tv_fused;
    tr->tv_scratch = tv_scratch;
    tr->tv_scratch_bytes = tv_scratch_bytes;
    tr->batch_count = (n_rays + cfg->batch_size - 1) / cfg->batch_size;
    tr->t = 0;
    tr->lr = cfg->learning_rate;
    tr->sh_lr = cfg->sh_learning_rate;
    tr->ev_used = 0;
    tr->march_ws = nullptr;
    tr->march_ws_bytes = 0;
    tr->binned_ok = false;
    tr->order_buf = nullptr;
    tr->order_ws = nullptr;
    tr->order_ws_bytes = 0;
    tr->order_mode = cfg->batch_order;
    tr->sched.waves_per_block = cfg->waves_per_block > 0 ? cfg->waves_per_block : 4;
    tr->sched.persistent = cfg->persistent;
    tr->sched.gradient = cfg->gradient;
    tr->sched.grad_waves_per_block = 0;
    tr->tuned = !cfg->autotune;
    tr->prop_on = false;
    tr->prop_ws = nullptr;
    tr->prop_logits = nullptr;
    tr->prop_masks = tr->prop_targets = nullptr;
    tr->prop_valid = nullptr;
    tr->prop_loss = nullptr;
    const bool logged = cfg->fused && cfg->log_samples > 0;
    if (!logged && (cfg->autotune || cfg->batch_order || cfg->persistent || cfg->gradient ||
                    (cfg->waves_per_block && cfg->waves_per_block != 4))) {
        set_error("pxl_trainer_create: batch orders and schedules need the logged step "
                  "(fused with log_samples > 0)");
        return create_failed(tr);
    }
    if (cfg->gradient == PXL_GRAD_CHUNKED && cfg->persistent) {
        set_error("pxl_trainer_create: the chunk-parallel gradient has no persistent form");
        return create_failed(tr);
    }
    if (cfg->gradient < PXL_GRAD_PER_RAY || cfg->gradient > PXL_GRAD_BINNED) {
        set_error("pxl_trainer_create: unknown gradient schedule %d", cfg->gradient);
        return create_failed(tr);
    }
    if (logged) {
        // Sized for every gradient schedule, the brick-binned one included
        // where the grid allows it.
        tr->march_ws_bytes =
            pxl_train_rays_workspace_bytes_grid(grid, cfg->batch_size, cfg->log_samples);
        tr->binned_ok = tr->march_ws_bytes >
                        pxl_train_rays_workspace_bytes(cfg->batch_size, cfg->log_samples);
        if (cfg->gradient == PXL_GRAD_BINNED && !tr->binned_ok) {
            set_error("pxl_trainer_create: the brick-binned gradient needs 2 to 1025 vertices "
                      "per side");
            return create_failed(tr);
        }
        if (tr->march_ws_bytes == 0 ||
            hipMalloc(&tr->march_ws, tr->march_ws_bytes) != hipSuccess) {
            set_error("pxl_trainer_create: cannot allocate a %zu-byte sample log for %d samples "
                      "per ray",
                      tr->march_ws_bytes, cfg->log_samples);
            return create_failed(tr);
        }
        if (cfg->autotune || cfg->batch_order) {
            tr->order_ws_bytes = pxl_batch_order_scratch_bytes(cfg->batch_size);
            if (hipMalloc(&tr->order_buf, (size_t)cfg->batch_size * sizeof(int32_t)) !=
                    hipSuccess ||
                hipMalloc(&tr->order_ws, tr->order_ws_bytes) != hipSuccess) {
                set_error("pxl_trainer_create: cannot allocate the batch-order buffers");
                return create_failed(tr);
            }
        }
    }
    return tr;
}

// The render call of one batch under a given order and schedule; with the
// proposal on, every prop_interval-th batch (or every batch when not
// pruning) runs unpruned, returns targets and trains the proposal, the
// others run with the proposal's bin masks.
static int32_t render_batch(pxl_trainer *tr, int64_t start, int64_t len, int32_t order,
                            const pxl_schedule &sched, bool use_prop = false,
                            hipEvent_t grad_wait = nullptr) {
    const pxl_train_config &c = tr->cfg;
    const int32_t *si = tr->order;
    int64_t s0 = start;
    if (order != 0) {
        const int32_t st = pxl_batch_order(&tr->grid, &tr->rays, tr->order, start, len, order,
                                           tr->order_buf, tr->order_ws, tr->order_ws_bytes,
                                           tr->stream);
        if (st != PXL_OK) return st;
        si = tr->order_buf;
        s0 = 0;
    }
    TrainExtra x;
    x.workspace = tr->march_ws;
    x.workspace_bytes = tr->march_ws_bytes;
    x.max_samples = c.log_samples;
    x.sched = &sched;
    x.grad_wait = grad_wait;
    if (!use_prop || !tr->prop_on)
        return train_rays_ex(&tr->grid, &tr->optim, &tr->rays, tr->losses, si, s0, len,
                             c.mask_coefficient, c.huber_delta, tr->stream, x);
    int32_t st = pxl_proposal_features(&tr->grid, &tr->rays, si, s0, len, tr->prop_ws,
                                       tr->prop_ws_bytes, tr->stream);
    if (st != PXL_OK) return st;
    const bool learn = !tr->prop_prune || tr->t % tr->prop_interval == 0;
    if (learn) {
        st = pxl_train_rays_prop(&tr->grid, &tr->optim, &tr->rays, tr->losses, si, s0, len,
                                 c.mask_coefficient, c.huber_delta, tr->march_ws,
                                 tr->march_ws_bytes, c.log_samples, &sched, nullptr,
                                 tr->prop_targets, tr->prop_valid, tr->prop_wmin, tr->stream);
        if (st != PXL_OK) return st;
        tr->prop_t++;
        return pxl_proposal_train_step(&tr->prop, len, tr->prop_targets, tr->prop_valid, len,
                                       tr->prop_pos_weight, tr->prop_lr, tr->prop_t, tr->prop_ws,
                                       tr->prop_ws_bytes, tr->prop_loss, tr->stream);
    }
    st = pxl_proposal_forward(&tr->prop, len, tr->prop_logits, tr->prop_ws, tr->prop_ws_bytes,
                              tr->stream);
    if (st == PXL_OK)
        st = pxl_proposal_masks(tr->prop_logits, len, tr->prop_threshold, tr->prop_dilate,
                                tr->prop_masks, tr->stream);
    if (st != PXL_OK) return st;
    return pxl_train_rays_prop(&tr->grid, &tr->optim, &tr->rays, tr->losses, si, s0, len,
                               c.mask_coefficient, c.huber_delta, tr->march_ws,
                               tr->march_ws_bytes, c.log_samples, &sched, tr->prop_masks,
                               nullptr, nullptr, 0.0, tr->stream);
}

extern "C" int32_t pxl_trainer_set_proposal(pxl_trainer *tr, const pxl_proposal *p,
                                            int32_t interval, int32_t prune, float threshold,
                                            int32_t dilate, double lr, float pos_weight,
                                            double target_wmin) {
    clear_error();
    PXL_REQUIRE(tr && p && p->weights && p->master && p->m && p->v && p->grad,
                "pxl_trainer_set_proposal: null argument");
    PXL_REQUIRE(tr->march_ws != nullptr,
                "pxl_trainer_set_proposal: needs the logged step (fused, log_samples > 0)");
    PXL_REQUIRE(interval >= 1 && dilate >= 0 && dilate < 64 && lr >= 0.0,
                "pxl_trainer_set_proposal: bad interval / dilate / lr");
    const int64_t B = tr->cfg.batch_size;
    if (!tr->prop_ws) {
        tr->prop_ws_bytes = pxl_proposal_workspace_bytes(B);
        if (hipMalloc(&tr->prop_ws, tr->prop_ws_bytes) != hipSuccess ||
            hipMalloc((void **)&tr->prop_logits, (size_t)B * 64 * sizeof(float)) != hipSuccess ||
            hipMalloc((void **)&tr->prop_masks, (size_t)B * sizeof(uint64_t)) != hipSuccess ||
            hipMalloc((void **)&tr->prop_targets, (size_t)B * sizeof(uint64_t)) != hipSuccess ||
            hipMalloc((void **)&tr->prop_valid, (size_t)B) != hipSuccess ||
            hipMalloc((void **)&tr->prop_loss, sizeof(double)) != hipSuccess) {
            set_error("pxl_trainer_set_proposal: cannot allocate the proposal buffers");
            return PXL_ENOMEM;
        }
    }
    tr->prop = *p;
    tr->prop_interval = interval;
    tr->prop_prune = prune;
    tr->prop_threshold = threshold;
    tr->prop_dilate = dilate;
    tr->prop_lr = lr;
    tr->prop_pos_weight = pos_weight;
    tr->prop_wmin = target_wmin;
    tr->prop_t = 0;
    tr->prop_on = true;
    return PXL_OK;
}

extern "C" int32_t pxl_trainer_proposal_loss(pxl_trainer *tr, double *bce_sum) {
    PXL_REQUIRE(tr && bce_sum && tr->prop_on, "pxl_trainer_proposal_loss: no proposal");
    if (hipMemcpyAsync(bce_sum, tr->prop_loss, sizeof(double), hipMemcpyDeviceToHost,
                       tr->stream) != hipSuccess ||
        hipStreamSynchronize(tr->stream) != hipSuccess) {
        set_error("pxl_trainer_proposal_loss: copy failed");
        return PXL_EHIP;
    }
    return PXL_OK;
}

// Times every batch order x schedule twice on the first batch (HIP events
// on the trainer's stream), keeps the fastest, and zeroes the gradient the
// trial runs added.
static int32_t autotune(pxl_trainer *tr, int64_t start, int64_t len) {
    struct Cand {
        int32_t order;
        pxl_schedule s;
    };
    std::vector<Cand> cands;
    // Waves per workgroup of the forward and (separately) of the gradient
    // kernel; persistent launches share one shape; the chunk-parallel
    // gradient with 4 waves per workgroup.  The brick-binned gradient is not
    // a candidate: it measured 3x (256^3 caged) to 10x (dense 128^3) slower
    // than the per-ray walk; cfg.gradient selects it without autotuning.
    for (int32_t order : {0, (int32_t)PXL_ORDER_MORTON, (int32_t)PXL_ORDER_LENGTH_MORTON})
        for (int32_t wpb : {1, 2, 4}) {
            for (int32_t gwpb : {1, 2, 4})
                cands.push_back({order, {wpb, 0, PXL_GRAD_PER_RAY, gwpb}});
            cands.push_back({order, {wpb, 1, PXL_GRAD_PER_RAY, 0}});
            cands.push_back({order, {wpb, 0, PXL_GRAD_CHUNKED, 4}});
        }
    constexpr int kReps = 2;
    std::vector<hipEvent_t> ev(cands.size() * kReps * 2, nullptr);
    int32_t st = PXL_OK;
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) st = PXL_EHIP;
    for (size_t i = 0; st == PXL_OK && i < cands.size() * kReps; i++) {
        const Cand &cd = cands[i % cands.size()];
        (void)hipEventRecord(ev[2 * i], tr->stream);
        st = render_batch(tr, start, len, cd.order, cd.s);
        (void)hipEventRecord(ev[2 * i + 1], tr->stream);
    }
    if (st == PXL_OK && hipStreamSynchronize(tr->stream) != hipSuccess) st = PXL_EHIP;
    if (st == PXL_OK) {
        std::vector<float> best(cands.size(), 1e30f);
        for (size_t i = 0; i < cands.size() * kReps; i++) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]);
            best[i % cands.size()] = std::min(best[i % cands.size()], ms);
        }
        size_t k = 0;
        for (size_t i = 1; i < cands.size(); i++)
            if (best[i] < best[k]) k = i;
        tr->order_mode = cands[k].order;
        tr->sched = cands[k].s;
        const int stride = grad_stride(&tr->grid, &tr->optim);
        if (hipMemsetAsync(tr->optim.grad, 0,
                           (size_t)tr->grid.n_vox * (size_t)stride * sizeof(float),
                           tr->stream) != hipSuccess)
            st = PXL_EHIP;
    }
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    if (st == PXL_EHIP) set_error("pxl_trainer_step: autotuning failed");
    tr->tuned = true;
    return st;
}

extern "C" int32_t pxl_trainer_time_render(pxl_trainer *tr, int32_t max_steps) {
    clear_error();
    PXL_REQUIRE(tr && max_steps >= 0, "pxl_trainer_time_render: bad argument");
    for (hipEvent_t e : tr->ev) (void)hipEventDestroy(e);
    tr->ev.assign((size_t)max_steps * 2, nullptr);
    tr->ev_beside.assign((size_t)max_steps, 0);
    for (auto &e : tr->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            set_error("pxl_trainer_time_render: hipEventCreate failed");
            return PXL_EHIP;
        }
    }
    tr->ev_used = 0;
    return PXL_OK;
}

// One batch (Render.cu:399-500).
extern "C" int32_t pxl_trainer_step(pxl_trainer *tr) {
    PXL_REQUIRE(tr != nullptr, "pxl_trainer_step: null trainer");
    clear_error();
    const pxl_train_config &c = tr->cfg;
    const int64_t b = tr->t % tr->batch_count;
    const int64_t epoch = tr->t / tr->batch_count;
    const int64_t start = b * c.batch_size;
    const int64_t len = std::min<int64_t>(c.batch_size, tr->n_rays - start);
    // Regulariser every batch when dense, every interval-th batch when
    // sparse (Render.cu:469).
    const bool do_tv = !tr->grid.sparse || b % c.variance_interval == 0;
    // With the logged step (schedule already chosen, no proposal) TV goes to
    // the side stream ahead of the render, and only the gradient kernels of
    // that step wait for it; its gradient then enters the planes before the
    // render's instead of after: the same sum in another order.  Grids the
    // logged step does not serve write gradients from the forward kernel
    // itself, so their TV stays on the trainer's stream.
    const bool tv_side = do_tv && tr->tv_overlap && c.fused && tr->march_ws && tr->tuned &&
                         !tr->prop_on && len > 0 && logged_step_serves(&tr->grid);
    const bool timed = tr->ev_used * 2 + 1 < (int64_t)tr->ev.size();
    int32_t st;
    if (tv_side) {
        if (hipEventRecord(tr->ev_step, tr->stream) != hipSuccess ||
            hipStreamWaitEvent(tr->side, tr->ev_step, 0) != hipSuccess) {
            set_error("pxl_trainer_step: event ordering failed");
            return PXL_EHIP;
        }
        st = RenderUtils::addVarianceLoss(tr->grid, tr->optim, c.variance_coefficient,
                                          c.sh_coefficient, c.variance_type, c.neighbour_nums,
                                          c.optimize_step, tr->tv_scratch,
                                          tr->tv_scratch_bytes, tr->side);
        if (st != PXL_OK) return st;
        if (hipEventRecord(tr->ev_tv, tr->side) != hipSuccess) {
            set_error("pxl_trainer_step: hipEventRecord failed");
            return PXL_EHIP;
        }
        tr->side_used = true;
    }
    if (c.fused && tr->march_ws && !tr->tuned) {
        st = autotune(tr, start, len);
        if (st != PXL_OK) return st;
    }
    if (timed) (void)hipEventRecord(tr->ev[tr->ev_used * 2], tr->stream);
    if (c.fused && tr->march_ws) {
        st = render_batch(tr, start, len, tr->order_mode, tr->sched, true,
                          tv_side ? tr->ev_tv : nullptr);
    } else if (c.fused) {
        st = pxl_train_rays(&tr->grid, &tr->optim, &tr->rays, tr->losses, tr->order, start, len,
                            c.mask_coefficient, c.huber_delta, tr->stream);
    } else {
        st = RenderUtils::renderRays(tr->rays, tr->grid, tr->order, start, len, tr->stream);
        if (st == PXL_OK)
            st = RenderUtils::renderRays_withGradients_RGB(tr->optim, tr->rays, tr->grid,
                                                           tr->losses, tr->order, start, len,
                                                           c.mask_coefficient, c.huber_delta,
                                                           tr->stream);
    }
    // Whatever the launch did, nothing after it on this stream touches the
    // gradient planes before TV is done.
    if (tv_side && hipStreamWaitEvent(tr->stream, tr->ev_tv, 0) != hipSuccess) {
        set_error("pxl_trainer_step: hipStreamWaitEvent failed");
        return PXL_EHIP;
    }
    if (st != PXL_OK) return st;
    if (timed) {
        (void)hipEventRecord(tr->ev[tr->ev_used * 2 + 1], tr->stream);
        tr->ev_beside[tr->ev_used] = tv_side ? 1 : 0;
        tr->ev_used++;
    }
    // Batch loss on the GPU in FP64 instead of a D2H copy and a host loop
    // (Render.cu:443-449), on the side stream once the render is done.
    if (tr->side) {
        if (hipEventRecord(tr->ev_rendered, tr->stream) != hipSuccess ||
            hipStreamWaitEvent(tr->side, tr->ev_rendered, 0) != hipSuccess) {
            set_error("pxl_trainer_step: event ordering failed");
            return PXL_EHIP;
        }
        st = pxl_loss_reduce_f64(tr->losses, len, tr->loss_out, tr->side);
        if (st != PXL_OK) return st;
        if (hipEventRecord(tr->ev_reduced, tr->side) != hipSuccess) {
            set_error("pxl_trainer_step: hipEventRecord failed");
            return PXL_EHIP;
        }
        tr->side_used = true;
    } else {
        st = pxl_loss_reduce_f64(tr->losses, len, tr->loss_out, tr->stream);
        if (st != PXL_OK) return st;
    }
    const int32_t time_step = (int32_t)(epoch * tr->batch_count + b + 1);
    if (do_tv && tr->tv_fused) {
        // One pass: the stepped records go to the other buffer, which then
        // becomes the grid's.
        void *other = tr->grid.params == tr->params_home ? tr->params_alt
                                                         : (void *)tr->params_home;
        st = pxl_tv_adam_step(&tr->grid, &tr->optim, c.variance_coefficient, c.sh_coefficient,
                              c.optimize_step, tr->tv_scratch, tr->tv_scratch_bytes, tr->lr,
                              tr->sh_lr, time_step, other, tr->stream);
        if (st != PXL_OK) return st;
        tr->grid.params = reinterpret_cast<float *>(other);
    } else {
        if (do_tv && !tv_side) {
            st = RenderUtils::addVarianceLoss(tr->grid, tr->optim, c.variance_coefficient,
                                              c.sh_coefficient, c.variance_type,
                                              c.neighbour_nums, c.optimize_step, tr->tv_scratch,
                                              tr->tv_scratch_bytes, tr->stream);
            if (st != PXL_OK) return st;
        }
        st = Optimizer::updateRGBScene(tr->optim, tr->grid, tr->lr, tr->sh_lr, time_step,
                                       tr->stream);
        if (st != PXL_OK) return st;
    }
    // The caller's stream holds all of this batch's work, the loss
    // reduction included (it ran beside TV and Adam): the next render, a
    // synchronisation of the stream or a read of losses / loss_out after it
    // see the finished reduction.
    if (tr->side && hipStreamWaitEvent(tr->stream, tr->ev_reduced, 0) != hipSuccess) {
        set_error("pxl_trainer_step: hipStreamWaitEvent failed");
        return PXL_EHIP;
    }
    // LR decay (Render.cu:499-500).
    const double decay = std::pow(2.0, 1.0 / (double)tr->batch_count / ((double)epoch + 1.0));
    tr->lr /= decay;
    tr->sh_lr /= decay;
    tr->t++;
    return PXL_OK;
}

// Average duration of the timed render launches; synchronises on them.
extern "C" int32_t pxl_trainer_render_ms_split(pxl_trainer *tr, double *avg_ms, int64_t *count,
                                               double *alone_ms, int64_t *alone_count) {
    PXL_REQUIRE(tr && avg_ms && count, "pxl_trainer_render_ms: null argument");
    double tot = 0.0, tot_alone = 0.0;
    int64_t n_alone = 0;
    for (int64_t i = 0; i < tr->ev_used; i++) {
        if (hipEventSynchronize(tr->ev[2 * i + 1]) != hipSuccess) {
            set_error("pxl_trainer_render_ms: hipEventSynchronize failed");
            return PXL_EHIP;
        }
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, tr->ev[2 * i], tr->ev[2 * i + 1]) != hipSuccess) {
            set_error("pxl_trainer_render_ms: hipEventElapsedTime failed");
            return PXL_EHIP;
        }
        tot += ms;
        if (!tr->ev_beside[i]) {
            tot_alone += ms;
            n_alone++;
        }
    }
    *count = tr->ev_used;
    *avg_ms = tr->ev_used ? tot / (double)tr->ev_used : 0.0;
    if (alone_ms) *alone_ms = n_alone ? tot_alone / (double)n_alone : 0.0;
    if (alone_count) *alone_count = n_alone;
    return PXL_OK;
}

extern "C" int32_t pxl_trainer_render_ms(pxl_trainer *tr, double *avg_ms, int64_t *count) {
    return pxl_trainer_render_ms_split(tr, avg_ms, count, nullptr, nullptr);
}

// Copies the last batch's FP64 MSE and PSNR to the host; PXL_ENAN when the
// loss is not finite (the reference aborts training, Render.cu:450-457).
extern "C" int32_t pxl_trainer_loss(pxl_trainer *tr, double *mse, double *psnr) {
    PXL_REQUIRE(tr && mse && psnr, "pxl_trainer_loss: null argument");
    double h[2];
    if ((tr->side_used && hipStreamWaitEvent(tr->stream, tr->ev_reduced, 0) != hipSuccess) ||
        hipMemcpyAsync(h, tr->loss_out, sizeof(h), hipMemcpyDeviceToHost, tr->stream) !=
            hipSuccess ||
        hipStreamSynchronize(tr->stream) != hipSuccess) {
        set_error("pxl_trainer_loss: copy failed");
        return PXL_EHIP;
    }
    *mse = h[0];
    *psnr = h[1];
    if (std::isnan(h[0])) {
        set_error("pxl_trainer_loss: batch loss is NaN");
        return PXL_ENAN;
    }
    return PXL_OK;
}

extern "C" int32_t pxl_trainer_state(const pxl_trainer *tr, int64_t *t, double *lr,
                                     double *sh_lr) {
    PXL_REQUIRE(tr != nullptr, "pxl_trainer_state: null trainer");
    if (t) *t = tr->t;
    if (lr) *lr = tr->lr;
    if (sh_lr) *sh_lr = tr->sh_lr;
    return PXL_OK;
}

extern "C" int32_t pxl_trainer_sync(pxl_trainer *tr) {
    PXL_REQUIRE(tr != nullptr, "pxl_trainer_sync: null trainer");
    clear_error();
    if (tr->params_alt == nullptr || tr->grid.params == tr->params_home) return PXL_OK;
    if (hipMemcpyAsync(tr->params_home, tr->grid.params, tr->params_bytes,
                       hipMemcpyDeviceToDevice, tr->stream) != hipSuccess) {
        set_error("pxl_trainer_sync: hipMemcpyAsync of %zu bytes failed", tr->params_bytes);
        return PXL_EHIP;
    }
    tr->grid.params = tr->params_home;
    return PXL_OK;
}

extern "C" int32_t pxl_trainer_grid(const pxl_trainer *tr, pxl_grid *out) {
    PXL_REQUIRE(tr != nullptr && out != nullptr, "pxl_trainer_grid: null argument");
    *out = tr->grid;
    return PXL_OK;
}

extern "C" void pxl_trainer_destroy(pxl_trainer *tr) {
    if (!tr) return;
    (void)pxl_trainer_sync(tr);
    for (hipEvent_t e : tr->ev) (void)hipEventDestroy(e);
    (void)hipStreamSynchronize(tr->stream);
    if (tr->side) {
        (void)hipStreamSynchronize(tr->side);
        for (hipEvent_t e : {tr->ev_rendered, tr->ev_reduced, tr->ev_step, tr->ev_tv})
            (void)hipEventDestroy(e);
        (void)hipStreamDestroy(tr->side);
    }
    if (tr->params_alt) (void)hipFree(tr->params_alt);
    if (tr->march_ws) (void)hipFree(tr->march_ws);
    if (tr->order_buf) (void)hipFree(tr->order_buf);
    if (tr->order_ws) (void)hipFree(tr->order_ws);
    for (void *q : {tr->prop_ws, (void *)tr->prop_logits, (void *)tr->prop_masks,
                    (void *)tr->prop_targets, (void *)tr->prop_valid, (void *)tr->prop_loss})
        if (q) (void)hipFree(q);
    delete tr;
}

extern "C" int32_t pxl_trainer_schedule(const pxl_trainer *tr, int32_t *batch_order,
                                        pxl_schedule *schedule) {
    PXL_REQUIRE(tr != nullptr, "pxl_trainer_schedule: null trainer");
    if (batch_order) *batch_order = tr->order_mode;
    if (schedule) *schedule = tr->sched;
    return PXL_OK;
}

// Band b of the per-band SH strategy (Render.cu:282-315): the colour uses
// the coefficients below (b+1)^2; band mode restricts gradients and Adam to
// the band's own coefficients; the optimiser state of the band starts at
// zero (pxl_optim_init_band); learning rates and the time step restart.
extern "C" int32_t pxl_trainer_set_band(pxl_trainer *tr, int32_t band, int32_t band_mode,
                                        int32_t density_initialized) {
    clear_error();
    PXL_REQUIRE(tr != nullptr, "pxl_trainer_set_band: null trainer");
    const int32_t S = tr->grid.sh_dim;
    PXL_REQUIRE(band >= 0 && (band + 1) * (band + 1) <= S,
                "pxl_trainer_set_band: band %d beyond the grid's %d coefficients", band, S);
    int32_t st = pxl_optim_init_band(&tr->grid, &tr->optim, band, band_mode, density_initialized,
                                     tr->stream);
    if (st != PXL_OK) return st;
    tr->grid.sh_active = (band + 1) * (band + 1);
    tr->optim.sh_first = band_mode ? band * band : 0;
    tr->lr = tr->cfg.learning_rate;
    tr->sh_lr = tr->cfg.sh_learning_rate;
    tr->t = 0;
    return PXL_OK;
}

extern "C" int32_t pxl_trainer_learn_level(pxl_trainer *tr, const pxl_level_config *lc) {
    clear_error();
    PXL_REQUIRE(tr && lc, "pxl_trainer_learn_level: null argument");
    PXL_REQUIRE(lc->epochs >= 0 && lc->sh_band >= 0 &&
                    (lc->sh_band + 1) * (lc->sh_band + 1) == tr->grid.sh_dim,
                "pxl_trainer_learn_level: sh_band %d does not match sh_dim %d", lc->sh_band,
                tr->grid.sh_dim);
    // First band (Render.cu:267-276).
    int32_t start_band = lc->initial_from_file ? (lc->upsampled ? 0 : lc->available_sh_band + 1) : 0;
    const bool below_target = lc->sh_for_target_only && lc->current_hierarchy != lc->minimum_hierarchy;
    if (!lc->hierarchy_band_strategy) start_band = below_target ? 0 : lc->sh_band;
    PXL_REQUIRE(start_band >= 0, "pxl_trainer_learn_level: bad start band %d", start_band);
    const bool band_mode = lc->hierarchy_band_strategy || below_target;
    for (int32_t band = start_band; band <= lc->sh_band; band++) {
        const bool density_initialized = !(lc->initial_from_file && band == lc->available_sh_band + 1);
        int32_t st;
        if (band_mode) {
            st = pxl_trainer_set_band(tr, band, 1, density_initialized ? 1 : 0);
        } else {
            // initialAll: every coefficient, every moment from zero.
            st = pxl_trainer_set_band(tr, band, 0, 1);
        }
        if (st != PXL_OK) return st;
        const int64_t steps = (int64_t)lc->epochs * tr->batch_count;
        for (int64_t i = 0; i < steps; i++) {
            st = pxl_trainer_step(tr);
            if (st != PXL_OK) return st;
            if (lc->check_nan || i + 1 == steps) {
                double mse = 0.0, psnr = 0.0;
                st = pxl_trainer_loss(tr, &mse, &psnr);
                if (st != PXL_OK) return st;
            }
        }
    }
    // The level's records end in the caller's buffer (upsampling and
    // export read them there).
    return pxl_trainer_sync(tr);
}

extern "C" int32_t pxl_hierarchy_plan(int32_t hierarchy, int32_t minimum_hierarchy,
                                      int32_t epochs, int64_t batch_size, int32_t shuffle_rays,
                                      int32_t fix_batch_size, int32_t to_sparse_hierarchy,
                                      pxl_level *out, int32_t capacity, int32_t *count) {
    clear_error();
    PXL_REQUIRE(count != nullptr, "pxl_hierarchy_plan: null count");
    PXL_REQUIRE(minimum_hierarchy <= hierarchy,
                "pxl_hierarchy_plan: minimum_hierarchy %d exceeds hierarchy %d",
                minimum_hierarchy, hierarchy);
    PXL_REQUIRE(fix_batch_size || hierarchy <= 3,
                "pxl_hierarchy_plan: batch doubling (batch << (3 - level)) needs levels <= 3");
    PXL_REQUIRE(minimum_hierarchy >= 0 || fix_batch_size || 3 - minimum_hierarchy < 62,
                "pxl_hierarchy_plan: level %d out of range", minimum_hierarchy);
    const int32_t n = hierarchy - minimum_hierarchy + 1;
    *count = n;
    if (!out) return PXL_OK;
    PXL_REQUIRE(capacity >= n, "pxl_hierarchy_plan: %d levels, capacity %d", n, capacity);
    // VolumeFieldRender.cu:459, 476-495, 549-557.
    int32_t ep = epochs + 3 - hierarchy;
    for (int32_t k = 0; k < n; k++) {
        const int32_t lvl = hierarchy - k;
        ep = shuffle_rays ? ep + 1 : 10;
        pxl_level &L = out[k];
        L.level = lvl;
        L.epochs = ep;
        L.batch_size = fix_batch_size ? batch_size : batch_size << (3 - lvl);
        L.upsample = lvl > minimum_hierarchy
                         ? (lvl <= minimum_hierarchy - to_sparse_hierarchy ? 2 : 1)
                         : 0;
        L.reserved = 0;
    }
    return PXL_OK;
}
