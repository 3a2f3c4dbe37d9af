// Regulariser gradients (variance / TV loss), the Adam step and the FP64
// batch-loss reduction.
#include "common.hpp"

namespace pxl {
namespace {

struct TvArgs {
    const float *params;
    float *grad;
    const uint32_t *occ;
    const uint32_t *io;
    const uint32_t *c2d;
    int W, H, D, stride, P;
    int sparse;
    int64_t n_vox;
    float lam_d, lam_sh;
};

__device__ __forceinline__ bool cell_occupied(const TvArgs &a, int cx, int cy, int cz) {
    const uint32_t cell = (uint32_t)cz * (uint32_t)(a.W - 1) * (uint32_t)(a.H - 1) +
                          (uint32_t)cy * (uint32_t)(a.W - 1) + (uint32_t)cx;
    return (a.occ[cell >> 5] >> (cell & 31)) & 1u;
}

__device__ __forceinline__ uint32_t record_of(const TvArgs &a, int x, int y, int z) {
    const uint32_t d = (uint32_t)z * (uint32_t)a.W * (uint32_t)a.H + (uint32_t)y * (uint32_t)a.W +
                       (uint32_t)x;
    return a.sparse ? a.io[d] : d;
}

// addVarianceLossGPU (RenderUtils.cu:2469-2547) as a race-free gather.  The
// reference visits every cell (every occupied cell when sparse), takes the
// mean mu of its 8 corners and adds lambda * (u_corner - mu) to each
// corner's gradient with unsynchronised +=.  Here each record gathers the
// same terms from the (up to 8) cells that have its vertex as a corner:
// one group of LPV lanes per record, one lane per parameter, 27-neighbour
// values loaded once.
template <int LPV>
__global__ __launch_bounds__(256) void k_tv_disparity(const TvArgs a) {
    constexpr int kGroups = 256 / LPV;
    const int g = threadIdx.x / LPV, p = threadIdx.x % LPV;
    const int64_t r = (int64_t)blockIdx.x * kGroups + g;
    if (r >= a.n_vox) return;
    if (p >= a.P) return;
    const float lam = p == 0 ? a.lam_d : a.lam_sh;
    if (lam == 0.0f) return;
    const uint32_t d = a.sparse ? a.c2d[r] : (uint32_t)r;
    const int x = (int)(d % (uint32_t)a.W);
    const int y = (int)((d / (uint32_t)a.W) % (uint32_t)a.H);
    const int z = (int)(d / ((uint32_t)a.W * (uint32_t)a.H));

    // Neighbour values u[(dz+1)*9 + (dy+1)*3 + (dx+1)].
    float u[27];
#pragma unroll
    for (int i = 0; i < 27; i++) {
        const int dx = i % 3 - 1, dy = (i / 3) % 3 - 1, dz = i / 9 - 1;
        const int nx = x + dx, ny = y + dy, nz = z + dz;
        float val = 0.0f;
        if (nx >= 0 && nx < a.W && ny >= 0 && ny < a.H && nz >= 0 && nz < a.D)
            val = a.params[(size_t)record_of(a, nx, ny, nz) * (size_t)a.stride + p];
        u[i] = val;
    }
    const float self = a.params[(size_t)r * (size_t)a.stride + p];
    float tsum = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const int ax = c & 1, ay = (c >> 1) & 1, az = c >> 2;
        const int cx = x - ax, cy = y - ay, cz = z - az;
        if (cx < 0 || cx > a.W - 2 || cy < 0 || cy > a.H - 2 || cz < 0 || cz > a.D - 2) continue;
        if (a.sparse && !cell_occupied(a, cx, cy, cz)) continue;
        float s = 0.0f;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int jx = j & 1, jy = (j >> 1) & 1, jz = j >> 2;
            s += u[(jz - az + 1) * 9 + (jy - ay + 1) * 3 + (jx - ax + 1)];
        }
        tsum += self - s * 0.125f;
    }
    a.grad[(size_t)r * (size_t)a.stride + p] += lam * tsum;
}

// Caged grids: a cell corner that is not in cage reads and writes record 0
// (index_offsets = 0, SceneUtils.cu:1106-1113).  Those (cell, corner) pairs
// are not found by the gather above; this kernel adds them into record 0.
// One thread per cell; waves without such a cell exit after one ballot.
__global__ __launch_bounds__(256) void k_tv_alias0(const TvArgs a, int64_t ncell) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int n_alias = 0;
    uint32_t rec[8];
    const uint32_t d0 = a.c2d[0];
    if (c < ncell) {
        const int64_t cw = a.W - 1, ch = a.H - 1;
        const int x = (int)(c % cw), y = (int)((c / cw) % ch), z = (int)(c / (cw * ch));
        if ((a.occ[c >> 5] >> (c & 31)) & 1u) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int jx = j & 1, jy = (j >> 1) & 1, jz = j >> 2;
                const uint32_t d = (uint32_t)(z + jz) * (uint32_t)a.W * (uint32_t)a.H +
                                   (uint32_t)(y + jy) * (uint32_t)a.W + (uint32_t)(x + jx);
                rec[j] = a.io[d];
                if (rec[j] == 0u && d != d0) n_alias++;
            }
        }
    }
    if (__ballot(n_alias > 0) == 0) return;
    for (int p = 0; p < a.P; p++) {
        const float lam = p == 0 ? a.lam_d : a.lam_sh;
        if (lam == 0.0f) continue;
        float contrib = 0.0f;
        if (n_alias > 0) {
            float s = 0.0f;
#pragma unroll
            for (int j = 0; j < 8; j++) s += a.params[(size_t)rec[j] * (size_t)a.stride + p];
            contrib = (float)n_alias * lam * (a.params[p] - s * 0.125f);
        }
        contrib = wave_sum_f(contrib);
        if ((threadIdx.x & 63) == 0 && contrib != 0.0f) unsafeAtomicAdd(a.grad + p, contrib);
    }
}

// totalVarianceLoss, 6-neighbour branch (RenderUtils.cu:2660-2763), dense
// grids: interior vertices, per parameter plane
// grad += lambda / (sqrt(sum (u - u_n)^2) / 6 + 1e-6) * sum (u - u_n);
// a neighbour whose linear index is 0 counts as the vertex itself.
template <int LPV>
__global__ __launch_bounds__(256) void k_tv_neighbour6(const TvArgs a) {
    constexpr int kGroups = 256 / LPV;
    const int g = threadIdx.x / LPV, p = threadIdx.x % LPV;
    const int64_t r = (int64_t)blockIdx.x * kGroups + g;
    if (r >= a.n_vox || p >= a.P) return;
    const float lam = p == 0 ? a.lam_d : a.lam_sh;
    if (lam == 0.0f) return;
    const int x = (int)(r % a.W), y = (int)((r / a.W) % a.H), z = (int)(r / ((int64_t)a.W * a.H));
    if (x == 0 || x == a.W - 1 || y == 0 || y == a.H - 1 || z == 0 || z == a.D - 1) return;
    const float self = a.params[(size_t)r * a.stride + p];
    const int off[6][3] = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    float sq = 0.0f, dis = 0.0f;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const uint32_t n = record_of(a, x + off[i][0], y + off[i][1], z + off[i][2]);
        const float un = n != 0u ? a.params[(size_t)n * a.stride + p] : self;
        sq += (self - un) * (self - un);
        dis += self - un;
    }
    const float s = sqrtf(sq) / 6.0f;
    a.grad[(size_t)r * a.stride + p] += (float)((double)lam / ((double)s + 1e-6) * (double)dis);
}

// updateRGBSceneGPU_ADAM (Optimizer.cu:359-415).  One thread per 4
// consecutive parameters of a record; the float/double mix of each update
// follows the reference expression by expression.
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ prm, float *__restrict__ grd,
                                              float *__restrict__ mom, float *__restrict__ var,
                                              const float *__restrict__ alr, int64_t n_vox,
                                              int stride, int P, int Q, double lr, double sh_lr,
                                              double bc1, double bc2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_vox * Q) return;
    const int64_t r = i / Q;
    const int q = (int)(i % Q);
    const size_t off = (size_t)r * stride + 4 * q;
    float4 th = *reinterpret_cast<const float4 *>(prm + off);
    float4 g4 = *reinterpret_cast<const float4 *>(grd + off);
    float4 m4 = *reinterpret_cast<const float4 *>(mom + off);
    float4 v4 = *reinterpret_cast<const float4 *>(var + off);
    float *tp = &th.x, *gp = &g4.x, *mp = &m4.x, *vp = &v4.x;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1.e-6f;
    const double lr_d = alr ? lr * (double)alr[r] : lr;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int p = 4 * q + e;
        if (p >= P) break;
        const float gg = gp[e];
        const float mm = (float)((double)(b1 * mp[e]) + (1.0 - (double)b1) * (double)gg);
        const float vv =
            (float)((double)(b2 * vp[e]) + (1.0 - (double)b2) * (double)gg * (double)gg);
        const float cm = (float)((double)mm / bc1);
        const float cv = (float)((double)vv / bc2);
        const double upd = (p == 0 ? lr_d : sh_lr) * (double)cm / (double)(sqrtf(cv) + eps);
        const double nt = (double)tp[e] - upd;
        tp[e] = p == 0 ? (nt < 0.0 ? 0.0f : (float)nt) : (float)nt;
        mp[e] = mm;
        vp[e] = vv;
        gp[e] = 0.0f;
    }
    *reinterpret_cast<float4 *>(prm + off) = th;
    *reinterpret_cast<float4 *>(grd + off) = g4;
    *reinterpret_cast<float4 *>(mom + off) = m4;
    *reinterpret_cast<float4 *>(var + off) = v4;
}

__global__ __launch_bounds__(256) void k_loss_reduce(const double *__restrict__ l, int64_t n,
                                                     double *__restrict__ out) {
    __shared__ double part[4];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += l[i] * l[i];
    s = wave_sum_d(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double tot = part[0] + part[1] + part[2] + part[3];
        const double mse = tot / (double)n;
        out[0] = mse;
        out[1] = 10.0 * log10(65025.0 / mse);
    }
}

int32_t tv_check(const pxl_grid *g, const pxl_optim *o, const char *who) {
    PXL_REQUIRE(g && o && o->grad && g->params, "%s: null argument", who);
    PXL_REQUIRE(g->W >= 2 && g->H >= 2 && g->D >= 2, "%s: grid too small", who);
    PXL_REQUIRE(g->sh_dim >= 1 && g->sh_dim <= 21, "%s: bad sh_dim %d", who, g->sh_dim);
    PXL_REQUIRE(g->stride >= 1 + 3 * g->sh_dim && g->stride % 4 == 0, "%s: bad stride %d", who,
                g->stride);
    PXL_REQUIRE(g->sparse || g->n_vox == (int64_t)g->W * g->H * g->D,
                "%s: dense grid needs n_vox = W*H*D", who);
    PXL_REQUIRE(!g->sparse || (g->occupancy && g->index_offsets && g->compact_to_dense),
                "%s: sparse grid needs occupancy, index_offsets and compact_to_dense", who);
    PXL_REQUIRE(g->n_vox > 0, "%s: empty grid", who);
    return PXL_OK;
}

}  // namespace
}  // namespace pxl

using namespace pxl;

extern "C" size_t pxl_tv_scratch_bytes(const pxl_grid *grid) {
    (void)grid;
    return 0;
}

extern "C" int32_t pxl_add_variance_loss(const pxl_grid *grid, const pxl_optim *optim,
                                         double lambda_density, double lambda_sh, int32_t type,
                                         int32_t neighbours, int32_t optimize_step,
                                         void *scratch, size_t scratch_bytes, void *stream) {
    (void)scratch;
    (void)scratch_bytes;
    clear_error();
    int32_t st = tv_check(grid, optim, "pxl_add_variance_loss");
    if (st != PXL_OK) return st;
    PXL_REQUIRE(type == PXL_TV_DISPARITY || type == PXL_TV_NEIGHBOUR,
                "pxl_add_variance_loss: unsupported type %d", type);
    TvArgs a;
    a.params = grid->params;
    a.grad = optim->grad;
    a.occ = grid->occupancy;
    a.io = grid->index_offsets;
    a.c2d = grid->compact_to_dense;
    a.W = grid->W;
    a.H = grid->H;
    a.D = grid->D;
    a.stride = grid->stride;
    a.P = 1 + 3 * grid->sh_dim;
    a.sparse = grid->sparse;
    a.n_vox = grid->n_vox;
    a.lam_d = (float)lambda_density;
    hipStream_t s = as_stream(stream);
    const bool wide = a.P > 32;
    if (type == PXL_TV_DISPARITY) {
        a.lam_sh = optimize_step <= 1 ? (float)lambda_sh : 0.0f;
        const int groups = wide ? 4 : 8;
        const unsigned blocks = (unsigned)((a.n_vox + groups - 1) / groups);
        if (wide) hipLaunchKernelGGL(k_tv_disparity<64>, dim3(blocks), dim3(256), 0, s, a);
        else hipLaunchKernelGGL(k_tv_disparity<32>, dim3(blocks), dim3(256), 0, s, a);
        if (a.sparse) {
            const int64_t ncell = (int64_t)(a.W - 1) * (a.H - 1) * (a.D - 1);
            hipLaunchKernelGGL(k_tv_alias0, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0,
                               s, a, ncell);
        }
    } else {
        PXL_REQUIRE(neighbours == 6, "pxl_add_variance_loss: only the 6-neighbour TV is built");
        PXL_REQUIRE(!grid->sparse, "pxl_add_variance_loss: 6-neighbour TV needs a dense grid");
        a.lam_sh = (float)lambda_sh;
        const int groups = wide ? 4 : 8;
        const unsigned blocks = (unsigned)((a.n_vox + groups - 1) / groups);
        if (wide) hipLaunchKernelGGL(k_tv_neighbour6<64>, dim3(blocks), dim3(256), 0, s, a);
        else hipLaunchKernelGGL(k_tv_neighbour6<32>, dim3(blocks), dim3(256), 0, s, a);
    }
    return check_launch("pxl_add_variance_loss");
}

extern "C" int32_t pxl_adam_step(const pxl_grid *grid, const pxl_optim *optim, double lr,
                                 double sh_lr, int32_t t, void *stream) {
    clear_error();
    PXL_REQUIRE(grid && optim && grid->params && optim->grad && optim->m && optim->v,
                "pxl_adam_step: null argument");
    PXL_REQUIRE(t >= 1, "pxl_adam_step: time step %d must be >= 1", t);
    PXL_REQUIRE(grid->stride >= 1 + 3 * grid->sh_dim && grid->stride % 4 == 0 && grid->sh_dim >= 1,
                "pxl_adam_step: bad stride %d / sh_dim %d", grid->stride, grid->sh_dim);
    PXL_REQUIRE(grid->n_vox >= 0, "pxl_adam_step: bad n_vox");
    PXL_REQUIRE(((uintptr_t)grid->params | (uintptr_t)optim->grad | (uintptr_t)optim->m |
                 (uintptr_t)optim->v) % 16 == 0,
                "pxl_adam_step: parameter and state buffers must be 16-byte aligned");
    if (grid->n_vox == 0) return PXL_OK;
    const int P = 1 + 3 * grid->sh_dim;
    const int Q = (P + 3) / 4;
    // Bias corrections with float pow as in the reference (Optimizer.cu:384-385).
    const double bc1 = 1.0 - (double)powf(0.9f, (float)t);
    const double bc2 = 1.0 - (double)powf(0.999f, (float)t);
    const int64_t total = grid->n_vox * Q;
    const float *alr = grid->sparse ? nullptr : optim->adaptive_lr;
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       as_stream(stream), grid->params, optim->grad, optim->m, optim->v, alr,
                       grid->n_vox, grid->stride, P, Q, lr, sh_lr, bc1, bc2);
    return check_launch("pxl_adam_step");
}

extern "C" int32_t pxl_loss_reduce_f64(const double *losses, int64_t len, double *out,
                                       void *stream) {
    clear_error();
    PXL_REQUIRE(losses && out, "pxl_loss_reduce_f64: null argument");
    PXL_REQUIRE(len > 0, "pxl_loss_reduce_f64: empty batch");
    hipLaunchKernelGGL(k_loss_reduce, dim3(1), dim3(256), 0, as_stream(stream), losses, len, out);
    return check_launch("pxl_loss_reduce_f64");
}
