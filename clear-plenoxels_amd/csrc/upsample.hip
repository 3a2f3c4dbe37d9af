// Coarse-to-fine upsampling of a grid to a finer one: upsampleScene +
// upsample_trilinear (SceneUtils.cu:250-485), the step initSelfByUpsample
// (SceneUtils.cu:487-657) takes between hierarchy levels.  Destination
// vertex (x, y, z) samples the source at x * W / W2 (float), trilinearly,
// with the reference's corner order and float accumulation, so the result
// is the reference's to the bit.  A caged destination is filled record by
// record through compact_to_dense; its occupancy and index come from the
// thresholded density (pxl_upsample_density) through pxl_build_occupancy and
// pxl_build_sparse_index, as the reference prunes its dense temporary.
#include "common.hpp"

namespace pxl {
namespace {

struct UpArgs {
    const float *src;  // elements of the source storage type
    int sstride, sW, sH, sD, sS, ssparse;
    const uint32_t *socc, *sio;
    float *dst;        // FP32 records, or the density plane
    int dstride, dW, dH, dD, dS;
    const uint32_t *dc2d;  // caged destination: record -> dense vertex
    int64_t n;             // destination records (or vertices)
    float threshold;
    int ncopy;             // SH coefficients per channel carried over
};

// getIndex (SceneUtils.cu:250-260): the inputs are unsigned, so only the
// upper clamp applies.
__device__ __forceinline__ uint32_t up_index(uint32_t x, uint32_t y, uint32_t z, int W, int H,
                                             int D) {
    x = x < (uint32_t)(W - 1) ? x : (uint32_t)(W - 1);
    y = y < (uint32_t)(H - 1) ? y : (uint32_t)(H - 1);
    z = z < (uint32_t)(D - 1) ? z : (uint32_t)(D - 1);
    return z * (uint32_t)W * (uint32_t)H + y * (uint32_t)W + x;
}

struct UpSample {
    uint32_t b0, b1, b2;
    double fx, fy, fz;
    bool live;
};

// Sample position of destination vertex d (upsampleScene :428-433, :448)
// and the source cell test of upsample_trilinear (:270-278).
__device__ __forceinline__ UpSample up_locate(const UpArgs &a, uint32_t d) {
    const uint32_t x = d % (uint32_t)a.dW;
    const uint32_t y = (d / (uint32_t)a.dW) % (uint32_t)a.dH;
    const uint32_t z = d / ((uint32_t)a.dW * (uint32_t)a.dH);
    const float c0 = (float)x * 1.f * (float)a.sW / (float)a.dW;
    const float c1 = (float)y * 1.f * (float)a.sH / (float)a.dH;
    const float c2 = (float)z * 1.f * (float)a.sD / (float)a.dD;
    UpSample s;
    s.b0 = (uint32_t)c0;
    s.b1 = (uint32_t)c1;
    s.b2 = (uint32_t)c2;
    s.fx = (double)(c0 - (float)s.b0);
    s.fy = (double)(c1 - (float)s.b1);
    s.fz = (double)(c2 - (float)s.b2);
    s.live = true;
    if (a.ssparse) {
        const uint32_t cell = up_index(s.b0, s.b1, s.b2, a.sW - 1, a.sH - 1, a.sD - 1);
        s.live = (a.socc[cell >> 5] >> (cell & 31)) & 1u;
    }
    return s;
}

// Corner c in the reference's visiting order (0,0,0) (0,1,0) (1,0,0)
// (1,1,0) (0,0,1) (0,1,1) (1,0,1) (1,1,1): dx = c >> 1 & 1, dy = c & 1,
// dz = c >> 2.
__device__ __forceinline__ uint32_t up_corner(const UpArgs &a, const UpSample &s, int c,
                                              double &w) {
    const int dx = (c >> 1) & 1, dy = c & 1, dz = c >> 2;
    w = (dx ? s.fx : 1. - s.fx) * (dy ? s.fy : 1. - s.fy) * (dz ? s.fz : 1. - s.fz);
    const uint32_t v = up_index(s.b0 + dx, s.b1 + dy, s.b2 + dz, a.sW, a.sH, a.sD);
    return a.ssparse ? a.sio[v] : v;
}

// One group of LPV lanes per destination record, lane q = destination
// element q: the eight corner loads of a group read the same source records
// (one coalesced line each).
template <int LPV, int STORE>
__global__ __launch_bounds__(256) void k_upsample(const UpArgs a) {
    constexpr int kGroups = 256 / LPV;
    const int q = threadIdx.x % LPV;
    const int64_t r = (int64_t)xcd_slice_block(blockIdx.x, gridDim.x) * kGroups + threadIdx.x / LPV;
    if (r >= a.n || q >= a.dstride) return;
    const uint32_t d = a.dc2d ? a.dc2d[r] : (uint32_t)r;
    const UpSample s = up_locate(a, d);
    // Source element of destination element q, or -1 when it stays zero.
    int sq = -1;
    if (q == 0) {
        sq = 0;
    } else if (q < 1 + 3 * a.dS) {
        const int ch = (q - 1) / a.dS, i = (q - 1) % a.dS;
        if (i < a.ncopy) sq = 1 + ch * a.sS + i;
    }
    float acc = 0.f;
    if (s.live && sq >= 0) {
        float v[8];
        double w[8];
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const uint32_t rec = up_corner(a, s, c, w[c]);
            v[c] = load_elem<STORE>(a.src, (size_t)rec * (size_t)a.sstride + (size_t)sq);
        }
#pragma unroll
        for (int c = 0; c < 8; c++) acc = (float)((double)acc + w[c] * (double)v[c]);
        if (q == 0 && acc < a.threshold) acc = 0.f;
    }
    a.dst[(size_t)r * (size_t)a.dstride + q] = acc;
}

// Thresholded density of every destination vertex (the dense temporary the
// reference prunes, SceneUtils.cu:540-546); thread per vertex.
template <int STORE>
__global__ __launch_bounds__(256) void k_upsample_density(const UpArgs a) {
    const int64_t d = (int64_t)xcd_slice_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (d >= a.n) return;
    const UpSample s = up_locate(a, (uint32_t)d);
    float acc = 0.f;
    if (s.live) {
        float v[8];
        double w[8];
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const uint32_t rec = up_corner(a, s, c, w[c]);
            v[c] = load_elem<STORE>(a.src, (size_t)rec * (size_t)a.sstride);
        }
#pragma unroll
        for (int c = 0; c < 8; c++) acc = (float)((double)acc + w[c] * (double)v[c]);
        if (acc < a.threshold) acc = 0.f;
    }
    a.dst[d] = acc;
}

int32_t check_src(const pxl_grid *src, const char *what) {
    PXL_REQUIRE(src && src->params, "%s: null source grid", what);
    PXL_REQUIRE(src->W >= 2 && src->H >= 2 && src->D >= 2, "%s: source dims %dx%dx%d", what,
                src->W, src->H, src->D);
    PXL_REQUIRE(src->sh_dim >= 1 && src->stride >= 1 + 3 * src->sh_dim,
                "%s: source stride %d < 1 + 3 * %d", what, src->stride, src->sh_dim);
    PXL_REQUIRE(src->storage == PXL_STORE_F32 || src->storage == PXL_STORE_BF16 ||
                    src->storage == PXL_STORE_F16,
                "%s: unknown source storage %d", what, src->storage);
    PXL_REQUIRE(!src->sparse || (src->occupancy && src->index_offsets),
                "%s: a caged source needs occupancy and index_offsets", what);
    return PXL_OK;
}

void fill_src(const pxl_grid *src, UpArgs &a) {
    a.src = src->params;
    a.sstride = src->stride;
    a.sW = src->W;
    a.sH = src->H;
    a.sD = src->D;
    a.sS = src->sh_dim;
    a.ssparse = src->sparse;
    a.socc = src->occupancy;
    a.sio = src->index_offsets;
}

template <typename K>
void launch_store(int storage, K f32, K bf16, K f16, unsigned blocks, hipStream_t st,
                  const UpArgs &a) {
    K k = storage == PXL_STORE_BF16 ? bf16 : (storage == PXL_STORE_F16 ? f16 : f32);
    hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, st, a);
}

}  // namespace
}  // namespace pxl

using namespace pxl;

extern "C" int32_t pxl_upsample_density(const pxl_grid *src, int32_t W2, int32_t H2, int32_t D2,
                                        float threshold, float *density, void *stream) {
    clear_error();
    if (check_src(src, "pxl_upsample_density") != PXL_OK) return PXL_EINVAL;
    PXL_REQUIRE(density, "pxl_upsample_density: null output");
    PXL_REQUIRE(W2 >= 1 && H2 >= 1 && D2 >= 1 &&
                    (uint64_t)W2 * (uint64_t)H2 * (uint64_t)D2 <= 0xFFFFFFFFull,
                "pxl_upsample_density: destination dims %dx%dx%d", W2, H2, D2);
    UpArgs a{};
    fill_src(src, a);
    a.dst = density;
    a.dW = W2;
    a.dH = H2;
    a.dD = D2;
    a.n = (int64_t)W2 * H2 * D2;
    a.threshold = threshold;
    const unsigned blocks = (unsigned)((a.n + 255) / 256);
    using K = void (*)(UpArgs);
    launch_store<K>(src->storage, k_upsample_density<PXL_STORE_F32>,
                    k_upsample_density<PXL_STORE_BF16>, k_upsample_density<PXL_STORE_F16>, blocks,
                    as_stream(stream), a);
    return check_launch("pxl_upsample_density");
}

extern "C" int32_t pxl_upsample(const pxl_grid *src, const pxl_grid *dst, float threshold,
                                int32_t low_band_only, void *stream) {
    clear_error();
    if (check_src(src, "pxl_upsample") != PXL_OK) return PXL_EINVAL;
    PXL_REQUIRE(dst && dst->params, "pxl_upsample: null destination grid");
    PXL_REQUIRE(dst->storage == PXL_STORE_F32, "pxl_upsample: the destination must be FP32");
    PXL_REQUIRE(dst->W >= 1 && dst->H >= 1 && dst->D >= 1 &&
                    (uint64_t)dst->W * (uint64_t)dst->H * (uint64_t)dst->D <= 0xFFFFFFFFull,
                "pxl_upsample: destination dims %dx%dx%d", dst->W, dst->H, dst->D);
    const int ncopy = low_band_only ? 1 : src->sh_dim;
    PXL_REQUIRE(dst->sh_dim >= ncopy,
                "pxl_upsample: destination keeps %d SH coefficients, %d are carried over",
                dst->sh_dim, ncopy);
    PXL_REQUIRE(dst->stride >= 1 + 3 * dst->sh_dim && dst->stride <= 64,
                "pxl_upsample: destination stride %d (needs 1 + 3S .. 64)", dst->stride);
    const int64_t dense_n = (int64_t)dst->W * dst->H * dst->D;
    PXL_REQUIRE(dst->sparse ? (dst->compact_to_dense != nullptr && dst->n_vox >= 0 &&
                               dst->n_vox <= dense_n)
                            : dst->n_vox == dense_n,
                "pxl_upsample: destination of %lld records (dense %lld, caged needs "
                "compact_to_dense)",
                (long long)dst->n_vox, (long long)dense_n);
    if (dst->n_vox == 0) return PXL_OK;
    UpArgs a{};
    fill_src(src, a);
    a.dst = dst->params;
    a.dstride = dst->stride;
    a.dW = dst->W;
    a.dH = dst->H;
    a.dD = dst->D;
    a.dS = dst->sh_dim;
    a.dc2d = dst->sparse ? dst->compact_to_dense : nullptr;
    a.n = dst->n_vox;
    a.threshold = threshold;
    a.ncopy = ncopy;
    using K = void (*)(UpArgs);
    const hipStream_t st = as_stream(stream);
    if (dst->stride <= 32) {
        const unsigned blocks = (unsigned)((a.n + 7) / 8);
        launch_store<K>(src->storage, k_upsample<32, PXL_STORE_F32>,
                        k_upsample<32, PXL_STORE_BF16>, k_upsample<32, PXL_STORE_F16>, blocks, st,
                        a);
    } else {
        const unsigned blocks = (unsigned)((a.n + 3) / 4);
        launch_store<K>(src->storage, k_upsample<64, PXL_STORE_F32>,
                        k_upsample<64, PXL_STORE_BF16>, k_upsample<64, PXL_STORE_F16>, blocks, st,
                        a);
    }
    return check_launch("pxl_upsample");
}
