// Batch ordering for the uneven-ray schedules: the rays of one batch sorted
// by the Morton code of where they enter the grid, optionally under a
// longest-first bucket of their geometric length.  The reference draws
// batches as consecutive chunks of one shuffle (Render.cu:399-406); a sorted
// batch holds the same rays, so every per-ray result and the batch loss are
// unchanged, and the gradient differs only by the order of float atomics.
#include <rocprim/device/device_radix_sort.hpp>

#include "common.hpp"

namespace pxl {
namespace {

// Spreads the low 10 bits of v to every third bit.
__device__ __forceinline__ uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x3FFu;
    x = (x | (x << 16)) & 0x30000FFull;
    x = (x | (x << 8)) & 0x300F00Full;
    x = (x | (x << 4)) & 0x30C30C3ull;
    x = (x | (x << 2)) & 0x9249249ull;
    return x;
}

__device__ __forceinline__ uint32_t quant10(double p, int dim) {
    const double q = p / (double)(dim > 1 ? dim - 1 : 1) * 1023.0;
    return (uint32_t)(q < 0.0 ? 0.0 : (q > 1023.0 ? 1023.0 : q));
}

// Key of batch entry b: bits 0-29 Morton code of the entry point
// o + max(t_near, 0) d (10 bits per axis over the grid), bits 30-33 the
// length bucket (mode 2: 15 - min(15, bit length of the sample positions in
// [t_near, t_far)), so longer rays sort first); rays that miss the box last.
__global__ void k_order_keys(const pxl_rays r, const int32_t *__restrict__ sample_index,
                             int64_t start, int64_t len, int W, int H, int D, float step,
                             int mode, uint64_t *__restrict__ keys, int32_t *__restrict__ vals) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= len) return;
    const int32_t ray = sample_index[start + b];
    vals[b] = ray;
    if (ray < 0 || (int64_t)ray >= r.n) {
        keys[b] = 1ull << 40;
        return;
    }
    const double tn = r.near_t[ray], tf = r.far_t[ray];
    if (tn > tf) {
        keys[b] = 1ull << 40;
        return;
    }
    const double t0 = tn > 0.0 ? tn : 0.0;
    const double px = r.ox[ray] + t0 * (double)r.dx[ray];
    const double py = r.oy[ray] + t0 * (double)r.dy[ray];
    const double pz = r.oz[ray] + t0 * (double)r.dz[ray];
    uint64_t key = spread3(quant10(px, W)) | (spread3(quant10(py, H)) << 1) |
                   (spread3(quant10(pz, D)) << 2);
    if (mode == PXL_ORDER_LENGTH_MORTON) {
        const double n = (tf - tn) / (double)step;
        const uint32_t steps = n < 1.0 ? 0u : (n > 4.0e9 ? 0xFFFFFFFFu : (uint32_t)n);
        const int bits = steps == 0 ? 0 : 32 - __clz(steps);
        key |= (uint64_t)(15 - (bits < 15 ? bits : 15)) << 30;
    }
    keys[b] = key;
}

struct OrderLayout {
    size_t keys_in, keys_out, vals_in, temp, temp_bytes, total;
};
inline size_t align256(size_t x) { return (x + 255) / 256 * 256; }

OrderLayout order_layout(int64_t len) {
    OrderLayout L;
    size_t o = 0;
    L.keys_in = o;
    o += align256((size_t)len * 8);
    L.keys_out = o;
    o += align256((size_t)len * 8);
    L.vals_in = o;
    o += align256((size_t)len * 4);
    L.temp = o;
    L.temp_bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, L.temp_bytes, (const uint64_t *)nullptr,
                                    (uint64_t *)nullptr, (const int32_t *)nullptr,
                                    (int32_t *)nullptr, (size_t)len, 0u, 41u);
    o += align256(L.temp_bytes);
    L.total = o;
    return L;
}

}  // namespace
}  // namespace pxl

using namespace pxl;

extern "C" size_t pxl_batch_order_scratch_bytes(int64_t len) {
    if (len <= 0) return 0;
    return order_layout(len).total;
}

extern "C" int32_t pxl_batch_order(const pxl_grid *grid, const pxl_rays *rays,
                                   const int32_t *sample_index, int64_t start, int64_t len,
                                   int32_t mode, int32_t *out_index, void *scratch,
                                   size_t scratch_bytes, void *stream) {
    clear_error();
    PXL_REQUIRE(grid && rays && ((sample_index && out_index) || len == 0),
                "pxl_batch_order: null argument");
    PXL_REQUIRE(start >= 0 && len >= 0, "pxl_batch_order: negative start/len");
    PXL_REQUIRE(mode == PXL_ORDER_MORTON || mode == PXL_ORDER_LENGTH_MORTON,
                "pxl_batch_order: unknown mode %d", mode);
    PXL_REQUIRE(rays->ox && rays->oy && rays->oz && rays->dx && rays->dy && rays->dz &&
                    rays->near_t && rays->far_t,
                "pxl_batch_order: ray geometry planes must be non-null");
    PXL_REQUIRE(grid->W >= 2 && grid->H >= 2 && grid->D >= 2 && grid->step > 0.0f,
                "pxl_batch_order: bad grid");
    if (len == 0) return PXL_OK;
    const OrderLayout L = order_layout(len);
    PXL_REQUIRE(scratch && scratch_bytes >= L.total && ((uintptr_t)scratch & 255) == 0,
                "pxl_batch_order: scratch of %zu bytes (256-byte aligned, %zu needed)",
                scratch_bytes, L.total);
    char *w = static_cast<char *>(scratch);
    uint64_t *kin = reinterpret_cast<uint64_t *>(w + L.keys_in);
    uint64_t *kout = reinterpret_cast<uint64_t *>(w + L.keys_out);
    int32_t *vin = reinterpret_cast<int32_t *>(w + L.vals_in);
    const hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_order_keys, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, *rays,
                       sample_index, start, len, grid->W, grid->H, grid->D, grid->step, mode, kin,
                       vin);
    size_t tb = L.temp_bytes;
    if (rocprim::radix_sort_pairs(w + L.temp, tb, kin, kout, vin, out_index, (size_t)len, 0u, 41u,
                                  st) != hipSuccess) {
        set_error("pxl_batch_order: radix sort failed");
        return PXL_EHIP;
    }
    return check_launch("pxl_batch_order");
}
