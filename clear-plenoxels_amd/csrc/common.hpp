// Shared device helpers and error plumbing of libpxl.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>

#include "pxl.h"

namespace pxl {

// Per-thread last-error text returned by pxl_last_error().
void set_error(const char *fmt, ...);
void clear_error();

// Checks the launch that was just issued; records the HIP error text.
int32_t check_launch(const char *what);

// Library options (pxl_set_option): the eight-sample marcher for the fixed
// rule (else one sample per wave), and the TV gathers' 64-bit address form.
bool option_group_marcher();
bool option_tv_addr64();
int64_t option_claim_grid();

// Floats per record of the optimiser's moments / master and of its
// gradient (pxl_optim: 0 = the grid's stride, resp. the moments' stride).
inline int optim_stride(const pxl_grid *g, const pxl_optim *o) {
    return o && o->stride > 0 ? o->stride : g->stride;
}
inline int grad_stride(const pxl_grid *g, const pxl_optim *o) {
    return o && o->grad_stride > 0 ? o->grad_stride : optim_stride(g, o);
}

// Extra arguments of the train entry points (pxl_train_rays_ws / _sched /
// _prop), and the event the gradient kernels wait for (the trainer runs the
// variance loss, which writes the same gradient planes, on another stream).
struct TrainExtra {
    void *workspace = nullptr;
    size_t workspace_bytes = 0;
    int32_t max_samples = 0;
    const pxl_schedule *sched = nullptr;
    const uint64_t *bin_mask = nullptr;
    uint64_t *bin_targets = nullptr;
    uint8_t *target_valid = nullptr;
    double target_wmin = 0.0;
    hipEvent_t grad_wait = nullptr;
};
int32_t train_rays_ex(const pxl_grid *grid, const pxl_optim *optim, const pxl_rays *rays,
                      double *losses, const int32_t *sample_index, int64_t start, int64_t len,
                      double mask_coef, float huber_delta, void *stream, const TrainExtra &x);
// Whether the train step of this grid runs as the two-kernel logged step.
bool logged_step_serves(const pxl_grid *g);

#define PXL_REQUIRE(cond, ...)                 \
    do {                                       \
        if (!(cond)) {                         \
            ::pxl::set_error(__VA_ARGS__);     \
            return PXL_EINVAL;                 \
        }                                      \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

constexpr int kWave = 64;

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

// getSceneIndex (RenderUtils.cu:3-13): clamp each coordinate to
// [0, dim-1], linear z-y-x index.
__device__ __forceinline__ uint32_t scene_index(int x, int y, int z, int W,
                                                int H, int D) {
    x = clampi(x, 0, W - 1);
    y = clampi(y, 0, H - 1);
    z = clampi(z, 0, D - 1);
    return (uint32_t)z * (uint32_t)W * (uint32_t)H + (uint32_t)y * (uint32_t)W +
           (uint32_t)x;
}

// Wave-uniform broadcast of a lane's 32-bit value.
__device__ __forceinline__ float readlane_f(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double readlane_d(double v, int lane) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ int uniform_i(int v) {
    return __builtin_amdgcn_readfirstlane(v);
}

// The two wave halves' values through gfx950's v_permlane32_swap (a VALU
// operation, no LDS round trip): lo = v of lane L & 31, hi = v of lane
// (L & 31) + 32, in every lane L.
__device__ __forceinline__ void halves_f(float v, float &lo, float &hi) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false,
                                                    false);
    lo = __uint_as_float(r[0]);
    hi = __uint_as_float(r[1]);
}
// v of lane L ^ 32.
__device__ __forceinline__ float xor32_f(float v) {
    float lo, hi;
    halves_f(v, lo, hi);
    return (threadIdx.x & 32) ? lo : hi;
}

// Lane moves by DPP (VALU, no LDS round trip) that pair lane i with lane
// i ^ 1, i ^ 2 (quad permutes), 7 - i within each 8 (row_half_mirror: a
// lane of the other quad of the 8) and i ^ 8 (row_ror:8 in 16-lane rows);
// gfx950's v_permlane16_swap / v_permlane32_swap pair i with i ^ 16 and
// i ^ 32.  Butterfly sums built on them add the same pairs as the xor
// shuffles, so every lane holds the same bits.
constexpr int kDppXor1 = 0xB1;     // quad_perm [1, 0, 3, 2]
constexpr int kDppXor2 = 0x4E;     // quad_perm [2, 3, 0, 1]
constexpr int kDppMirror8 = 0x141; // row_half_mirror
constexpr int kDppXor8 = 0x128;    // row_ror:8
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(
        (long long)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo));
}
// DPP move whose lanes without a source in range (or in rows outside ROWS)
// receive `old`: row_shr:n = 0x110 + n (within 16-lane rows), wave_shr:1 =
// 0x138, row_bcast15 = 0x142 (lane 15 of each row to the next row),
// row_bcast31 = 0x143 (lane 31 to rows 2 and 3).
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double dpp_or_d(double v, double old) {
    const uint64_t u = (uint64_t)__double_as_longlong(v), o = (uint64_t)__double_as_longlong(old);
    const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)o, (int)(uint32_t)u, CTRL, ROWS, 0xF,
                                               false);
    const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(o >> 32), (int)(uint32_t)(u >> 32),
                                               CTRL, ROWS, 0xF, false);
    return __longlong_as_double(
        (long long)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo));
}

// Inclusive FP64 scans over the 64 lanes of a wave: Hillis-Steele steps of
// 1, 2, 4, 8 lanes inside each 16-lane row, then the row totals carried to
// the rows after them (row_bcast15 into rows 1 and 3, row_bcast31 into
// rows 2 and 3).  Lanes without a partner take the identity.
__device__ __forceinline__ double wave_scan_add64(double v) {
    v += dpp_or_d<0x111>(v, 0.0);
    v += dpp_or_d<0x112>(v, 0.0);
    v += dpp_or_d<0x114>(v, 0.0);
    v += dpp_or_d<0x118>(v, 0.0);
    v += dpp_or_d<0x142, 0xA>(v, 0.0);
    v += dpp_or_d<0x143, 0xC>(v, 0.0);
    return v;
}
__device__ __forceinline__ double wave_scan_mul64(double v) {
    v *= dpp_or_d<0x111>(v, 1.0);
    v *= dpp_or_d<0x112>(v, 1.0);
    v *= dpp_or_d<0x114>(v, 1.0);
    v *= dpp_or_d<0x118>(v, 1.0);
    v *= dpp_or_d<0x142, 0xA>(v, 1.0);
    v *= dpp_or_d<0x143, 0xC>(v, 1.0);
    return v;
}
// v of lane L - 1 (lane 0: `first`).
__device__ __forceinline__ double wave_shr1_d(double v, double first) {
    return dpp_or_d<0x138>(v, first);
}

// v of lane L ^ 16 and L ^ 32.
__device__ __forceinline__ uint32_t xor16_u(uint32_t v) {
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (threadIdx.x & 16) ? r[0] : r[1];
}
__device__ __forceinline__ uint32_t xor32_u(uint32_t v) {
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (threadIdx.x & 32) ? r[0] : r[1];
}
__device__ __forceinline__ double xor16_d(double v) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint64_t r = ((uint64_t)xor16_u((uint32_t)(u >> 32)) << 32) | xor16_u((uint32_t)u);
    return __longlong_as_double((long long)r);
}
__device__ __forceinline__ double xor32_d(double v) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint64_t r = ((uint64_t)xor32_u((uint32_t)(u >> 32)) << 32) | xor32_u((uint32_t)u);
    return __longlong_as_double((long long)r);
}

// DPP row rotate (16-lane rows): ctrl 0x120 + n = row_ror:n.
template <int N>
__device__ __forceinline__ float row_ror_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(
        0, __float_as_int(v), 0x120 + N, 0xF, 0xF, false));
}
template <int N>
__device__ __forceinline__ double row_ror_d(double v) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, 0x120 + N, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), 0x120 + N, 0xF, 0xF, false);
    return __longlong_as_double(
        (long long)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo));
}

// Sum over the 16 lanes of each DPP row; every lane of the row gets the sum.
__device__ __forceinline__ float row_sum_f(float v) {
    v += row_ror_f<8>(v);
    v += row_ror_f<4>(v);
    v += row_ror_f<2>(v);
    v += row_ror_f<1>(v);
    return v;
}
__device__ __forceinline__ double row_sum_d(double v) {
    v += row_ror_d<8>(v);
    v += row_ror_d<4>(v);
    v += row_ror_d<2>(v);
    v += row_ror_d<1>(v);
    return v;
}

// Wave-wide sums through cross-lane shuffles (used off the hot loop).
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ float wave_sum_f(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One element of a record in the grid's storage type (PXL_STORE_*), widened
// to float.
template <int STORE>
__device__ __forceinline__ float load_elem(const float *base, size_t idx) {
    if (STORE == PXL_STORE_F32) return base[idx];
    const uint16_t h = reinterpret_cast<const uint16_t *>(base)[idx];
    if (STORE == PXL_STORE_BF16) return __uint_as_float((uint32_t)h << 16);
    return __half2float(__ushort_as_half(h));
}

// Workgroups are dispatched round-robin over the 8 XCDs (block i -> XCD
// i mod 8).  Give XCD x a contiguous slice of the launch instead, so work
// items that are neighbours in launch order (rays of a sorted batch, records
// of one grid region) share one XCD's L2 for their gathers and atomics.  A
// bijection on [0, gridDim.x).
__device__ __forceinline__ unsigned xcd_slice_block(unsigned blk, unsigned nb) {
    constexpr unsigned kXcd = 8;
    if (nb < kXcd) return blk;
    const unsigned per = nb / kXcd, rem = nb % kXcd;
    const unsigned x = blk % kXcd, idx = blk / kXcd;
    return x < rem ? x * (per + 1) + idx : rem * (per + 1) + (x - rem) * per + idx;
}

}  // namespace pxl
