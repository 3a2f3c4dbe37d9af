// Learned proposal: a per-ray MLP that predicts which of 64 equal bins of
// [t_near, t_far) hold samples with visible weight, so the marcher can skip
// the others.  No counterpart in the reference (a build-only addition); the
// plain occupancy marcher stays the fallback and the reference its error is
// measured against.
//
// Network: 128 input features (positional encoding of the grid entry point
// and of the direction, 126 used) -> 64 -> 64 -> 64 bin logits, ReLU between
// layers.  Weights are BF16 with an FP32 master copy; activations and
// backward signals are rounded to BF16 for the matrix cores, products
// accumulate in FP32 (gfx950's v_mfma_f32_16x16x32_bf16, K = 32 per
// instruction).  Training is binary cross entropy per bin against the bins
// that the grid's own samples fill, with an FP32 Adam.
//
// Two fused kernels carry the network: k_prop_fwd (inference: the three
// layers of a 64-row tile per workgroup, logits out) and k_prop_train (the
// forward, the loss and the backward of a tile, each workgroup summing its
// tiles' weight and bias gradients in registers; k_prop_reduce then adds the
// workgroups' partial gradients in a fixed order).  Both are persistent
// grids of resident workgroups that stage the weights in LDS once.
#include <algorithm>

#include "common.hpp"

namespace pxl {
namespace {

constexpr int kIn = PXL_PROP_IN;       // 128 (126 features + 2 zero)
constexpr int kW = PXL_PROP_WIDTH;     // 64
constexpr int kBins = PXL_PROP_BINS;   // 64
constexpr int kW1 = kIn * kW, kW2 = kW * kW, kW3 = kW * kBins;
constexpr int kWeights = kW1 + kW2 + kW3;             // 16384
constexpr int kParams = kWeights + kW + kW + kBins;   // + 3 bias vectors

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ------------------------------------------------------------ features --

// Features of batch entry b: the grid entry point o + max(t_near, 0) d
// mapped to [-1, 1]^3 and the direction, each as
// [x, sin(2^i pi x), cos(2^i pi x) for i < 10] per component (63 + 63),
// then two zeros.  BF16, row b of a [len][128] matrix.
__global__ void k_prop_features(const pxl_rays r, const int32_t *__restrict__ sample_index,
                                int64_t start, int64_t len, int64_t rows, float sx, float sy,
                                float sz, uint16_t *__restrict__ X) {
    const int64_t b = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x / 64);
    const int lane = threadIdx.x & 63;
    if (b >= rows) return;
    const int32_t ray = b < len ? sample_index[start + b] : -1;  // tile padding: zero row
    float v[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (ray >= 0 && (int64_t)ray < r.n) {
        const float t0 = fmaxf(r.near_t[ray], 0.0f);
        const float dx = r.dx[ray], dy = r.dy[ray], dz = r.dz[ray];
        v[0] = (r.ox[ray] + t0 * dx) * sx - 1.0f;
        v[1] = (r.oy[ray] + t0 * dy) * sy - 1.0f;
        v[2] = (r.oz[ray] + t0 * dz) * sz - 1.0f;
        v[3] = dx;
        v[4] = dy;
        v[5] = dz;
    }
    // Two features per lane: 128 per row.
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int f = 2 * lane + h;  // feature index
        float out = 0.0f;
        if (f < 126) {
            const int g = f / 63, q = f % 63;  // 0: position, 1: direction
            if (q < 3) {
                out = v[3 * g + q];
            } else {
                const int i = (q - 3) / 6, rem = (q - 3) % 6, c = rem % 3;
                const float x = v[3 * g + c] * (float)(1 << i) * 3.14159265358979f;
                out = rem < 3 ? sinf(x) : cosf(x);
            }
        }
        X[(size_t)b * kIn + f] = f2bf(ray >= 0 ? out : 0.0f);
    }
}

// ------------------------------------------------------------ network --

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// MFMA 16x16x32 BF16 operand layout (lane l, g = l / 16): A[row l % 16]
// [k 8g .. 8g+7], B[k 8g .. 8g+7][col l % 16] as 16-byte vectors; the
// accumulator holds D[rows 4g .. 4g+3][col l % 16].
__device__ __forceinline__ v8bf ld8(const uint16_t *p) {
    return __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4 *>(p));
}
__device__ __forceinline__ v4f mma32(const v8bf &a, const v8bf &b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// LDS row strides (BF16 elements): 8 pad elements per row so that the 16
// lanes of a lane group, reading 16 rows at one column offset, hit distinct
// banks; row starts stay 16-byte aligned.
constexpr int kS128 = kIn + 8;   // rows of 128
constexpr int kS64 = kW + 8;     // rows of 64
constexpr int kTile = 64;        // rows per workgroup tile (4 waves x 16)

// Forward weights transposed in LDS (B operands: B[k][n] = W[k][n] read as
// WT[n][k .. k+7]) and the biases.
struct PropLds {
    uint16_t w1t[kW * kS128];
    uint16_t w2t[kW * kS64];
    uint16_t w3t[kBins * kS64];
    float b[3 * kW];
};
__device__ __forceinline__ void stage_weights(const uint16_t *__restrict__ w,
                                              const float *__restrict__ bias, PropLds &L) {
    for (int i = threadIdx.x; i < kW1; i += blockDim.x) {
        const int k = i / kW, n = i % kW;
        L.w1t[n * kS128 + k] = w[i];
    }
    for (int i = threadIdx.x; i < kW2; i += blockDim.x) {
        const int k = i / kW, n = i % kW;
        L.w2t[n * kS64 + k] = w[kW1 + i];
        L.w3t[n * kS64 + k] = w[kW1 + kW2 + i];
    }
    for (int i = threadIdx.x; i < 3 * kW; i += blockDim.x) L.b[i] = bias[i];
}

// One layer of a wave's 16 rows: acc[n] (output columns 16n .. 16n+15) =
// A (16 x K, row-major at a, row stride sa) x W (K x 64, transposed at wt).
template <int K>
__device__ __forceinline__ void layer(const uint16_t *a, int sa, const uint16_t *wt, int swt,
                                      v4f (&acc)[4]) {
    const int lane = threadIdx.x & 63, lr = lane & 15, g = lane >> 4;
#pragma unroll
    for (int n = 0; n < 4; n++) acc[n] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < K / 32; kb++) {
        const v8bf av = ld8(a + lr * sa + 32 * kb + 8 * g);
#pragma unroll
        for (int n = 0; n < 4; n++)
            acc[n] = mma32(av, ld8(wt + (16 * n + lr) * swt + 32 * kb + 8 * g), acc[n]);
    }
}

// relu(acc + bias) as BF16 into a wave's 16 rows at h (row stride sh), and
// (when ht) transposed into ht[col][rows] (row stride sht, first row r0);
// rows of the tile at or beyond len are zero.
__device__ __forceinline__ void relu_out(const v4f (&acc)[4], const float *bias, int64_t row0,
                                         int64_t len, uint16_t *h, int sh, uint16_t *ht, int sht,
                                         int r0) {
    const int lane = threadIdx.x & 63, lr = lane & 15, g = lane >> 4;
#pragma unroll
    for (int n = 0; n < 4; n++) {
        const int col = 16 * n + lr;
        uint16_t v[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int row = 4 * g + j;
            const float x = row0 + row < len ? fmaxf(acc[n][j] + bias[col], 0.0f) : 0.0f;
            v[j] = f2bf(x);
            h[row * sh + col] = v[j];
        }
        if (ht) {
            const uint2 pk = make_uint2((uint32_t)v[0] | ((uint32_t)v[1] << 16),
                                        (uint32_t)v[2] | ((uint32_t)v[3] << 16));
            *reinterpret_cast<uint2 *>(ht + col * sht + r0 + 4 * g) = pk;
        }
    }
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inference: logits Z[rows][64] (FP32) of the featured rows X[rows][128],
// one 64-row tile per workgroup at a time (wave w: rows 16w .. 16w+15).
__global__ __launch_bounds__(256) void k_prop_fwd(const uint16_t *__restrict__ X,
                                                  const uint16_t *__restrict__ w,
                                                  const float *__restrict__ bias, int64_t len,
                                                  int64_t rows, float *__restrict__ Z) {
    __shared__ PropLds L;
    __shared__ __align__(16) uint16_t hbuf[4][16 * kS64];
    stage_weights(w, bias, L);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, g = lane >> 4;
    uint16_t *h = hbuf[wave];
    const int64_t tiles = rows / kTile;
    for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const int64_t row0 = t * kTile + 16 * wave;
        v4f acc[4];
        layer<kIn>(X + row0 * kIn, kIn, L.w1t, kS128, acc);
        relu_out(acc, L.b, row0, len, h, kS64, nullptr, 0, 0);
        wave_sync();
        layer<kW>(h, kS64, L.w2t, kS64, acc);
        wave_sync();  // every lane has read h before it is overwritten
        relu_out(acc, L.b + kW, row0, len, h, kS64, nullptr, 0, 0);
        wave_sync();
        layer<kW>(h, kS64, L.w3t, kS64, acc);
        wave_sync();
#pragma unroll
        for (int n = 0; n < 4; n++) {
            const int col = 16 * n + lr;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int64_t r = row0 + 4 * g + j;
                if (r < len) Z[r * kBins + col] = acc[n][j] + L.b[2 * kW + col];
            }
        }
    }
}

// Training.  Per tile: the forward (H1, H2 kept in LDS transposed for the
// weight gradients), the loss and dZ = bf16(scale w (sigmoid(z) - y)), then
// dW3 += H2^T dZ, G2 = dZ W3^T, dH2 = bf16(G2 [H2 > 0]), dW2 += H1^T dH2,
// G1 = dH2 W2^T, dH1 = bf16(G1 [H1 > 0]), dW1 += X^T dH1, and the bias
// gradients as column sums of dZ, dH2, dH1.  Weight gradient blocks of
// 16 x 16 are split over the waves (dW1: wave w owns rows 32w .. 32w+31,
// dW2 and dW3 rows 16w .. 16w+15) and accumulate over the workgroup's tiles
// in registers; at the end each workgroup writes its partial gradient and
// loss to part[blockIdx.x].
struct TrainLds {
    uint16_t xt[kIn * (kTile + 8)];    // X^T [feature][row]
    uint16_t h1t[kW * (kTile + 8)];    // H1^T [unit][row]
    uint16_t h2t[kW * (kTile + 8)];    // H2^T
    uint16_t rm[kTile * kS64];         // row-major tile: H1, H2, dZ, dH2 in turn
    uint16_t cm[kW * (kTile + 8)];     // column-major tile: dZ, dH2, dH1 in turn
    float db[3 * kW];                  // bias gradients of this workgroup's tiles
};
constexpr int kST = kTile + 8;  // row stride of the transposed tiles

__global__ __launch_bounds__(256) void k_prop_train(
    const uint16_t *__restrict__ X, const uint16_t *__restrict__ w,
    const float *__restrict__ bias, int64_t len, int64_t rows,
    const uint64_t *__restrict__ targets, const uint8_t *__restrict__ valid, float pos_weight,
    float scale, float *__restrict__ part, double *__restrict__ part_loss) {
    __shared__ PropLds L;
    __shared__ TrainLds T;
    stage_weights(w, bias, L);
    for (int i = threadIdx.x; i < 3 * kW; i += blockDim.x) T.db[i] = 0.0f;
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, g = lane >> 4;
    const uint16_t *W2 = w + kW1, *W3 = w + kW1 + kW2;  // row-major [in][out]
    v4f dw1[2][4], dw2[4], dw3[4];
#pragma unroll
    for (int n = 0; n < 4; n++) {
        dw1[0][n] = dw1[1][n] = dw2[n] = dw3[n] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    }
    double loss = 0.0;
    const int64_t tiles = rows / kTile;
    const int r0 = 16 * wave;  // this wave's first row within the tile
    for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const int64_t row0 = t * kTile + r0;
        // X rows of this wave, transposed into LDS for dW1.
#pragma unroll
        for (int kb = 0; kb < kIn / 32; kb++) {
            const uint4 u = *reinterpret_cast<const uint4 *>(X + (row0 + lr) * kIn + 32 * kb + 8 * g);
            const uint32_t e[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
            for (int q = 0; q < 8; q++)
                T.xt[(32 * kb + 8 * g + q) * kST + r0 + lr] = (uint16_t)(e[q >> 1] >> (16 * (q & 1)));
        }
        v4f acc[4];
        uint16_t *rm = T.rm + r0 * kS64;
        layer<kIn>(X + row0 * kIn, kIn, L.w1t, kS128, acc);
        relu_out(acc, L.b, row0, len, rm, kS64, T.h1t, kST, r0);
        wave_sync();
        layer<kW>(rm, kS64, L.w2t, kS64, acc);
        wave_sync();
        relu_out(acc, L.b + kW, row0, len, rm, kS64, T.h2t, kST, r0);
        wave_sync();
        layer<kW>(rm, kS64, L.w3t, kS64, acc);
        wave_sync();
        // Loss and dZ (k_logits_bce's arithmetic).
        double lrow[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int n = 0; n < 4; n++) {
            const int col = 16 * n + lr;
            uint16_t v[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int64_t r = row0 + 4 * g + j;
                const bool live = r < len;
                const float z = live ? acc[n][j] + L.b[2 * kW + col] : 0.0f;
                const bool use = live && (valid == nullptr || valid[r] != 0);
                const float y = use ? (float)((targets[r] >> col) & 1ull) : 0.0f;
                const float p = 1.0f / (1.0f + expf(-z));
                const float wt = y > 0.0f ? pos_weight : 1.0f;
                v[j] = f2bf(use ? scale * wt * (p - y) : 0.0f);
                rm[(4 * g + j) * kS64 + col] = v[j];
                if (use)
                    lrow[j] += (double)wt * ((double)fmaxf(z, 0.0f) - (double)z * y +
                                             (double)log1pf(expf(-fabsf(z))));
            }
            *reinterpret_cast<uint2 *>(T.cm + col * kST + r0 + 4 * g) =
                make_uint2((uint32_t)v[0] | ((uint32_t)v[1] << 16),
                           (uint32_t)v[2] | ((uint32_t)v[3] << 16));
        }
        loss += lrow[0] + lrow[1] + lrow[2] + lrow[3];
        __syncthreads();  // dZ (cm) and H2^T of every wave are in LDS
        // dW3 rows 16w..16w+15 (hidden units), all 64 bins; db3.
#pragma unroll
        for (int ks = 0; ks < kTile / 32; ks++) {
            const v8bf a = ld8(T.h2t + (16 * wave + lr) * kST + 32 * ks + 8 * g);
#pragma unroll
            for (int n = 0; n < 4; n++)
                dw3[n] = mma32(a, ld8(T.cm + (16 * n + lr) * kST + 32 * ks + 8 * g), dw3[n]);
        }
        if (threadIdx.x < kBins) {
            float s = 0.0f;
            for (int r = 0; r < kTile; r++) s += bf2f(T.cm[threadIdx.x * kST + r]);
            T.db[2 * kW + threadIdx.x] += s;
        }
        // G2 = dZ W3^T (this wave's rows), B[k = bin][n = unit] = W3[unit][bin].
        v4f gacc[4];
#pragma unroll
        for (int n = 0; n < 4; n++) gacc[n] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < kBins / 32; kb++) {
            const v8bf a = ld8(rm + lr * kS64 + 32 * kb + 8 * g);
#pragma unroll
            for (int n = 0; n < 4; n++)
                gacc[n] = mma32(a, ld8(W3 + (16 * n + lr) * kBins + 32 * kb + 8 * g), gacc[n]);
        }
        __syncthreads();  // dZ in cm and rm fully read
        // dH2 = bf16(G2 [H2 > 0]) into rm (row-major) and cm (column-major).
#pragma unroll
        for (int n = 0; n < 4; n++) {
            const int col = 16 * n + lr;
            uint16_t v[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint16_t hv = T.h2t[col * kST + r0 + 4 * g + j];
                v[j] = f2bf(hv != 0 && (hv & 0x8000u) == 0 ? gacc[n][j] : 0.0f);
                rm[(4 * g + j) * kS64 + col] = v[j];
            }
            *reinterpret_cast<uint2 *>(T.cm + col * kST + r0 + 4 * g) =
                make_uint2((uint32_t)v[0] | ((uint32_t)v[1] << 16),
                           (uint32_t)v[2] | ((uint32_t)v[3] << 16));
        }
        __syncthreads();
        // dW2 rows 16w..16w+15, db2.
#pragma unroll
        for (int ks = 0; ks < kTile / 32; ks++) {
            const v8bf a = ld8(T.h1t + (16 * wave + lr) * kST + 32 * ks + 8 * g);
#pragma unroll
            for (int n = 0; n < 4; n++)
                dw2[n] = mma32(a, ld8(T.cm + (16 * n + lr) * kST + 32 * ks + 8 * g), dw2[n]);
        }
        if (threadIdx.x < kW) {
            float s = 0.0f;
            for (int r = 0; r < kTile; r++) s += bf2f(T.cm[threadIdx.x * kST + r]);
            T.db[kW + threadIdx.x] += s;
        }
        // G1 = dH2 W2^T, B[k = out][n = in] = W2[in][out].
#pragma unroll
        for (int n = 0; n < 4; n++) gacc[n] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < kW / 32; kb++) {
            const v8bf a = ld8(rm + lr * kS64 + 32 * kb + 8 * g);
#pragma unroll
            for (int n = 0; n < 4; n++)
                gacc[n] = mma32(a, ld8(W2 + (16 * n + lr) * kW + 32 * kb + 8 * g), gacc[n]);
        }
        __syncthreads();  // dH2 in cm fully read
        // dH1 = bf16(G1 [H1 > 0]) into cm.
#pragma unroll
        for (int n = 0; n < 4; n++) {
            const int col = 16 * n + lr;
            uint16_t v[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint16_t hv = T.h1t[col * kST + r0 + 4 * g + j];
                v[j] = f2bf(hv != 0 && (hv & 0x8000u) == 0 ? gacc[n][j] : 0.0f);
            }
            *reinterpret_cast<uint2 *>(T.cm + col * kST + r0 + 4 * g) =
                make_uint2((uint32_t)v[0] | ((uint32_t)v[1] << 16),
                           (uint32_t)v[2] | ((uint32_t)v[3] << 16));
        }
        __syncthreads();
        // dW1 rows 32w..32w+31 (features), db1.
#pragma unroll
        for (int ks = 0; ks < kTile / 32; ks++) {
#pragma unroll
            for (int m = 0; m < 2; m++) {
                const v8bf a = ld8(T.xt + (32 * wave + 16 * m + lr) * kST + 32 * ks + 8 * g);
#pragma unroll
                for (int n = 0; n < 4; n++)
                    dw1[m][n] = mma32(a, ld8(T.cm + (16 * n + lr) * kST + 32 * ks + 8 * g),
                                      dw1[m][n]);
            }
        }
        if (threadIdx.x < kW) {
            float s = 0.0f;
            for (int r = 0; r < kTile; r++) s += bf2f(T.cm[threadIdx.x * kST + r]);
            T.db[threadIdx.x] += s;
        }
        __syncthreads();  // the next tile overwrites every LDS tile
    }
    // This workgroup's partial gradient, in the gradient's layout.
    float *pg = part + (size_t)blockIdx.x * kParams;
#pragma unroll
    for (int n = 0; n < 4; n++) {
        const int col = 16 * n + lr;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int r = 4 * g + j;
            pg[(32 * wave + r) * kW + col] = dw1[0][n][j];
            pg[(32 * wave + 16 + r) * kW + col] = dw1[1][n][j];
            pg[kW1 + (16 * wave + r) * kW + col] = dw2[n][j];
            pg[kW1 + kW2 + (16 * wave + r) * kBins + col] = dw3[n][j];
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * kW; i += blockDim.x) pg[kWeights + i] = T.db[i];
    // Loss: per-wave sums (each row's 64 bins over 16 lanes and 4 n), then
    // the waves in order.
    __shared__ double lw[4];
    loss = wave_sum_d(loss);
    if (lane == 0) lw[wave] = loss;
    __syncthreads();
    if (threadIdx.x == 0) part_loss[blockIdx.x] = lw[0] + lw[1] + lw[2] + lw[3];
}

// Gradient = sum of the workgroups' partials; loss too.  A block of 256
// threads covers 64 parameters: thread (q, i) sums the partials of
// workgroups q, q + 4, ... of parameter i, then the four sums are added in
// order (a fixed order for a given number of workgroups).
__global__ __launch_bounds__(256) void k_prop_reduce(const float *__restrict__ part,
                                                     const double *__restrict__ part_loss,
                                                     int groups, float *__restrict__ grad,
                                                     double *__restrict__ loss_out) {
    __shared__ float red[4][64];
    const int q = threadIdx.x >> 6, il = threadIdx.x & 63;
    const int i = blockIdx.x * 64 + il;
    float s = 0.0f;
    if (i < kParams) {
        int c = q;
        for (; c + 12 < groups; c += 16) {
            const float a = part[(size_t)c * kParams + i], b = part[(size_t)(c + 4) * kParams + i],
                        d = part[(size_t)(c + 8) * kParams + i],
                        e = part[(size_t)(c + 12) * kParams + i];
            s += a;
            s += b;
            s += d;
            s += e;
        }
        for (; c < groups; c += 4) s += part[(size_t)c * kParams + i];
    }
    red[q][il] = s;
    __syncthreads();
    if (q == 0 && i < kParams) grad[i] = ((red[0][il] + red[1][il]) + red[2][il]) + red[3][il];
    if (blockIdx.x == 0 && threadIdx.x == 0 && loss_out) {
        double l = 0.0;
        for (int c = 0; c < groups; c++) l += part_loss[c];
        loss_out[0] = l;
    }
}

// Adam on the FP32 master (weights then biases), BF16 copy of the weights.
__global__ void k_prop_adam(float *__restrict__ master, const float *__restrict__ g,
                            float *__restrict__ m, float *__restrict__ v,
                            uint16_t *__restrict__ w16, double lr, double c1, double c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const double step = lr * ((double)mi / c1) / (sqrt((double)vi / c2) + 1e-8);
    const float p = (float)((double)master[i] - step);
    master[i] = p;
    if (i < kWeights) w16[i] = f2bf(p);
}

// Deterministic initialisation: weights uniform in +-sqrt(6 / fan_in)
// (He-uniform) from a counter-based hash, biases 0, moments 0.
__global__ void k_prop_init(float *__restrict__ master, float *__restrict__ m,
                            float *__restrict__ v, uint16_t *__restrict__ w16, uint64_t seed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    float p = 0.0f;
    if (i < kWeights) {
        const int fan_in = i < kW1 ? 126 : kW;
        const bool pad = i < kW1 && (i / kW) >= 126;  // rows of the two zero features
        uint64_t x = seed * 0x9E3779B97F4A7C15ull + (uint64_t)i * 0xBF58476D1CE4E5B9ull;
        x ^= x >> 31;
        x *= 0x94D049BB133111EBull;
        x ^= x >> 29;
        const float u = (float)((x >> 40) & 0xFFFFFFull) / 16777216.0f;  // [0, 1)
        p = pad ? 0.0f : (2.0f * u - 1.0f) * sqrtf(6.0f / (float)fan_in);
    }
    master[i] = p;
    m[i] = 0.0f;
    v[i] = 0.0f;
    if (i < kWeights) w16[i] = f2bf(p);
}

// Allowed-bin masks: bin j is kept when its probability exceeds threshold,
// dilated by `dilate` bins on each side.
__global__ void k_prop_masks(const float *__restrict__ Z, int64_t len, float threshold,
                             int dilate, uint64_t *__restrict__ masks) {
    const int64_t r = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x / 64);
    const int j = threadIdx.x & 63;
    if (r >= len) return;
    const float p = 1.0f / (1.0f + expf(-Z[r * 64 + j]));
    uint64_t m = __ballot(p > threshold);
    uint64_t d = m;
    for (int s = 1; s <= dilate; s++) d |= (m << s) | (m >> s);
    if (j == 0) masks[r] = d;
}

// -------------------------------------------------------------- host side --

// Proposal workspace: feature rows X [rows][128] BF16 (rows = len rounded
// up to the 64-row tile), logits [rows][64] FP32, the training kernel's
// per-workgroup partial gradients and losses.
struct PropLayout {
    size_t X, Z, part, ploss, total;
    int64_t rows;
};
inline size_t a256(size_t x) { return (x + 255) / 256 * 256; }
constexpr int kMaxGroups = 512;  // workgroups of the persistent kernels, at most

PropLayout prop_layout(int64_t len) {
    PropLayout L;
    L.rows = (len + kTile - 1) / kTile * kTile;
    size_t o = 0;
    L.X = o;
    o += a256((size_t)L.rows * kIn * 2);
    L.Z = o;
    o += a256((size_t)L.rows * kBins * 4);
    L.part = o;
    o += a256((size_t)kMaxGroups * kParams * 4);
    L.ploss = o;
    o += a256((size_t)kMaxGroups * 8);
    L.total = o;
    return L;
}

// Workgroups of a persistent kernel: those resident on the device at once,
// at most one per tile and kMaxGroups.
template <typename K>
unsigned prop_groups(K kernel, int64_t tiles) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess ||
        cus <= 0 || per_cu <= 0)
        cus = 256, per_cu = 1;
    int64_t g = (int64_t)cus * per_cu;
    g = std::min<int64_t>(g, tiles);
    g = std::min<int64_t>(g, kMaxGroups);
    return (unsigned)std::max<int64_t>(g, 1);
}

int32_t check_prop(const pxl_proposal *p, const char *who) {
    PXL_REQUIRE(p && p->weights && p->master && p->m && p->v && p->grad, "%s: null proposal buffer",
                who);
    return PXL_OK;
}

}  // namespace
}  // namespace pxl

using namespace pxl;

extern "C" size_t pxl_proposal_workspace_bytes(int64_t len) {
    if (len <= 0) return 0;
    return prop_layout(len).total;
}

extern "C" int32_t pxl_proposal_init(const pxl_proposal *p, uint64_t seed, void *stream) {
    clear_error();
    int32_t st = check_prop(p, "pxl_proposal_init");
    if (st != PXL_OK) return st;
    hipLaunchKernelGGL(k_prop_init, dim3((kParams + 255) / 256), dim3(256), 0, as_stream(stream),
                       p->master, p->m, p->v, p->weights, seed);
    return check_launch("pxl_proposal_init");
}

extern "C" int32_t pxl_proposal_features(const pxl_grid *grid, const pxl_rays *rays,
                                         const int32_t *sample_index, int64_t start, int64_t len,
                                         void *workspace, size_t workspace_bytes, void *stream) {
    clear_error();
    PXL_REQUIRE(grid && rays && sample_index && workspace, "pxl_proposal_features: null argument");
    PXL_REQUIRE(start >= 0 && len > 0, "pxl_proposal_features: bad start/len");
    const PropLayout L = prop_layout(len);
    PXL_REQUIRE(workspace_bytes >= L.total && ((uintptr_t)workspace & 255) == 0,
                "pxl_proposal_features: workspace of %zu bytes (%zu needed, 256-byte aligned)",
                workspace_bytes, L.total);
    char *w = static_cast<char *>(workspace);
    uint16_t *X = (uint16_t *)(w + L.X);
    const hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_prop_features, dim3((unsigned)((L.rows + 3) / 4)), dim3(256), 0, st,
                       *rays, sample_index, start, len, L.rows, 2.0f / (float)(grid->W - 1),
                       2.0f / (float)(grid->H - 1), 2.0f / (float)(grid->D - 1), X);
    return check_launch("pxl_proposal_features");
}

extern "C" int32_t pxl_proposal_forward(const pxl_proposal *p, int64_t len, float *logits,
                                        void *workspace, size_t workspace_bytes, void *stream) {
    clear_error();
    int32_t st = check_prop(p, "pxl_proposal_forward");
    if (st != PXL_OK) return st;
    PXL_REQUIRE(len > 0 && workspace, "pxl_proposal_forward: bad len / null workspace");
    const PropLayout L = prop_layout(len);
    PXL_REQUIRE(workspace_bytes >= L.total, "pxl_proposal_forward: workspace too small");
    char *w = static_cast<char *>(workspace);
    const hipStream_t s = as_stream(stream);
    float *Z = logits ? logits : (float *)(w + L.Z);
    const int64_t tiles = L.rows / kTile;
    hipLaunchKernelGGL(k_prop_fwd, dim3(prop_groups(k_prop_fwd, tiles)), dim3(256), 0, s,
                       (const uint16_t *)(w + L.X), (const uint16_t *)p->weights,
                       (const float *)(p->master + kWeights), len, L.rows, Z);
    return check_launch("pxl_proposal_forward");
}

extern "C" int32_t pxl_proposal_train_step(const pxl_proposal *p, int64_t len,
                                           const uint64_t *targets, const uint8_t *valid,
                                           int64_t n_valid, float pos_weight, double lr,
                                           int32_t t, void *workspace, size_t workspace_bytes,
                                           double *loss_out, void *stream) {
    clear_error();
    int32_t st = check_prop(p, "pxl_proposal_train_step");
    if (st != PXL_OK) return st;
    PXL_REQUIRE(len > 0 && workspace && targets && t >= 1,
                "pxl_proposal_train_step: bad len / null targets / t < 1");
    const PropLayout L = prop_layout(len);
    PXL_REQUIRE(workspace_bytes >= L.total, "pxl_proposal_train_step: workspace too small");
    char *w = static_cast<char *>(workspace);
    const hipStream_t s = as_stream(stream);
    const float scale = 1.0f / (float)((n_valid > 0 ? n_valid : len) * 64);
    const int64_t tiles = L.rows / kTile;
    const unsigned groups = prop_groups(k_prop_train, tiles);
    float *part = (float *)(w + L.part);
    double *ploss = (double *)(w + L.ploss);
    hipLaunchKernelGGL(k_prop_train, dim3(groups), dim3(256), 0, s, (const uint16_t *)(w + L.X),
                       (const uint16_t *)p->weights, (const float *)(p->master + kWeights), len,
                       L.rows, targets, valid, pos_weight, scale, part, ploss);
    hipLaunchKernelGGL(k_prop_reduce, dim3((kParams + 63) / 64), dim3(256), 0, s, part, ploss,
                       (int)groups, p->grad, loss_out);
    if (lr > 0.0) {
        const double c1 = 1.0 - pow(0.9, (double)t), c2 = 1.0 - pow(0.999, (double)t);
        hipLaunchKernelGGL(k_prop_adam, dim3((kParams + 255) / 256), dim3(256), 0, s, p->master,
                           p->grad, p->m, p->v, p->weights, lr, c1, c2);
    }
    return check_launch("pxl_proposal_train_step");
}

extern "C" int32_t pxl_proposal_masks(const float *logits, int64_t len, float threshold,
                                      int32_t dilate, uint64_t *masks, void *stream) {
    clear_error();
    PXL_REQUIRE(logits && masks && len > 0 && dilate >= 0 && dilate < 64,
                "pxl_proposal_masks: bad argument");
    hipLaunchKernelGGL(k_prop_masks, dim3((unsigned)((len + 3) / 4)), dim3(256), 0,
                       as_stream(stream), logits, len, threshold, dilate, masks);
    return check_launch("pxl_proposal_masks");
}
