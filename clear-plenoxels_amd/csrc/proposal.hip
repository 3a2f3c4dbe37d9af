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
// accumulate in FP32 (v_mfma_f32_16x16x16_bf16).  Training is binary cross
// entropy per bin against the bins that the grid's own samples fill, with an
// FP32 Adam.
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

// ----------------------------------------------------------------- GEMM --

// C[M][64] = op(A)[M][K] x op(B)[K][64] in FP32, BF16 operands.  A is
// [M][K] row-major, or [K][M] when TA; B is [K][64] row-major, or [64][K]
// when TB.  A workgroup of 4 waves owns 64 rows; wave w owns rows
// 16w .. 16w+15 and all 64 columns (4 MFMA accumulators).  blockIdx.y
// selects a K chunk of kchunk; its partial product goes to
// C + blockIdx.y * M * 64 (split-K, summed by k_sum_partials).
// MFMA 16x16x16 operand layout: lane l holds A[row l%16][k 4(l/16)..+3],
// B[k 4(l/16)..+3][col l%16], and gets D[rows 4(l/16)..+3][col l%16].
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm_bf16(const uint16_t *__restrict__ A,
                                                   const uint16_t *__restrict__ B,
                                                   float *__restrict__ C, int M, int K,
                                                   int kchunk) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row0 = blockIdx.x * 64 + wave * 16;
    if (row0 >= M) return;
    const int kb = blockIdx.y * kchunk, ke = min(K, kb + kchunk);
    const int lr = lane & 15, lk = (lane >> 4) * 4;
    v4f acc[4];
#pragma unroll
    for (int n = 0; n < 4; n++) acc[n] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = kb; k0 < ke; k0 += 16) {
        v4s a;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int m = row0 + lr, k = k0 + lk + j;
            a[j] = (short)(TA ? A[(size_t)k * M + m] : A[(size_t)m * K + k]);
        }
#pragma unroll
        for (int n = 0; n < 4; n++) {
            v4s bb;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int col = n * 16 + lr, k = k0 + lk + j;
                bb[j] = (short)(TB ? B[(size_t)col * K + k] : B[(size_t)k * 64 + col]);
            }
            acc[n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc[n], 0, 0, 0);
        }
    }
    float *Cp = C + (size_t)blockIdx.y * (size_t)M * 64;
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
        for (int j = 0; j < 4; j++)
            Cp[(size_t)(row0 + lk + j) * 64 + n * 16 + lr] = acc[n][j];
}

// out[i] = sum over c < chunks of part[c * n + i] (+ out[i] when accumulate).
__global__ void k_sum_partials(const float *__restrict__ part, int chunks, int n,
                               float *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int c = 0; c < chunks; c++) s += part[(size_t)c * n + i];
    out[i] = s;
}

// ------------------------------------------------------------ epilogues --

// H = bf16(relu(C + bias)); rows >= len (tile padding) are zero.
__global__ void k_bias_relu(const float *__restrict__ Cm, const float *__restrict__ bias,
                            int64_t len, int64_t rows, uint16_t *__restrict__ H) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * 64) return;
    const int64_t r = i / 64;
    const float v = r < len ? fmaxf(Cm[i] + bias[i & 63], 0.0f) : 0.0f;
    H[i] = f2bf(v);
}

// Logits z = C + bias (FP32).  Training: p = sigmoid(z), target bit y of
// the row's mask, weight pos_weight for y = 1; dZ = bf16(scale w (p - y)),
// scale = 1 / (valid rows * 64); the row's BCE summed into loss_rows.
__global__ void k_logits_bce(const float *__restrict__ Cm, const float *__restrict__ bias,
                             int64_t len, int64_t rows, float *__restrict__ Z,
                             const uint64_t *__restrict__ targets,
                             const uint8_t *__restrict__ valid, float pos_weight, float scale,
                             uint16_t *__restrict__ dZ, double *__restrict__ loss_rows) {
    const int64_t r = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x / 64);
    const int j = threadIdx.x & 63;
    if (r >= rows) return;
    const int64_t i = r * 64 + j;
    const bool live = r < len;
    const float z = live ? Cm[i] + bias[j] : 0.0f;
    if (Z && live) Z[i] = z;
    if (!dZ) return;
    const bool use = live && (valid == nullptr || valid[r] != 0);
    const float y = use ? (float)((targets[r] >> j) & 1ull) : 0.0f;
    const float p = 1.0f / (1.0f + expf(-z));
    const float w = y > 0.0f ? pos_weight : 1.0f;
    dZ[i] = f2bf(use ? scale * w * (p - y) : 0.0f);
    // BCE with logits: max(z,0) - z y + log1p(exp(-|z|)), weighted.
    double l = use ? (double)w * ((double)fmaxf(z, 0.0f) - (double)z * y +
                                  (double)log1pf(expf(-fabsf(z))))
                   : 0.0;
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
    if (j == 0 && loss_rows) loss_rows[r] = l;
}

// dH = bf16(G * [H > 0]).
__global__ void k_relu_back(const float *__restrict__ G, const uint16_t *__restrict__ H,
                            int64_t n, uint16_t *__restrict__ dH) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dH[i] = f2bf(H[i] != 0 && (H[i] & 0x8000u) == 0 ? G[i] : 0.0f);
}

// Column sums of a [rows][64] BF16 matrix (bias gradients): block c sums
// rows [c * chunk_rows, (c+1) * chunk_rows) into part[c][64]; k_sum_partials
// adds the chunks.
__global__ void k_colsum(const uint16_t *__restrict__ D, int64_t rows, int chunk_rows,
                         float *__restrict__ part) {
    __shared__ float red[4][64];
    const int j = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * chunk_rows;
    const int64_t r1 = min(rows, r0 + chunk_rows);
    float s = 0.0f;
    for (int64_t r = r0 + w; r < r1; r += 4) s += bf2f(D[r * 64 + j]);
    red[w][j] = s;
    __syncthreads();
    if (w == 0) part[(size_t)blockIdx.x * 64 + j] = red[0][j] + red[1][j] + red[2][j] + red[3][j];
}

// Sum of per-row losses (one block, FP64).
__global__ void k_sum_rows(const double *__restrict__ l, int64_t n, double *__restrict__ out) {
    __shared__ double part[4];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += l[i];
    s = wave_sum_d(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = part[0] + part[1] + part[2] + part[3];
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

struct PropLayout {
    size_t X, C, H1, H2, Z, dZ, dH, part, loss, total;
    int64_t rows;
    int chunks;
};
inline size_t a256(size_t x) { return (x + 255) / 256 * 256; }
constexpr int kChunkRows = 1024;  // split-K chunk of the weight-gradient GEMMs

PropLayout prop_layout(int64_t len) {
    PropLayout L;
    L.rows = (len + 63) / 64 * 64;
    L.chunks = (int)((L.rows + kChunkRows - 1) / kChunkRows);
    size_t o = 0;
    L.X = o;
    o += a256((size_t)L.rows * kIn * 2);
    L.C = o;
    o += a256((size_t)L.rows * 64 * 4);
    L.H1 = o;
    o += a256((size_t)L.rows * 64 * 2);
    L.H2 = o;
    o += a256((size_t)L.rows * 64 * 2);
    L.Z = o;
    o += a256((size_t)L.rows * 64 * 4);
    L.dZ = o;
    o += a256((size_t)L.rows * 64 * 2);
    L.dH = o;
    o += a256((size_t)L.rows * 64 * 2);
    L.part = o;
    o += a256((size_t)L.chunks * kIn * 64 * 4);
    L.loss = o;
    o += a256((size_t)L.rows * 8 + 8);
    L.total = o;
    return L;
}

void gemm(bool ta, bool tb, const uint16_t *A, const uint16_t *B, float *C, int M, int K,
          int chunks, hipStream_t st) {
    // Chunks are whole 64-row tiles, so no 16-deep K step crosses a chunk.
    const int kchunk = ((K + chunks - 1) / chunks + 63) / 64 * 64;
    const dim3 grid((unsigned)((M + 63) / 64), (unsigned)chunks);
    if (!ta && !tb) hipLaunchKernelGGL((k_gemm_bf16<false, false>), grid, dim3(256), 0, st, A, B, C, M, K, kchunk);
    else if (!ta && tb) hipLaunchKernelGGL((k_gemm_bf16<false, true>), grid, dim3(256), 0, st, A, B, C, M, K, kchunk);
    else hipLaunchKernelGGL((k_gemm_bf16<true, false>), grid, dim3(256), 0, st, A, B, C, M, K, kchunk);
}

// Bias gradient: column sums of D in chunks of kChunkRows rows, then added.
void colsum(const uint16_t *D, const PropLayout &L, float *part, float *out, hipStream_t st) {
    hipLaunchKernelGGL(k_colsum, dim3((unsigned)L.chunks), dim3(256), 0, st, D, L.rows,
                       kChunkRows, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, st, part, L.chunks, 64, out);
}

int32_t check_prop(const pxl_proposal *p, const char *who) {
    PXL_REQUIRE(p && p->weights && p->master && p->m && p->v && p->grad, "%s: null proposal buffer",
                who);
    return PXL_OK;
}

// Forward through the three layers; leaves H1, H2 (BF16) and the last
// layer's FP32 product in C.
void forward(const pxl_proposal *p, const char *w, const PropLayout &L, int64_t len,
             hipStream_t st) {
    const uint16_t *X = (const uint16_t *)(w + L.X);
    float *C = (float *)(w + L.C);
    uint16_t *H1 = (uint16_t *)(w + L.H1), *H2 = (uint16_t *)(w + L.H2);
    const uint16_t *W1 = p->weights, *W2 = W1 + kW1, *W3 = W2 + kW2;
    const float *b1 = p->master + kWeights, *b2 = b1 + kW, *b3 = b2 + kW;
    (void)b3;
    const int M = (int)L.rows;
    const unsigned eb = (unsigned)((L.rows * 64 + 255) / 256);
    gemm(false, false, X, W1, C, M, kIn, 1, st);
    hipLaunchKernelGGL(k_bias_relu, dim3(eb), dim3(256), 0, st, C, b1, len, L.rows, H1);
    gemm(false, false, H1, W2, C, M, kW, 1, st);
    hipLaunchKernelGGL(k_bias_relu, dim3(eb), dim3(256), 0, st, C, b2, len, L.rows, H2);
    gemm(false, false, H2, W3, C, M, kW, 1, st);
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
    forward(p, w, L, len, s);
    const float *b3 = p->master + kWeights + 2 * kW;
    float *Z = logits ? logits : (float *)(w + L.Z);
    hipLaunchKernelGGL(k_logits_bce, dim3((unsigned)((L.rows + 3) / 4)), dim3(256), 0, s,
                       (const float *)(w + L.C), b3, len, L.rows, Z, nullptr, nullptr, 1.0f,
                       1.0f, nullptr, nullptr);
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
    forward(p, w, L, len, s);
    const int M = (int)L.rows;
    const uint16_t *X = (const uint16_t *)(w + L.X);
    float *C = (float *)(w + L.C);
    uint16_t *H1 = (uint16_t *)(w + L.H1), *H2 = (uint16_t *)(w + L.H2);
    uint16_t *dZ = (uint16_t *)(w + L.dZ), *dH = (uint16_t *)(w + L.dH);
    float *part = (float *)(w + L.part);
    double *lrow = (double *)(w + L.loss);
    const uint16_t *W2 = p->weights + kW1, *W3 = W2 + kW2;
    const float *b3 = p->master + kWeights + 2 * kW;
    float *g = p->grad;
    const float scale = 1.0f / (float)((n_valid > 0 ? n_valid : len) * 64);
    hipLaunchKernelGGL(k_logits_bce, dim3((unsigned)((L.rows + 3) / 4)), dim3(256), 0, s, C, b3,
                       len, L.rows, (float *)(w + L.Z), targets, valid, pos_weight, scale, dZ,
                       lrow);
    const unsigned n64 = (unsigned)((L.rows * 64 + 255) / 256);
    // Layer 3: dW3 = H2^T dZ, db3 = colsum(dZ), G = dZ W3^T.
    gemm(true, false, H2, dZ, part, kW, M, L.chunks, s);
    hipLaunchKernelGGL(k_sum_partials, dim3(kW3 / 256), dim3(256), 0, s, part, L.chunks, kW3,
                       g + kW1 + kW2);
    colsum(dZ, L, part, g + kWeights + 2 * kW, s);
    gemm(false, true, dZ, W3, C, M, kBins, 1, s);
    hipLaunchKernelGGL(k_relu_back, dim3(n64), dim3(256), 0, s, C, H2, L.rows * 64, dH);
    // Layer 2: dW2 = H1^T dH2, db2, G = dH2 W2^T.
    gemm(true, false, H1, dH, part, kW, M, L.chunks, s);
    hipLaunchKernelGGL(k_sum_partials, dim3(kW2 / 256), dim3(256), 0, s, part, L.chunks, kW2,
                       g + kW1);
    colsum(dH, L, part, g + kWeights + kW, s);
    gemm(false, true, dH, W2, C, M, kW, 1, s);
    hipLaunchKernelGGL(k_relu_back, dim3(n64), dim3(256), 0, s, C, H1, L.rows * 64, dZ);
    // Layer 1: dW1 = X^T dH1, db1.
    gemm(true, false, X, dZ, part, kIn, M, L.chunks, s);
    hipLaunchKernelGGL(k_sum_partials, dim3(kW1 / 256), dim3(256), 0, s, part, L.chunks, kW1, g);
    colsum(dZ, L, part, g + kWeights, s);
    if (loss_out)
        hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(256), 0, s, lrow, L.rows, loss_out);
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
