// Scene structure: cell occupancy bitfield, in-cage flags, the caged
// (sparse) index, compaction and plane import/export.
#include "common.hpp"

namespace pxl {
namespace {

__global__ void k_fill_u8(uint8_t *p, int64_t n, uint8_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// pruneScene + occupiedVoxel (SceneUtils.cu:800-881), one thread per cell
// instead of one per (x, y) column.  A wave's 64 consecutive cells are packed
// by ballot into two 32-bit words of the bitfield.  In-cage marking keeps
// the reference's rule: the occupied cell's base vertex, plus the z = D-1
// and y = H-1 faces, plus - for x = W-2 - the vertex at linear index
// z*W*H + H*W + x + 1, which is what SceneUtils.cu:870-873 writes.
__global__ __launch_bounds__(256) void k_occupancy(const float *__restrict__ dens, int W, int H,
                                                   int D, float tau, uint32_t *__restrict__ occ,
                                                   uint8_t *__restrict__ in_cage,
                                                   unsigned long long *n_occ, int64_t ncell,
                                                   int64_t nwords) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool o = false;
    if (c < ncell) {
        const int64_t cw = W - 1, ch = H - 1;
        const int x = (int)(c % cw);
        const int y = (int)((c / cw) % ch);
        const int z = (int)(c / (cw * ch));
        const int64_t WH = (int64_t)W * H;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int dx = j & 1, dy = (j >> 1) & 1, dz = j >> 2;
            const int64_t v = (int64_t)(z + dz) * WH + (int64_t)(y + dy) * W + (x + dx);
            o = o || (dens[v] > tau);
        }
        if (o) {
            in_cage[(int64_t)z * WH + (int64_t)y * W + x] = 1;
            if (z == D - 2) in_cage[(int64_t)(D - 1) * WH + (int64_t)y * W + x] = 1;
            if (y == H - 2) in_cage[(int64_t)z * WH + (int64_t)(H - 1) * W + x] = 1;
            if (x == W - 2) in_cage[(int64_t)z * WH + WH + x + 1] = 1;
        }
    }
    const uint64_t bal = __ballot(o);
    if ((threadIdx.x & 63) == 0) {
        const int64_t w0 = (c >> 5);  // c is the wave's first cell, a multiple of 64
        if (w0 < nwords) occ[w0] = (uint32_t)bal;
        if (w0 + 1 < nwords) occ[w0 + 1] = (uint32_t)(bal >> 32);
        if (n_occ && bal) atomicAdd(n_occ, (unsigned long long)__popcll(bal));
    }
}

// One thread per 8x8x8-cell brick; a wave's 64 bricks are packed by ballot.
__global__ __launch_bounds__(256) void k_bricks(const uint32_t *__restrict__ occ, int W, int H,
                                                int D, int nbx, int nby, int64_t nbricks,
                                                int64_t nwords, uint32_t *__restrict__ bricks) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool any = false;
    if (b < nbricks) {
        const int bx = (int)(b % nbx), by = (int)((b / nbx) % nby), bz = (int)(b / ((int64_t)nbx * nby));
        const int cw = W - 1, ch = H - 1, cd = D - 1;
        for (int z = bz * 8; z < min(bz * 8 + 8, cd) && !any; z++)
            for (int y = by * 8; y < min(by * 8 + 8, ch) && !any; y++)
                for (int x = bx * 8; x < min(bx * 8 + 8, cw); x++) {
                    const int64_t c = ((int64_t)z * ch + y) * cw + x;
                    if ((occ[c >> 5] >> (c & 31)) & 1u) {
                        any = true;
                        break;
                    }
                }
    }
    const uint64_t bal = __ballot(any);
    if ((threadIdx.x & 63) == 0) {
        const int64_t w0 = b >> 5;
        if (w0 < nwords) bricks[w0] = (uint32_t)bal;
        if (w0 + 1 < nwords) bricks[w0 + 1] = (uint32_t)(bal >> 32);
    }
}

// The brick map dilated by one brick: bit b set when brick b or any of its
// 26 neighbours (within the grid) has an occupied cell.  A ray segment
// shorter than a brick side that starts in brick b stays inside b and its
// neighbours, so a clear dilated bit clears the whole segment.
__global__ __launch_bounds__(256) void k_bricks_dilate(const uint32_t *__restrict__ bricks,
                                                       int nbx, int nby, int nbz,
                                                       int64_t nbricks, int64_t nwords,
                                                       uint32_t *__restrict__ dilated) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool any = false;
    if (b < nbricks) {
        const int bx = (int)(b % nbx), by = (int)((b / nbx) % nby), bz = (int)(b / ((int64_t)nbx * nby));
        for (int z = max(bz - 1, 0); z <= min(bz + 1, nbz - 1) && !any; z++)
            for (int y = max(by - 1, 0); y <= min(by + 1, nby - 1) && !any; y++)
                for (int x = max(bx - 1, 0); x <= min(bx + 1, nbx - 1); x++) {
                    const int64_t n = ((int64_t)z * nby + y) * nbx + x;
                    if ((bricks[n >> 5] >> (n & 31)) & 1u) {
                        any = true;
                        break;
                    }
                }
    }
    const uint64_t bal = __ballot(any);
    if ((threadIdx.x & 63) == 0) {
        const int64_t w0 = b >> 5;
        if (w0 < nwords) dilated[w0] = (uint32_t)bal;
        if (w0 + 1 < nwords) dilated[w0 + 1] = (uint32_t)(bal >> 32);
    }
}

constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 16;
constexpr int64_t kScanTile = (int64_t)kScanThreads * kScanPerThread;

// 16 consecutive flags of one thread, as 0/1 bits.
__device__ __forceinline__ uint32_t load_flags16(const uint8_t *f, int64_t base, int64_t n) {
    uint32_t bits = 0;
    if (base + 16 <= n) {
        const uint4 q = *reinterpret_cast<const uint4 *>(f + base);
        const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int bb = 0; bb < 4; bb++)
                if ((w[i] >> (8 * bb)) & 0xFFu) bits |= 1u << (i * 4 + bb);
    } else {
        for (int i = 0; i < 16; i++)
            if (base + i < n && f[base + i]) bits |= 1u << i;
    }
    return bits;
}

// Block-wide exclusive scan of one value per thread (256 threads).
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *lds_wave,
                                                    uint32_t &total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t inc = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(inc, off, 64);
        if (lane >= off) inc += u;
    }
    if (lane == 63) lds_wave[wid] = inc;
    __syncthreads();
    uint32_t wbase = 0;
    total = 0;
    for (int i = 0; i < kScanThreads / 64; i++) {
        const uint32_t t = lds_wave[i];
        if (i < wid) wbase += t;
        total += t;
    }
    __syncthreads();
    return wbase + inc - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_count(const uint8_t *__restrict__ f,
                                                             int64_t n,
                                                             uint32_t *__restrict__ block_sums) {
    __shared__ uint32_t lw[kScanThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPerThread;
    const uint32_t c = __popc(load_flags16(f, base, n));
    uint32_t total;
    (void)block_excl_scan(c, lw, total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Exclusive scan of the per-tile counts by a single block.
__global__ __launch_bounds__(kScanThreads) void k_scan_blocks(uint32_t *__restrict__ sums,
                                                              int64_t nb,
                                                              uint32_t *__restrict__ total_out) {
    __shared__ uint32_t lw[kScanThreads / 64];
    uint32_t carry = 0;
    for (int64_t base = 0; base < nb; base += kScanThreads) {
        const int64_t i = base + threadIdx.x;
        const uint32_t v = i < nb ? sums[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan(v, lw, tot);
        if (i < nb) sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total_out = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_write(
    const uint8_t *__restrict__ f, int64_t n, const uint32_t *__restrict__ block_off,
    uint32_t *__restrict__ io, uint32_t *__restrict__ c2d) {
    __shared__ uint32_t lw[kScanThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPerThread;
    const uint32_t bits = load_flags16(f, base, n);
    uint32_t total;
    uint32_t run = block_excl_scan(__popc(bits), lw, total) + block_off[blockIdx.x];
    uint32_t out[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        if ((bits >> i) & 1u) {
            out[i] = run;
            if (c2d) c2d[run] = (uint32_t)(base + i);
            run++;
        } else {
            out[i] = 0u;
        }
    }
    if (base + 16 <= n) {
        uint4 *dst = reinterpret_cast<uint4 *>(io + base);
#pragma unroll
        for (int q = 0; q < 4; q++)
            dst[q] = make_uint4(out[4 * q], out[4 * q + 1], out[4 * q + 2], out[4 * q + 3]);
    } else {
        for (int i = 0; i < 16; i++)
            if (base + i < n) io[base + i] = out[i];
    }
}

__global__ void k_compact(const float4 *__restrict__ dense, int q_per_rec,
                          const uint32_t *__restrict__ c2d, int64_t n_vox,
                          float4 *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_vox * q_per_rec) return;
    const int64_t r = i / q_per_rec;
    const int q = (int)(i % q_per_rec);
    out[i] = dense[(int64_t)c2d[r] * q_per_rec + q];
}

__global__ void k_to16(const float *__restrict__ rec32, int s32, int64_t n, uint16_t *rec16,
                       int s16, int storage) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * s16) return;
    const int64_t r = i / s16;
    const int e = (int)(i % s16);
    uint16_t h = 0;
    if (e < s32) {
        const float f = rec32[r * s32 + e];
        if (storage == PXL_STORE_BF16) {
            uint32_t u = __float_as_uint(f);
            if ((u & 0x7f800000u) == 0x7f800000u) {
                h = (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
            } else {
                u += 0x7fffu + ((u >> 16) & 1u);
                h = (uint16_t)(u >> 16);
            }
        } else {
            h = __half_as_ushort(__float2half_rn(f));
        }
    }
    rec16[i] = h;
}

constexpr int kMaxPlanes = 80;
struct PlaneTable {
    const float *p[kMaxPlanes];
};

__global__ void k_import(PlaneTable t, int np, int64_t n_vox, float *__restrict__ rec,
                         int stride) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_vox) return;
    float *r = rec + v * stride;
    for (int i = 0; i < stride; i++) r[i] = i < np ? t.p[i][v] : 0.0f;
}

__global__ void k_export(const float *__restrict__ rec, int stride, int64_t n_vox, PlaneTable t,
                         int np) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_vox) return;
    const float *r = rec + v * stride;
    for (int i = 0; i < np; i++) const_cast<float *>(t.p[i])[v] = r[i];
}

}  // namespace
}  // namespace pxl

using namespace pxl;

extern "C" int32_t pxl_build_occupancy(const float *density, int32_t W, int32_t H, int32_t D,
                                       float threshold, uint32_t *occupancy, uint8_t *in_cage,
                                       uint64_t *n_occupied, void *stream) {
    clear_error();
    PXL_REQUIRE(density && occupancy && in_cage, "pxl_build_occupancy: null argument");
    PXL_REQUIRE(W >= 2 && H >= 2 && D >= 2, "pxl_build_occupancy: grid %dx%dx%d too small", W,
                H, D);
    const int64_t nv = (int64_t)W * H * D;
    const int64_t ncell = (int64_t)(W - 1) * (H - 1) * (D - 1);
    const int64_t nwords = (ncell + 31) / 32;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s, in_cage,
                       nv, (uint8_t)0);
    if (n_occupied)
        hipLaunchKernelGGL(k_fill_u8, dim3(1), dim3(64), 0, s,
                           reinterpret_cast<uint8_t *>(n_occupied), (int64_t)8, (uint8_t)0);
    const unsigned blocks = (unsigned)((ncell + 255) / 256);
    hipLaunchKernelGGL(k_occupancy, dim3(blocks), dim3(256), 0, s, density, W, H, D, threshold,
                       occupancy, in_cage, reinterpret_cast<unsigned long long *>(n_occupied),
                       ncell, nwords);
    return check_launch("pxl_build_occupancy");
}

extern "C" int64_t pxl_brick_words(int32_t W, int32_t H, int32_t D) {
    if (W < 2 || H < 2 || D < 2) return 0;
    const int64_t n = (int64_t)((W - 1 + 7) / 8) * ((H - 1 + 7) / 8) * ((D - 1 + 7) / 8);
    return 2 * ((n + 31) / 32);
}

extern "C" int32_t pxl_build_brick_occupancy(const uint32_t *occupancy, int32_t W, int32_t H,
                                             int32_t D, uint32_t *bricks, void *stream) {
    clear_error();
    PXL_REQUIRE(occupancy && bricks, "pxl_build_brick_occupancy: null argument");
    PXL_REQUIRE(W >= 2 && H >= 2 && D >= 2, "pxl_build_brick_occupancy: grid too small");
    const int nbx = (W - 1 + 7) / 8, nby = (H - 1 + 7) / 8, nbz = (D - 1 + 7) / 8;
    const int64_t nb = (int64_t)nbx * nby * nbz;
    const int64_t nwords = (nb + 31) / 32;
    hipLaunchKernelGGL(k_bricks, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0,
                       as_stream(stream), occupancy, W, H, D, nbx, nby, nb, nwords, bricks);
    hipLaunchKernelGGL(k_bricks_dilate, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0,
                       as_stream(stream), bricks, nbx, nby, nbz, nb, nwords, bricks + nwords);
    return check_launch("pxl_build_brick_occupancy");
}

extern "C" size_t pxl_sparse_index_scratch_bytes(int64_t n_vertices) {
    if (n_vertices <= 0) return 0;
    const int64_t nb = (n_vertices + kScanTile - 1) / kScanTile;
    return (size_t)(((nb * 4) + 255) / 256 * 256);
}

extern "C" int32_t pxl_build_sparse_index(const uint8_t *in_cage, int32_t W, int32_t H,
                                          int32_t D, uint32_t *index_offsets,
                                          uint32_t *compact_to_dense, uint32_t *in_cage_nums,
                                          void *scratch, size_t scratch_bytes, void *stream) {
    clear_error();
    PXL_REQUIRE(in_cage && index_offsets && in_cage_nums, "pxl_build_sparse_index: null argument");
    PXL_REQUIRE(W >= 1 && H >= 1 && D >= 1, "pxl_build_sparse_index: bad dims");
    PXL_REQUIRE(((uintptr_t)in_cage & 15) == 0 && ((uintptr_t)index_offsets & 15) == 0,
                "pxl_build_sparse_index: in_cage and index_offsets must be 16-byte aligned");
    const int64_t n = (int64_t)W * H * D;
    PXL_REQUIRE(n < ((int64_t)1 << 32), "pxl_build_sparse_index: too many vertices");
    const size_t need = pxl_sparse_index_scratch_bytes(n);
    if (scratch_bytes < need || (need && !scratch)) {
        set_error("pxl_build_sparse_index: scratch of %zu bytes, need %zu", scratch_bytes, need);
        return PXL_ENOMEM;
    }
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    uint32_t *sums = reinterpret_cast<uint32_t *>(scratch);
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_scan_count, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in_cage, n,
                       sums);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kScanThreads), 0, s, sums, nb, in_cage_nums);
    hipLaunchKernelGGL(k_scan_write, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in_cage, n,
                       sums, index_offsets, compact_to_dense);
    return check_launch("pxl_build_sparse_index");
}

extern "C" int32_t pxl_grid_compact(const float *dense_records, int32_t stride,
                                    const uint32_t *compact_to_dense, int64_t n_vox,
                                    float *sparse_records, void *stream) {
    clear_error();
    PXL_REQUIRE(dense_records && compact_to_dense && sparse_records,
                "pxl_grid_compact: null argument");
    PXL_REQUIRE(stride > 0 && stride % 4 == 0, "pxl_grid_compact: bad stride %d", stride);
    PXL_REQUIRE(n_vox >= 0, "pxl_grid_compact: bad n_vox");
    PXL_REQUIRE(((uintptr_t)dense_records | (uintptr_t)sparse_records) % 16 == 0,
                "pxl_grid_compact: record buffers must be 16-byte aligned");
    if (n_vox == 0) return PXL_OK;
    const int q = stride / 4;
    const int64_t total = n_vox * q;
    hipLaunchKernelGGL(k_compact, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       as_stream(stream), reinterpret_cast<const float4 *>(dense_records), q,
                       compact_to_dense, n_vox, reinterpret_cast<float4 *>(sparse_records));
    return check_launch("pxl_grid_compact");
}

extern "C" int32_t pxl_grid_import_planes(const float *const *planes, int32_t n_planes,
                                          int64_t n_vox, float *records, int32_t stride,
                                          void *stream) {
    clear_error();
    PXL_REQUIRE(planes && records, "pxl_grid_import_planes: null argument");
    PXL_REQUIRE(n_planes > 0 && n_planes <= kMaxPlanes && n_planes <= stride,
                "pxl_grid_import_planes: bad plane count %d (stride %d)", n_planes, stride);
    PlaneTable t;
    memset(&t, 0, sizeof(t));
    for (int i = 0; i < n_planes; i++) {
        PXL_REQUIRE(planes[i] != nullptr, "pxl_grid_import_planes: plane %d is null", i);
        t.p[i] = planes[i];
    }
    if (n_vox <= 0) return PXL_OK;
    hipLaunchKernelGGL(k_import, dim3((unsigned)((n_vox + 255) / 256)), dim3(256), 0,
                       as_stream(stream), t, n_planes, n_vox, records, stride);
    return check_launch("pxl_grid_import_planes");
}

extern "C" int32_t pxl_grid_export_planes(const float *records, int32_t stride, int64_t n_vox,
                                          float *const *planes, int32_t n_planes,
                                          void *stream) {
    clear_error();
    PXL_REQUIRE(planes && records, "pxl_grid_export_planes: null argument");
    PXL_REQUIRE(n_planes > 0 && n_planes <= kMaxPlanes && n_planes <= stride,
                "pxl_grid_export_planes: bad plane count %d (stride %d)", n_planes, stride);
    PlaneTable t;
    memset(&t, 0, sizeof(t));
    for (int i = 0; i < n_planes; i++) {
        PXL_REQUIRE(planes[i] != nullptr, "pxl_grid_export_planes: plane %d is null", i);
        t.p[i] = planes[i];
    }
    if (n_vox <= 0) return PXL_OK;
    hipLaunchKernelGGL(k_export, dim3((unsigned)((n_vox + 255) / 256)), dim3(256), 0,
                       as_stream(stream), records, stride, n_vox, t, n_planes);
    return check_launch("pxl_grid_export_planes");
}

extern "C" int32_t pxl_records_to_16bit(const float *rec32, int32_t stride32, int64_t n_vox,
                                        void *rec16, int32_t stride16, int32_t storage,
                                        void *stream) {
    clear_error();
    PXL_REQUIRE(rec32 && rec16, "pxl_records_to_16bit: null argument");
    PXL_REQUIRE(storage == PXL_STORE_BF16 || storage == PXL_STORE_F16,
                "pxl_records_to_16bit: storage must be PXL_STORE_BF16 or PXL_STORE_F16");
    PXL_REQUIRE(stride32 > 0 && stride16 > 0 && n_vox >= 0, "pxl_records_to_16bit: bad sizes");
    if (n_vox == 0) return PXL_OK;
    const int64_t total = n_vox * stride16;
    hipLaunchKernelGGL(k_to16, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       as_stream(stream), rec32, stride32, n_vox,
                       reinterpret_cast<uint16_t *>(rec16), stride16, storage);
    return check_launch("pxl_records_to_16bit");
}
