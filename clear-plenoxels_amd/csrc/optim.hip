// Regulariser gradients (variance / TV loss), the Adam step and the FP64
// batch-loss reduction.
#include <cstdlib>

#include "common.hpp"

namespace pxl {
namespace {

struct TvArgs {
    const float *params;  // elements of the storage type
    float *grad;
    int gstride;          // floats per gradient record
    const uint32_t *occ;
    const uint32_t *io;
    const uint32_t *c2d;
    const uint8_t *cage;  // in-cage flag per dense vertex
    int W, H, D, stride, P;
    int S, sh_lo, sh_hi;  // SH terms for coefficients [sh_lo, sh_hi) of each channel
    int sparse;
    int64_t n_vox;
    float lam_d, lam_sh;
};

// The Adam step fused into a TV pass (pxl_tv_adam_step): the records the
// TV reads (TvArgs::params) stay as they are and the updated ones go to
// `out` (same layout), so no lane reads a neighbour another lane has
// already stepped.  Optimiser records (grad, m, v, master) hold ostride
// floats; q4 = 4 ceil(P / 4), the elements the Adam kernel visits.
struct AdamArgs {
    void *out;
    float *m, *v, *master;
    const float *alr;
    int ostride, q4;
    double lr, sh_lr, bc1, bc2;
};

// updateRGBSceneGPU_ADAM's update of element p (Optimizer.cu:380-412), the
// float/double mix expression by expression: false (nothing changes) for an
// SH coefficient outside the optimised band [sh_lo, sh_hi).
__device__ __forceinline__ bool adam_elem(int p, int S, int sh_lo, int sh_hi, double lr_d,
                                          double sh_lr, double bc1, double bc2, float &th,
                                          float &g, float &m, float &v) {
    if (p > 0 && ((p - 1) % S < sh_lo || (p - 1) % S >= sh_hi)) return false;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1.e-6f;
    const float gg = g;
    const float mm = (float)((double)(b1 * m) + (1.0 - (double)b1) * (double)gg);
    const float vv = (float)((double)(b2 * v) + (1.0 - (double)b2) * (double)gg * (double)gg);
    const float cm = (float)((double)mm / bc1);
    const float cv = (float)((double)vv / bc2);
    const double upd = (p == 0 ? lr_d : sh_lr) * (double)cm / (double)(sqrtf(cv) + eps);
    const double nt = (double)th - upd;
    th = p == 0 ? (nt < 0.0 ? 0.0f : (float)nt) : (float)nt;
    m = mm;
    v = vv;
    g = 0.0f;
    return true;
}

// Coefficient of parameter p: the density's, or the SH one for the
// coefficients of the band being optimised (startSHIdx .. (optimizingBand +
// 1)^2, RenderUtils.cu:2480-2493, 2594), else 0.
__device__ __forceinline__ float tv_lambda(const TvArgs &a, int p) {
    if (p == 0) return a.lam_d;
    const int k = (p - 1) % a.S;
    return (k >= a.sh_lo && k < a.sh_hi) ? a.lam_sh : 0.0f;
}

template <int STORE>
__device__ __forceinline__ float ld(const float *base, size_t idx) {
    if (STORE == PXL_STORE_F32) return base[idx];
    const uint16_t h = reinterpret_cast<const uint16_t *>(base)[idx];
    if (STORE == PXL_STORE_BF16) return __uint_as_float((uint32_t)h << 16);
    return __half2float(__ushort_as_half(h));
}

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

template <int STORE>
__device__ __forceinline__ float widen16(uint16_t h) {
    if (STORE == PXL_STORE_BF16) return __uint_as_float((uint32_t)h << 16);
    return __half2float(__ushort_as_half(h));
}
template <int STORE>
__device__ __forceinline__ uint16_t narrow16(float v) {
    if (STORE == PXL_STORE_BF16) return to_bf16(v);
    return __half_as_ushort(__float2half_rn(v));
}

// Streaming store (non-temporal hint) of the Adam moments: no kernel reads
// them before the next Adam step, so they need not stay in the caches.
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void st_stream(float *p, const float4 &v) {
    const f4v w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<f4v *>(p));
}

__device__ __forceinline__ void st_stream1(float *p, float v) { __builtin_nontemporal_store(v, p); }

// One element of the fused Adam step: lane p of record r loads it before
// the TV gathers (AdamElem: parameter, gradient, moments) and stores it once
// the TV sum is known.  Elements the Adam kernel does not update keep their
// value in `out`: SH coefficients outside the band, padding in [P, q4)
// (rounded through the FP32 value, as the Adam kernel writes it back) and
// the rest of the record (copied bit for bit).
struct AdamElem {
    float th, g, m, v;
};

template <int STORE>
__device__ __forceinline__ AdamElem adam_load(const TvArgs &a, const AdamArgs &ad, int64_t r,
                                              int p) {
    AdamElem e{0.0f, 0.0f, 0.0f, 0.0f};
    const size_t po = (size_t)r * (size_t)a.stride + (size_t)p;
    const size_t o = (size_t)r * (size_t)ad.ostride + (size_t)p;
    if (p < a.P) {
        e.g = a.grad[(size_t)r * (size_t)a.gstride + (size_t)p];
        e.m = ad.m[o];
        e.v = ad.v[o];
    }
    if (STORE == PXL_STORE_F32) {
        e.th = a.params[po];
    } else if (p < ad.q4) {
        e.th = ad.master != nullptr ? ad.master[o] : ld<STORE>(a.params, po);
    } else {
        e.th = __uint_as_float((uint32_t)reinterpret_cast<const uint16_t *>(a.params)[po]);
    }
    return e;
}

template <int STORE>
__device__ __forceinline__ void adam_store(const TvArgs &a, const AdamArgs &ad, int64_t r, int p,
                                           AdamElem e, bool tvl, float lam, double tsum) {
    const size_t po = (size_t)r * (size_t)a.stride + (size_t)p;
    if (p < a.P) {
        const size_t o = (size_t)r * (size_t)ad.ostride + (size_t)p;
        const size_t go = (size_t)r * (size_t)a.gstride + (size_t)p;
        if (tvl) e.g += (float)((double)lam * tsum);
        const double lr_d = ad.alr != nullptr ? ad.lr * (double)ad.alr[r] : ad.lr;
        if (adam_elem(p, a.S, a.sh_lo, a.sh_hi, lr_d, ad.sh_lr, ad.bc1, ad.bc2, e.th, e.g, e.m,
                      e.v)) {
            a.grad[go] = e.g;
            st_stream1(ad.m + o, e.m);
            st_stream1(ad.v + o, e.v);
            if (STORE != PXL_STORE_F32 && ad.master != nullptr) ad.master[o] = e.th;
        } else if (tvl) {
            a.grad[go] = e.g;
        }
    }
    if (STORE == PXL_STORE_F32)
        reinterpret_cast<float *>(ad.out)[po] = e.th;
    else if (p < ad.q4)
        reinterpret_cast<uint16_t *>(ad.out)[po] = narrow16<STORE>(e.th);
    else
        reinterpret_cast<uint16_t *>(ad.out)[po] = (uint16_t)__float_as_uint(e.th);
}

__device__ __forceinline__ bool cell_occupied(const TvArgs &a, int cx, int cy, int cz) {
    const uint32_t cell = (uint32_t)cz * (uint32_t)(a.W - 1) * (uint32_t)(a.H - 1) +
                          (uint32_t)cy * (uint32_t)(a.W - 1) + (uint32_t)cx;
    return (a.occ[cell >> 5] >> (cell & 31)) & 1u;
}

// Record of vertex (x, y, z); an index beyond the grid's records (a caged
// index that does not belong to this grid) reads record 0.
__device__ __forceinline__ uint32_t record_of(const TvArgs &a, int x, int y, int z) {
    const uint32_t d = (uint32_t)z * (uint32_t)a.W * (uint32_t)a.H + (uint32_t)y * (uint32_t)a.W +
                       (uint32_t)x;
    if (!a.sparse) return d;
    const uint32_t r = a.io[d];
    return (int64_t)r < a.n_vox ? r : 0u;
}

// Cells around a vertex that contain its neighbour i = (dz+1)*9 + (dy+1)*3 +
// (dx+1): bit c = cell (x - (c&1), y - (c>>1&1), z - (c>>2)) contains it
// when dx is -ax or 1-ax, and likewise for y and z.
__device__ __forceinline__ uint32_t tv_cells_of(int i) {
    const int dx = i % 3 - 1, dy = (i / 3) % 3 - 1, dz = i / 9 - 1;
    uint32_t cells = 0u;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const int ax = c & 1, ay = (c >> 1) & 1, az = c >> 2;
        if ((dx == -ax || dx == 1 - ax) && (dy == -ay || dy == 1 - ay) &&
            (dz == -az || dz == 1 - az))
            cells |= 1u << c;
    }
    return cells;
}

// addVarianceLossGPU (RenderUtils.cu:2469-2547) as a race-free gather.  The
// reference visits every cell (every occupied cell when sparse), takes the
// mean mu of its 8 corners and adds lambda * (u_corner - mu) to each
// corner's gradient with unsynchronised +=.  Here each record gathers the
// same terms from the (up to 8) cells that have its vertex as a corner:
// one group of LPV lanes per record, one lane per parameter; a neighbour
// value is loaded only when an occupied cell around the vertex contains it.
//
// TV sum of one record from its 27 neighbour records (lane i of the group
// holds neighbour i's record in nrec) and the mask of occupied cells around
// it (bit c = cell (x - (c&1), y - (c>>1&1), z - (c>>2))); lane q's
// element.  The gradient takes lambda times it.
template <int LPV, int STORE>
__device__ __forceinline__ double tv_record_sum(const TvArgs &a, int64_t r, int q, uint32_t nrec,
                                                uint32_t cmask) {
    // Every lane issues all 27 loads unconditionally, so they are in flight
    // together (a load under a branch is waited for at the branch's end).
    // A neighbour outside the occupied cells reads the record itself.
    float v[27];
#pragma unroll
    for (int i = 0; i < 27; i++) {
        const uint32_t cells = tv_cells_of(i);
        const uint32_t ri = (uint32_t)__shfl((int)nrec, i, LPV);
        const bool used = (cells & cmask) != 0u;
        v[i] = ld<STORE>(a.params, (size_t)(used ? ri : (uint32_t)r) * (size_t)a.stride + q);
    }
    // sum over occupied cells c of (u - mean_c) = n u - (1/8) sum_i m_i u_i,
    // n the occupied cells around the vertex, m_i those that also contain
    // neighbour i (m_13 = n).  Sums run in FP64: a vertex value and the mean
    // of a cell around it nearly cancel.
    double msum = 0.0;
#pragma unroll
    for (int i = 0; i < 27; i++)
        msum = fma((double)__builtin_popcount(tv_cells_of(i) & cmask), (double)v[i], msum);
    return (double)__builtin_popcount(cmask) * (double)v[13] - 0.125 * msum;
}

// Neighbour records and cell mask of record r, computed by its group: lane
// i < 27 looks up neighbour i = (dz+1)*9 + (dy+1)*3 + (dx+1), lane c < 8
// tests cell c.
template <int LPV>
__device__ __forceinline__ uint32_t tv_neighbours(const TvArgs &a, int64_t r, int p,
                                                  uint32_t &cmask) {
    const uint32_t d = a.sparse ? a.c2d[r] : (uint32_t)r;
    const int x = (int)(d % (uint32_t)a.W);
    const int y = (int)((d / (uint32_t)a.W) % (uint32_t)a.H);
    const int z = (int)(d / ((uint32_t)a.W * (uint32_t)a.H));
    uint32_t nrec = 0u;
    if (p < 27) {
        const int nx = x + p % 3 - 1, ny = y + (p / 3) % 3 - 1, nz = z + p / 9 - 1;
        if (nx >= 0 && nx < a.W && ny >= 0 && ny < a.H && nz >= 0 && nz < a.D)
            nrec = record_of(a, nx, ny, nz);
    }
    bool cocc = false;
    if (p < 8) {
        const int cx = x - (p & 1), cy = y - ((p >> 1) & 1), cz = z - (p >> 2);
        cocc = cx >= 0 && cx <= a.W - 2 && cy >= 0 && cy <= a.H - 2 && cz >= 0 &&
               cz <= a.D - 2 && (!a.sparse || cell_occupied(a, cx, cy, cz));
    }
    const int seg = (int)(threadIdx.x & 63) & ~(LPV - 1);
    cmask = (uint32_t)(__ballot(cocc) >> seg) & 0xFFu;
    return nrec;
}

// One element of the storage type at a byte offset from base.
template <int STORE>
__device__ __forceinline__ float ld_at(const char *base, uint32_t off) {
    if (STORE == PXL_STORE_F32) return *reinterpret_cast<const float *>(base + off);
    const uint16_t h = *reinterpret_cast<const uint16_t *>(base + off);
    if (STORE == PXL_STORE_BF16) return __uint_as_float((uint32_t)h << 16);
    return __half2float(__ushort_as_half(h));
}

// One element of the storage type through a buffer resource: byte offset
// voff (per lane) + soff (wave-uniform).
template <int STORE>
__device__ __forceinline__ float ld_buf(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff) {
    if (STORE == PXL_STORE_F32)
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, (int)voff, (int)soff, 0));
    const uint16_t h = __builtin_amdgcn_raw_buffer_load_b16(rs, (int)voff, (int)soff, 0);
    if (STORE == PXL_STORE_BF16) return __uint_as_float((uint32_t)h << 16);
    return __half2float(__ushort_as_half(h));
}

// TV sum of an interior vertex of a dense grid (lane q's element): all 8 cells around it
// exist, so m_i = 2^(zero components of (dx, dy, dz)) and
//   n u - (1/8) sum_i m_i u_i = 7 u - F / 2 - E / 4 - C / 8
// with F, E, C the sums over the 6 face, 12 edge and 8 corner neighbours.
// Neighbour i lies at a grid-uniform record offset from the vertex's first
// neighbour r0 = r - (1 + W + W H).  OFF32 (the records span less than
// 4 GiB): the loads go through a buffer resource with the lane's byte
// offset from r0 in a vector register, shared by all 27 loads, and the
// neighbour's offset from r0 in a scalar register.
template <int STORE, bool OFF32>
__device__ __forceinline__ double tv_interior_sum(const TvArgs &a, int64_t r, int q) {
    constexpr uint32_t kElem = STORE == PXL_STORE_F32 ? 4u : 2u;
    const int64_t plane = (int64_t)a.W * a.H;
    const int64_t r0 = r - (1 + a.W + plane);
    float v[27];
    if (OFF32) {
        const uint32_t rec_bytes = (uint32_t)a.stride * kElem;
        const uint32_t bytes = (uint32_t)a.n_vox * rec_bytes;
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float *>(a.params), (short)0, (int)bytes, 0x00020000);
        const uint32_t lane_off = (uint32_t)r0 * rec_bytes + (uint32_t)q * kElem;
#pragma unroll
        for (int i = 0; i < 27; i++) {
            const uint32_t off = (uint32_t)(i % 3) + (uint32_t)((i / 3) % 3) * (uint32_t)a.W +
                                 (uint32_t)(i / 9) * (uint32_t)plane;
            v[i] = ld_buf<STORE>(rs, lane_off, off * rec_bytes);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 27; i++) {
            const int64_t off = (int64_t)(i % 3) + (int64_t)((i / 3) % 3) * a.W +
                                (int64_t)(i / 9) * plane;
            v[i] = ld<STORE>(a.params, (size_t)(r0 + off) * (size_t)a.stride + q);
        }
    }
    double f = 0.0, e = 0.0, c = 0.0;
#pragma unroll
    for (int i = 0; i < 27; i++) {
        const int nz = (i % 3 != 1) + ((i / 3) % 3 != 1) + (i / 9 != 1);
        if (nz == 1) f += (double)v[i];
        else if (nz == 2) e += (double)v[i];
        else if (nz == 3) c += (double)v[i];
    }
    return 7.0 * (double)v[13] - (0.5 * f + 0.25 * e + 0.125 * c);
}

// One vertex of addVarianceLossGPU's gather (record r) by its group of LPV
// lanes.
template <int LPV, int STORE, bool OFF32, bool FUSED>
__device__ __forceinline__ void tv_general_item(const TvArgs &a, const AdamArgs &ad, int64_t r,
                                                int g, int p) {
    const float lam = tv_lambda(a, p);
    const bool mine = p < a.P && lam != 0.0f;
    // A lane without a TV term reads element 0 (the line is read anyway).
    const int q = mine ? p : 0;
    AdamElem ae{0.0f, 0.0f, 0.0f, 0.0f};
    if (FUSED && p < a.stride) ae = adam_load<STORE>(a, ad, r, p);
    const uint32_t d = (uint32_t)r;
    const uint32_t x = d % (uint32_t)a.W, y = (d / (uint32_t)a.W) % (uint32_t)a.H;
    const uint32_t z = d / ((uint32_t)a.W * (uint32_t)a.H);
    const bool interior = x - 1u < (uint32_t)a.W - 2u && y - 1u < (uint32_t)a.H - 2u &&
                          z - 1u < (uint32_t)a.D - 2u;
    double tsum = 0.0;
    bool tvl = false;
    if (interior) {
        tsum = tv_interior_sum<STORE, OFF32>(a, r, q);
        tvl = mine;
    } else {
        uint32_t cmask;
        const uint32_t nrec = tv_neighbours<LPV>(a, r, p, cmask);
        // No occupied cell around the vertex: no TV term.
        if (cmask != 0u) {
            tsum = tv_record_sum<LPV, STORE>(a, r, q, nrec, cmask);
            tvl = mine;
        }
    }
    if (!FUSED) {
        if (tvl) a.grad[(size_t)r * (size_t)a.gstride + p] += (float)((double)lam * tsum);
        return;
    }
    if (p < a.stride) adam_store<STORE>(a, ad, r, p, ae, tvl, lam, tsum);
}

// Dense grids, marched along z.  Along each axis a vertex has a cell below
// (x >= 1) and one above (x <= W - 2); the cells around it that contain a
// neighbour at offset (dx, dy, dz) number m = wx(dx) wy(dy) wz(dz) with
// w(-1) = [cell below], w(+1) = [cell above], w(0) = their sum, and
// n = wx(0) wy(0) wz(0) cells surround the vertex, so its gradient is
//   n u - B / 8,  B = sum_z' wz(z' - z) Bxy(z'),  Bxy(z') = sum wx wy u
// over the 3 x 3 vertices around (x, y) in plane z'.  Inside the grid the
// weights are [1, 2, 1] on every axis and n = 8.  One group of LPV lanes
// (one per parameter) walks column (x, y) over kTvChunk vertices, keeping
// Bxy of the planes below and at z: 9 loads per vertex instead of 27.
// Columns are numbered x fastest, then y, then the z chunk; every column of
// the grid (W, H >= 3), its faces included, is one of them.  The 3 x 3
// block loaded per plane stays inside the grid: on a face it is shifted one
// column (row) inwards, and the neighbour it then holds past the vertex
// takes weight 0.  Loads go through a buffer resource (the records span
// less than 4 GiB): lane offset in a vector register, neighbour offset in a
// scalar one.  FUSED: each vertex's record then takes its Adam step into
// ad.out.
constexpr int kTvChunk = 128;

__device__ __forceinline__ double tv_blur_xy(const float (&v)[9]) {
    const double e = ((double)v[1] + (double)v[3]) + ((double)v[5] + (double)v[7]);
    const double c = ((double)v[0] + (double)v[2]) + ((double)v[6] + (double)v[8]);
    return 4.0 * (double)v[4] + 2.0 * e + c;
}

// Bxy of a column on a face of the grid from its block v (row-major, x
// fastest), shifted by sx (sy): -1 on the face x = 0 (y = 0), +1 on the
// face x = W-1 (y = H-1), 0 inside, so the vertex's own column (row) in
// the block is 1 + sx (1 + sy).  u: the vertex's own value.
__device__ __forceinline__ double tv_blur_face(const float (&v)[9], int sx, int sy, double &u) {
    double c[3], r[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double a0 = (double)v[3 * i], a1 = (double)v[3 * i + 1], a2 = (double)v[3 * i + 2];
        c[i] = sx < 0 ? a0 : (sx > 0 ? a2 : a1);
        r[i] = sx == 0 ? 2.0 * a1 + (a0 + a2) : c[i] + a1;
    }
    u = sy < 0 ? c[0] : (sy > 0 ? c[2] : c[1]);
    if (sy == 0) return 2.0 * r[1] + (r[0] + r[2]);
    return (sy < 0 ? r[0] : r[2]) + r[1];
}

template <int LPV, int STORE, bool FUSED>
__global__ __launch_bounds__(256) void k_tv_zmarch(const TvArgs a, const AdamArgs ad,
                                                   int64_t n_cols) {
    constexpr int kGroups = 256 / LPV;
    constexpr uint32_t kElem = STORE == PXL_STORE_F32 ? 4u : 2u;
    const int g = threadIdx.x / LPV, p = threadIdx.x % LPV;
    const int64_t col = (int64_t)xcd_slice_block(blockIdx.x, gridDim.x) * kGroups + g;
    if (col >= n_cols) return;
    const float lam = tv_lambda(a, p);
    const bool tvl = p < a.P && lam != 0.0f;
    // Alone, only the lanes with a TV term work; fused with Adam, every lane
    // of the record.
    if (FUSED ? p >= a.stride : !tvl) return;
    const int x = (int)(col % a.W), y = (int)((col / a.W) % a.H);
    const int z0 = (int)(col / ((int64_t)a.W * a.H)) * kTvChunk;
    const int z1 = min(z0 + kTvChunk, a.D);  // vertices z0 .. z1-1
    const int sx = x == 0 ? -1 : (x == a.W - 1 ? 1 : 0);
    const int sy = y == 0 ? -1 : (y == a.H - 1 ? 1 : 0);
    const bool inner = sx == 0 && sy == 0;
    const uint32_t rec_bytes = (uint32_t)a.stride * kElem;
    const uint32_t plane_bytes = (uint32_t)a.W * (uint32_t)a.H * rec_bytes;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float *>(a.params), (short)0, (int)((uint32_t)a.n_vox * rec_bytes),
        0x00020000);
    // Byte offset of the block's first vertex (x - 1 - sx, y - 1 - sy, z')
    // and of the 3 x 3 block from it.
    const uint32_t lane0 =
        ((uint32_t)(y - 1 - sy) * (uint32_t)a.W + (uint32_t)(x - 1 - sx)) * rec_bytes +
        (uint32_t)p * kElem;
    uint32_t nb[9];
#pragma unroll
    for (int i = 0; i < 9; i++) nb[i] = ((uint32_t)(i / 3) * (uint32_t)a.W + (uint32_t)(i % 3)) * rec_bytes;
    const size_t gplane = (size_t)a.W * a.H * (size_t)a.gstride;
    float *const gcol = a.grad + ((size_t)y * a.W + x) * (size_t)a.gstride + p;
    const int64_t rcol = (int64_t)y * a.W + x, rplane = (int64_t)a.W * a.H;
    auto plane = [&](int zz, float (&v)[9]) {
#pragma unroll
        for (int i = 0; i < 9; i++) v[i] = ld_buf<STORE>(rs, lane0 + (uint32_t)zz * plane_bytes, nb[i]);
    };
    // Bxy of a plane's block, and the vertex's own value in it.
    auto blur = [&](const float (&v)[9], double &c) {
        if (inner) {
            c = (double)v[4];
            return tv_blur_xy(v);
        }
        return tv_blur_face(v, sx, sy, c);
    };
    double b_lo = 0.0, b_mid = 0.0, u = 0.0;
    if (tvl) {
        float v[9];
        if (z0 >= 1) {
            plane(z0 - 1, v);
            b_lo = blur(v, u);
        }
        plane(z0, v);
        b_mid = blur(v, u);
    }
    // Ring of three slots: the slot of vertex s holds plane s + 1 and the
    // gradient of s (fused: the Adam element of s), loaded three steps
    // before it is used (indices clamped to the chunk: in bounds, unused).
    // A step uses its slot, then refills it for vertex s + 3, so two steps'
    // loads stay in flight behind it.
    float pv[3][9], pg[3];
    AdamElem pa[3];
    auto fill = [&](int k, int s) {
        const int sc = min(s, z1 - 1);
        if (FUSED)
            pa[k] = adam_load<STORE>(a, ad, (int64_t)sc * rplane + rcol, p);
        else
            pg[k] = gcol[(size_t)sc * gplane];
        if (tvl) plane(min(s + 1, a.D - 1), pv[k]);
    };
    fill(0, z0);
    fill(1, z0 + 1);
    fill(2, z0 + 2);
    // Cells around the vertex in its xy plane's two directions.
    const double nxy = (double)((sx == 0 ? 2 : 1) * (sy == 0 ? 2 : 1));
    auto step = [&](int k, int s, bool refill) {
        double tsum = 0.0;
        if (tvl) {
            // Cells below and above along z (plane s + 1 exists when zp).
            const int zm = s >= 1, zp = s <= a.D - 2;
            double u_hi = 0.0;
            const double b_hi = zp ? blur(pv[k], u_hi) : 0.0;
            if (inner && zm && zp)
                tsum = 8.0 * u - 0.125 * (b_lo + 2.0 * b_mid + b_hi);
            else
                tsum = nxy * (double)(zm + zp) * u -
                       0.125 * ((double)(zm + zp) * b_mid + ((zm ? b_lo : 0.0) + b_hi));
            b_lo = b_mid;
            b_mid = b_hi;
            u = u_hi;
        }
        if (FUSED)
            adam_store<STORE>(a, ad, (int64_t)s * rplane + rcol, p, pa[k], tvl, lam, tsum);
        else
            gcol[(size_t)s * gplane] = pg[k] + (float)((double)lam * tsum);
        if (refill) fill(k, s + 3);
    };
    int s = z0;
    for (; s + 2 < z1; s += 3) {
        step(0, s, true);
        step(1, s + 1, true);
        step(2, s + 2, true);
    }
    if (s < z1) step(0, s, false);
    if (s + 1 < z1) step(1, s + 1, false);
}

// addVarianceLossGPU as a gather, looking the neighbours up on the fly
// (dense grids, where a record's neighbours are arithmetic), over all
// n_items = n_vox vertices: interior ones take tv_interior_sum, the rest the
// general gather (grids k_tv_zmarch does not serve: 64-bit offsets, or a
// side shorter than 3).  FUSED: the group then runs the Adam step of its
// record.
template <int LPV, int STORE, bool OFF32, bool FUSED>
__global__ __launch_bounds__(256) void k_tv_disparity(const TvArgs a, const AdamArgs ad,
                                                      int64_t n_items) {
    constexpr int kGroups = 256 / LPV;
    const int g = threadIdx.x / LPV, p = threadIdx.x % LPV;
    const int64_t item = (int64_t)xcd_slice_block(blockIdx.x, gridDim.x) * kGroups + g;
    if (item >= n_items) return;  // the whole group leaves together
    tv_general_item<LPV, STORE, OFF32, FUSED>(a, ad, item, g, p);
}

// Caged grids: the neighbour records and the cell mask depend only on the
// cage, so pxl_tv_prepare stores them, 32 words per record, and each TV call
// reads them with one coalesced load instead of the compact-to-dense, index
// and occupancy lookups.  Words 0-26: the record to read for neighbour i
// (the record itself when no occupied cell around the vertex contains
// neighbour i); word 27: the occupied-cell mask; words 28-31: the
// multiplicities m_i (occupied cells containing both the vertex and
// neighbour i, 0..8) as nibbles, m_i in word 28 + i/8 at bit 4 (i % 8).
constexpr int kTableWords = 32;

template <int LPV>
__global__ __launch_bounds__(256) void k_tv_table_build(const TvArgs a, uint32_t *table) {
    constexpr int kGroups = 256 / LPV;
    const int g = threadIdx.x / LPV, p = threadIdx.x % LPV;
    const int64_t r = (int64_t)blockIdx.x * kGroups + g;
    if (r >= a.n_vox) return;  // the whole group leaves together
    uint32_t cmask;
    const uint32_t nrec = tv_neighbours<LPV>(a, r, p, cmask);
    const uint32_t m = p < 27 ? (uint32_t)__builtin_popcount(tv_cells_of(p) & cmask) : 0u;
    uint32_t nib = 0u;
#pragma unroll
    for (int k = 0; k < 8; k++)
        nib |= (uint32_t)__shfl((int)m, ((p - 28) & 3) * 8 + k, LPV) << (4 * k);
    uint32_t word;
    if (p < 27) word = m != 0u ? nrec : (uint32_t)r;
    else if (p == 27) word = cmask;
    else word = nib;
    if (p < kTableWords) table[(size_t)r * kTableWords + p] = word;
}

// TV gradient of record r from its prepared table row.  OFF24: every record
// index is below 2^24 and every byte offset into the records below 2^32, so
// an address is one 24-bit multiply-add on the records' base pointer.
// FUSED: the group then runs the Adam step of record r into ad.out.
template <int LPV, int STORE, bool OFF24, bool FUSED>
__global__ __launch_bounds__(256) void k_tv_disparity_table(const TvArgs a, const AdamArgs ad,
                                                            const uint32_t *__restrict__ table) {
    constexpr int kGroups = 256 / LPV;
    constexpr uint32_t kElem = STORE == PXL_STORE_F32 ? 4u : 2u;
    const int g = threadIdx.x / LPV, p = threadIdx.x % LPV;
    const int64_t r = (int64_t)xcd_slice_block(blockIdx.x, gridDim.x) * kGroups + g;
    if (r >= a.n_vox) return;
    AdamElem ae{0.0f, 0.0f, 0.0f, 0.0f};
    if (FUSED && p < a.stride) ae = adam_load<STORE>(a, ad, r, p);
    uint32_t e = p < kTableWords ? table[(size_t)r * kTableWords + p] : 0u;
    const uint32_t cmask = (uint32_t)__shfl((int)e, 27, LPV) & 0xFFu;
    const float lam = tv_lambda(a, p);
    const bool mine = p < a.P && lam != 0.0f;
    double tsum = 0.0;
    bool tvl = false;
    // No occupied cell around the vertex: no TV term.
    if (cmask != 0u) {
        // A record index beyond the grid (scratch not prepared for this
        // cage) reads record 0 instead of faulting.
        if (p < 27 && (int64_t)e >= a.n_vox) e = 0u;
        // Every lane issues all 27 loads unconditionally, so they are in
        // flight together; a lane without a parameter reads element 0.
        const uint32_t q = mine ? (uint32_t)p : 0u;
        const char *base = reinterpret_cast<const char *>(a.params);
        const uint32_t rec_bytes = (uint32_t)a.stride * kElem;
        float v[27];
#pragma unroll
        for (int i = 0; i < 27; i++) {
            const uint32_t ri = (uint32_t)__shfl((int)e, i, LPV);
            if (OFF24)
                v[i] = ld_at<STORE>(base, (uint32_t)__umul24(ri, rec_bytes) + q * kElem);
            else
                v[i] = ld<STORE>(a.params, (size_t)ri * (size_t)a.stride + q);
        }
        uint32_t mw[4];
#pragma unroll
        for (int j = 0; j < 4; j++) mw[j] = (uint32_t)__shfl((int)e, 28 + j, LPV);
        if (mine) {
            // sum over occupied cells c of (u - mean_c) = n u - (1/8) sum_i
            // m_i u_i (m_13 = n), in FP64: a vertex value and the mean of a
            // cell around it nearly cancel.
            double msum = 0.0;
#pragma unroll
            for (int i = 0; i < 27; i++)
                msum = fma((double)((mw[i >> 3] >> (4 * (i & 7))) & 15u), (double)v[i], msum);
            tsum = (double)__builtin_popcount(cmask) * (double)v[13] - 0.125 * msum;
            tvl = true;
        }
    }
    if (!FUSED) {
        if (tvl) a.grad[(size_t)r * (size_t)a.gstride + p] += (float)((double)lam * tsum);
        return;
    }
    if (p < a.stride) adam_store<STORE>(a, ad, r, p, ae, tvl, lam, tsum);
}

// Caged grids: a cell corner that is not in cage reads and writes record 0
// (index_offsets = 0, SceneUtils.cu:1106-1113).  Those (cell, corner) pairs
// are not found by the gather above.  Their sum into record 0 is
//   lambda * (N * u_0 - sum_c n_c * mu_c) = lambda * (N * u_0 - sum_r w_r * u_r)
// where n_c counts the aliased corners of occupied cell c, N = sum_c n_c and
// w_r = sum over (c, corner j) with record(j) = r of n_c / 8.  N and w depend
// only on the cage, so pxl_tv_prepare computes them once; each TV call then
// needs one weighted reduction over the records with w_r != 0.
struct AliasHeader {
    uint32_t magic;
    uint32_t n_alias;
    int64_t n_vox;
    uint32_t n_list;  // records with w_r != 0 (the compact list)
};
constexpr uint32_t kAliasMagic = 0x70786131u;
constexpr int kAliasBlocks = 1024;
constexpr size_t kAliasHdrBytes = 256;

__host__ __device__ inline size_t alias_weights_off() { return kAliasHdrBytes; }
__host__ __device__ inline size_t alias_partials_off(int64_t n_vox) {
    return (kAliasHdrBytes + (size_t)n_vox * 4 + 255) / 256 * 256;
}
__host__ __device__ inline size_t tv_table_off(int64_t n_vox) {
    return alias_partials_off(n_vox) + (size_t)kAliasBlocks * 64 * sizeof(double);
}
// After the neighbour table: the records with w_r != 0 in record order,
// (record, w_r) pairs, and the per-block counts of the compaction.
constexpr int kListPerBlock = 4096;  // 256 threads x 16 records
__host__ __device__ inline size_t alias_list_off(int64_t n_vox) {
    return tv_table_off(n_vox) + (size_t)n_vox * kTableWords * sizeof(uint32_t);
}
__host__ __device__ inline size_t alias_counts_off(int64_t n_vox) {
    return alias_list_off(n_vox) + ((size_t)n_vox * 8 + 255) / 256 * 256;
}
__host__ __device__ inline size_t tv_scratch_total(int64_t n_vox) {
    const size_t nb = (size_t)((n_vox + kListPerBlock - 1) / kListPerBlock);
    return alias_counts_off(n_vox) + (nb * sizeof(uint32_t) + 255) / 256 * 256;
}

// Compaction of the records with w_r != 0, in record order (so the
// weighted sums below run in a fixed order).  Thread t of block b looks at
// records b * 4096 + 16 t .. + 15.
__device__ __forceinline__ uint32_t alias_flags16(const float *w, int64_t n, int64_t base) {
    uint32_t f = 0u;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const int64_t r = base + k;
        if (r < n && w[r] != 0.0f) f |= 1u << k;
    }
    return f;
}

__global__ __launch_bounds__(256) void k_alias_count(const float *w, int64_t n,
                                                     uint32_t *counts) {
    __shared__ uint32_t part[4];
    const int64_t base = (int64_t)blockIdx.x * kListPerBlock + 16 * (int64_t)threadIdx.x;
    const uint32_t c = wave_sum_u32((uint32_t)__builtin_popcount(alias_flags16(w, n, base)));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// One block: exclusive scan of the block counts in place, total to the header.
__global__ __launch_bounds__(1024) void k_alias_scan(uint32_t *counts, int64_t nb,
                                                     AliasHeader *hdr) {
    __shared__ uint32_t lds[1024];
    uint32_t carry = 0u;
    for (int64_t c0 = 0; c0 < nb; c0 += 1024) {
        const int64_t i = c0 + threadIdx.x;
        const uint32_t v = i < nb ? counts[i] : 0u;
        lds[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {  // inclusive Hillis-Steele scan
            const uint32_t x = (int)threadIdx.x >= o ? lds[threadIdx.x - o] : 0u;
            __syncthreads();
            lds[threadIdx.x] += x;
            __syncthreads();
        }
        if (i < nb) counts[i] = carry + lds[threadIdx.x] - v;
        carry += lds[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) hdr->n_list = carry;
}

__global__ __launch_bounds__(256) void k_alias_list(const float *w, int64_t n,
                                                    const uint32_t *offsets, uint2 *list) {
    __shared__ uint32_t wsum[4];
    const int64_t base = (int64_t)blockIdx.x * kListPerBlock + 16 * (int64_t)threadIdx.x;
    const uint32_t f = alias_flags16(w, n, base);
    const uint32_t c = (uint32_t)__builtin_popcount(f);
    // Exclusive prefix of c over the block, in thread order.
    uint32_t incl = c;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t x = (uint32_t)__shfl_up((int)incl, o, 64);
        if (lane >= o) incl += x;
    }
    if (lane == 63) wsum[threadIdx.x >> 6] = incl;
    __syncthreads();
    uint32_t before = 0u;
    for (int k = 0; k < (int)(threadIdx.x >> 6); k++) before += wsum[k];
    uint32_t at = offsets[blockIdx.x] + before + incl - c;
#pragma unroll
    for (int k = 0; k < 16; k++)
        if ((f >> k) & 1u) {
            const int64_t r = base + k;
            list[at++] = make_uint2((uint32_t)r, __float_as_uint(w[r]));
        }
}

__global__ __launch_bounds__(256) void k_alias_weights(const TvArgs a, int64_t ncell,
                                                       AliasHeader *hdr, float *w) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    if (!((a.occ[c >> 5] >> (c & 31)) & 1u)) return;
    const uint32_t d0 = a.c2d[0];
    const int64_t cw = a.W - 1, ch = a.H - 1;
    const int x = (int)(c % cw), y = (int)((c / cw) % ch), z = (int)(c / (cw * ch));
    uint32_t rec[8];
    int n = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int jx = j & 1, jy = (j >> 1) & 1, jz = j >> 2;
        const uint32_t d = (uint32_t)(z + jz) * (uint32_t)a.W * (uint32_t)a.H +
                           (uint32_t)(y + jy) * (uint32_t)a.W + (uint32_t)(x + jx);
        rec[j] = a.io[d];
        if ((int64_t)rec[j] >= a.n_vox) rec[j] = 0u;  // not this grid's index: record 0
        if (rec[j] == 0u && d != d0) n++;
    }
    if (n == 0) return;
#pragma unroll
    for (int j = 0; j < 8; j++) atomicAdd(w + rec[j], (float)n * 0.125f);
    atomicAdd(&hdr->n_alias, (uint32_t)n);
}

// Per-block partial sums of w_r * u_r[p] over the compact list of records
// with w_r != 0; LPV lanes per list entry, one per parameter.
template <int LPV, int STORE>
__global__ __launch_bounds__(256) void k_alias_partial(const TvArgs a, const AliasHeader *hdr,
                                                       const uint2 *list, double *partials) {
    constexpr int kGroups = 256 / LPV;
    __shared__ double red[256];
    const int g = threadIdx.x / LPV, p = threadIdx.x % LPV;
    double acc = 0.0;
    // Scratch not prepared for this grid: no list (k_alias_final adds nothing
    // either); an entry naming a record beyond the grid reads record 0.
    const bool ok = hdr->magic == kAliasMagic && hdr->n_vox == a.n_vox;
    const int64_t n = ok ? min((int64_t)hdr->n_list, a.n_vox) : 0;
    if (p < a.P) {
        // Four entries per step; an entry past the end reads entry 0 with
        // weight 0, so every step's loads are in flight together.
        const int64_t step = (int64_t)gridDim.x * kGroups;
        for (int64_t e0 = (int64_t)blockIdx.x * kGroups + g; e0 < n; e0 += 4 * step) {
            uint2 ent[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int64_t e = e0 + k * step;
                ent[k] = list[e < n ? e : 0];
                if (e >= n) ent[k].y = 0u;
            }
            float ur[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t rec = (int64_t)ent[k].x < a.n_vox ? ent[k].x : 0u;
                ur[k] = ld<STORE>(a.params, (size_t)rec * a.stride + p);
            }
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (ent[k].y != 0u) acc += (double)__uint_as_float(ent[k].y) * (double)ur[k];
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < LPV) {
        double t = 0.0;
        for (int i = 0; i < kGroups; i++) t += red[i * LPV + threadIdx.x];
        partials[(size_t)blockIdx.x * 64 + threadIdx.x] = t;
    }
}

// One block of 256 threads per parameter p (blockIdx.x): thread t sums the
// partials of blocks t, t + 256, t + 512, t + 768 (loads in flight
// together), then the block folds the 256 sums in LDS in a fixed order.
template <int STORE>
__global__ __launch_bounds__(256) void k_alias_final(const TvArgs a, const AliasHeader *hdr,
                                                     const double *partials) {
    static_assert(kAliasBlocks == 4 * 256, "four partials per thread");
    __shared__ double red[256];
    const int p = blockIdx.x, t = threadIdx.x;
    double v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = partials[(size_t)(t + 256 * k) * 64 + p];
    red[t] = (v[0] + v[1]) + (v[2] + v[3]);
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    if (t != 0 || p >= a.P) return;
    const float lam = tv_lambda(a, p);
    if (lam == 0.0f || hdr->magic != kAliasMagic || hdr->n_vox != a.n_vox) return;
    const double u0 = (double)ld<STORE>(a.params, (size_t)p);
    a.grad[p] += (float)((double)lam * ((double)hdr->n_alias * u0 - red[0]));
}

__global__ void k_alias_init(AliasHeader *hdr, float *w, int64_t n_vox) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_vox) w[i] = 0.0f;
    if (i == 0) {
        hdr->magic = kAliasMagic;
        hdr->n_alias = 0u;
        hdr->n_vox = n_vox;
    }
}

// totalVarianceLoss, 6-neighbour branch (RenderUtils.cu:2660-2763): for an
// interior vertex and each parameter plane,
//   grad += lambda / (sqrt(sum (u - u_n)^2) / 6 + 1e-6) * sum (u - u_n)
// over the 6 face neighbours; a neighbour whose index (record, when caged)
// is 0 counts as the vertex itself.  One group of LPV lanes (one per
// parameter) per record; a dense record is its vertex.  A caged record r > 0
// belongs to one vertex, c2d[r]; the reference tests the in-cage flag array
// (indexed by dense vertex) at the record number r (RenderUtils.cu:2669),
// which is kept.  Record 0 (every vertex outside the cage, and the first one
// in it) is k_tv_neighbour6_rec0's.  Only vertices whose column y W + x is
// below col_limit take part (the launch of the reference's kernel covers
// that many columns).
__device__ __forceinline__ double tv6_term(float self, const float (&un)[6], float lam) {
    double sq = 0.0, dis = 0.0;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const double d = (double)self - (double)un[i];
        sq += d * d;
        dis += d;
    }
    const double sd = sqrt(sq) / 6.0;
    return (double)lam / (sd + 1e-6) * dis;
}

template <int LPV, int STORE, bool SPARSE>
__global__ __launch_bounds__(256) void k_tv_neighbour6(const TvArgs a, int64_t col_limit) {
    constexpr int kGroups = 256 / LPV;
    const int g = threadIdx.x / LPV, p = threadIdx.x % LPV;
    const int64_t r = (int64_t)blockIdx.x * kGroups + g;
    if (r >= a.n_vox || p >= a.P || (SPARSE && r == 0)) return;
    const float lam = tv_lambda(a, p);
    if (lam == 0.0f) return;
    const uint32_t d = SPARSE ? a.c2d[r] : (uint32_t)r;
    const int x = (int)(d % (uint32_t)a.W), y = (int)((d / (uint32_t)a.W) % (uint32_t)a.H);
    const int z = (int)(d / ((uint32_t)a.W * (uint32_t)a.H));
    if (x == 0 || x == a.W - 1 || y == 0 || y == a.H - 1 || z == 0 || z == a.D - 1) return;
    if ((int64_t)y * a.W + x >= col_limit) return;
    if (SPARSE && !a.cage[r]) return;
    const float self = ld<STORE>(a.params, (size_t)r * a.stride + p);
    const int off[6][3] = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    // Unconditional loads (index 0 reads the vertex itself), all in flight
    // together.
    float un[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const uint32_t n = record_of(a, x + off[i][0], y + off[i][1], z + off[i][2]);
        un[i] = ld<STORE>(a.params, (size_t)(n != 0u ? n : (uint32_t)r) * a.stride + p);
    }
    a.grad[(size_t)r * a.gstride + p] += (float)tv6_term(self, un, lam);
}

// Record 0 of a caged grid under the 6-neighbour term: every interior
// vertex whose record is 0 adds its term, with u = record 0, when the
// in-cage flag of dense vertex 0 is set (the reference's test at record
// number 0).  Thread per dense vertex; a vertex whose 6 neighbours all read
// record 0 adds 0.  The adds into record 0 are float atomics (the reference
// races on them).
template <int STORE>
__global__ __launch_bounds__(256) void k_tv_neighbour6_rec0(const TvArgs a, int64_t col_limit) {
    if (!a.cage[0]) return;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nv = (int64_t)a.W * a.H * a.D;
    if (v >= nv) return;
    const int x = (int)(v % a.W), y = (int)((v / a.W) % a.H), z = (int)(v / ((int64_t)a.W * a.H));
    if (x == 0 || x == a.W - 1 || y == 0 || y == a.H - 1 || z == 0 || z == a.D - 1) return;
    if ((int64_t)y * a.W + x >= col_limit) return;
    if (record_of(a, x, y, z) != 0u) return;
    const int off[6][3] = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    uint32_t nrec[6];
    bool any = false;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        nrec[i] = record_of(a, x + off[i][0], y + off[i][1], z + off[i][2]);
        any = any || nrec[i] != 0u;
    }
    if (!any) return;
    for (int p = 0; p < a.P; p++) {
        const float lam = tv_lambda(a, p);
        if (lam == 0.0f) continue;
        const float self = ld<STORE>(a.params, (size_t)p);
        float un[6];
#pragma unroll
        for (int i = 0; i < 6; i++)
            un[i] = nrec[i] != 0u ? ld<STORE>(a.params, (size_t)nrec[i] * a.stride + p) : self;
        atomicAdd(a.grad + p, (float)tv6_term(self, un, lam));
    }
}

// updateRGBSceneGPU_ADAM (Optimizer.cu:359-415).  One thread per 4
// consecutive parameters of a record; the float/double mix of each update
// follows the reference expression by expression.  With 16-bit storage the
// update runs on the FP32 master copy when there is one (else on the
// widened stored value) and the grid receives the rounded result.
template <int STORE>
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ prm, float *__restrict__ grd,
                                              float *__restrict__ mom, float *__restrict__ var,
                                              float *__restrict__ master,
                                              const float *__restrict__ alr, int64_t n_vox,
                                              int pstride, int ostride, int gstride, int P,
                                              int Q, int S, int sh_lo, int sh_hi, double lr,
                                              double sh_lr, double bc1, double bc2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_vox * Q) return;
    const int64_t r = i / Q;
    const int q = (int)(i % Q);
    const size_t off = (size_t)r * ostride + 4 * q;   // moments, master
    const size_t goff = (size_t)r * gstride + 4 * q;  // gradient
    const size_t poff = (size_t)r * pstride + 4 * q;
    float4 th;
    if (STORE == PXL_STORE_F32) {
        th = *reinterpret_cast<const float4 *>(prm + poff);
    } else if (master != nullptr) {
        th = *reinterpret_cast<const float4 *>(master + off);
    } else {
        // Four 16-bit values in one 8-byte load (the 16-bit stride is a
        // multiple of 8 elements, so 4 q is 8-byte aligned).
        const uint2 h = *reinterpret_cast<const uint2 *>(reinterpret_cast<const uint16_t *>(prm) +
                                                         poff);
        th.x = widen16<STORE>((uint16_t)(h.x & 0xFFFFu));
        th.y = widen16<STORE>((uint16_t)(h.x >> 16));
        th.z = widen16<STORE>((uint16_t)(h.y & 0xFFFFu));
        th.w = widen16<STORE>((uint16_t)(h.y >> 16));
    }
    float4 g4 = *reinterpret_cast<const float4 *>(grd + goff);
    float4 m4 = *reinterpret_cast<const float4 *>(mom + off);
    float4 v4 = *reinterpret_cast<const float4 *>(var + off);
    float *tp = &th.x, *gp = &g4.x, *mp = &m4.x, *vp = &v4.x;
    const double lr_d = alr ? lr * (double)alr[r] : lr;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int p = 4 * q + e;
        if (p >= P) break;
        // SH coefficients outside the optimised band keep their state
        // (Optimizer.cu:393-412).
        (void)adam_elem(p, S, sh_lo, sh_hi, lr_d, sh_lr, bc1, bc2, tp[e], gp[e], mp[e], vp[e]);
    }
    if (STORE == PXL_STORE_F32) {
        *reinterpret_cast<float4 *>(prm + poff) = th;
    } else {
        if (master != nullptr) *reinterpret_cast<float4 *>(master + off) = th;
        // One 8-byte store of the four rounded values (elements past P hold
        // what was loaded, so they are written back unchanged or from the
        // master's padding).
        uint2 h;
        h.x = (uint32_t)narrow16<STORE>(th.x) | ((uint32_t)narrow16<STORE>(th.y) << 16);
        h.y = (uint32_t)narrow16<STORE>(th.z) | ((uint32_t)narrow16<STORE>(th.w) << 16);
        *reinterpret_cast<uint2 *>(reinterpret_cast<uint16_t *>(prm) + poff) = h;
    }
    *reinterpret_cast<float4 *>(grd + goff) = g4;
    st_stream(mom + off, m4);
    st_stream(var + off, v4);
}

// Optimiser state of a new band (bandMemoryManage / initialAll,
// Optimizer.cu:11-64, 161-187): element p of every record of grad, m and v
// becomes 0 when `all`, or when it is the density and `density`, or an SH
// coefficient k in [lo, hi) of any channel.
// Element (record r, element p < width = max(stride, gstride)) per thread:
// moments of `stride` floats, gradient of `gstride`; padding (p >= P) is
// cleared with the whole records only.
__global__ __launch_bounds__(256) void k_init_band(float *grad, float *m, float *v,
                                                   int64_t n_vox, int stride, int gstride,
                                                   int width, int P, int S, int lo, int hi,
                                                   int density, int all) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_vox * width) return;
    const int64_t r = i / width;
    const int p = (int)(i % width);
    bool clear = all != 0;
    if (p == 0) clear = clear || density != 0;
    else if (p < P) clear = clear || ((p - 1) % S >= lo && (p - 1) % S < hi);
    if (!clear) return;
    if (p < gstride) grad[(size_t)r * gstride + p] = 0.0f;
    if (p < stride) {
        m[(size_t)r * stride + p] = 0.0f;
        v[(size_t)r * stride + p] = 0.0f;
    }
}

__global__ __launch_bounds__(1024) void k_loss_reduce(const double *__restrict__ l, int64_t n,
                                                      double *__restrict__ out) {
    __shared__ double part[16];
    // Four independent accumulators per lane keep several loads in flight.
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    const int64_t T = blockDim.x;
    int64_t i = threadIdx.x;
    for (; i + 3 * T < n; i += 4 * T) {
        const double a = l[i], b = l[i + T], c = l[i + 2 * T], d = l[i + 3 * T];
        s0 += a * a;
        s1 += b * b;
        s2 += c * c;
        s3 += d * d;
    }
    for (; i < n; i += T) s0 += l[i] * l[i];
    double s = wave_sum_d((s0 + s1) + (s2 + s3));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) tot += part[w];
        const double mse = tot / (double)n;
        out[0] = mse;
        out[1] = 10.0 * log10(65025.0 / mse);
    }
}

int32_t tv_check(const pxl_grid *g, const pxl_optim *o, const char *who) {
    PXL_REQUIRE(g && o && o->grad && g->params, "%s: null argument", who);
    PXL_REQUIRE(g->W >= 2 && g->H >= 2 && g->D >= 2, "%s: grid too small", who);
    PXL_REQUIRE(g->sh_dim >= 1 && g->sh_dim <= 21, "%s: bad sh_dim %d", who, g->sh_dim);
    PXL_REQUIRE(g->storage == PXL_STORE_F32 || g->storage == PXL_STORE_BF16 ||
                    g->storage == PXL_STORE_F16,
                "%s: unknown storage %d", who, g->storage);
    PXL_REQUIRE(g->stride >= 1 + 3 * g->sh_dim && g->stride % (g->storage ? 8 : 4) == 0,
                "%s: bad stride %d", who, g->stride);
    PXL_REQUIRE(g->storage == PXL_STORE_F32 || o->stride > 0,
                "%s: 16-bit grids need an explicit optimiser stride", who);
    PXL_REQUIRE(grad_stride(g, o) >= 1 + 3 * g->sh_dim && grad_stride(g, o) % 4 == 0,
                "%s: bad gradient stride %d", who, grad_stride(g, o));
    PXL_REQUIRE(g->sparse || g->n_vox == (int64_t)g->W * g->H * g->D,
                "%s: dense grid needs n_vox = W*H*D", who);
    PXL_REQUIRE(!g->sparse || (g->occupancy && g->index_offsets && g->compact_to_dense),
                "%s: sparse grid needs occupancy, index_offsets and compact_to_dense", who);
    PXL_REQUIRE(g->n_vox > 0, "%s: empty grid", who);
    return PXL_OK;
}

// The PXL_OPT_TV_ADDR64 option forces the 64-bit address form of the TV
// gathers, which otherwise serves only grids beyond 2^24 records or 4 GiB
// (both are tested).
bool offsets64_forced() { return option_tv_addr64(); }

// The disparity TV of the grid into the gradient, or (ad != nullptr) fused
// with the Adam step into ad->out.  Caged grids add the aliased corners'
// sum into record 0 first (k_alias_partial / k_alias_final), so that both
// forms add record 0's two terms in the same order.
template <int STORE, bool FUSED>
void tv_launch_as(const TvArgs &a, const AdamArgs &ad, void *scratch, hipStream_t s) {
    const bool wide = a.P > 32;
    const int groups = wide ? 4 : 8;
    const unsigned blocks = (unsigned)((a.n_vox + groups - 1) / groups);
    if (a.sparse) {
        char *base = reinterpret_cast<char *>(scratch);
        const AliasHeader *hdr = reinterpret_cast<const AliasHeader *>(base);
        const uint2 *list = reinterpret_cast<const uint2 *>(base + alias_list_off(a.n_vox));
        double *part = reinterpret_cast<double *>(base + alias_partials_off(a.n_vox));
        if (wide)
            hipLaunchKernelGGL((k_alias_partial<64, STORE>), dim3(kAliasBlocks), dim3(256), 0, s,
                               a, hdr, list, part);
        else
            hipLaunchKernelGGL((k_alias_partial<32, STORE>), dim3(kAliasBlocks), dim3(256), 0, s,
                               a, hdr, list, part);
        hipLaunchKernelGGL(k_alias_final<STORE>, dim3((unsigned)a.P), dim3(256), 0, s, a, hdr,
                           part);
        const uint32_t *table = reinterpret_cast<const uint32_t *>(base + tv_table_off(a.n_vox));
        constexpr uint64_t kElem = STORE == PXL_STORE_F32 ? 4u : 2u;
        const bool off24 = !offsets64_forced() && a.n_vox <= (int64_t)(1u << 24) &&
                           (uint64_t)a.stride * kElem < (1u << 24) &&
                           (uint64_t)a.n_vox * (uint64_t)a.stride * kElem <= 0xFFFFFFFFull;
        if (wide && off24)
            hipLaunchKernelGGL((k_tv_disparity_table<64, STORE, true, FUSED>), dim3(blocks),
                               dim3(256), 0, s, a, ad, table);
        else if (wide)
            hipLaunchKernelGGL((k_tv_disparity_table<64, STORE, false, FUSED>), dim3(blocks),
                               dim3(256), 0, s, a, ad, table);
        else if (off24)
            hipLaunchKernelGGL((k_tv_disparity_table<32, STORE, true, FUSED>), dim3(blocks),
                               dim3(256), 0, s, a, ad, table);
        else
            hipLaunchKernelGGL((k_tv_disparity_table<32, STORE, false, FUSED>), dim3(blocks),
                               dim3(256), 0, s, a, ad, table);
        return;
    }
    constexpr uint64_t kElem = STORE == PXL_STORE_F32 ? 4u : 2u;
    const bool off32 = !offsets64_forced() &&
                       (uint64_t)a.n_vox * (uint64_t)a.stride * kElem <= 0xFFFFFFFFull;
    // 32-bit offsets: every vertex marched along z; otherwise every vertex
    // gathered.
    if (off32 && a.W >= 3 && a.H >= 3 && a.D >= 3) {
        const int64_t n_cols = (int64_t)a.W * a.H * ((a.D + kTvChunk - 1) / kTvChunk);
        const unsigned nb = (unsigned)((n_cols + groups - 1) / groups);
        if (wide)
            hipLaunchKernelGGL((k_tv_zmarch<64, STORE, FUSED>), dim3(nb), dim3(256), 0, s, a, ad,
                               n_cols);
        else
            hipLaunchKernelGGL((k_tv_zmarch<32, STORE, FUSED>), dim3(nb), dim3(256), 0, s, a, ad,
                               n_cols);
        return;
    }
    const unsigned ib = (unsigned)((a.n_vox + groups - 1) / groups);
    if (wide && off32)
        hipLaunchKernelGGL((k_tv_disparity<64, STORE, true, FUSED>), dim3(ib), dim3(256), 0, s, a,
                           ad, a.n_vox);
    else if (wide)
        hipLaunchKernelGGL((k_tv_disparity<64, STORE, false, FUSED>), dim3(ib), dim3(256), 0, s,
                           a, ad, a.n_vox);
    else if (off32)
        hipLaunchKernelGGL((k_tv_disparity<32, STORE, true, FUSED>), dim3(ib), dim3(256), 0, s, a,
                           ad, a.n_vox);
    else
        hipLaunchKernelGGL((k_tv_disparity<32, STORE, false, FUSED>), dim3(ib), dim3(256), 0, s,
                           a, ad, a.n_vox);
}

template <int STORE>
void tv_launch(const TvArgs &a, const AdamArgs *ad, void *scratch, hipStream_t s) {
    if (ad != nullptr)
        tv_launch_as<STORE, true>(a, *ad, scratch, s);
    else
        tv_launch_as<STORE, false>(a, AdamArgs{}, scratch, s);
}

// The 6-neighbour term over the columns y W + x < col_limit.
template <int STORE>
void tv6_launch(const TvArgs &a, int64_t col_limit, hipStream_t s) {
    const bool wide = a.P > 32;
    const int groups = wide ? 4 : 8;
    const unsigned blocks = (unsigned)((a.n_vox + groups - 1) / groups);
    if (a.sparse) {
        if (wide)
            hipLaunchKernelGGL((k_tv_neighbour6<64, STORE, true>), dim3(blocks), dim3(256), 0, s, a,
                               col_limit);
        else
            hipLaunchKernelGGL((k_tv_neighbour6<32, STORE, true>), dim3(blocks), dim3(256), 0, s, a,
                               col_limit);
        const int64_t nv = (int64_t)a.W * a.H * a.D;
        hipLaunchKernelGGL(k_tv_neighbour6_rec0<STORE>, dim3((unsigned)((nv + 255) / 256)),
                           dim3(256), 0, s, a, col_limit);
    } else if (wide) {
        hipLaunchKernelGGL((k_tv_neighbour6<64, STORE, false>), dim3(blocks), dim3(256), 0, s, a,
                           col_limit);
    } else {
        hipLaunchKernelGGL((k_tv_neighbour6<32, STORE, false>), dim3(blocks), dim3(256), 0, s, a,
                           col_limit);
    }
}

}  // namespace
}  // namespace pxl

using namespace pxl;

extern "C" size_t pxl_tv_scratch_bytes(const pxl_grid *grid) {
    if (!grid || !grid->sparse || grid->n_vox <= 0) return 0;
    return tv_scratch_total(grid->n_vox);
}

static void fill_tv_args(const pxl_grid *grid, const pxl_optim *optim, TvArgs &a) {
    a.params = grid->params;
    a.grad = optim ? optim->grad : nullptr;
    a.gstride = grad_stride(grid, optim);
    a.occ = grid->occupancy;
    a.io = grid->index_offsets;
    a.c2d = grid->compact_to_dense;
    a.cage = grid->in_cage;
    a.W = grid->W;
    a.H = grid->H;
    a.D = grid->D;
    a.stride = grid->stride;
    a.P = 1 + 3 * grid->sh_dim;
    a.S = grid->sh_dim;
    a.sh_lo = optim ? optim->sh_first : 0;
    a.sh_hi = grid->sh_active > 0 ? grid->sh_active : grid->sh_dim;
    a.sparse = grid->sparse;
    a.n_vox = grid->n_vox;
    a.lam_d = 0.0f;
    a.lam_sh = 0.0f;
}

extern "C" int32_t pxl_tv_prepare(const pxl_grid *grid, void *scratch, size_t scratch_bytes,
                                  void *stream) {
    clear_error();
    PXL_REQUIRE(grid && grid->sparse && grid->occupancy && grid->index_offsets &&
                    grid->compact_to_dense && grid->n_vox > 0,
                "pxl_tv_prepare: needs a non-empty sparse grid");
    const size_t need = pxl_tv_scratch_bytes(grid);
    if (!scratch || scratch_bytes < need || ((uintptr_t)scratch & 255)) {
        set_error("pxl_tv_prepare: scratch of %zu bytes (256-byte aligned) needed, got %zu",
                  need, scratch_bytes);
        return PXL_ENOMEM;
    }
    TvArgs a;
    fill_tv_args(grid, nullptr, a);
    char *base = reinterpret_cast<char *>(scratch);
    AliasHeader *hdr = reinterpret_cast<AliasHeader *>(base);
    float *w = reinterpret_cast<float *>(base + alias_weights_off());
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_alias_init, dim3((unsigned)((a.n_vox + 255) / 256)), dim3(256), 0, s,
                       hdr, w, a.n_vox);
    const int64_t ncell = (int64_t)(a.W - 1) * (a.H - 1) * (a.D - 1);
    hipLaunchKernelGGL(k_alias_weights, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s, a,
                       ncell, hdr, w);
    const int64_t nb = (a.n_vox + kListPerBlock - 1) / kListPerBlock;
    uint32_t *counts = reinterpret_cast<uint32_t *>(base + alias_counts_off(a.n_vox));
    uint2 *list = reinterpret_cast<uint2 *>(base + alias_list_off(a.n_vox));
    hipLaunchKernelGGL(k_alias_count, dim3((unsigned)nb), dim3(256), 0, s, w, a.n_vox, counts);
    hipLaunchKernelGGL(k_alias_scan, dim3(1), dim3(1024), 0, s, counts, nb, hdr);
    hipLaunchKernelGGL(k_alias_list, dim3((unsigned)nb), dim3(256), 0, s, w, a.n_vox, counts,
                       list);
    uint32_t *table = reinterpret_cast<uint32_t *>(base + tv_table_off(a.n_vox));
    hipLaunchKernelGGL(k_tv_table_build<32>, dim3((unsigned)((a.n_vox + 7) / 8)), dim3(256), 0, s,
                       a, table);
    return check_launch("pxl_tv_prepare");
}

extern "C" int32_t pxl_add_variance_loss(const pxl_grid *grid, const pxl_optim *optim,
                                         double lambda_density, double lambda_sh, int32_t type,
                                         int32_t neighbours, int32_t optimize_step,
                                         void *scratch, size_t scratch_bytes, void *stream) {
    clear_error();
    int32_t st = tv_check(grid, optim, "pxl_add_variance_loss");
    if (st != PXL_OK) return st;
    PXL_REQUIRE(type == PXL_TV_DISPARITY || type == PXL_TV_NEIGHBOUR || type == PXL_TV_BOTH,
                "pxl_add_variance_loss: unsupported type %d", type);
    const bool disparity = type == PXL_TV_DISPARITY || type == PXL_TV_BOTH;
    const bool nbr6 = type == PXL_TV_NEIGHBOUR || type == PXL_TV_BOTH;
    if (nbr6) {
        PXL_REQUIRE(neighbours == 6, "pxl_add_variance_loss: only the 6-neighbour TV is built");
        PXL_REQUIRE(!grid->sparse || grid->in_cage,
                    "pxl_add_variance_loss: the 6-neighbour TV of a caged grid needs in_cage");
    }
    if (disparity && grid->sparse) {
        const size_t need = pxl_tv_scratch_bytes(grid);
        if (!scratch || scratch_bytes < need || ((uintptr_t)scratch & 255)) {
            set_error("pxl_add_variance_loss: sparse grids need %zu bytes of scratch prepared "
                      "by pxl_tv_prepare, got %zu", need, scratch_bytes);
            return PXL_ENOMEM;
        }
    }
    TvArgs a;
    fill_tv_args(grid, optim, a);
    PXL_REQUIRE(a.sh_hi <= a.S && a.sh_lo >= 0 && a.sh_lo < a.sh_hi,
                "pxl_add_variance_loss: band [%d, %d) outside the grid's %d coefficients",
                a.sh_lo, a.sh_hi, a.S);
    a.lam_d = (float)lambda_density;
    hipStream_t s = as_stream(stream);
    // The 6-neighbour term first (VarianceType 2 runs it, then the disparity
    // term; RenderUtils.cu:2840-2845), over the columns its launch covers:
    // all W x H for type 0, ceil((W-1)/16) x ceil((H-1)/16) blocks of 256 for
    // type 2.
    if (nbr6) {
        a.lam_sh = (float)lambda_sh;
        const int64_t limit =
            type == PXL_TV_NEIGHBOUR
                ? (int64_t)grid->W * grid->H
                : (int64_t)((grid->W - 1 + 15) / 16) * ((grid->H - 1 + 15) / 16) * 256;
        switch (grid->storage) {
            case PXL_STORE_F32: tv6_launch<PXL_STORE_F32>(a, limit, s); break;
            case PXL_STORE_BF16: tv6_launch<PXL_STORE_BF16>(a, limit, s); break;
            default: tv6_launch<PXL_STORE_F16>(a, limit, s); break;
        }
    }
    if (disparity) {
        // SH terms only for optimize_step <= 1 (RenderUtils.cu:2493).
        a.lam_sh = optimize_step <= 1 ? (float)lambda_sh : 0.0f;
        switch (grid->storage) {
            case PXL_STORE_F32: tv_launch<PXL_STORE_F32>(a, nullptr, scratch, s); break;
            case PXL_STORE_BF16: tv_launch<PXL_STORE_BF16>(a, nullptr, scratch, s); break;
            default: tv_launch<PXL_STORE_F16>(a, nullptr, scratch, s); break;
        }
    }
    return check_launch("pxl_add_variance_loss");
}

extern "C" int32_t pxl_adam_step(const pxl_grid *grid, const pxl_optim *optim, double lr,
                                 double sh_lr, int32_t t, void *stream) {
    clear_error();
    PXL_REQUIRE(grid && optim && grid->params && optim->grad && optim->m && optim->v,
                "pxl_adam_step: null argument");
    PXL_REQUIRE(t >= 1, "pxl_adam_step: time step %d must be >= 1", t);
    PXL_REQUIRE(grid->sh_dim >= 1, "pxl_adam_step: bad sh_dim %d", grid->sh_dim);
    PXL_REQUIRE(grid->storage == PXL_STORE_F32 || grid->storage == PXL_STORE_BF16 ||
                    grid->storage == PXL_STORE_F16,
                "pxl_adam_step: unknown storage %d", grid->storage);
    const int P = 1 + 3 * grid->sh_dim;
    const int pstride = grid->stride;
    const int ostride = optim_stride(grid, optim), gstride = grad_stride(grid, optim);
    PXL_REQUIRE(grid->storage == PXL_STORE_F32 || optim->stride > 0,
                "pxl_adam_step: 16-bit grids need an explicit optimiser stride");
    PXL_REQUIRE(pstride >= P && pstride % (grid->storage ? 8 : 4) == 0 && ostride >= P &&
                    ostride % 4 == 0 && gstride >= P && gstride % 4 == 0,
                "pxl_adam_step: bad strides %d / %d / %d", pstride, ostride, gstride);
    PXL_REQUIRE(((uintptr_t)grid->params | (uintptr_t)optim->grad | (uintptr_t)optim->m |
                 (uintptr_t)optim->v | (uintptr_t)optim->master) % 16 == 0,
                "pxl_adam_step: parameter and state buffers must be 16-byte aligned");
    PXL_REQUIRE(grid->n_vox >= 0, "pxl_adam_step: bad n_vox");
    if (grid->n_vox == 0) return PXL_OK;
    const int Q = (P + 3) / 4;
    const int sh_hi = grid->sh_active > 0 ? grid->sh_active : grid->sh_dim;
    PXL_REQUIRE(sh_hi <= grid->sh_dim && optim->sh_first >= 0 && optim->sh_first < sh_hi,
                "pxl_adam_step: band [%d, %d) outside the grid's %d coefficients", optim->sh_first,
                sh_hi, grid->sh_dim);
    // Bias corrections with float pow as in the reference (Optimizer.cu:384-385).
    const double bc1 = 1.0 - (double)powf(0.9f, (float)t);
    const double bc2 = 1.0 - (double)powf(0.999f, (float)t);
    const int64_t total = grid->n_vox * Q;
    const float *alr = grid->sparse ? nullptr : optim->adaptive_lr;
    float *master = grid->storage == PXL_STORE_F32 ? nullptr : optim->master;
    const dim3 gdim((unsigned)((total + 255) / 256));
    hipStream_t s = as_stream(stream);
#define PXL_ADAM(ST)                                                                            \
    hipLaunchKernelGGL(k_adam<ST>, gdim, dim3(256), 0, s, grid->params, optim->grad, optim->m, \
                       optim->v, master, alr, grid->n_vox, pstride, ostride, gstride, P, Q,     \
                       grid->sh_dim,                                                            \
                       optim->sh_first, sh_hi, lr, sh_lr, bc1, bc2)
    switch (grid->storage) {
        case PXL_STORE_F32: PXL_ADAM(PXL_STORE_F32); break;
        case PXL_STORE_BF16: PXL_ADAM(PXL_STORE_BF16); break;
        default: PXL_ADAM(PXL_STORE_F16); break;
    }
#undef PXL_ADAM
    return check_launch("pxl_adam_step");
}

extern "C" int32_t pxl_tv_adam_step(const pxl_grid *grid, const pxl_optim *optim,
                                    double lambda_density, double lambda_sh,
                                    int32_t optimize_step, void *scratch, size_t scratch_bytes,
                                    double lr, double sh_lr, int32_t t, void *params_out,
                                    void *stream) {
    clear_error();
    int32_t st = tv_check(grid, optim, "pxl_tv_adam_step");
    if (st != PXL_OK) return st;
    PXL_REQUIRE(optim->m && optim->v && params_out, "pxl_tv_adam_step: null argument");
    PXL_REQUIRE(t >= 1, "pxl_tv_adam_step: time step %d must be >= 1", t);
    const int P = 1 + 3 * grid->sh_dim;
    const size_t elem = grid->storage == PXL_STORE_F32 ? 4u : 2u;
    const size_t bytes = (size_t)grid->n_vox * (size_t)grid->stride * elem;
    const uintptr_t in0 = (uintptr_t)grid->params, out0 = (uintptr_t)params_out;
    PXL_REQUIRE(out0 + bytes <= in0 || in0 + bytes <= out0,
                "pxl_tv_adam_step: params_out overlaps the grid's records");
    PXL_REQUIRE(grid->stride <= (P > 32 ? 64 : 32),
                "pxl_tv_adam_step: record stride %d above %d", grid->stride, P > 32 ? 64 : 32);
    const int ostride = optim_stride(grid, optim), gs = grad_stride(grid, optim);
    PXL_REQUIRE(ostride >= P && ostride % 4 == 0 && gs >= P && gs % 4 == 0,
                "pxl_tv_adam_step: bad optimiser strides %d / %d", ostride, gs);
    if (grid->sparse) {
        const size_t need = pxl_tv_scratch_bytes(grid);
        if (!scratch || scratch_bytes < need || ((uintptr_t)scratch & 255)) {
            set_error("pxl_tv_adam_step: sparse grids need %zu bytes of scratch prepared by "
                      "pxl_tv_prepare, got %zu", need, scratch_bytes);
            return PXL_ENOMEM;
        }
    }
    TvArgs a;
    fill_tv_args(grid, optim, a);
    PXL_REQUIRE(a.sh_hi <= a.S && a.sh_lo >= 0 && a.sh_lo < a.sh_hi,
                "pxl_tv_adam_step: band [%d, %d) outside the grid's %d coefficients", a.sh_lo,
                a.sh_hi, a.S);
    a.lam_d = (float)lambda_density;
    // SH terms only for optimize_step <= 1 (RenderUtils.cu:2493).
    a.lam_sh = optimize_step <= 1 ? (float)lambda_sh : 0.0f;
    AdamArgs ad;
    ad.out = params_out;
    ad.m = optim->m;
    ad.v = optim->v;
    ad.master = grid->storage == PXL_STORE_F32 ? nullptr : optim->master;
    ad.alr = grid->sparse ? nullptr : optim->adaptive_lr;
    ad.ostride = ostride;
    ad.q4 = (P + 3) / 4 * 4;
    ad.lr = lr;
    ad.sh_lr = sh_lr;
    // Bias corrections with float pow as in the reference (Optimizer.cu:384-385).
    ad.bc1 = 1.0 - (double)powf(0.9f, (float)t);
    ad.bc2 = 1.0 - (double)powf(0.999f, (float)t);
    hipStream_t s = as_stream(stream);
    switch (grid->storage) {
        case PXL_STORE_F32: tv_launch<PXL_STORE_F32>(a, &ad, scratch, s); break;
        case PXL_STORE_BF16: tv_launch<PXL_STORE_BF16>(a, &ad, scratch, s); break;
        default: tv_launch<PXL_STORE_F16>(a, &ad, scratch, s); break;
    }
    return check_launch("pxl_tv_adam_step");
}

extern "C" int32_t pxl_loss_reduce_f64(const double *losses, int64_t len, double *out,
                                       void *stream) {
    clear_error();
    PXL_REQUIRE(losses && out, "pxl_loss_reduce_f64: null argument");
    PXL_REQUIRE(len > 0, "pxl_loss_reduce_f64: empty batch");
    hipLaunchKernelGGL(k_loss_reduce, dim3(1), dim3(1024), 0, as_stream(stream), losses, len, out);
    return check_launch("pxl_loss_reduce_f64");
}

extern "C" int32_t pxl_optim_init_band(const pxl_grid *grid, const pxl_optim *optim,
                                       int32_t band, int32_t band_mode,
                                       int32_t density_initialized, void *stream) {
    clear_error();
    PXL_REQUIRE(grid && optim && optim->grad && optim->m && optim->v,
                "pxl_optim_init_band: null argument");
    PXL_REQUIRE(grid->sh_dim >= 1 && grid->n_vox >= 0, "pxl_optim_init_band: bad grid");
    PXL_REQUIRE(band >= 0 && (band + 1) * (band + 1) <= grid->sh_dim,
                "pxl_optim_init_band: band %d beyond the grid's %d coefficients", band,
                grid->sh_dim);
    const int P = 1 + 3 * grid->sh_dim;
    const int stride = optim_stride(grid, optim), gstride = grad_stride(grid, optim);
    PXL_REQUIRE(stride >= P && gstride >= P, "pxl_optim_init_band: strides %d / %d below %d",
                stride, gstride, P);
    PXL_REQUIRE(grid->storage == PXL_STORE_F32 || optim->stride > 0,
                "pxl_optim_init_band: 16-bit grids need an explicit optimiser stride");
    const int width = stride > gstride ? stride : gstride;
    const int64_t total = grid->n_vox * (int64_t)width;
    if (total == 0) return PXL_OK;
    const int all = band_mode ? 0 : 1;
    const int density = band == 0 || !density_initialized ? 1 : 0;
    hipLaunchKernelGGL(k_init_band, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       as_stream(stream), optim->grad, optim->m, optim->v, grid->n_vox, stride,
                       gstride, width, P, grid->sh_dim, band * band, (band + 1) * (band + 1),
                       density, all);
    return check_launch("pxl_optim_init_band");
}
