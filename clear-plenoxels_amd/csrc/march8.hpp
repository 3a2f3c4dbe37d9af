// Fixed-step marcher with eight samples per wavefront (included by
// render.hip, inside its anonymous namespace; uses RenderArgs and the
// helpers defined there).
//
// Same maths and the same evaluated-sample set as k_render under the fixed
// rule t_k = fma(k, step, t_near) (RenderUtils.cu:1674-1707 with a constant
// step); only the work mapping differs:
//   * occupancy probe: each of the 64 lanes tests one position
//     k0 + lane; a ballot gives the occupied positions of the next 64, so
//     empty space costs one probe per 64 positions;
//   * gather: lane = (slot s, corner j), s = lane / 8.  The next eight
//     occupied positions go to slots 0..7; lane (s, j) loads corner j's whole
//     record (7 x 16-byte loads for 28 FP32 parameters, all independent) and
//     forms its trilinear share of the raw density and of the three raw SH
//     colours; three xor-shuffles sum the 8 corners of a slot;
//   * compositing: transmittance before each slot is an FP64 exclusive
//     product-scan of (1 - alpha) over the slots; opacity and colour
//     prefixes are sum-scans, so the per-sample scalar maths runs once per
//     8 samples of wave time instead of once per sample;
//   * backward: the gradient of the current cell's 8 corners lives in
//     registers with lanes over parameters (4 values per lane); the slots'
//     scalars (trilinear fractions, dL/dsigma, dL/dcolour) are broadcast with
//     readlane and accumulated in order.  Leaving a cell keeps the corners the
//     next cell shares (a face when the ray steps to a neighbour) and adds the
//     rest to the gradient planes, two records per atomic instruction, so a
//     record gets one global atomic per parameter per visit of the ray
//     instead of one per sample.
// The backward recomputes every forward quantity; render.hip is compiled
// with -ffp-contract=off (FMAs written out) so both passes produce the same
// bits and C - C_i stays exact.

template <typename T>
__device__ __forceinline__ T shfl_up64(T v, int off) {
    return __shfl_up(v, off, 64);
}
template <typename T>
__device__ __forceinline__ T shfl_xor64(T v, int off) {
    return __shfl_xor(v, off, 64);
}

// Sum over the 8 corner lanes of a slot (lanes 8s .. 8s+7); all get it.
__device__ __forceinline__ float corner_sum(float v) {
    v += dpp_f<kDppXor1>(v);
    v += dpp_f<kDppXor2>(v);
    v += dpp_f<kDppMirror8>(v);
    return v;
}
// Sum over the 8 slots of values replicated across each slot's lanes.
__device__ __forceinline__ double slot_sum(double v) {
    v += dpp_d<kDppXor8>(v);
    v += xor16_d(v);
    v += xor32_d(v);
    return v;
}
// Inclusive scans over slots of values replicated across each slot's 8
// lanes, in FP64: the second half of each 16-lane row adds the first
// (row_shr:8), then the row totals are carried as in wave_scan_add64.
__device__ __forceinline__ double slot_scan_add(double v) {
    v += dpp_or_d<0x118>(v, 0.0);
    v += dpp_or_d<0x142, 0xA>(v, 0.0);
    v += dpp_or_d<0x143, 0xC>(v, 0.0);
    return v;
}
__device__ __forceinline__ double slot_scan_mul(double v) {
    v *= dpp_or_d<0x118>(v, 1.0);
    v *= dpp_or_d<0x142, 0xA>(v, 1.0);
    v *= dpp_or_d<0x143, 0xC>(v, 1.0);
    return v;
}
// v of the slot before (slot 0: `first`): the first half of each row from
// the previous row's lane 15, the second half from the first half.
__device__ __forceinline__ double slot_prev(double v, double first, int lane) {
    const double from_half = dpp_or_d<0x118>(v, first);
    const double from_row = dpp_or_d<0x142, 0xE>(v, first);
    return (lane & 8) ? from_half : from_row;
}

// Record of P parameters as floats, loaded with 16-byte accesses.
template <int P, int STORE>
__device__ __forceinline__ void load_record(const float *params, size_t rec, int stride,
                                            float (&th)[P]) {
    if (STORE == PXL_STORE_F32) {
        const float4 *src = reinterpret_cast<const float4 *>(params + rec * (size_t)stride);
#pragma unroll
        for (int q = 0; q < (P + 3) / 4; q++) {
            const float4 v = src[q];
            if (4 * q + 0 < P) th[4 * q + 0] = v.x;
            if (4 * q + 1 < P) th[4 * q + 1] = v.y;
            if (4 * q + 2 < P) th[4 * q + 2] = v.z;
            if (4 * q + 3 < P) th[4 * q + 3] = v.w;
        }
    } else {
        const uint4 *src = reinterpret_cast<const uint4 *>(
            reinterpret_cast<const uint16_t *>(params) + rec * (size_t)stride);
#pragma unroll
        for (int q = 0; q < (P + 7) / 8; q++) {
            const uint4 v = src[q];
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int i = 8 * q + e;
                if (i < P) {
                    const uint16_t h = (uint16_t)((w[e >> 1] >> (16 * (e & 1))) & 0xFFFFu);
                    th[i] = STORE == PXL_STORE_BF16 ? __uint_as_float((uint32_t)h << 16)
                                                    : __half2float(__ushort_as_half(h));
                }
            }
        }
    }
}

// Loss and clipped residuals of one ray from its forward totals
// (renderRays_withGradients_RGBGPU, RenderUtils.cu:2149-2193: Huber-clipped
// colour residuals, the mask term gated on A < 1 and the 1e-3 log-odds
// opacity prior).  False when the loss
// is NaN (the reference then skips the ray and books a loss of 255).
struct RayResid {
    double r0, r1, r2, mterm, loss;
};
__device__ __forceinline__ bool ray_residuals(const RenderArgs &a, int ray, double C0, double C1,
                                              double C2, double A, RayResid &o) {
    double r0 = C0 - (double)a.rays.gt_r[ray] / 255.0;
    double r1 = C1 - (double)a.rays.gt_g[ray] / 255.0;
    double r2 = C2 - (double)a.rays.gt_b[ray] / 255.0;
    o.loss = (fabs(r0) + fabs(r1) + fabs(r2)) * 255.0 / 3.0;
    if (isnan(o.loss)) return false;
    const double hd = (double)a.huber;
    const double hi = hd - 0.5 * hd * hd, lo = 0.5 * hd * hd - hd;
    o.r0 = r0 > hd ? hi : (r0 < -hd ? lo : r0);
    o.r1 = r1 > hd ? hi : (r1 < -hd ? lo : r1);
    o.r2 = r2 > hd ? hi : (r2 < -hd ? lo : r2);
    const int gm = a.rays.gt_mask[ray];
    double mterm = 0.0;
    if (gm != 0 && A < 1.0) mterm = ((double)gm / 255.0 - A) * a.mask_coef;
    double wpre = A > 1e-6 ? A : 1e-6;
    wpre = wpre < 1.0 - 1e-6 ? wpre : 1.0 - 1e-6;
    o.mterm = mterm + (log(wpre) - log(1.0 - wpre)) * 1e-3;
    return true;
}

// Per-sample gradient scalars (RenderUtils.cu:2229-2241): dL/draw-colour per
// channel (g0..g2, before the SH basis and trilinear weight) and dL/draw-
// density (gs, before the trilinear weight).  ws = T_i alpha_i, Tn =
// T_{i+1}, Ai and H* the opacity / colour prefixes including sample i.
// R: the type the activations were evaluated in (float under the fixed
// rule, double under the adaptive rule).
template <typename R>
__device__ __forceinline__ void sample_grads(const RayResid &rr, double C0, double C1,
                                             double C2, double A, double step, double ws,
                                             double Tn, double Ai, double H0, double H1,
                                             double H2, R col0, R col1, R col2, R sraw,
                                             float &g0, float &g1, float &g2, double &gs) {
    const double kc = (double)0.6667f;
    g0 = (float)(kc * ws * rr.r0 * (double)((R)0.1 * col0 * ((R)1 - col0)));
    g1 = (float)(kc * ws * rr.r1 * (double)((R)0.1 * col1 * ((R)1 - col1)));
    g2 = (float)(kc * ws * rr.r2 * (double)((R)0.1 * col2 * ((R)1 - col2)));
    const double ds = (double)dsoftplus_shift<R>(sraw);
    gs = rr.r0 * (Tn * (double)col0 - (C0 - H0)) + rr.r1 * (Tn * (double)col1 - (C1 - H1)) +
         rr.r2 * (Tn * (double)col2 - (C2 - H2));
    gs *= 0.6667 * step * ds;
    gs += step * ds * (A - Ai - Tn) * rr.mterm;
}

// Leaves the current cell for one displaced by (dx, dy, dz) cells.  Lane
// (lh, pp) holds in acc[h] parameter pp of corner j = 2h + lh = (jx, jy, jz).
// Old corner j sits at the new cell's corner j - d when that is in {0,1}^3:
// its value moves there; every other old corner is added to the gradient
// planes.  New corners not reached by a move start at zero.
// Record 0 is also where every corner outside the cage of a caged grid
// lands (index_offsets 0, SceneUtils.cu:1106-1113), so it collects a flush
// from a large share of all cells: those go to acc0, a per-lane register
// added to the gradient once per ray (flush_record0), instead of each being
// an atomic on the same cache line.
// `live`: this lane's parameter takes a gradient (the density, or an SH
// coefficient of the optimised band); the other lanes hold zeros.
__device__ __forceinline__ void retire_cell(const RenderArgs &a, float (&acc)[4],
                                            const uint32_t (&rec)[4], double &acc0, int dx,
                                            int dy, int dz, int lh, int pp, bool live) {
    const auto in01 = [](int v) { return v == 0 || v == 1; };
#pragma unroll
    for (int h = 0; h < 4; h++) {
        const bool kept = in01(lh - dx) && in01((h & 1) - dy) && in01((h >> 1) - dz);
        if (!kept && rec[h] == 0u) acc0 += (double)acc[h];
        else if (!kept && live)
            unsafeAtomicAdd(a.grad + (size_t)rec[h] * (size_t)a.gstride + pp, acc[h]);
    }
    // New corner j' = 2h' + lh takes old corner j' + d: register
    // (h' & 1) + dy + 2 ((h' >> 1) + dz), lane half lh + dx.
    const bool x_ok = in01(lh + dx);
    float moved[4];
#pragma unroll
    for (int h = 0; h < 4; h++) {
        const int yy = (h & 1) + dy, zz = (h >> 1) + dz;
        const bool yz_ok = in01(yy) && in01(zz);
        const int src = yz_ok ? yy + 2 * zz : 0;
        float v = src == 0 ? acc[0] : (src == 1 ? acc[1] : (src == 2 ? acc[2] : acc[3]));
        if (dx != 0) v = xor32_f(v);
        moved[h] = (yz_ok && x_ok) ? v : 0.0f;
    }
#pragma unroll
    for (int h = 0; h < 4; h++) acc[h] = moved[h];
}

// Adds the record-0 sum of batch entry b (both lane halves): into the
// entry's FP64 slot of the logged step (a.rec0, folded into the gradient in
// a fixed order afterwards, so the result does not depend on the order in
// which rays finish), else with one float atomic.
__device__ __forceinline__ void flush_record0(const RenderArgs &a, double acc0, int lh, int pp,
                                              bool live, int64_t b) {
    const double v = acc0 + __shfl_xor(acc0, 32, 64);
    if (lh != 0 || !live || v == 0.0) return;
    if (a.rec0 != nullptr) unsafeAtomicAdd(a.rec0 + (size_t)pp * (size_t)a.len + (size_t)b, v);
    else unsafeAtomicAdd(a.grad + pp, (float)v);
}

// Gradient walk of one ray's logged samples (log_grad_ray).  Lane (hf, ch,
// j) = lane bits (5, 4..3, 2..0) holds, for corner j of the current cell and
// channel ch (density, R, G, B), half hf's sum of weight x channel
// coefficient over the samples in the cell (the halves take alternate
// samples of a run).  A chunk's samples are first staged in LDS, kWalkRows
// at a time: per sample row its 32 (corner, channel) products, written by
// the sample's lane, and its 8 corner records; the walk then sums row after
// row with one LDS read per lane, and on a cell change flushes the corners
// the new cell does not share (parameter p of a corner takes its channel's
// sum times Y_k) and moves the shared ones (retire).
constexpr int kWalkRows = 32;
constexpr int kWalkStride = 33;  // floats per row: the sample lanes' writes hit distinct banks
constexpr size_t kWalkLdsBytes = (size_t)kWalkRows * kWalkStride * 4 + (size_t)kWalkRows * 8 * 4;

// A cell change by d = (dx, dy, dz) (each clamped to [-2, 2]; 2, 2, 2: the
// end of the walk) for the gradient walk: bits 0-7 the old corners that
// leave (j - d outside {0,1}^3), bits 8-15 the new corners taken from an
// old one (j + d inside {0,1}^3), bits 16-20 the lane offset dx + 2 dy +
// 4 dz + 16 of that old corner.
__device__ __forceinline__ uint32_t cell_move(int dx, int dy, int dz) {
    const auto in01 = [](int v) { return v == 0 || v == 1; };
    uint32_t m = (uint32_t)(dx + 2 * dy + 4 * dz + 16) << 16;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
        if (!(in01(cx - dx) && in01(cy - dy) && in01(cz - dz))) m |= 1u << c;
        if (in01(cx + dx) && in01(cy + dy) && in01(cz + dz)) m |= 1u << (8 + c);
    }
    return m;
}

struct GradWalk {
    float *q;      // [kWalkRows][kWalkStride]
    uint32_t *r;   // [kWalkRows][8]
    float acc;     // this lane's half sum for (corner, channel)
    uint32_t rec;  // record of this lane's corner of the current cell
    bool has_cur;
    int cx, cy, cz;  // the current cell (wave-uniform)

    __device__ __forceinline__ void init(float *lq, uint32_t *lr) {
        q = lq;
        r = lr;
        acc = 0.0f;
        rec = 0u;
        has_cur = false;
        cx = cy = cz = 0;
    }

    // Leaves the current cell for one displaced by d: move = cell_move(d)
    // (wave-uniform).  Old corner j stays as corner j - d of the new cell
    // when that is in {0,1}^3; the others (bits of rmask) are added to the
    // gradient, two corners per pass (half hf takes the pair's corner hf,
    // lane pp its parameter).  Record 0 (where every corner outside the cage
    // of a caged grid lands, SceneUtils.cu:1106-1113) collects into acc0,
    // added once per ray (flush_record0).
    __device__ __forceinline__ void retire(const RenderArgs &a, uint32_t move, int cj, int cch,
                                           int hf, int pp, int pch, bool live, float ysel, int P,
                                           double &acc0) {
        // The corner's sum over both halves (the same in both).
        float lo, hi;
        halves_f(acc, lo, hi);
        const float tot = lo + hi;
        uint32_t rmask = move & 0xFFu;
        while (rmask != 0u) {
            const int ja = __builtin_ctz(rmask);
            rmask &= rmask - 1u;
            const int jb = rmask != 0u ? __builtin_ctz(rmask) : -1;
            if (jb >= 0) rmask &= rmask - 1u;
            const int jsel = hf ? jb : ja;
            const float v = __shfl(tot, (jsel < 0 ? 0 : jsel) + 8 * pch, 64);
            const uint32_t ra = (uint32_t)__builtin_amdgcn_readlane((int)rec, ja);
            const uint32_t rb = (uint32_t)__builtin_amdgcn_readlane((int)rec, jb < 0 ? ja : jb);
            const uint32_t rc = hf ? rb : ra;
            if (jsel >= 0 && pp < P) {
                const float val = pp == 0 ? v : v * ysel;
                if (rc == 0u) acc0 += (double)val;
                else if (live) unsafeAtomicAdd(a.grad + (size_t)rc * (size_t)a.gstride + pp, val);
            }
        }
        // New corner cj is old corner cj + d (lane cj + dx + 2 dy + 4 dz)
        // when bit cj of the keep mask is set.
        const bool ok = (move >> (8 + cj)) & 1u;
        const int src = cj + (int)((move >> 16) & 31u) - 16 + 8 * cch;
        const float mv = __shfl(tot, ok ? src : 0, 64);
        acc = (ok && hf == 0) ? mv : 0.0f;
    }

    // The m (<= 64) samples of one chunk, sample i in lane i: its cell, the
    // fractions within it, the four channel coefficients and its 8 corner
    // records (q0, q1).
    __device__ __forceinline__ void walk(const RenderArgs &a, int lane, int m, bool act, int bx,
                                         int by, int bz, float fx, float fy, float fz, float gs,
                                         float g0, float g1, float g2, const uint4 &q0,
                                         const uint4 &q1, int cj, int cch, int hf, int pp,
                                         int pch, bool live, float ysel, int P, double &acc0) {
        // A sample starts a new cell when its cell differs from the previous
        // sample's (for the chunk's first, the cell carried in).
        int ux = __shfl_up(bx, 1, 64), uy = __shfl_up(by, 1, 64), uz = __shfl_up(bz, 1, 64);
        bool prev = true;
        if (lane == 0) {
            ux = cx;
            uy = cy;
            uz = cz;
            prev = has_cur;
        }
        const bool changed = act && !(prev && bx == ux && by == uy && bz == uz);
        const uint32_t move = cell_move(clampi(bx - ux, -2, 2), clampi(by - uy, -2, 2),
                                        clampi(bz - uz, -2, 2));
        const uint64_t chg = __ballot(changed);
        float wj[8];
        {
            const float x0 = 1.0f - fx, y0 = 1.0f - fy, z0 = 1.0f - fz;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const float wyz = ((j >> 1) & 1 ? fy : y0) * ((j >> 2) ? fz : z0);
                wj[j] = ((j & 1) ? fx : x0) * wyz;
            }
        }
        const float cc[4] = {gs, g0, g1, g2};
        for (int base = 0; base < m; base += kWalkRows) {
            const int n = min(kWalkRows, m - base);
            if (act && lane >= base && lane < base + kWalkRows) {
                const int row = lane - base;
#pragma unroll
                for (int ch = 0; ch < 4; ch++)
#pragma unroll
                    for (int j = 0; j < 8; j++) q[row * kWalkStride + 8 * ch + j] = wj[j] * cc[ch];
                *reinterpret_cast<uint4 *>(r + row * 8) = q0;
                *reinterpret_cast<uint4 *>(r + row * 8 + 4) = q1;
            }
            // The rows are read by other lanes of this wave.
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const uint64_t rc = (chg >> base) & ((1ull << n) - 1ull);
            int s = 0;
            while (s < n) {
                if ((rc >> s) & 1ull) {
                    if (has_cur)
                        retire(a, (uint32_t)__builtin_amdgcn_readlane((int)move, base + s), cj,
                               cch, hf, pp, pch, live, ysel, P, acc0);
                    has_cur = true;
                    cx = __builtin_amdgcn_readlane(bx, base + s);
                    cy = __builtin_amdgcn_readlane(by, base + s);
                    cz = __builtin_amdgcn_readlane(bz, base + s);
                    rec = r[s * 8 + cj];
                }
                const uint64_t later = rc >> (s + 1);
                const int e = later != 0ull ? min(n, s + 1 + __builtin_ctzll(later)) : n;
                // Two rows per step in flight (each half takes every other row).
                int t = s + hf;
                for (; t + 2 < e; t += 4)
                    acc += q[t * kWalkStride + (lane & 31)] + q[(t + 2) * kWalkStride + (lane & 31)];
                if (t < e) acc += q[t * kWalkStride + (lane & 31)];
                s = e;
            }
            // Every lane's reads of these rows are done before the next
            // rows overwrite them.
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
};

// Batch entry of the next ray a wave of a persistent launch processes.  The
// batch is cut into 8 slices, one per XCD (a block's XCD taken as blockIdx
// mod 8, the dispatch order); a wave claims entries of its XCD's slice from
// that slice's counter and, once the slice is used up, of the next slices in
// turn.  Returns len when every slice is used up; `slice` carries the slice
// the wave is on between calls (start: blockIdx mod 8, `visited` 0).
__device__ __forceinline__ int64_t claim_ray(unsigned long long *ctr, int64_t len, int &slice,
                                             int &visited) {
    while (visited < 8) {
        const int64_t lo = len * slice / 8, hi = len * (slice + 1) / 8;
        unsigned long long got = 0;
        if ((threadIdx.x & 63) == 0) got = atomicAdd(ctr + slice, 1ull);
        const int64_t b = lo + (int64_t)__builtin_amdgcn_readfirstlane((int)got);
        if (b < hi) return b;
        slice = (slice + 1) & 7;
        visited++;
    }
    return len;
}

template <int S, bool SPARSE, int MODE, int STORE>
__device__ __forceinline__ void march8_ray(const RenderArgs &a, const int64_t b);

// Workspace word counting the batch entries longer than the sample log
// (after the 8 + 8 claim counters of the persistent launches).
constexpr int kRedoCounter = 16;

// One wave per ray: either the block's waves take consecutive batch entries
// of an XCD-sliced block order, or (a.persist) a fixed grid of waves claims
// entries until the batch is done.
// The forward + log kernel waits on gathers: five waves per SIMD (at most
// 96 VGPRs, met without spills) keep more rays in flight; the backward
// modes keep their registers.
template <int S, bool SPARSE, int MODE, int STORE, bool PERSIST = false>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(MODE == kLog ? 5 : 1))) void k_march8(const RenderArgs a) {
    if (PERSIST) {
        int slice = blockIdx.x & 7, visited = 0;
        for (int64_t b = claim_ray(a.work, a.len, slice, visited); b < a.len;
             b = claim_ray(a.work, a.len, slice, visited))
            march8_ray<S, SPARSE, MODE, STORE>(a, b);
    } else {
        const int64_t b = (int64_t)xcd_slice_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) +
                          (threadIdx.x >> 6);
        if (b < a.len) march8_ray<S, SPARSE, MODE, STORE>(a, b);
    }
}

// The fused marcher over the batch entries the logged step could not hold
// (the list k_march8<kLog> appends to): a small grid of waves striding over
// the list, which is empty in most batches.
template <int S, bool SPARSE, int STORE>
__global__ __launch_bounds__(256) void k_march8_redo(const RenderArgs a) {
    const int64_t n = (int64_t)a.work[kRedoCounter];
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < n; i += nw)
        march8_ray<S, SPARSE, kFused, STORE>(a, uniform_i(a.redo_list[i]));
}

template <int S, bool SPARSE, int MODE, int STORE>
__device__ __forceinline__ void march8_ray(const RenderArgs &a, const int64_t b) {
    constexpr int P = 1 + 3 * S;
    const int lane = threadIdx.x & 63;
    const int slot = lane >> 3, corner = lane & 7;
    const int ray = uniform_i(a.sample_index[a.start + b]);
    const bool do_fwd = MODE == kFwd || MODE == kFused || MODE == kTrace || MODE == kLog;
    const bool do_bwd = MODE == kBwd || MODE == kFused;
    const bool logging = MODE == kLog;
    // Second fused launch of the logged step: only the rays the log could
    // not hold.
    if (MODE == kFused && a.log_redo && uniform_i(a.log_count[b]) != -1) return;
    // The entry's record-0 slots start at zero (every logged entry, whatever
    // path its gradient takes afterwards).
    if (logging && lane < 32) a.rec0[(size_t)lane * (size_t)a.len + (size_t)b] = 0.0;
    if (ray < 0 || (int64_t)ray >= a.rays.n) {
        if ((do_bwd || logging) && lane == 0) a.losses[b] = 0.0;
        if (logging && lane == 0) a.log_count[b] = 0;
        if (MODE == kTrace && lane == 0) {
            a.trace[2 * b] = 0u;
            a.trace[2 * b + 1] = 2166136261u;
        }
        return;
    }
    const double ox = a.rays.ox[ray], oy = a.rays.oy[ray], oz = a.rays.oz[ray];
    const float fdx = a.rays.dx[ray], fdy = a.rays.dy[ray], fdz = a.rays.dz[ray];
    const double dx = fdx, dy = fdy, dz = fdz;
    const double tnear = a.rays.near_t[ray], tfar = a.rays.far_t[ray];
    const double step = (double)a.step;
    const float fstep = a.step;
    // Positions per coarse probe step: (cstride - 1) * step < 8 cells.
    const int cstride = step > 0.0 && step < 8.0 ? (int)(8.0 / step) : 1;

    if (tnear > tfar) {
        if (do_fwd && lane == 0) {
            a.rays.acc_r[ray] = 0.0f;
            a.rays.acc_g[ray] = 0.0f;
            a.rays.acc_b[ray] = 0.0f;
            a.rays.acc_w[ray] = 0.0f;
            a.rays.acc_density[ray] = 0.0f;
        }
        if (MODE == kTrace && lane == 0) {
            a.trace[2 * b] = 0u;
            a.trace[2 * b + 1] = 2166136261u;
        }
        if ((do_bwd || logging) && lane == 0) a.losses[b] = 0.0;
        if (logging && lane == 0) a.log_count[b] = 0;
        return;
    }

    // Basis of the coefficients that enter the colour (k < sh_active).
    float Y[S];
#pragma unroll
    for (int k = 0; k < S; k++) Y[k] = k < a.sh_active ? sh_basis(k, fdx, fdy, fdz) : 0.0f;
    // Proposal mask of this batch entry (all bins when none is given).
    uint64_t bmask = ~0ull;
    if (a.bin_mask) {
        const uint64_t m = a.bin_mask[b];
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)m);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(m >> 32));
        bmask = ((uint64_t)hi << 32) | (uint64_t)lo;
    }

    // Gradient layout (backward): lane (lh = lane / 32, pp = lane % 32) holds
    // in acc[h] parameter pp of corner 2h + lh of the current cell; ysel is
    // the SH basis value parameter pp multiplies, pch its colour channel.
    const int lh = lane >> 5, pp = lane & 31;
    const int pch = pp == 0 ? -1 : (pp - 1) / S;
    const bool live = pp == 0 || (pp < P && (pp - 1) % S >= a.sh_first && (pp - 1) % S < a.sh_active);
    float ysel = 0.0f;
#pragma unroll
    for (int k = 0; k < S; k++) ysel = (pp >= 1 && (pp - 1) % S == k) ? Y[k] : ysel;
    if (!live) ysel = 0.0f;

    // Backward coefficients (set after the forward / from the ray planes).
    double C0 = 0.0, C1 = 0.0, C2 = 0.0, A = 0.0;
    RayResid rr = {0.0, 0.0, 0.0, 0.0, 0.0};
    int logged = 0;  // samples written to the log (logging mode)

    // pass 0 = forward, pass 1 = backward re-march.
    for (int pass = do_fwd ? 0 : 1; pass < (do_bwd ? 2 : 1); pass++) {
        const bool bwd = pass == 1;
        if (bwd) {
            if (MODE == kBwd) {
                C0 = a.rays.acc_r[ray];
                C1 = a.rays.acc_g[ray];
                C2 = a.rays.acc_b[ray];
                A = a.rays.acc_w[ray];
            }
            if (!ray_residuals(a, ray, C0, C1, C2, A, rr)) {
                if (lane == 0) a.losses[b] = 255.0;
                return;
            }
            if (lane == 0) a.losses[b] = rr.loss;
        }

        // Running totals before the current group (uniform, FP64).
        double T = 1.0, Acc = 0.0, Cr = 0.0, Cg = 0.0, Cb = 0.0, sig = 0.0;
        // Backward: the current cell (base coordinates), its corner records
        // in the acc layout, and the accumulated gradient.
        bool has_cur = false;
        int cur_x = 0, cur_y = 0, cur_z = 0;
        uint32_t cur_rec[4] = {0u, 0u, 0u, 0u};
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        double acc0 = 0.0;
        uint32_t hcount = 0, hash = 2166136261u;
        int k0 = 0;
        uint64_t pending = 0;  // occupied positions of [k0, k0 + 64) not yet done
        bool more = true;      // positions beyond k0 + 64 may remain
        // At the ray's start and after a probe that found nothing, a coarse
        // probe runs first: lane l tests position k0 + cstride * l against
        // the dilated brick map.  The positions from there to the next
        // coarse one lie within 8 cells, so inside that position's brick or
        // a neighbour of it: a clear dilated bit clears them all, and the
        // fine probe starts at the first coarse position whose bit is set.
        bool coarse = SPARSE && a.dbricks != nullptr;
        while (true) {
            if (pending == 0) {
                if (!more) break;
                if (SPARSE && coarse) {
                    const int kc = k0 + cstride * lane;
                    const double tc = fma((double)kc, step, tnear);
                    const bool cvalid = tc < tfar && kc < kMaxSteps;
                    bool hit = false;
                    if (cvalid) {
                        const int cx = clampi(base_of(fma(tc, dx, ox)), 0, a.W - 2),
                                  cy = clampi(base_of(fma(tc, dy, oy)), 0, a.H - 2),
                                  cz = clampi(base_of(fma(tc, dz, oz)), 0, a.D - 2);
                        const uint32_t br = (uint32_t)(cz >> 3) * (uint32_t)a.nbx * (uint32_t)a.nby +
                                            (uint32_t)(cy >> 3) * (uint32_t)a.nbx + (uint32_t)(cx >> 3);
                        hit = (a.dbricks[br >> 5] >> (br & 31)) & 1u;
                    }
                    const uint64_t hits = __ballot(hit);
                    if (hits == 0ull) {
                        // Every position up to the ray's end is empty, or on
                        // to the next 64 coarse positions.
                        if (__ballot(cvalid) != ~0ull) break;
                        k0 += 64 * cstride;
                        continue;
                    }
                    k0 += cstride * __builtin_ctzll(hits);
                    coarse = false;
                }
                // Probe 64 positions, one per lane.
                const int k = k0 + lane;
                const double t = fma((double)k, step, tnear);
                const bool valid = t < tfar && k < kMaxSteps;
                bool occ = valid;
                if (SPARSE && valid) {
                    const int bx = base_of(fma(t, dx, ox)), by = base_of(fma(t, dy, oy)),
                              bz = base_of(fma(t, dz, oz));
                    const int cx = clampi(bx, 0, a.W - 2), cy = clampi(by, 0, a.H - 2),
                              cz = clampi(bz, 0, a.D - 2);
                    // The cell bit alone: the coarse probe has already
                    // stepped over the empty bricks, and one load here keeps
                    // a probe to one round trip.
                    const uint32_t cell = (uint32_t)cz * (uint32_t)(a.W - 1) * (uint32_t)(a.H - 1) +
                                          (uint32_t)cy * (uint32_t)(a.W - 1) + (uint32_t)cx;
                    occ = (a.occ[cell >> 5] >> (cell & 31)) & 1u;
                }
                if (bmask != ~0ull && occ) occ = (bmask >> bin_of(k, step, tnear, tfar)) & 1ull;
                pending = __ballot(occ);
                more = __ballot(valid) == ~0ull;
                if (pending == 0) {
                    k0 += 64;
                    coarse = SPARSE && a.dbricks != nullptr;
                    continue;
                }
            }
            // Next (up to) eight occupied positions -> slots.
            int ks[8];
            int n = 0;
#pragma unroll
            for (int s = 0; s < 8; s++) {
                ks[s] = -1;
                if (pending != 0) {
                    const int bit = __builtin_ctzll(pending);
                    pending &= pending - 1;
                    ks[s] = k0 + bit;
                    n++;
                }
            }
            if (pending == 0) k0 += 64;
            int kslot = ks[0];
#pragma unroll
            for (int s = 1; s < 8; s++) kslot = slot == s ? ks[s] : kslot;
            const bool active = slot < n;

            // Sample position, base cell and this lane's corner.
            const double t = fma((double)(active ? kslot : 0), step, tnear);
            const double px = fma(t, dx, ox), py = fma(t, dy, oy), pz = fma(t, dz, oz);
            const int bx = base_of(px), by = base_of(py), bz = base_of(pz);
            const int jx = corner & 1, jy = (corner >> 1) & 1, jz = corner >> 2;
            const float fx = (float)(px - (double)bx), fy = (float)(py - (double)by),
                        fz = (float)(pz - (double)bz);
            const float w = (jx ? fx : 1.0f - fx) * (jy ? fy : 1.0f - fy) * (jz ? fz : 1.0f - fz);
            uint32_t rec = scene_index(bx + jx, by + jy, bz + jz, a.W, a.H, a.D);
            if (SPARSE) rec = min(a.io[rec], (uint32_t)(a.n_vox - 1));

            if (MODE == kTrace && !bwd) {
                for (int s = 0; s < n; s++) {
                    const int src = 8 * s;
                    hcount++;
                    hash = fnv_mix(hash, (uint32_t)ks[s]);
                    hash = fnv_mix(hash, (uint32_t)__shfl(bx, src, 64));
                    hash = fnv_mix(hash, (uint32_t)__shfl(by, src, 64));
                    hash = fnv_mix(hash, (uint32_t)__shfl(bz, src, 64));
                }
            }

            float th[P];
            load_record<P, STORE>(a.params, active ? rec : 0u, a.stride, th);
            // This corner's share: density and SH-weighted colours.
            float part[4];
            part[0] = w * th[0];
#pragma unroll
            for (int ch = 0; ch < 3; ch++) {
                float c = 0.0f;
#pragma unroll
                for (int k = 0; k < S; k++) c = fmaf(Y[k], th[1 + ch * S + k], c);
                part[1 + ch] = w * c;
            }
            const float sraw = corner_sum(part[0]);
            const float cr0 = corner_sum(part[1]);
            const float cr1 = corner_sum(part[2]);
            const float cr2 = corner_sum(part[3]);

            const float sigma = active ? softplus_shift<float>(sraw) : 0.0f;
            const float col0 = sigmoid01<float>(cr0), col1 = sigmoid01<float>(cr1),
                        col2 = sigmoid01<float>(cr2);
            const float alpha = active ? one_minus_exp_neg<float>(sigma * fstep) : 0.0f;
            // Transmittance before each slot: exclusive product-scan.
            const double keep = 1.0 - (double)alpha;
            const double incl = slot_scan_mul(keep);
            const double excl = slot_prev(incl, 1.0, lane);
            const double Ts = T * excl;
            const double ws = Ts * (double)alpha;
            const double total_keep = readlane_d(incl, 63);

            if (!bwd) {
                if (logging) {
                    // Lane (s, j): record of corner j, and word j of the
                    // sample's (sraw, col0, col1, col2, k) row.
                    const int i = logged + slot;
                    if (active && i < a.log_cap) {
                        const size_t at = ((size_t)b * (size_t)a.log_cap + (size_t)i) * 8 + corner;
                        a.log_rec[at] = rec;
                        if (corner < 5)
                            a.log_val[at] = corner == 0   ? sraw
                                            : corner == 1 ? col0
                                            : corner == 2 ? col1
                                            : corner == 3 ? col2
                                                          : __int_as_float(kslot);
                    }
                    logged += n;
                }
                // Forward totals: xor-sums over the slots of one corner column
                // (every column holds the same per-slot values).
                Acc += slot_sum(ws);
                Cr += slot_sum(ws * (double)col0);
                Cg += slot_sum(ws * (double)col1);
                Cb += slot_sum(ws * (double)col2);
                sig += slot_sum((double)sigma);
                T *= total_keep;
            } else {
                // Prefix sums up to and including each slot.
                const double Ai = Acc + slot_scan_add(ws);
                const double H0 = Cr + slot_scan_add(ws * (double)col0);
                const double H1 = Cg + slot_scan_add(ws * (double)col1);
                const double H2 = Cb + slot_scan_add(ws * (double)col2);
                const double Tn = Ts * keep;
                float g0, g1, g2;
                double gs;
                sample_grads(rr, C0, C1, C2, A, step, ws, Tn, Ai, H0, H1, H2, col0, col1, col2,
                             sraw, g0, g1, g2, gs);
                // Accumulate the slots in order.  A slot in the current cell
                // adds to acc; a slot in another cell first retires the
                // current one: corners the two cells share (same vertex, so
                // the same record) move to their new corner position, the
                // others go to the gradient planes with one float atomic per
                // parameter.  All control flow here is wave-uniform.
                const float gsf = (float)gs;
#pragma unroll
                for (int s2 = 0; s2 < 8; s2++) {
                    if (s2 >= n) break;
                    const int sb = 8 * s2;
                    const int nx = __builtin_amdgcn_readlane(bx, sb);
                    const int ny = __builtin_amdgcn_readlane(by, sb);
                    const int nz = __builtin_amdgcn_readlane(bz, sb);
                    if (!has_cur || nx != cur_x || ny != cur_y || nz != cur_z) {
                        if (has_cur) retire_cell(a, acc, cur_rec, acc0, nx - cur_x,
                                                 ny - cur_y, nz - cur_z, lh, pp, live);
                        has_cur = true;
                        cur_x = nx;
                        cur_y = ny;
                        cur_z = nz;
#pragma unroll
                        for (int h = 0; h < 4; h++)
                            cur_rec[h] = (uint32_t)__shfl((int)rec, sb + 2 * h + lh, 64);
                    }
                    const float sfx = readlane_f(fx, sb), sfy = readlane_f(fy, sb),
                                sfz = readlane_f(fz, sb);
                    const float sgs = readlane_f(gsf, sb), sg0 = readlane_f(g0, sb),
                                sg1 = readlane_f(g1, sb), sg2 = readlane_f(g2, sb);
                    const float coef =
                        pch < 0 ? sgs : (pch == 0 ? sg0 : (pch == 1 ? sg1 : sg2)) * ysel;
                    const float wx = lh ? sfx : 1.0f - sfx;
#pragma unroll
                    for (int h = 0; h < 4; h++) {
                        const float wyz = ((h & 1) ? sfy : 1.0f - sfy) * ((h >> 1) ? sfz : 1.0f - sfz);
                        acc[h] = fmaf(wx * wyz, coef, acc[h]);
                    }
                }
                Acc = readlane_d(Ai, 63);
                Cr = readlane_d(H0, 63);
                Cg = readlane_d(H1, 63);
                Cb = readlane_d(H2, 63);
                T *= total_keep;
            }
        }

        if (bwd && has_cur) {
            // Retire the last cell: a displacement of 2 shares no corner.
            retire_cell(a, acc, cur_rec, acc0, 2, 2, 2, lh, pp, live);
            flush_record0(a, acc0, lh, pp, live, b);
        }

        if (!bwd) {
            if (a.white_bg) {
                Cr += 1.0 - Acc;
                Cg += 1.0 - Acc;
                Cb += 1.0 - Acc;
            }
            C0 = Cr;
            C1 = Cg;
            C2 = Cb;
            A = Acc;
            if (lane == 0) {
                a.rays.acc_r[ray] = (float)Cr;
                a.rays.acc_g[ray] = (float)Cg;
                a.rays.acc_b[ray] = (float)Cb;
                a.rays.acc_w[ray] = (float)Acc;
                a.rays.acc_density[ray] = (float)sig;
                if (MODE == kTrace) {
                    a.trace[2 * b] = hcount;
                    a.trace[2 * b + 1] = hash;
                }
            }
            if (logging) {
                if (logged <= a.log_cap) {
                    // Whole ray logged: the loss here, the gradient in
                    // k_log_grad.
                    const bool ok = ray_residuals(a, ray, C0, C1, C2, A, rr);
                    if (lane == 0) {
                        a.losses[b] = ok ? rr.loss : 255.0;
                        a.log_count[b] = ok ? logged : 0;
                        a.log_tot[4 * b + 0] = C0;
                        a.log_tot[4 * b + 1] = C1;
                        a.log_tot[4 * b + 2] = C2;
                        a.log_tot[4 * b + 3] = A;
                    }
                    return;
                }
                // Longer than the log: the fused marcher redoes this ray
                // (k_march8_redo walks the list of such entries).
                if (lane == 0) {
                    a.log_count[b] = -1;
                    a.redo_list[atomicAdd(a.work + kRedoCounter, 1ull)] = (int32_t)b;
                }
                return;
            }
        }
    }
}

// Gradient of the rays k_march8 logged (second kernel of the two-kernel
// fused step).  One wave per batch entry; 64 logged samples at a time, one
// per lane: sigma, alpha and the colours come from the logged raw values
// through the same float functions as the forward, transmittance and the
// opacity / colour prefixes are FP64 wave scans, the gradient scalars are
// those of sample_grads.  The scattering then runs as in the marcher's own
// backward (register accumulators, corner carry), but with every input
// already in registers: no global load follows an atomic until the next 64
// samples, so the atomics never hold up the wave.  chunk >= 0: only the
// logged samples 64 chunk .. 64 chunk + 63, starting from the prefix
// k_log_prefix stored for them (chunk-parallel schedule).
// ADAPT: the log of the adaptive rule's marcher (k_adapt8): per sample 8
// FP64 words (raw density, three raw colours, the position t), activations
// in FP64.
template <int S, bool TGT, bool ADAPT>
__device__ __forceinline__ void log_grad_ray(const RenderArgs &a, const int64_t b,
                                             float *lds_q, uint32_t *lds_r, const int chunk = -1);

// The logged raw values of sample row i of entry b and the activations of
// the rule: sigma / alpha / colours (float under the fixed rule, FP64 under
// the adaptive one), and the sample's position t.
template <bool ADAPT>
struct LoggedSample {
    using R = typename std::conditional<ADAPT, double, float>::type;
    R sraw, col0, col1, col2, alpha;
    double t;
    int k;
};
template <bool ADAPT>
__device__ __forceinline__ LoggedSample<ADAPT> read_logged(const RenderArgs &a, size_t row,
                                                           bool act, double step, double tnear) {
    using R = typename LoggedSample<ADAPT>::R;
    LoggedSample<ADAPT> o;
    if (ADAPT) {
        const double *v = reinterpret_cast<const double *>(a.log_val) + row * 8;
        o.sraw = (R)v[0];
        o.col0 = (R)sigmoid01<double>(v[1]);
        o.col1 = (R)sigmoid01<double>(v[2]);
        o.col2 = (R)sigmoid01<double>(v[3]);
        o.t = v[4];
        o.k = 0;
    } else {
        const float4 v = *reinterpret_cast<const float4 *>(a.log_val + row * 8);
        o.k = __float_as_int(a.log_val[row * 8 + 4]);
        o.sraw = (R)v.x;
        o.col0 = (R)v.y;
        o.col1 = (R)v.z;
        o.col2 = (R)v.w;
        o.t = fma((double)(act ? o.k : 0), step, tnear);
    }
    const R sigma = act ? softplus_shift<R>(o.sraw) : (R)0;
    o.alpha = act ? one_minus_exp_neg<R>(sigma * (R)step) : (R)0;
    return o;
}

// TGT: also write the proposal targets (a.bin_targets, a.target_valid).
// Each wave's gradient-walk rows in the block's dynamic LDS (launches pass
// kWalkLdsBytes per wave).
__device__ __forceinline__ void walk_lds(float *&q, uint32_t *&r) {
    extern __shared__ __align__(16) unsigned char walk_lds_base[];
    unsigned char *w = walk_lds_base + (size_t)(threadIdx.x >> 6) * kWalkLdsBytes;
    q = reinterpret_cast<float *>(w);
    r = reinterpret_cast<uint32_t *>(w + (size_t)kWalkRows * kWalkStride * 4);
}

template <int S, bool PERSIST = false, bool TGT = false, bool ADAPT = false>
__global__ __launch_bounds__(256) void k_log_grad(const RenderArgs a) {
    float *lq;
    uint32_t *lr;
    walk_lds(lq, lr);
    if (PERSIST) {
        int slice = blockIdx.x & 7, visited = 0;
        for (int64_t b = claim_ray(a.work + 8, a.len, slice, visited); b < a.len;
             b = claim_ray(a.work + 8, a.len, slice, visited))
            log_grad_ray<S, TGT, ADAPT>(a, b, lq, lr);
    } else {
        const int64_t b = (int64_t)xcd_slice_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) +
                          (threadIdx.x >> 6);
        if (b < a.len) log_grad_ray<S, TGT, ADAPT>(a, b, lq, lr);
    }
}

// Transmittance and the opacity / colour prefixes (T, A, C0, C1, C2) in
// front of each 64-sample chunk of a ray's log after the first, for the
// chunk-parallel schedule: the same float activations and FP64 wave scans,
// in the same order, as log_grad_ray, so every chunk starts from the values
// the one-wave-per-ray kernel carries into it.  One wave per batch entry;
// entries with at most 64 logged samples have nothing to store.
template <int S, bool ADAPT = false>
__global__ __launch_bounds__(256) void k_log_prefix(const RenderArgs a) {
    const int64_t b = (int64_t)xcd_slice_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) +
                      (threadIdx.x >> 6);
    if (b >= a.len) return;
    const int lane = threadIdx.x & 63;
    const int cnt = uniform_i(a.log_count[b]);
    const double step = (double)a.step;
    double T = 1.0, Acc = 0.0, Cr = 0.0, Cg = 0.0, Cb = 0.0;
    for (int c0 = 0; c0 + 64 < cnt; c0 += 64) {
        const size_t row = (size_t)b * (size_t)a.log_cap + (size_t)(c0 + lane);
        const LoggedSample<ADAPT> v = read_logged<ADAPT>(a, row, true, step, 0.0);
        const double keep = 1.0 - (double)v.alpha;
        const double incl = wave_scan_mul64(keep);
        const double excl = wave_shr1_d(incl, 1.0);
        const double ws = T * excl * (double)v.alpha;
        const double Ai = Acc + wave_scan_add64(ws);
        const double H0 = Cr + wave_scan_add64(ws * (double)v.col0);
        const double H1 = Cg + wave_scan_add64(ws * (double)v.col1);
        const double H2 = Cb + wave_scan_add64(ws * (double)v.col2);
        T *= readlane_d(incl, 63);
        Acc = readlane_d(Ai, 63);
        Cr = readlane_d(H0, 63);
        Cg = readlane_d(H1, 63);
        Cb = readlane_d(H2, 63);
        double *pre = a.log_pre + ((size_t)b * (size_t)a.log_chunks + (size_t)(c0 / 64 + 1)) * 5;
        if (lane < 5)
            pre[lane] = lane == 0 ? T : (lane == 1 ? Acc : (lane == 2 ? Cr : (lane == 3 ? Cg : Cb)));
    }
}

// Chunk-parallel gradient: one wave per (batch entry, 64-sample chunk of its
// log), entry = item / log_chunks; the corner accumulators are flushed at
// the end of each chunk.  For batches too small to fill the chip, where the
// longest rays' chunks would otherwise run one after another.
template <int S, bool ADAPT = false>
__global__ __launch_bounds__(256) void k_log_grad_chunked(const RenderArgs a, int64_t n_items) {
    const int64_t item = (int64_t)xcd_slice_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) +
                         (threadIdx.x >> 6);
    if (item >= n_items) return;
    float *lq;
    uint32_t *lr;
    walk_lds(lq, lr);
    log_grad_ray<S, false, ADAPT>(a, item / a.log_chunks, lq, lr, (int)(item % a.log_chunks));
}

template <int S, bool TGT, bool ADAPT>
__device__ __forceinline__ void log_grad_ray(const RenderArgs &a, const int64_t b,
                                             float *lds_q, uint32_t *lds_r, const int chunk) {
    constexpr int P = 1 + 3 * S;
    const int lane = threadIdx.x & 63;
    const int cnt = uniform_i(a.log_count[b]);
    if (cnt <= 0) {  // nothing logged, or done by the marcher
        if (TGT && lane == 0) {
            a.bin_targets[b] = 0ull;
            a.target_valid[b] = cnt == 0 ? 1 : 0;
        }
        return;
    }
    const int c_begin = chunk < 0 ? 0 : 64 * chunk;
    if (c_begin >= cnt) return;
    const int c_end = chunk < 0 ? cnt : min(cnt, c_begin + 64);
    const int ray = uniform_i(a.sample_index[a.start + b]);
    const double ox = a.rays.ox[ray], oy = a.rays.oy[ray], oz = a.rays.oz[ray];
    const float fdx = a.rays.dx[ray], fdy = a.rays.dy[ray], fdz = a.rays.dz[ray];
    const double dx = fdx, dy = fdy, dz = fdz;
    const double tnear = a.rays.near_t[ray];
    const double step = (double)a.step;
    const double C0 = a.log_tot[4 * b + 0], C1 = a.log_tot[4 * b + 1],
                 C2 = a.log_tot[4 * b + 2], A = a.log_tot[4 * b + 3];
    RayResid rr;
    ray_residuals(a, ray, C0, C1, C2, A, rr);  // finite: the marcher logged it

    float Y[S];
#pragma unroll
    for (int k = 0; k < S; k++) Y[k] = sh_basis(k, fdx, fdy, fdz);
    // Flush lanes (hf, pp): parameter pp of the corner a flush pair gives
    // half hf; its channel (density, R, G, B) and SH factor.  An SH
    // coefficient's gradient is its channel's coefficient times Y_k, Y_k
    // fixed along the ray, so the walk sums per corner and channel only and
    // multiplies by Y_k when a record is flushed.
    const int hf = lane >> 5, pp = lane & 31;
    const int pch = pp == 0 ? 0 : min(3, 1 + (pp - 1) / S);
    const bool live = pp == 0 || (pp < P && (pp - 1) % S >= a.sh_first && (pp - 1) % S < a.sh_active);
    float ysel = 1.0f;
#pragma unroll
    for (int k = 0; k < S; k++) ysel = (pp >= 1 && (pp - 1) % S == k) ? Y[k] : ysel;
    if (!live) ysel = 0.0f;
    // Accumulator lanes: corner cj = (jx, jy, jz) and channel ch of lane & 31;
    // the two halves sum alternate samples of a run.
    const int cj = lane & 7, cch = (lane >> 3) & 3;
    GradWalk w;
    w.init(lds_q, lds_r);

    double T = 1.0, Acc = 0.0, Cr = 0.0, Cg = 0.0, Cb = 0.0;
    if (chunk > 0) {
        const double *pre = a.log_pre + ((size_t)b * (size_t)a.log_chunks + (size_t)chunk) * 5;
        T = pre[0];
        Acc = pre[1];
        Cr = pre[2];
        Cg = pre[3];
        Cb = pre[4];
    }
    double acc0 = 0.0;
    uint64_t tmask = 0ull;  // bins holding weight (proposal targets)
    const double tfar = a.rays.far_t[ray];
    for (int c0 = c_begin; c0 < c_end; c0 += 64) {
        const int m = min(64, c_end - c0);
        const bool act = lane < m;
        const size_t row = (size_t)b * (size_t)a.log_cap + (size_t)(c0 + (act ? lane : 0));
        const size_t at = row * 8;
        const LoggedSample<ADAPT> v = read_logged<ADAPT>(a, row, act, step, tnear);
        const int k = v.k;
        const uint4 q0 = *reinterpret_cast<const uint4 *>(a.log_rec + at);
        const uint4 q1 = *reinterpret_cast<const uint4 *>(a.log_rec + at + 4);
        const auto sraw = v.sraw, col0 = v.col0, col1 = v.col1, col2 = v.col2;
        const auto alpha = v.alpha;
        const double keep = 1.0 - (double)alpha;
        const double incl = wave_scan_mul64(keep);
        const double excl = wave_shr1_d(incl, 1.0);
        const double Ts = T * excl;
        const double ws = Ts * (double)alpha;
        const double Ai = Acc + wave_scan_add64(ws);
        const double H0 = Cr + wave_scan_add64(ws * (double)col0);
        const double H1 = Cg + wave_scan_add64(ws * (double)col1);
        const double H2 = Cb + wave_scan_add64(ws * (double)col2);
        const double Tn = Ts * keep;
        float g0, g1, g2;
        double gs;
        sample_grads(rr, C0, C1, C2, A, step, ws, Tn, Ai, H0, H1, H2, col0, col1, col2, sraw, g0,
                     g1, g2, gs);
        if (TGT) {
            uint64_t bit = act && ws > a.target_wmin ? 1ull << bin_of(k, step, tnear, tfar) : 0ull;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1)
                bit |= (uint64_t)__shfl_xor((unsigned long long)bit, off, 64);
            tmask |= bit;
        }
        // Position of the sample exactly as the marcher computed it.
        const double t = v.t;
        const double px = fma(t, dx, ox), py = fma(t, dy, oy), pz = fma(t, dz, oz);
        const int bx = base_of(px), by = base_of(py), bz = base_of(pz);
        const float fx = (float)(px - (double)bx), fy = (float)(py - (double)by),
                    fz = (float)(pz - (double)bz);
        w.walk(a, lane, m, act, bx, by, bz, fx, fy, fz, (float)gs, g0, g1, g2, q0, q1, cj, cch,
               hf, pp, pch, live, ysel, P, acc0);
        T *= readlane_d(incl, 63);
        Acc = readlane_d(Ai, 63);
        Cr = readlane_d(H0, 63);
        Cg = readlane_d(H1, 63);
        Cb = readlane_d(H2, 63);
    }
    if (w.has_cur) {
        w.retire(a, 0xFFu | (16u << 16), cj, cch, hf, pp, pch, live, ysel, P, acc0);
        flush_record0(a, acc0, hf, pp, live, b);
    }
    if (TGT && lane == 0) {
        a.bin_targets[b] = tmask;
        a.target_valid[b] = 1;
    }
}
