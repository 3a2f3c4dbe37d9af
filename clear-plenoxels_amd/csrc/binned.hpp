// Brick-binned gradient of the logged step (schedule gradient
// PXL_GRAD_BINNED; included by render.hip after march8.hpp, inside its
// anonymous namespace).
//
// The per-ray gradient walk (k_log_grad) adds each record a ray leaves to the
// gradient planes with global float atomics: ~2.9 row flushes per touched
// record per batch at 256^3, executed one by one at the memory side.  Here
// the same gradient is formed owner-side instead:
//
//   1. k_bin_samples (one wave per ray): the FP64 scans and per-sample
//      scalars of k_log_grad (dL/dsigma_raw, dL/dcolour_raw per channel),
//      written with the sample's trilinear fractions and cell over its
//      corner-record row of the log; the ray's SH basis per batch entry; a
//      count per bin key.  The key of a sample is its cell's brick of 8^3
//      vertices and a 3-bit class: bit a set when the cell is the last of
//      its brick along axis a, i.e. when corners of the cell belong to the
//      next brick along a.
//   2. k_bin_scan_sum / k_bin_scan_blocks / k_bin_scan_final: exclusive
//      prefix of the counts (the fill pointers) and the list of bricks that
//      own a corner of any sample (the targets).
//   3. k_bin_scatter: the log rows in key order.
//   4. k_bin_accum (one workgroup per target brick): the samples of the
//      brick's own keys and the neighbour keys whose cells reach into it (27
//      key ranges), each corner the brick owns added into an LDS row per
//      owned vertex; then one plain read-modify-write of each touched
//      record.  Every record has one owner, so no global atomics remain
//      except FP64 ones into the record-0 slots (where every corner outside
//      the cage of a caged grid lands, SceneUtils.cu:1106-1113).
//
// The sums are those of renderRays_withGradients_RGB's scatter
// (RenderUtils.cu:1074-1364, 2229-2256): per corner j of a sample,
// w_j * dL/dsigma_raw for the density and w_j * dL/dcolour_ch * Y_k for SH
// coefficient k of channel ch.  A corner clamped onto the grid's last vertex
// (a sample whose base cell is the last vertex along an axis) takes its
// two weights, (1 - f) + f, as one weight 1 on that vertex.

constexpr int kBinShift = 3;              // bricks of 8^3 vertices
constexpr int kBinVerts = 1 << (3 * kBinShift);
constexpr int kBinScanKeys = 8;           // keys per thread of the scan kernels
constexpr int kBinScanThreads = 1024;
constexpr int kBinScanBlockKeys = kBinScanKeys * kBinScanThreads;
constexpr int kBinYStride = 9;            // SH basis floats per batch entry
constexpr int kTargetCounter = kRedoCounter + 1;  // workspace word: target bricks
constexpr uint32_t kNoRecord = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t pack_cell(int cx, int cy, int cz) {
    return (uint32_t)cx | ((uint32_t)cy << 10) | ((uint32_t)cz << 20);
}
__device__ __forceinline__ void unpack_cell(uint32_t c, int &cx, int &cy, int &cz) {
    cx = (int)(c & 1023u);
    cy = (int)((c >> 10) & 1023u);
    cz = (int)((c >> 20) & 1023u);
}
// Bin key of a cell: brick (vertex bricks, z-y-x order) * 8 + class.
__device__ __forceinline__ uint32_t bin_key(const RenderArgs &a, uint32_t packed) {
    int cx, cy, cz;
    unpack_cell(packed, cx, cy, cz);
    const uint32_t brick =
        ((uint32_t)(cz >> kBinShift) * (uint32_t)a.vby + (uint32_t)(cy >> kBinShift)) *
            (uint32_t)a.vbx +
        (uint32_t)(cx >> kBinShift);
    const uint32_t cls = (uint32_t)((cx & 7) == 7) | ((uint32_t)((cy & 7) == 7) << 1) |
                         ((uint32_t)((cz & 7) == 7) << 2);
    return brick * 8u + cls;
}
// The cell whose corners carry a sample's weights and the fractions within
// it, along one axis of n vertices: the base cell, except that a base cell
// on the last vertex (both corners clamp onto it) becomes the cell below
// with fraction 1.
__device__ __forceinline__ void bin_axis(int base, float f, int n, int &c, float &fe) {
    if (base >= n - 1) {
        c = n - 2;
        fe = 1.0f;
    } else {
        c = base;
        fe = f;
    }
}

// Lanes with equal keys, one representative per group: `each(key, n, lead)`
// runs on lane `lead` for every distinct key among the active lanes (n lanes
// share it); returns the lead lane of this lane's group.
// `group` gets the mask of the lanes sharing this lane's key.
template <typename F>
__device__ __forceinline__ int for_each_key(bool act, uint32_t key, int lane, F each,
                                            uint64_t &group) {
    uint64_t todo = __ballot(act);
    int mine = lane;
    group = 0ull;
    while (todo != 0ull) {
        const int lead = __builtin_ctzll(todo);
        const uint32_t k0 = (uint32_t)__builtin_amdgcn_readlane((int)key, lead);
        const uint64_t same = __ballot(act && key == k0);
        if (act && key == k0) {
            mine = lead;
            group = same;
        }
        if (lane == lead) each(k0, (uint32_t)__popcll(same));
        todo &= ~same;
    }
    return mine;
}

// Stage 1: per-sample gradient scalars of one ray (as log_grad_ray, without
// the walk), the bin counts, the entry's SH basis; the proposal targets with
// TGT.  The sample's corner-record row of the log (8 words) becomes
// (dL/dsigma_raw, dL/dcolour_raw x 3, fx, fy, fz, packed cell).
template <int S, bool TGT, bool ADAPT>
__device__ __forceinline__ void bin_samples_ray(const RenderArgs &a, const int64_t b) {
    const int lane = threadIdx.x & 63;
    const int cnt = uniform_i(a.log_count[b]);
    if (cnt <= 0) {  // nothing logged, or done by the marcher
        if (TGT && lane == 0) {
            a.bin_targets[b] = 0ull;
            a.target_valid[b] = cnt == 0 ? 1 : 0;
        }
        return;
    }
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
    if (lane < S) a.gb_y[(size_t)b * kBinYStride + lane] = sh_basis(lane, fdx, fdy, fdz);

    double T = 1.0, Acc = 0.0, Cr = 0.0, Cg = 0.0, Cb = 0.0;
    uint64_t tmask = 0ull;
    const double tfar = a.rays.far_t[ray];
    for (int c0 = 0; c0 < cnt; c0 += 64) {
        const int m = min(64, cnt - c0);
        const bool act = lane < m;
        const size_t row = (size_t)b * (size_t)a.log_cap + (size_t)(c0 + (act ? lane : 0));
        const LoggedSample<ADAPT> v = read_logged<ADAPT>(a, row, act, step, tnear);
        const double keep = 1.0 - (double)v.alpha;
        const double incl = wave_scan_mul64(keep);
        const double excl = wave_shr1_d(incl, 1.0);
        const double Ts = T * excl;
        const double ws = Ts * (double)v.alpha;
        const double Ai = Acc + wave_scan_add64(ws);
        const double H0 = Cr + wave_scan_add64(ws * (double)v.col0);
        const double H1 = Cg + wave_scan_add64(ws * (double)v.col1);
        const double H2 = Cb + wave_scan_add64(ws * (double)v.col2);
        const double Tn = Ts * keep;
        float g0, g1, g2;
        double gs;
        sample_grads(rr, C0, C1, C2, A, step, ws, Tn, Ai, H0, H1, H2, v.col0, v.col1, v.col2,
                     v.sraw, g0, g1, g2, gs);
        if (TGT) {
            uint64_t bit = act && ws > a.target_wmin ? 1ull << bin_of(v.k, step, tnear, tfar) : 0ull;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1)
                bit |= (uint64_t)__shfl_xor((unsigned long long)bit, off, 64);
            tmask |= bit;
        }
        // Position exactly as the marcher computed it.
        const double t = v.t;
        const double px = fma(t, dx, ox), py = fma(t, dy, oy), pz = fma(t, dz, oz);
        const int bx = base_of(px), by = base_of(py), bz = base_of(pz);
        int cx, cy, cz;
        float fx, fy, fz;
        bin_axis(bx, (float)(px - (double)bx), a.W, cx, fx);
        bin_axis(by, (float)(py - (double)by), a.H, cy, fy);
        bin_axis(bz, (float)(pz - (double)bz), a.D, cz, fz);
        const uint32_t packed = pack_cell(cx, cy, cz);
        if (act) {
            uint4 *dst = reinterpret_cast<uint4 *>(a.log_rec + row * 8);
            dst[0] = make_uint4(__float_as_uint((float)gs), __float_as_uint(g0),
                                __float_as_uint(g1), __float_as_uint(g2));
            dst[1] = make_uint4(__float_as_uint(fx), __float_as_uint(fy), __float_as_uint(fz),
                                packed);
        }
        const uint32_t key = bin_key(a, packed);
        uint64_t group;
        for_each_key(
            act, key, lane, [&](uint32_t k0, uint32_t n) { atomicAdd(a.gb_count + k0, n); },
            group);
        T *= readlane_d(incl, 63);
        Acc = readlane_d(Ai, 63);
        Cr = readlane_d(H0, 63);
        Cg = readlane_d(H1, 63);
        Cb = readlane_d(H2, 63);
    }
    if (TGT && lane == 0) {
        a.bin_targets[b] = tmask;
        a.target_valid[b] = 1;
    }
}

template <int S, bool TGT, bool ADAPT>
__global__ __launch_bounds__(256) void k_bin_samples(const RenderArgs a) {
    const int64_t b = (int64_t)xcd_slice_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) +
                      (threadIdx.x >> 6);
    if (b < a.len) bin_samples_ray<S, TGT, ADAPT>(a, b);
}

__global__ __launch_bounds__(256) void k_bin_zero(uint32_t *p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Inclusive scan over the wave (u32).
__device__ __forceinline__ uint32_t wave_scan_u32(uint32_t v, int lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)v, off, 64);
        if (lane >= off) v += y;
    }
    return v;
}
// Exclusive scan over a block of kBinScanThreads threads; *total gets the
// block's sum.
__device__ __forceinline__ uint32_t block_scan_excl(uint32_t v, uint32_t *lds, uint32_t *total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    constexpr int kWaves = kBinScanThreads / 64;
    const uint32_t incl = wave_scan_u32(v, lane);
    if (lane == 63) lds[w] = incl;
    __syncthreads();
    if (w == 0) {
        const uint32_t x = lane < kWaves ? lds[lane] : 0u;
        const uint32_t xi = wave_scan_u32(x, lane);
        if (lane < kWaves) lds[kWaves + lane] = xi - x;
        if (lane == kWaves - 1) lds[2 * kWaves] = xi;
    }
    __syncthreads();
    const uint32_t r = lds[kWaves + w] + incl - v;
    *total = lds[2 * kWaves];
    __syncthreads();
    return r;
}

// Scan pass 1: the sum of each block's kBinScanBlockKeys counts.
__global__ __launch_bounds__(kBinScanThreads) void k_bin_scan_sum(const uint32_t *count,
                                                                  int64_t nkeys, uint32_t *blk) {
    __shared__ uint32_t lds[2 * (kBinScanThreads / 64) + 1];
    const int64_t k0 = (int64_t)blockIdx.x * kBinScanBlockKeys + (int64_t)threadIdx.x * kBinScanKeys;
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kBinScanKeys; i++)
        if (k0 + i < nkeys) s += count[k0 + i];
    uint32_t total;
    (void)block_scan_excl(s, lds, &total);
    if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

// Scan pass 2 (one block): exclusive prefix of the block sums, in place.
__global__ __launch_bounds__(kBinScanThreads) void k_bin_scan_blocks(uint32_t *blk, int64_t n) {
    __shared__ uint32_t lds[2 * (kBinScanThreads / 64) + 1];
    uint32_t carry = 0;
    for (int64_t base = 0; base < n; base += kBinScanThreads) {
        const int64_t i = base + threadIdx.x;
        const uint32_t v = i < n ? blk[i] : 0u;
        uint32_t total;
        const uint32_t e = block_scan_excl(v, lds, &total);
        if (i < n) blk[i] = carry + e;
        carry += total;
    }
}

// Key range j (0..26) of the samples that reach target brick (tx, ty, tz):
// source brick target - d for d in {0,1}^3 and the classes that contain d.
__device__ __forceinline__ bool bin_segment(const RenderArgs &a, int j, int tx, int ty, int tz,
                                            uint32_t &key) {
    int seen = 0;
    for (int dm = 0; dm < 8; dm++)
        for (int c = 0; c < 8; c++) {
            if ((c & dm) != dm) continue;
            if (seen++ != j) continue;
            const int sx = tx - (dm & 1), sy = ty - ((dm >> 1) & 1), sz = tz - (dm >> 2);
            if (sx < 0 || sy < 0 || sz < 0) return false;
            key = (((uint32_t)sz * (uint32_t)a.vby + (uint32_t)sy) * (uint32_t)a.vbx +
                   (uint32_t)sx) *
                      8u +
                  (uint32_t)c;
            return true;
        }
    return false;
}

// Scan pass 3: fill pointers (exclusive prefix of the counts); thread t of
// block k owns brick k * kBinScanThreads + t (its 8 keys) and appends it to
// the target list when any of its 27 key ranges holds a sample.
__global__ __launch_bounds__(kBinScanThreads) void k_bin_scan_final(const RenderArgs a,
                                                                    int64_t nkeys) {
    __shared__ uint32_t lds[2 * (kBinScanThreads / 64) + 1];
    const int64_t brick = (int64_t)blockIdx.x * kBinScanThreads + threadIdx.x;
    const int64_t k0 = brick * kBinScanKeys;
    uint32_t c[kBinScanKeys];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kBinScanKeys; i++) {
        c[i] = k0 + i < nkeys ? a.gb_count[k0 + i] : 0u;
        s += c[i];
    }
    uint32_t total;
    uint32_t run = a.gb_blk[blockIdx.x] + block_scan_excl(s, lds, &total);
#pragma unroll
    for (int i = 0; i < kBinScanKeys; i++) {
        if (k0 + i < nkeys) a.gb_fill[k0 + i] = run;
        run += c[i];
    }
    const int64_t nb = (int64_t)a.vbx * a.vby * a.vbz;
    bool has = false;
    if (brick < nb) {
        const int tx = (int)(brick % a.vbx), ty = (int)((brick / a.vbx) % a.vby),
                  tz = (int)(brick / ((int64_t)a.vbx * a.vby));
        has = s != 0u;
        for (int j = 8; j < 27 && !has; j++) {
            uint32_t key;
            if (bin_segment(a, j, tx, ty, tz, key)) has = a.gb_count[key] != 0u;
        }
    }
    const int lane = threadIdx.x & 63;
    const uint64_t hm = __ballot(has);
    if (hm != 0ull) {
        const int lead = __builtin_ctzll(hm);
        uint32_t base = 0;
        if (lane == lead)
            base = (uint32_t)atomicAdd(a.work + kTargetCounter, (unsigned long long)__popcll(hm));
        base = (uint32_t)__shfl((int)base, lead, 64);
        if (has) {
            const uint64_t below = lane == 0 ? 0ull : hm & (~0ull >> (64 - lane));
            a.gb_bricks[base + (uint32_t)__popcll(below)] = (uint32_t)brick;
        }
    }
}

// Stage 3: the log rows of one ray's samples go to their keys' ranges of
// bin_list (one returning atomic per distinct key per 64 samples).
__global__ __launch_bounds__(256) void k_bin_scatter(const RenderArgs a) {
    const int64_t b = (int64_t)xcd_slice_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) +
                      (threadIdx.x >> 6);
    if (b >= a.len) return;
    const int lane = threadIdx.x & 63;
    const int cnt = uniform_i(a.log_count[b]);
    for (int c0 = 0; c0 < cnt; c0 += 64) {
        const int m = min(64, cnt - c0);
        const bool act = lane < m;
        const size_t row = (size_t)b * (size_t)a.log_cap + (size_t)(c0 + (act ? lane : 0));
        const uint32_t key = bin_key(a, a.log_rec[row * 8 + 7]);
        uint32_t mybase = 0;
        uint64_t group;
        const int lead = for_each_key(
            act, key, lane, [&](uint32_t k0, uint32_t n) { mybase = atomicAdd(a.gb_fill + k0, n); },
            group);
        const uint32_t base = (uint32_t)__shfl((int)mybase, lead, 64);
        const uint64_t below = lane == 0 ? 0ull : ~0ull >> (64 - lane);
        const uint32_t rank = (uint32_t)__popcll(group & below);
        if (act) a.gb_list[base + rank] = (uint32_t)row;
    }
}

// Stage 4.  LDS floats per owned vertex: P rounded up to odd, so lanes
// walking vertex rows hit distinct banks.
template <int P>
constexpr int bin_row() {
    return P % 2 ? P : P + 1;
}
constexpr int kBinRound = 32;   // samples a wave stages per round
constexpr int kBinStage = 25;   // words per staged sample (odd stride)
constexpr int kBinWaves = 4;    // waves per workgroup of k_bin_accum
template <int S>
constexpr size_t bin_accum_lds() {
    return (size_t)kBinVerts * bin_row<1 + 3 * S>() * 4 +
           (size_t)kBinWaves * kBinRound * kBinStage * 4 + (size_t)kBinVerts * 4 + 64 * 4;
}

// One workgroup per target brick at a time (a grid of resident workgroups
// striding over the target list).  The brick's 512 owned vertices have one
// LDS row of P floats each; their records are looked up once.  A wave
// stages kBinRound samples (lane i: sample i of the round): the 8 corner
// weights, the channel scalars, the offset of corner 0 in the brick, which
// corners the brick owns, the SH basis of the sample's ray; then walks the
// staged samples two at a time (lane half = sample, lane & 31 = parameter)
// and adds w_j * (dL/dsigma_raw or dL/dcolour_ch * Y_k) to the owned corner
// rows with LDS float atomics.  At the end each touched record of the brick
// takes one float4 read-modify-write per 4 parameters (records 0: FP64
// atomics into the record-0 slot t mod len, folded in a fixed order
// afterwards).
template <int S, bool SPARSE>
__global__ __launch_bounds__(256) void k_bin_accum(const RenderArgs a) {
    constexpr int P = 1 + 3 * S, PS = bin_row<P>(), Q = (P + 3) / 4;
    extern __shared__ __align__(16) unsigned char bin_lds[];
    float *acc = reinterpret_cast<float *>(bin_lds);
    float *stage_all = acc + kBinVerts * PS;
    uint32_t *vrec = reinterpret_cast<uint32_t *>(stage_all + kBinWaves * kBinRound * kBinStage);
    uint32_t *seg_begin = vrec + kBinVerts;  // 32 words
    uint32_t *seg_pre = seg_begin + 32;      // 28 used
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    float *stage = stage_all + wv * kBinRound * kBinStage;
    const int h = lane >> 5, p = lane & 31;
    const int pch = p == 0 ? 0 : min(3, 1 + (p - 1) / S);
    const int kk = p == 0 ? 0 : (p - 1) % S;
    const bool live = p == 0 || (p < P && kk >= a.sh_first && kk < a.sh_active);
    const uint32_t ntg = (uint32_t)a.work[kTargetCounter];
    const int64_t cap = a.log_cap;
    const double inv_cap = 1.0 / (double)cap;
    for (uint32_t t = blockIdx.x; t < ntg; t += gridDim.x) {
        const uint32_t T = a.gb_bricks[t];
        const int tx = (int)(T % (uint32_t)a.vbx), ty = (int)((T / (uint32_t)a.vbx) % (uint32_t)a.vby),
                  tz = (int)(T / ((uint32_t)a.vbx * (uint32_t)a.vby));
        for (int i = tid; i < kBinVerts * PS; i += 256) acc[i] = 0.0f;
        for (int v = tid; v < kBinVerts; v += 256) {
            const int gx = tx * 8 + (v & 7), gy = ty * 8 + ((v >> 3) & 7), gz = tz * 8 + (v >> 6);
            uint32_t r = kNoRecord;
            if (gx < a.W && gy < a.H && gz < a.D) {
                r = scene_index(gx, gy, gz, a.W, a.H, a.D);
                if (SPARSE) r = min(a.io[r], (uint32_t)(a.n_vox - 1));
            }
            vrec[v] = r;
        }
        if (wv == 0) {
            uint32_t key = 0, beg = 0, len = 0;
            if (lane < 27 && bin_segment(a, lane, tx, ty, tz, key)) {
                len = a.gb_count[key];
                beg = a.gb_fill[key] - len;  // the scatter left the fill at the range's end
            }
            const uint32_t pre = wave_scan_u32(len, lane) - len;
            if (lane < 32) seg_begin[lane] = beg;
            if (lane < 28) seg_pre[lane] = pre;  // seg_pre[27]: the total
        }
        __syncthreads();
        const uint32_t total = seg_pre[27];
        for (uint32_t base = (uint32_t)wv * kBinRound; base < total;
             base += (uint32_t)(kBinWaves * kBinRound)) {
            const int n = (int)min((uint32_t)kBinRound, total - base);
            if (lane < n) {
                const uint32_t i = base + (uint32_t)lane;
                int s = 0;
#pragma unroll
                for (int j = 1; j < 27; j++) s = i >= seg_pre[j] ? j : s;
                const uint32_t row = a.gb_list[seg_begin[s] + (i - seg_pre[s])];
                const uint4 q0 = *reinterpret_cast<const uint4 *>(a.log_rec + (size_t)row * 8);
                const uint4 q1 = *reinterpret_cast<const uint4 *>(a.log_rec + (size_t)row * 8 + 4);
                int64_t eb = (int64_t)((double)row * inv_cap);
                if (eb * cap > (int64_t)row) eb--;
                else if ((eb + 1) * cap <= (int64_t)row) eb++;
                float y[S];
#pragma unroll
                for (int k = 0; k < S; k++) y[k] = a.gb_y[(size_t)eb * kBinYStride + k];
                int cx, cy, cz;
                unpack_cell(q1.w, cx, cy, cz);
                const float fx = __uint_as_float(q1.x), fy = __uint_as_float(q1.y),
                            fz = __uint_as_float(q1.z);
                const int lx = cx - tx * 8, ly = cy - ty * 8, lz = cz - tz * 8;
                const auto own01 = [](int l, int j) { return l + j >= 0 && l + j < 8; };
                float *st = stage + lane * kBinStage;
                const float x0 = 1.0f - fx, y0 = 1.0f - fy, z0 = 1.0f - fz;
                uint32_t own = 0u;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int jx = j & 1, jy = (j >> 1) & 1, jz = j >> 2;
                    const float wyz = (jy ? fy : y0) * (jz ? fz : z0);
                    st[j] = (jx ? fx : x0) * wyz;
                    if (own01(lx, jx) && own01(ly, jy) && own01(lz, jz)) own |= 1u << j;
                }
                st[8] = __uint_as_float(q0.x);
                st[9] = __uint_as_float(q0.y);
                st[10] = __uint_as_float(q0.z);
                st[11] = __uint_as_float(q0.w);
                st[12] = __int_as_float(lx + 8 * ly + 64 * lz);
                st[13] = __uint_as_float(own);
#pragma unroll
                for (int k = 0; k < S; k++) st[16 + k] = y[k];
            }
            // The staged rows are read by other lanes of this wave.
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (int s2 = 0; s2 < n; s2 += 2) {
                const int smp = s2 + h;
                if (smp < n && live) {
                    const float *st = stage + smp * kBinStage;
                    const float c = st[8 + pch];
                    const float val = p == 0 ? c : c * st[16 + kk];
                    const int v0 = __float_as_int(st[12]);
                    const uint32_t own = __float_as_uint(st[13]);
#pragma unroll
                    for (int j = 0; j < 8; j++)
                        if ((own >> j) & 1u) {
                            const int vj = v0 + (j & 1) + 8 * ((j >> 1) & 1) + 64 * (j >> 2);
                            __hip_atomic_fetch_add(acc + vj * PS + p, st[j] * val,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        }
                }
            }
            // Every lane's reads of the staged rows are done before the next
            // round overwrites them.
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        __syncthreads();
        // Write-back: item i = (vertex i / Q, float4 i % Q), kWb items of a
        // thread with their loads in flight together.
        constexpr int kItems = kBinVerts * Q;
        constexpr int kWb = 4;
        for (int i0 = 0; i0 < kItems; i0 += 256 * kWb) {
            float4 av[kWb], gv[kWb];
            uint32_t rr[kWb];
#pragma unroll
            for (int u = 0; u < kWb; u++) {
                const int idx = i0 + tid + 256 * u;
                rr[u] = kNoRecord;
                av[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                gv[u] = av[u];
                if (idx < kItems) {
                    const int v = idx / Q, q = idx - v * Q;
                    const float *ar = acc + v * PS + 4 * q;
                    av[u].x = ar[0];
                    av[u].y = 4 * q + 1 < P ? ar[1] : 0.0f;
                    av[u].z = 4 * q + 2 < P ? ar[2] : 0.0f;
                    av[u].w = 4 * q + 3 < P ? ar[3] : 0.0f;
                    const bool any = av[u].x != 0.0f || av[u].y != 0.0f || av[u].z != 0.0f ||
                                     av[u].w != 0.0f;
                    if (any) rr[u] = vrec[v];
                    if (rr[u] != kNoRecord && rr[u] != 0u)
                        gv[u] = *reinterpret_cast<const float4 *>(
                            a.grad + (size_t)rr[u] * (size_t)a.gstride + 4 * q);
                }
            }
#pragma unroll
            for (int u = 0; u < kWb; u++) {
                const int idx = i0 + tid + 256 * u;
                if (rr[u] == kNoRecord) continue;
                const int q = idx % Q;
                if (rr[u] == 0u) {
                    const size_t slot = (size_t)(t % (uint32_t)a.len);
                    const float e[4] = {av[u].x, av[u].y, av[u].z, av[u].w};
#pragma unroll
                    for (int c = 0; c < 4; c++)
                        if (4 * q + c < P && e[c] != 0.0f)
                            unsafeAtomicAdd(a.rec0 + (size_t)(4 * q + c) * (size_t)a.len + slot,
                                            (double)e[c]);
                } else {
                    float4 g = gv[u];
                    g.x += av[u].x;
                    g.y += av[u].y;
                    g.z += av[u].z;
                    g.w += av[u].w;
                    *reinterpret_cast<float4 *>(a.grad + (size_t)rr[u] * (size_t)a.gstride +
                                                4 * q) = g;
                }
            }
        }
        // The rows and the segment table are rebuilt for the next brick.
        __syncthreads();
    }
}
