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
//   * backward: lane (s, j) owns the gradient of corner j of sample s.
//     Consecutive samples in the same cell (same eight records) form a run;
//     each lane adds its 1 + 3S values into the LDS row (run, j), and a
//     finished run is flushed to the gradient planes with lanes over
//     parameters, two records per instruction.  The last run of a group stays
//     in LDS to merge with the next group, so a record gets one global atomic
//     per parameter per run instead of one per sample.
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
    v += shfl_xor64(v, 1);
    v += shfl_xor64(v, 2);
    v += shfl_xor64(v, 4);
    return v;
}
// Sum over the 8 slots of values replicated across each slot's lanes.
__device__ __forceinline__ double slot_sum(double v) {
    v += shfl_xor64(v, 8);
    v += shfl_xor64(v, 16);
    v += shfl_xor64(v, 32);
    return v;
}
// Inclusive scans over slots (stride 8 lanes) in FP64.
__device__ __forceinline__ double slot_scan_add(double v, int lane) {
    double y = shfl_up64(v, 8);
    if (lane >= 8) v += y;
    y = shfl_up64(v, 16);
    if (lane >= 16) v += y;
    y = shfl_up64(v, 32);
    if (lane >= 32) v += y;
    return v;
}
__device__ __forceinline__ double slot_scan_mul(double v, int lane) {
    double y = shfl_up64(v, 8);
    if (lane >= 8) v *= y;
    y = shfl_up64(v, 16);
    if (lane >= 16) v *= y;
    y = shfl_up64(v, 32);
    if (lane >= 32) v *= y;
    return v;
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

constexpr int kRuns = 9;  // LDS row blocks per wave: the kept run + 8 new

template <int S, bool SPARSE, int MODE, int STORE>
__global__ __launch_bounds__(256) void k_march8(const RenderArgs a) {
    constexpr int P = 1 + 3 * S;
    const int lane = threadIdx.x & 63;
    // Per-wave gradient rows [run][corner][32 floats] (backward only).
    constexpr int kRowFloats = (MODE == kBwd || MODE == kFused) ? kRuns * 8 * 32 : 1;
    __shared__ float lds_rows[4][kRowFloats];
    float *rows = lds_rows[threadIdx.x >> 6];
    if (MODE == kBwd || MODE == kFused) {
        for (int i = lane; i < kRowFloats; i += 64) rows[i] = 0.0f;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    const int slot = lane >> 3, corner = lane & 7;
    const int64_t b =
        (int64_t)xcd_slice_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
    if (b >= a.len) return;
    const int ray = uniform_i(a.sample_index[a.start + b]);
    const bool do_fwd = MODE == kFwd || MODE == kFused || MODE == kTrace;
    const bool do_bwd = MODE == kBwd || MODE == kFused;
    if (ray < 0 || (int64_t)ray >= a.rays.n) {
        if (do_bwd && lane == 0) a.losses[b] = 0.0;
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

    if (tnear > tfar) {
        if ((MODE == kFwd || MODE == kFused || MODE == kTrace) && lane == 0) {
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
        if (do_bwd && lane == 0) a.losses[b] = 0.0;
        return;
    }

    float Y[S];
#pragma unroll
    for (int k = 0; k < S; k++) Y[k] = sh_basis(k, fdx, fdy, fdz);

    // Backward coefficients (set after the forward / from the ray planes).
    double C0 = 0.0, C1 = 0.0, C2 = 0.0, A = 0.0;
    double r0 = 0.0, r1 = 0.0, r2 = 0.0, mterm = 0.0;

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
            r0 = C0 - (double)a.rays.gt_r[ray] / 255.0;
            r1 = C1 - (double)a.rays.gt_g[ray] / 255.0;
            r2 = C2 - (double)a.rays.gt_b[ray] / 255.0;
            const double loss = (fabs(r0) + fabs(r1) + fabs(r2)) * 255.0 / 3.0;
            if (isnan(loss)) {
                if (lane == 0) a.losses[b] = 255.0;
                return;
            }
            if (lane == 0) a.losses[b] = loss;
            const double hd = (double)a.huber;
            const double hi = hd - 0.5 * hd * hd, lo = 0.5 * hd * hd - hd;
            r0 = r0 > hd ? hi : (r0 < -hd ? lo : r0);
            r1 = r1 > hd ? hi : (r1 < -hd ? lo : r1);
            r2 = r2 > hd ? hi : (r2 < -hd ? lo : r2);
            const int gm = a.rays.gt_mask[ray];
            mterm = 0.0;
            if (gm != 0 && A < 1.0) mterm = ((double)gm / 255.0 - A) * a.mask_coef;
            double wpre = A > 1e-6 ? A : 1e-6;
            wpre = wpre < 1.0 - 1e-6 ? wpre : 1.0 - 1e-6;
            mterm += (log(wpre) - log(1.0 - wpre)) * 1e-3;
        }

        // Running totals before the current group (uniform, FP64).
        double T = 1.0, Acc = 0.0, Cr = 0.0, Cg = 0.0, Cb = 0.0, sig = 0.0;
        // Backward gradient merge state: whether a run is kept, its LDS row
        // block and (lane l) the record of its corner l & 7.
        bool keep_valid = false;
        int keep_base = 0;
        uint32_t keep_rec = 0u;
        uint32_t hcount = 0, hash = 2166136261u;
        int k0 = 0;
        uint64_t pending = 0;  // occupied positions of [k0, k0 + 64) not yet done
        bool more = true;      // positions beyond k0 + 64 may remain
        while (true) {
            if (pending == 0) {
                if (!more) break;
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
                    if (a.bricks != nullptr) {
                        const uint32_t br = (uint32_t)(cz >> 3) * (uint32_t)a.nbx * (uint32_t)a.nby +
                                            (uint32_t)(cy >> 3) * (uint32_t)a.nbx + (uint32_t)(cx >> 3);
                        occ = (a.bricks[br >> 5] >> (br & 31)) & 1u;
                    }
                    if (occ) {
                        const uint32_t cell = (uint32_t)cz * (uint32_t)(a.W - 1) * (uint32_t)(a.H - 1) +
                                              (uint32_t)cy * (uint32_t)(a.W - 1) + (uint32_t)cx;
                        occ = (a.occ[cell >> 5] >> (cell & 31)) & 1u;
                    }
                }
                pending = __ballot(occ);
                more = __ballot(valid) == ~0ull;
                if (pending == 0) {
                    k0 += 64;
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
            const double incl = slot_scan_mul(keep, lane);
            double excl = shfl_up64(incl, 8);
            if (lane < 8) excl = 1.0;
            const double Ts = T * excl;
            const double ws = Ts * (double)alpha;
            const double total_keep = __shfl(incl, 63, 64);

            if (!bwd) {
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
                const double Ai = Acc + slot_scan_add(ws, lane);
                const double H0 = Cr + slot_scan_add(ws * (double)col0, lane);
                const double H1 = Cg + slot_scan_add(ws * (double)col1, lane);
                const double H2 = Cb + slot_scan_add(ws * (double)col2, lane);
                const double Tn = Ts * keep;
                const double kc = (double)0.6667f;
                const float g0 = (float)(kc * ws * r0 * (double)(0.1f * col0 * (1.0f - col0)));
                const float g1 = (float)(kc * ws * r1 * (double)(0.1f * col1 * (1.0f - col1)));
                const float g2 = (float)(kc * ws * r2 * (double)(0.1f * col2 * (1.0f - col2)));
                const double ds = (double)dsoftplus_shift<float>(sraw);
                double gs = r0 * (Tn * (double)col0 - (C0 - H0)) + r1 * (Tn * (double)col1 - (C1 - H1)) +
                            r2 * (Tn * (double)col2 - (C2 - H2));
                gs *= 0.6667 * step * ds;
                gs += step * ds * (A - Ai - Tn) * mterm;
                // Runs of consecutive slots whose eight corner records are
                // identical.  Run 0 continues the run kept from the previous
                // group; a slot whose records differ from its predecessor's
                // (or from the kept ones) opens a new run.  The masks, run
                // numbers and run starts are wave-uniform.
                const uint32_t up_rec = (uint32_t)__shfl_up((int)rec, 8, 64);
                const uint32_t prev_rec = slot == 0 ? keep_rec : up_rec;
                const bool differs = active && (rec != prev_rec || (slot == 0 && !keep_valid));
                const uint64_t dmask = __ballot(differs);
                uint32_t starts = 0u;  // bit s: slot s opens a run
#pragma unroll
                for (int s2 = 0; s2 < 8; s2++)
                    starts |= ((dmask >> (8 * s2)) & 0xFFull) != 0 ? (1u << s2) : 0u;
                const int R = 1 + __builtin_popcount(starts);
                const int my_run = __builtin_popcount(starts & ((2u << slot) - 1u));
                // Merge this lane's gradient into its (run, corner) row.
                if (active) {
                    float *row = rows + ((keep_base + my_run) % kRuns) * (8 * 32) + corner * 32;
                    atomicAdd(row, (float)gs * w);
                    const float gch[3] = {g0 * w, g1 * w, g2 * w};
#pragma unroll
                    for (int ch = 0; ch < 3; ch++)
#pragma unroll
                        for (int k = 0; k < S; k++) atomicAdd(row + 1 + ch * S + k, gch[ch] * Y[k]);
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                // Flush runs 0 .. R-2 (run R-1 stays for the next group):
                // two rows per instruction, lanes over parameters.  Run r >= 1
                // starts at the r-th set bit of starts.
                uint32_t rest = starts;
                for (int r = 0; r < R - 1; r++) {
                    int src = -1;
                    if (r > 0) {
                        src = 8 * __builtin_ctz(rest);
                        rest &= rest - 1u;
                    }
                    if (r == 0 && !keep_valid) continue;  // nothing kept yet
#pragma unroll
                    for (int half = 0; half < 4; half++) {
                        const int j = 2 * half + (lane >> 5);
                        const int pp = lane & 31;
                        const uint32_t rj =
                            src < 0 ? (uint32_t)__shfl((int)keep_rec, j, 64)
                                    : (uint32_t)__shfl((int)rec, src + j, 64);
                        float *row = rows + ((keep_base + r) % kRuns) * (8 * 32) + j * 32;
                        if (pp < P) {
                            const float v = row[pp];
                            row[pp] = 0.0f;
                            unsafeAtomicAdd(a.grad + (size_t)rj * (size_t)a.gstride + pp, v);
                        }
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                // The last run becomes the kept one; lane l keeps the record
                // of corner l & 7.
                if (R > 1) {
                    const int fs = 31 - __builtin_clz(starts);
                    keep_rec = (uint32_t)__shfl((int)rec, 8 * fs + (lane & 7), 64);
                    keep_valid = true;
                    keep_base = (keep_base + R - 1) % kRuns;
                }
                Acc = __shfl(Ai, 63, 64);
                Cr = __shfl(H0, 63, 64);
                Cg = __shfl(H1, 63, 64);
                Cb = __shfl(H2, 63, 64);
                T *= total_keep;
            }
        }

        if (bwd && keep_valid) {
#pragma unroll
            for (int half = 0; half < 4; half++) {
                const int j = 2 * half + (lane >> 5);
                const int pp = lane & 31;
                const uint32_t rj = (uint32_t)__shfl((int)keep_rec, j, 64);
                float *row = rows + keep_base * (8 * 32) + j * 32;
                if (pp < P) {
                    const float v = row[pp];
                    row[pp] = 0.0f;
                    unsafeAtomicAdd(a.grad + (size_t)rj * (size_t)a.gstride + pp, v);
                }
            }
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
        }
    }
}
