// Forward marcher of the reference's adaptive rule with eight rays per
// wavefront (included by render.hip after march8.hpp, inside its anonymous
// namespace).
//
// The rule (volumeRender_RGB_train, RenderUtils.cu:1674-1707): t starts at
// t_near; an evaluated sample moves t by |step - w| (w = T alpha of that
// sample), a skipped one (empty cell) by step.  Each position depends on the
// previous sample's weight, so one ray's samples are sequential; the work
// mapping therefore puts eight rays in a wave instead of eight samples of
// one ray: lane = (ray slot g, corner j), g = lane / 8.  Per iteration the
// 8 lanes of a slot test the ray's next 8 positions at once (lane j reaches
// its position from t by j additions of step, the sequence the rule takes
// while it skips), and the slot moves to the first occupied one, or past
// all 8; positions in the cell of the ray's last sample need no occupancy
// lookup.  When a sample enters a new cell, lane (g, j) loads corner j's
// record with 16-byte loads and keeps its density and its three colour sums
// sum_k Y_k theta_k while the ray stays in the cell; each sample then weighs
// them trilinearly and three xor shuffles sum the 8 corners.  Weights, interpolation, activations and the
// compositing scalars are FP64, as in the one-sample kernel k_render, so
// the positions follow the FP64 oracle's.
//
// Modes: kFwd (the ray planes), kTrace (also the per-entry sample count and
// hash of pxl_render_fwd_trace) and kLog (also the sample log of the
// two-kernel step: per sample the 8 corner records and 8 FP64 words: raw
// density, three raw colours, t; read back by log_grad_ray<..., ADAPT>).

// Lanes per ray slot of k_adapt8 (4: sixteen rays per wave, two corners
// per lane; 8: eight rays, one corner per lane) and probe positions per lane
// (a probe tests kAdaptLanes * kProbePerLane consecutive positions, at most
// 32, with one round of occupancy loads).  Both are issue-bound trade-offs:
// on the 256^3 caged grid with 65536-ray batches (half the rays evaluate no
// sample, the longest ~400) the forward + gradient of a batch took 2.36 ms
// with 8 lanes and 1 position, 2.54 ms with 2 positions and 3.22 ms with 4;
// the train step's render stage took 2.78 ms with 4 lanes against 2.40 ms
// with 8: the extra positions' and corners' instructions cost more than the
// round trips and lanes they save.
constexpr int kAdaptLanes = 8;
constexpr int kProbePerLane = 1;

// Sum over the L lanes of a ray slot.
template <int L>
__device__ __forceinline__ double slot_sum_d(double v) {
    v += dpp_d<kDppXor1>(v);
    v += dpp_d<kDppXor2>(v);
    if (L == 8) v += dpp_d<kDppMirror8>(v);
    return v;
}

// Results of a slot's finished batch entry b (lane j == 0 writes): kind 1
// an out-of-range ray index, 2 a ray that missed the grid box, 0 a marched
// ray (colour totals, trace, and for the logged step the loss and log count,
// or the redo list when the ray outgrew the log).
template <int MODE>
__device__ __forceinline__ void adapt8_finish(const RenderArgs &a, int64_t b, int ray, int kind,
                                              double C0, double C1, double C2, double A,
                                              double sig, int logged, uint32_t hcount,
                                              uint32_t hash) {
    const bool logging = MODE == kLog;
    if (kind != 0) {
        if (kind == 2) {
            a.rays.acc_r[ray] = 0.0f;
            a.rays.acc_g[ray] = 0.0f;
            a.rays.acc_b[ray] = 0.0f;
            a.rays.acc_w[ray] = 0.0f;
            a.rays.acc_density[ray] = 0.0f;
        }
        if (MODE == kTrace) {
            a.trace[2 * b] = 0u;
            a.trace[2 * b + 1] = 2166136261u;
        }
        if (logging) {
            a.losses[b] = 0.0;
            a.log_count[b] = 0;
        }
        return;
    }
    if (a.white_bg) {
        C0 += 1.0 - A;
        C1 += 1.0 - A;
        C2 += 1.0 - A;
    }
    a.rays.acc_r[ray] = (float)C0;
    a.rays.acc_g[ray] = (float)C1;
    a.rays.acc_b[ray] = (float)C2;
    a.rays.acc_w[ray] = (float)A;
    a.rays.acc_density[ray] = (float)sig;
    if (MODE == kTrace) {
        a.trace[2 * b] = hcount;
        a.trace[2 * b + 1] = hash;
    }
    if (logging) {
        if (logged <= a.log_cap) {
            RayResid rr;
            const bool ok = ray_residuals(a, ray, C0, C1, C2, A, rr);
            a.losses[b] = ok ? rr.loss : 255.0;
            a.log_count[b] = ok ? logged : 0;
            a.log_tot[4 * b + 0] = C0;
            a.log_tot[4 * b + 1] = C1;
            a.log_tot[4 * b + 2] = C2;
            a.log_tot[4 * b + 3] = A;
        } else {
            // Longer than the log: the one-sample fused kernel redoes it.
            a.log_count[b] = -1;
            a.redo_list[atomicAdd(a.work + kRedoCounter, 1ull)] = (int32_t)b;
        }
    }
}

// Slots and work: the sample counts of one batch's rays spread widely (at
// 256^3 a ray evaluates 24 samples on average, the longest of 8 rays 83), so
// a slot whose ray has ended takes the next batch entry instead of idling
// until the wave's longest ray is done.  With claim counters (a.work[0..7],
// zeroed before the launch: the logged step) a wave claims entries for its
// free slots with one atomic add on the counter of its block's slice of the
// batch (block b runs on XCD b mod 8), moving to the next slice when one is
// used up; without them (forward and trace launches) each slot marches the
// one entry of its position in the grid.
//
// Two waves per SIMD (at most 256 VGPRs): left to itself the compiler
// overlaps the two corner record loads of a lane at 272 registers, one wave.
template <int S, bool SPARSE, int MODE, int L = kAdaptLanes, int UP = kProbePerLane>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_adapt8(
    const RenderArgs a) {
    constexpr int P = 1 + 3 * S;
    constexpr int U = 8 / L;     // corners per lane
    constexpr int NS = 64 / L;   // ray slots per wave
    constexpr int NP = L * UP;   // positions per probe
    constexpr uint32_t LM = (1u << L) - 1u;
    static_assert(L == 4 || L == 8, "four or eight lanes per ray slot");
    static_assert(UP >= 1 && NP <= 32, "at most 32 positions per probe");
    constexpr uint32_t kNone = (uint32_t)NP;
    const bool logging = MODE == kLog;
    const int lane = threadIdx.x & 63;
    const int g = lane / L, j = lane % L;
    const int gbase = L * g;
    const int64_t wave =
        (int64_t)xcd_slice_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const bool claiming = a.work != nullptr;
    int slice = (int)(blockIdx.x & 7u), visited = 0;
    bool first = true;
    const double step = (double)a.step;

    // The slot's batch entry and its ray.
    bool live = false, exhausted = false;
    int64_t b = 0;
    int ray = 0, kind = 0;
    // The ray's planes stay FP32 in registers (as stored) and widen where
    // used; the slot's SH basis (FP64) sits in LDS, read on a new cell only.
    float ox = 0.0f, oy = 0.0f, oz = 0.0f, dx = 0.0f, dy = 0.0f, dz = 0.0f, tnear = 0.0f,
          tfar = 0.0f;
    __shared__ double ylds[4][NS][S];
    double *const Y = ylds[threadIdx.x >> 6][g];
    double t = 0.0, T = 1.0, A = 0.0, C0 = 0.0, C1 = 0.0, C2 = 0.0, sig = 0.0;
    int k = 0, logged = 0;
    uint32_t hcount = 0, hash = 2166136261u;
    // The cell of the ray's last evaluated sample, this lane's corner
    // records (corners j + L u) and their densities and three SH colour sums
    // (the colour sums depend on the ray's direction: a new ray starts with
    // no cell).
    int cbx = INT32_MIN, cby = 0, cbz = 0;
    uint32_t crec[U];
    double q0[U], q1[U], q2[U], q3[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        crec[u] = 0u;
        q0[u] = q1[u] = q2[u] = q3[u] = 0.0;
    }
    // A sample found by the last probe and not yet evaluated: its t and cell.
    bool pending = false;
    double t_at = 0.0;
    int sbx = 0, sby = 0, sbz = 0;
    // Per lane: the last brick this lane found empty (sparse grids).
    uint32_t empty_brick = 0xFFFFFFFFu;
    while (true) {
        if (live && !pending && !(t < tfar && k < kMaxSteps)) {
            if (j == 0)
                adapt8_finish<MODE>(a, b, ray, kind, C0, C1, C2, A, sig, logged, hcount, hash);
            live = false;
        }
        // Free slots take the next batch entries (rank: the slot's place
        // among the wave's free slots).
        const uint64_t freem = __ballot(!live && !exhausted && j == 0);
        if (freem != 0ull) {
            const bool is_free = (freem >> gbase) & 1ull;
            const int rank = __builtin_popcountll(freem & ((1ull << gbase) - 1ull));
            int64_t nb = -1;
            if (claiming) {
                const int nfree = __builtin_popcountll(freem);
                int done = 0;
                while (done < nfree && visited < 8) {
                    const int64_t lo = a.len * slice / 8, hi = a.len * (slice + 1) / 8;
                    unsigned long long got = 0ull;
                    if (lane == 0) got = atomicAdd(a.work + slice, (unsigned long long)(nfree - done));
                    const uint32_t glo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)got);
                    const uint32_t ghi =
                        (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(got >> 32));
                    const int64_t base = lo + (int64_t)(((uint64_t)ghi << 32) | glo);
                    const int64_t room = hi - base;
                    const int avail =
                        room <= 0 ? 0 : (room < (int64_t)(nfree - done) ? (int)room : nfree - done);
                    if (is_free && rank >= done && rank < done + avail) nb = base + (rank - done);
                    done += avail;
                    if (done < nfree) {
                        slice = (slice + 1) & 7;
                        visited++;
                    }
                }
            } else if (first && wave * NS + g < a.len) {
                nb = wave * NS + g;
            }
            if (is_free) {
                if (nb < 0) {
                    exhausted = true;
                } else {
                    b = nb;
                    live = true;
                    ray = a.sample_index[a.start + b];
                    if (logging) {
                        // The entry's 32 record-0 slots start at zero.
                        for (int q = j; q < 32; q += L)
                            a.rec0[(size_t)q * (size_t)a.len + (size_t)b] = 0.0;
                    }
                    kind = 0;
                    tnear = 0.0f;
                    tfar = -1.0f;
                    if (ray < 0 || (int64_t)ray >= a.rays.n) {
                        kind = 1;
                    } else {
                        ox = a.rays.ox[ray];
                        oy = a.rays.oy[ray];
                        oz = a.rays.oz[ray];
                        dx = a.rays.dx[ray];
                        dy = a.rays.dy[ray];
                        dz = a.rays.dz[ray];
                        tnear = a.rays.near_t[ray];
                        tfar = a.rays.far_t[ray];
                        if (tnear > tfar) kind = 2;
                        for (int q = j; q < S; q += L)
                            Y[q] = q < a.sh_active ? sh_basis_d(q, dx, dy, dz) : 0.0;
                    }
                    cbx = INT32_MIN;
                    t = tnear;
                    T = 1.0;
                    A = C0 = C1 = C2 = sig = 0.0;
                    k = logged = 0;
                    hcount = 0u;
                    hash = 2166136261u;
                }
            }
        }
        first = false;
        // The basis written above is read by the slot's other lanes (LDS
        // operations of a wave complete in order; this keeps the compiler
        // from moving them).
        __builtin_amdgcn_wave_barrier();
        if (__ballot(live) == 0ull) break;
        // One probe per iteration for the slots without a pending sample, so
        // that a slot crossing empty space does not hold back the slots that
        // evaluate: the slot tests the ray's next NP positions, lane j the
        // positions j + L u, position p reached from t by p additions of
        // step, exactly the sequence the rule produces while it skips.
        const bool searching = live && !pending && t < tfar && k < kMaxSteps;
        if (__ballot(searching) != 0ull) {
            double tp[UP];
            {
                double tj = t;
#pragma unroll
                for (int q = 0; q < L - 1; q++)
                    if (q < j) tj += step;
                tp[0] = tj;
#pragma unroll
                for (int u = 1; u < UP; u++) {
#pragma unroll
                    for (int q = 0; q < L; q++) tj += step;
                    tp[u] = tj;
                }
            }
            // One step past the last position (meaningful on lane L - 1).
            const double tn = tp[UP - 1] + step;
            bool valid[UP];
            int qbx[UP], qby[UP], qbz[UP];
            uint32_t inmask = 0u, vmask = 0u;
#pragma unroll
            for (int u = 0; u < UP; u++) {
                const int p = j + L * u;
                valid[u] = searching && tp[u] < tfar && k + p < kMaxSteps;
                const double qx = fma(tp[u], (double)dx, (double)ox),
                             qy = fma(tp[u], (double)dy, (double)oy),
                             qz = fma(tp[u], (double)dz, (double)oz);
                qbx[u] = base_of(qx);
                qby[u] = base_of(qy);
                qbz[u] = base_of(qz);
                // A position in the cell of the ray's last sample is occupied.
                const bool incell = valid[u] && qbx[u] == cbx && qby[u] == cby && qbz[u] == cbz;
                inmask |= ((uint32_t)(__ballot(incell) >> gbase) & LM) << (L * u);
                vmask |= ((uint32_t)(__ballot(valid[u]) >> gbase) & LM) << (L * u);
            }
            // Only the positions in front of the first in-cell one look the
            // occupancy up, the brick word and the cell word together.
            const int fin = inmask ? __builtin_ctz(inmask) : (int)kNone;
            uint32_t omask = 0u;
#pragma unroll
            for (int u = 0; u < UP; u++) {
                const int p = j + L * u;
                bool occ = valid[u] && p <= fin;
                if (SPARSE && valid[u] && p < fin) {
                    const int cx = clampi(qbx[u], 0, a.W - 2), cy = clampi(qby[u], 0, a.H - 2),
                              cz = clampi(qbz[u], 0, a.D - 2);
                    const uint32_t cell = (uint32_t)cz * (uint32_t)(a.W - 1) * (uint32_t)(a.H - 1) +
                                          (uint32_t)cy * (uint32_t)(a.W - 1) + (uint32_t)cx;
                    uint32_t br = 0u;
                    bool skip = false;
                    if (a.bricks != nullptr) {
                        br = (uint32_t)(cz >> 3) * (uint32_t)a.nbx * (uint32_t)a.nby +
                             (uint32_t)(cy >> 3) * (uint32_t)a.nbx + (uint32_t)(cx >> 3);
                        skip = br == empty_brick;
                    }
                    if (skip) {
                        occ = false;
                    } else {
                        const uint32_t ow = a.occ[cell >> 5];
                        const uint32_t bw = a.bricks != nullptr ? a.bricks[br >> 5] : 0xFFFFFFFFu;
                        const bool bocc = (bw >> (br & 31)) & 1u;
                        if (!bocc) empty_brick = br;
                        occ = bocc && ((ow >> (cell & 31)) & 1u);
                    }
                }
                omask |= ((uint32_t)(__ballot(occ) >> gbase) & LM) << (L * u);
            }
            // First occupied position of the slot (NP: none), the valid ones
            // (a prefix: positions grow), the position whose t and cell the
            // slot takes: the sample, else the first invalid position, else
            // one step past the last position.
            const int f = omask ? __builtin_ctz(omask) : NP;
            const int nvalid = __builtin_popcount(vmask);
            const int sp = f < NP ? f : (nvalid < NP ? nvalid : NP - 1);
            double tc = tp[0];
            int xc = qbx[0], yc = qby[0], zc = qbz[0];
#pragma unroll
            for (int u = 1; u < UP; u++) {
                if (sp / L == u) {
                    tc = tp[u];
                    xc = qbx[u];
                    yc = qby[u];
                    zc = qbz[u];
                }
            }
            if (f == NP && nvalid == NP) tc = tn;
            const int src = gbase + sp % L;
            const double tsrc = __shfl(tc, src, 64);
            const int xb = __shfl(xc, src, 64), yb = __shfl(yc, src, 64), zb = __shfl(zc, src, 64);
            if (searching) {
                if (f < NP) {
                    // The sample: f skipped positions before it.
                    pending = true;
                    t_at = tsrc;
                    sbx = xb;
                    sby = yb;
                    sbz = zb;
                    k += f;
                } else {
                    // Every valid position skipped: on past them.
                    t = tsrc;
                    k += nvalid;
                }
            }
        }
        const bool eval = pending;
        pending = false;
        if (__ballot(eval) == 0ull) continue;
        if (eval && (sbx != cbx || sby != cby || sbz != cbz)) {
            // A new cell: this lane's corner records, their densities and
            // their three colour sums sum_k Y_k theta_k.
            cbx = sbx;
            cby = sby;
            cbz = sbz;
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int c = j + L * u;
                uint32_t r = scene_index(sbx + (c & 1), sby + ((c >> 1) & 1), sbz + (c >> 2), a.W,
                                         a.H, a.D);
                if (SPARSE) r = min(a.io[r], (uint32_t)(a.n_vox - 1));
                crec[u] = r;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                float th[P];
                load_record<P, PXL_STORE_F32>(a.params, crec[u], a.stride, th);
                double c[3] = {0.0, 0.0, 0.0};
#pragma unroll
                for (int q = 0; q < S; q++) {
                    const double y = Y[q];
#pragma unroll
                    for (int ch = 0; ch < 3; ch++)
                        c[ch] = fma(y, (double)th[1 + ch * S + q], c[ch]);
                }
                q0[u] = (double)th[0];
                q1[u] = c[0];
                q2[u] = c[1];
                q3[u] = c[2];
            }
        }
        // This lane's corners' share of the sample (zero for lanes without one).
        double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
        if (eval) {
            const double px = fma(t_at, (double)dx, (double)ox),
                         py = fma(t_at, (double)dy, (double)oy),
                         pz = fma(t_at, (double)dz, (double)oz);
            const double fx = px - (double)sbx, fy = py - (double)sby, fz = pz - (double)sbz;
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int c = j + L * u;
                const double w = ((c & 1) ? fx : 1.0 - fx) * (((c >> 1) & 1) ? fy : 1.0 - fy) *
                                 ((c >> 2) ? fz : 1.0 - fz);
                p0 += w * q0[u];
                p1 += w * q1[u];
                p2 += w * q2[u];
                p3 += w * q3[u];
            }
        }
        const double sraw = slot_sum_d<L>(p0);
        const double cr0 = slot_sum_d<L>(p1), cr1 = slot_sum_d<L>(p2), cr2 = slot_sum_d<L>(p3);
        // Activations spread over the slot's lanes, one exp each: lane 0 the
        // density, softplus(sraw - 13) = [x > 30: x] + log1p(exp(+-x)), lanes
        // 1-3 the colours, 1 / (1 + exp(-0.1 raw)) (the same operations as
        // softplus_shift<double> and sigmoid01<double>).
        const double xs = sraw - 13.0;
        const double craw = j == 1 ? cr0 : (j == 2 ? cr1 : cr2);
        const double e = exp(j == 0 ? (xs > 30.0 ? -xs : xs) : -0.1 * craw);
        const double act_v = j == 0 ? (xs > 30.0 ? xs + log1p(e) : log1p(e)) : 1.0 / (1.0 + e);
        const double sigma_s = __shfl(act_v, gbase, 64);
        const double col0_s = __shfl(act_v, gbase + 1, 64), col1_s = __shfl(act_v, gbase + 2, 64),
                     col2_s = __shfl(act_v, gbase + 3, 64);
        if (eval) {
            if (MODE == kTrace) {
                hcount++;
                hash = fnv_mix(hash, (uint32_t)k);
                hash = fnv_mix(hash, (uint32_t)sbx);
                hash = fnv_mix(hash, (uint32_t)sby);
                hash = fnv_mix(hash, (uint32_t)sbz);
            }
            const double sigma = sigma_s;
            const double col0 = col0_s, col1 = col1_s, col2 = col2_s;
            const double alpha = one_minus_exp_neg<double>(sigma * step);
            const double w = T * alpha;
            if (logging && logged < a.log_cap) {
                const size_t row = (size_t)b * (size_t)a.log_cap + (size_t)logged;
#pragma unroll
                for (int u = 0; u < U; u++) a.log_rec[row * 8 + j + L * u] = crec[u];
                for (int v = j; v < 5; v += L) {
                    const double val =
                        v == 0 ? sraw : (v == 1 ? cr0 : (v == 2 ? cr1 : (v == 3 ? cr2 : t_at)));
                    reinterpret_cast<double *>(a.log_val)[row * 8 + v] = val;
                }
            }
            logged++;
            A += w;
            C0 += w * col0;
            C1 += w * col1;
            C2 += w * col2;
            sig += sigma;
            T = T * (1.0 - alpha);
            t = t_at + fabs(step - w);
            ++k;
        }
    }
}
