// Forward marcher of the reference's adaptive rule with eight rays per
// wavefront (included by render.hip after march8.hpp, inside its anonymous
// namespace).
//
// The rule (volumeRender_RGB_train, RenderUtils.cu:1674-1707): t starts at
// t_near; an evaluated sample moves t by |step - w| (w = T alpha of that
// sample), a skipped one (empty cell) by step.  Each position depends on the
// previous sample's weight, so one ray's samples are sequential; the work
// mapping therefore puts eight rays in a wave instead of eight samples of
// one ray: lane = (ray slot g, corner j), g = lane / 8.  Per iteration each
// slot places its ray's next position, tests the cell (the occupancy and
// the corner records are kept while the ray stays in one cell), and for an
// evaluated sample lane (g, j) loads corner j's record with 16-byte loads
// and forms its share of the raw density and raw colours; three xor
// shuffles sum the 8 corners.  Weights, interpolation, activations and the
// compositing scalars are FP64, as in the one-sample kernel k_render, so
// the positions follow the FP64 oracle's.
//
// Modes: kFwd (the ray planes), kTrace (also the per-entry sample count and
// hash of pxl_render_fwd_trace) and kLog (also the sample log of the
// two-kernel step: per sample the 8 corner records and 8 FP64 words: raw
// density, three raw colours, t; read back by log_grad_ray<..., ADAPT>).

__device__ __forceinline__ double corner_sum_d(double v) {
    v += shfl_xor64(v, 1);
    v += shfl_xor64(v, 2);
    v += shfl_xor64(v, 4);
    return v;
}

template <int S, bool SPARSE, int MODE>
__global__ __launch_bounds__(256) void k_adapt8(const RenderArgs a) {
    constexpr int P = 1 + 3 * S;
    const bool logging = MODE == kLog;
    const int lane = threadIdx.x & 63;
    const int g = lane >> 3, j = lane & 7;
    const int jx = j & 1, jy = (j >> 1) & 1, jz = j >> 2;
    const int64_t wave =
        (int64_t)xcd_slice_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t b = wave * 8 + g;
    bool live = b < a.len;
    const int ray = live ? a.sample_index[a.start + b] : -1;
    if (logging && live) {
        // The entry's record-0 slots start at zero (4 per lane of the slot).
#pragma unroll
        for (int q = 0; q < 4; q++) a.rec0[(size_t)(4 * j + q) * (size_t)a.len + (size_t)b] = 0.0;
    }
    if (live && (ray < 0 || (int64_t)ray >= a.rays.n)) {
        if (j == 0) {
            if (logging) {
                a.losses[b] = 0.0;
                a.log_count[b] = 0;
            }
            if (MODE == kTrace) {
                a.trace[2 * b] = 0u;
                a.trace[2 * b + 1] = 2166136261u;
            }
        }
        live = false;
    }
    double ox = 0.0, oy = 0.0, oz = 0.0, dx = 0.0, dy = 0.0, dz = 0.0, tnear = 0.0, tfar = 0.0;
    float fdx = 0.0f, fdy = 0.0f, fdz = 0.0f;
    if (live) {
        ox = a.rays.ox[ray];
        oy = a.rays.oy[ray];
        oz = a.rays.oz[ray];
        fdx = a.rays.dx[ray];
        fdy = a.rays.dy[ray];
        fdz = a.rays.dz[ray];
        dx = fdx;
        dy = fdy;
        dz = fdz;
        tnear = a.rays.near_t[ray];
        tfar = a.rays.far_t[ray];
    }
    const double step = (double)a.step;
    double Y[S];
#pragma unroll
    for (int k = 0; k < S; k++) Y[k] = k < a.sh_active ? sh_basis_d(k, fdx, fdy, fdz) : 0.0;

    double t = tnear, T = 1.0, A = 0.0, C0 = 0.0, C1 = 0.0, C2 = 0.0, sig = 0.0;
    int k = 0, logged = 0;
    uint32_t hcount = 0, hash = 2166136261u;
    // The cell of the last position and what it holds: occupied, and this
    // lane's corner record.
    int cbx = INT32_MIN, cby = 0, cbz = 0;
    bool cocc = false;
    uint32_t crec = 0u;
    // Last brick found empty (sparse grids), so positions in it skip the
    // cell lookups.
    uint32_t empty_brick = 0xFFFFFFFFu;
    while (true) {
        const bool act = live && t < tfar && k < kMaxSteps;
        if (__ballot(act) == 0ull) break;
        const double px = fma(t, dx, ox), py = fma(t, dy, oy), pz = fma(t, dz, oz);
        const int bx = base_of(px), by = base_of(py), bz = base_of(pz);
        if (act && (bx != cbx || by != cby || bz != cbz)) {
            cbx = bx;
            cby = by;
            cbz = bz;
            cocc = true;
            if (SPARSE) {
                const int cx = clampi(bx, 0, a.W - 2), cy = clampi(by, 0, a.H - 2),
                          cz = clampi(bz, 0, a.D - 2);
                if (a.bricks != nullptr) {
                    const uint32_t br = (uint32_t)(cz >> 3) * (uint32_t)a.nbx * (uint32_t)a.nby +
                                        (uint32_t)(cy >> 3) * (uint32_t)a.nbx + (uint32_t)(cx >> 3);
                    if (br == empty_brick) {
                        cocc = false;
                    } else {
                        cocc = (a.bricks[br >> 5] >> (br & 31)) & 1u;
                        if (!cocc) empty_brick = br;
                    }
                }
                if (cocc) {
                    const uint32_t cell = (uint32_t)cz * (uint32_t)(a.W - 1) * (uint32_t)(a.H - 1) +
                                          (uint32_t)cy * (uint32_t)(a.W - 1) + (uint32_t)cx;
                    cocc = (a.occ[cell >> 5] >> (cell & 31)) & 1u;
                }
            }
            if (cocc) {
                crec = scene_index(bx + jx, by + jy, bz + jz, a.W, a.H, a.D);
                if (SPARSE) crec = min(a.io[crec], (uint32_t)(a.n_vox - 1));
            }
        }
        const bool eval = act && cocc;
        // This lane's corner share (zero for lanes without a sample).
        double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
        if (eval) {
            float th[P];
            load_record<P, PXL_STORE_F32>(a.params, crec, a.stride, th);
            const double fx = px - (double)bx, fy = py - (double)by, fz = pz - (double)bz;
            const double w = (jx ? fx : 1.0 - fx) * (jy ? fy : 1.0 - fy) * (jz ? fz : 1.0 - fz);
            double c[3];
#pragma unroll
            for (int ch = 0; ch < 3; ch++) {
                double v = 0.0;
#pragma unroll
                for (int q = 0; q < S; q++) v = fma(Y[q], (double)th[1 + ch * S + q], v);
                c[ch] = v;
            }
            p0 = w * (double)th[0];
            p1 = w * c[0];
            p2 = w * c[1];
            p3 = w * c[2];
        }
        const double sraw = corner_sum_d(p0);
        const double cr0 = corner_sum_d(p1), cr1 = corner_sum_d(p2), cr2 = corner_sum_d(p3);
        if (eval) {
            if (MODE == kTrace) {
                hcount++;
                hash = fnv_mix(hash, (uint32_t)k);
                hash = fnv_mix(hash, (uint32_t)bx);
                hash = fnv_mix(hash, (uint32_t)by);
                hash = fnv_mix(hash, (uint32_t)bz);
            }
            const double sigma = softplus_shift<double>(sraw);
            const double col0 = sigmoid01<double>(cr0), col1 = sigmoid01<double>(cr1),
                         col2 = sigmoid01<double>(cr2);
            const double alpha = one_minus_exp_neg<double>(sigma * step);
            const double w = T * alpha;
            if (logging && logged < a.log_cap) {
                const size_t row = (size_t)b * (size_t)a.log_cap + (size_t)logged;
                a.log_rec[row * 8 + j] = crec;
                if (j < 5) {
                    const double v = j == 0 ? sraw : (j == 1 ? cr0 : (j == 2 ? cr1 : (j == 3 ? cr2 : t)));
                    reinterpret_cast<double *>(a.log_val)[row * 8 + j] = v;
                }
            }
            logged++;
            A += w;
            C0 += w * col0;
            C1 += w * col1;
            C2 += w * col2;
            sig += sigma;
            T = T * (1.0 - alpha);
            t += fabs(step - w);
        } else if (act) {
            t += step;
        }
        if (act) ++k;
    }
    if (!live) return;
    if (tnear > tfar) {
        // Missed the grid box: nothing marched (the loop did not run).
        if (j == 0) {
            a.rays.acc_r[ray] = 0.0f;
            a.rays.acc_g[ray] = 0.0f;
            a.rays.acc_b[ray] = 0.0f;
            a.rays.acc_w[ray] = 0.0f;
            a.rays.acc_density[ray] = 0.0f;
            if (MODE == kTrace) {
                a.trace[2 * b] = 0u;
                a.trace[2 * b + 1] = 2166136261u;
            }
            if (logging) {
                a.losses[b] = 0.0;
                a.log_count[b] = 0;
            }
        }
        return;
    }
    if (a.white_bg) {
        C0 += 1.0 - A;
        C1 += 1.0 - A;
        C2 += 1.0 - A;
    }
    if (j == 0) {
        a.rays.acc_r[ray] = (float)C0;
        a.rays.acc_g[ray] = (float)C1;
        a.rays.acc_b[ray] = (float)C2;
        a.rays.acc_w[ray] = (float)A;
        a.rays.acc_density[ray] = (float)sig;
        if (MODE == kTrace) {
            a.trace[2 * b] = hcount;
            a.trace[2 * b + 1] = hash;
        }
    }
    if (logging && j == 0) {
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
