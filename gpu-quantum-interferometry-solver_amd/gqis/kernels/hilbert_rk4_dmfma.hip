// RK4 on the full density matrix in FP64 with the commutator product X = H_eff rho on the FP64
// matrix cores (v_mfma_f64_16x16x4_f64).  Model structure, RHS and output semantics are those of
// hilbert_rk4.hip (H = H0 + diag(v(ParX, ParY, t)), diagonal or single-entry collapse operators,
// diagonal observable; gqis/cuda_solvers.py:470-517, gqis/solver.py:2126-2152).
//
// X is the real product Hbig (2N x 2N) x Rbig (2N x N) per point.  A wavefront covers 16 columns
// of Rbig per column tile (N = 8: two points, N = 16: one point, N = 32: one point in two column
// tiles) and the 2N rows in 2N / 16 row tiles.  The instruction's lane layout (checked on gfx950):
// A: lane l holds A[l % 16][l / 16]; B: lane l holds B[l / 16][l % 16]; the 16 x 16 result: lane
// l, register r holds D[4 r + l / 16][l % 16].  So lane (c, g) = (l % 16, l / 16) holds, in row
// tile T, the rows 16 T + 4 r + g (r < 4) of column c, and at the K step (T, r) it supplies as B
// exactly its own value of that row: with K visited in the order (T, r) the output of one RHS is
// the operand of the next with no data movement.  The A operand of step (T, r) for output tile
// To, lane (i, g), is Hbig'[16 To + i][16 T + 4 r + g] of Hbig with rows and columns in that
// order (packed on the host: GQ_OFF_DPACK, [To][T][r][lane], staged in LDS).
//
// Rows are ordered so that each lane holds real and imaginary parts of the same elements: slot
// m = 4 T + r of lane group g is Re rho[g N/4 + m][j] for m < N/4 and Im rho[g N/4 + m - N/4][j]
// otherwise.  Y^+ needs Y transposed: a per-wavefront LDS tile with padded rows.

typedef double gq_f64x4 __attribute__((ext_vector_type(4)));

#define HN GQ_HN
#define HTR (HN / 8)             // 16-row tiles of Rbig (2N rows)
#define HTC (HN == 32 ? 2 : 1)   // 16-column tiles per wavefront
#define HPW (HN == 8 ? 2 : 1)    // points per wavefront
#define HQ (HN / 4)              // complex elements per lane and column tile
#define HV (HN / 2)              // state values per lane and column tile
#define HWAVES (GQ_BLOCK / 64)
#define HLD (HN + 1)             // padded LDS row
// N = 32 keeps 4 x 32 doubles of RK4 state per lane: there the Y parts and dephasing weights
// of the second RHS pass are read back from LDS instead of held in registers
#define HLOWREG (HN == 32)

extern "C" __global__ __launch_bounds__(GQ_BLOCK) void GQ_KERNEL(gq_hilbert_rk4_dmfma)(const gq_sweep_args a) {
    __shared__ double s_apack[HTR * HTR * 4 * 64];
    __shared__ double s_y[HWAVES][2][HPW * HN * HLD];
    __shared__ double s_diag[HWAVES][HPW * HN];
    __shared__ double s_dep[HN * HN];           // dephasing weights sum_k d_k(i) d_k(j)
    __shared__ double s_jw[HN * GQ_JMAX];       // population transfer into a: weights |l|^2 ...
    __shared__ int s_jb[HN * GQ_JMAX];          // ... from the populations b (sparse rows)
    __shared__ double s_red[HWAVES][64];
#if GQ_NMONO
    // general jump operators (one nonzero per row): a per-wavefront copy of the state and the
    // operators' row tables, (L rho L^+)_ij = l_i conj(l_j) rho[col_i][col_j]
    __shared__ double s_xm[HWAVES][2][HPW * HN * HLD];
    __shared__ int s_mcol[GQ_NMONO * HN];
    __shared__ double s_mre[GQ_NMONO * HN];
    __shared__ double s_mim[GQ_NMONO * HN];
#endif

    const double* __restrict__ cst = (const double*)a.consts;   // packed operator constants
    for (int e = threadIdx.x; e < HTR * HTR * 4 * 64; e += GQ_BLOCK) s_apack[e] = cst[GQ_OFF_DPACK + e];
    for (int e = threadIdx.x; e < HN * HN; e += GQ_BLOCK) s_dep[e] = cst[GQ_OFF_DEP + e];
    for (int e = threadIdx.x; e < HN * GQ_JMAX; e += GQ_BLOCK) {
        s_jw[e] = cst[GQ_OFF_JW + e];
        s_jb[e] = (int)cst[GQ_OFF_JB + e];
    }
#if GQ_NMONO
    for (int e = threadIdx.x; e < GQ_NMONO * HN; e += GQ_BLOCK) {
        s_mcol[e] = (int)cst[GQ_OFF_MCOL + e];
        s_mre[e] = cst[GQ_OFF_MRE + e];
        s_mim[e] = cst[GQ_OFF_MIM + e];
    }
#endif
    __syncthreads();

    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int c = lane & 15, g = lane >> 4;
    const int pw = (HPW == 2) ? (c >> 3) : 0;           // point of this lane within the wave
    int jc[HTC];                                        // column of each column tile
#pragma unroll
    for (int ct = 0; ct < HTC; ++ct) jc[ct] = (16 * ct + c) % HN;
    const long long q = ((long long)blockIdx.x * HWAVES + wave) * HPW + pw;
    const bool valid = q < a.point_count;
    const long long qc = valid ? q : 0;
    const gq_point pt = gq_locate(a, qc);
    const double px = ((const double*)a.par_x)[pt.ix];
    const double py = ((const double*)a.par_y)[pt.iy];
    double* __restrict__ ybuf_r = s_y[wave][0] + pw * HN * HLD;
    double* __restrict__ ybuf_i = s_y[wave][1] + pw * HN * HLD;
    double* __restrict__ dbuf = s_diag[wave] + pw * HN;
#if GQ_NMONO
    double* __restrict__ xm_r = s_xm[wave][0] + pw * HN * HLD;
    double* __restrict__ xm_i = s_xm[wave][1] + pw * HN * HLD;
    int mcj[HTC][GQ_NMONO];      // the general jump operators at the lane's columns
    double mlr[HTC][GQ_NMONO], mli[HTC][GQ_NMONO];
#pragma unroll
    for (int ct = 0; ct < HTC; ++ct)
#pragma unroll
        for (int kk = 0; kk < GQ_NMONO; ++kk) {
            mcj[ct][kk] = s_mcol[kk * HN + jc[ct]];
            mlr[ct][kk] = s_mre[kk * HN + jc[ct]];
            mli[ct][kk] = s_mim[kk * HN + jc[ct]];
        }
#endif

    // ---- initial state: rho0 values (reduced layout) or |0><0| ------------------------------
    double y[HTC][HV];
#pragma unroll
    for (int ct = 0; ct < HTC; ++ct) {
        const int j = jc[ct];
#pragma unroll
        for (int u = 0; u < HQ; ++u) {
            const int i = g * HQ + u;
            double re = 0.0, im = 0.0;
#if GQ_RHO0_VALUES
            const double* __restrict__ r0 = (const double*)a.rho0 + qc * (HN * HN - 1);
            if (i == j) {
                if (i < HN - 1) {
                    re = r0[i];
                } else {
                    double s = 0.0;
                    for (int k = 0; k < HN - 1; ++k) s += r0[k];
                    re = 1.0 - s;
                }
            } else {
                const int lo = i < j ? i : j, hi = i < j ? j : i;
                const int uu = lo * HN - (lo * (lo + 1)) / 2 + (hi - lo - 1);
                re = r0[HN - 1 + 2 * uu];
                im = (i < j ? 1.0 : -1.0) * r0[HN - 1 + 2 * uu + 1];
            }
#else
            re = (i == 0 && j == 0) ? 1.0 : 0.0;
#endif
            y[ct][u] = re;
            y[ct][u + HQ] = im;
        }
    }

    // per-row diagonal drive coefficients v_i = sum_k coef[k][i] g_k(px, py, t)
    double vco[GQ_NBASIS][HQ];
#pragma unroll
    for (int k = 0; k < GQ_NBASIS; ++k)
#pragma unroll
        for (int u = 0; u < HQ; ++u) vco[k][u] = cst[GQ_OFF_VCOEF + k * HN + g * HQ + u];

    // Per-lane dissipator constants: dephasing weights of the lane's elements and, for the lane
    // holding a diagonal element (at most one: column tile ctd, slot md), its jump list and
    // observable weight, so the RHS has no per-element branch.
#if !HLOWREG
    double depw[HTC][HQ];
#endif
    int md[HTC];
    int ctd = -1, jd = 0;
#pragma unroll
    for (int ct = 0; ct < HTC; ++ct) {
        md[ct] = -1;
#pragma unroll
        for (int u = 0; u < HQ; ++u) {
            const int i = g * HQ + u;
#if !HLOWREG
            depw[ct][u] = s_dep[i * HN + jc[ct]];
#endif
            if (i == jc[ct]) {
                md[ct] = u;
                ctd = ct;
                jd = jc[ct];
            }
        }
    }
    const double od = (ctd >= 0) ? cst[GQ_OFF_OBS + jd] : 0.0;
    double jwl[GQ_JMAX];
    int jbl[GQ_JMAX];
#pragma unroll
    for (int e = 0; e < GQ_JMAX; ++e) {
        jwl[e] = (ctd >= 0) ? s_jw[jd * GQ_JMAX + e] : 0.0;
        jbl[e] = (ctd >= 0) ? s_jb[jd * GQ_JMAX + e] : 0;
    }

    const double dt = a.dt;
    const double dt2 = dt * 0.5;
    const double dt6 = dt / 6.0;
    const int num_steps = a.num_steps;
    double v[HQ];

    auto stage_values = [&](int row) {
        const double* __restrict__ tab = (const double*)a.time_table;
        double TT[GQ_NATOMS_ALLOC];
#pragma unroll
        for (int k = 0; k < GQ_NATOMS; ++k) TT[k] = tab[row * GQ_NATOMS + k];
        double gb[GQ_NBASIS];
        gq_basis(px, py, TT, cst, gb);
#pragma unroll
        for (int u = 0; u < HQ; ++u) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < GQ_NBASIS; ++k) s = fma(vco[k][u], gb[k], s);
            v[u] = s;
        }
    };

    // the lane's diagonal element (population) of x, 0 without one
    auto diag_of = [&](const double (*x)[HV]) -> double {
        double xd = 0.0;
#pragma unroll
        for (int ct = 0; ct < HTC; ++ct)
#pragma unroll
            for (int u = 0; u < HQ; ++u) xd = (u == md[ct]) ? x[ct][u] : xd;
        return xd;
    };

    auto rhs = [&](const double (*x)[HV], double (*k)[HV]) {
        gq_f64x4 acc[HTR][HTC];
#pragma unroll
        for (int t = 0; t < HTR; ++t)
#pragma unroll
            for (int ct = 0; ct < HTC; ++ct) acc[t][ct] = (gq_f64x4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int T = 0; T < HTR; ++T)
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int to = 0; to < HTR; ++to) {
                    const double av = s_apack[((to * HTR + T) * 4 + r) * 64 + lane];
#pragma unroll
                    for (int ct = 0; ct < HTC; ++ct)
                        acc[to][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, x[ct][4 * T + r],
                                                                           acc[to][ct], 0, 0, 0);
                }
#if !HLOWREG
        double yr[HTC][HQ], yi[HTC][HQ];
#endif
#pragma unroll
        for (int ct = 0; ct < HTC; ++ct) {
            const int j = jc[ct];
#pragma unroll
            for (int u = 0; u < HQ; ++u) {
                const double xr = acc[u >> 2][ct][u & 3];
                const double xi = acc[(u + HQ) >> 2][ct][(u + HQ) & 3];
                // Y = -i X - i v rho
                const double ry = fma(v[u], x[ct][u + HQ], xi);
                const double iy = -fma(v[u], x[ct][u], xr);
#if !HLOWREG
                yr[ct][u] = ry;
                yi[ct][u] = iy;
#endif
                const int i = g * HQ + u;
                ybuf_r[i * HLD + j] = ry;
                ybuf_i[i * HLD + j] = iy;
#if GQ_NMONO
                xm_r[i * HLD + j] = x[ct][u];
                xm_i[i * HLD + j] = x[ct][u + HQ];
#endif
            }
        }
        if (ctd >= 0) dbuf[jd] = diag_of(x);
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        double jsum = 0.0;
#pragma unroll
        for (int e = 0; e < GQ_JMAX; ++e) jsum = fma(jwl[e], dbuf[jbl[e]], jsum);
#pragma unroll
        for (int ct = 0; ct < HTC; ++ct) {
            const int j = jc[ct];
#pragma unroll
            for (int u = 0; u < HQ; ++u) {
                const int i = g * HQ + u;
#if HLOWREG
                const double ry = ybuf_r[i * HLD + j], iy = ybuf_i[i * HLD + j];
                const double dw = s_dep[i * HN + j];
#else
                const double ry = yr[ct][u], iy = yi[ct][u];
                const double dw = depw[ct][u];
#endif
                double dr = ry + ybuf_r[j * HLD + i];
                double di = iy - ybuf_i[j * HLD + i];
                dr = fma(dw, x[ct][u], dr);
                di = fma(dw, x[ct][u + HQ], di);
#if GQ_NMONO
#pragma unroll
                for (int kk = 0; kk < GQ_NMONO; ++kk) {
                    const int ci = s_mcol[kk * HN + i];
                    if (ci >= 0 && mcj[ct][kk] >= 0) {
                        // w = l_i conj(l_j), times rho[col_i][col_j]
                        const double lr = s_mre[kk * HN + i], li = s_mim[kk * HN + i];
                        const double wr = fma(lr, mlr[ct][kk], li * mli[ct][kk]);
                        const double wi = fma(li, mlr[ct][kk], -lr * mli[ct][kk]);
                        const double pr = xm_r[ci * HLD + mcj[ct][kk]];
                        const double pim = xm_i[ci * HLD + mcj[ct][kk]];
                        dr = fma(wr, pr, fma(-wi, pim, dr));
                        di = fma(wr, pim, fma(wi, pr, di));
                    }
                }
#endif
                k[ct][u] = (u == md[ct]) ? dr + jsum : dr;
                k[ct][u + HQ] = di;
            }
        }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    };

    // sum over the lanes of each point (HPW points per wave: lanes with c / 8 == pw for N = 8,
    // all 64 lanes otherwise), valid in the point's leader lane; every lane takes part
    const bool leader = (g == 0) && (HPW == 1 ? c == 0 : (c & 7) == 0);
    auto point_sum = [&](double part) -> double {
        s_red[wave][lane] = part;
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        double s = 0.0;
        if (leader)
            for (int l = 0; l < 64; ++l)
                if (HPW == 1 || (((l & 15) >> 3) == pw)) s += s_red[wave][l];
        return s;
    };
    bool bad_trace = false;
    (void)bad_trace;

    stage_values(0);
    double k[HTC][HV];
    rhs(y, k);
    double avg = 0.0;
    for (int step = 0; step < num_steps; ++step) {
        stage_values(2 * step + 1);
        double yt[HTC][HV], acc[HTC][HV];
#pragma unroll
        for (int ct = 0; ct < HTC; ++ct)
#pragma unroll
            for (int i = 0; i < HV; ++i) {
                yt[ct][i] = fma(dt2, k[ct][i], y[ct][i]);
                acc[ct][i] = k[ct][i];
            }
        rhs(yt, k);
#pragma unroll
        for (int ct = 0; ct < HTC; ++ct)
#pragma unroll
            for (int i = 0; i < HV; ++i) {
                yt[ct][i] = fma(dt2, k[ct][i], y[ct][i]);
                acc[ct][i] = fma(2.0, k[ct][i], acc[ct][i]);
            }
        rhs(yt, k);
#pragma unroll
        for (int ct = 0; ct < HTC; ++ct)
#pragma unroll
            for (int i = 0; i < HV; ++i) {
                yt[ct][i] = fma(dt, k[ct][i], y[ct][i]);
                acc[ct][i] = fma(2.0, k[ct][i], acc[ct][i]);
            }
        stage_values(2 * step + 2);
        rhs(yt, k);
#pragma unroll
        for (int ct = 0; ct < HTC; ++ct)
#pragma unroll
            for (int i = 0; i < HV; ++i) {
                acc[ct][i] += k[ct][i];
                y[ct][i] = fma(dt6, acc[ct][i], y[ct][i]);
            }
        if (step + 1 < num_steps) rhs(y, k);
#if GQ_OUT_MEAN
        if (step >= a.warmup_steps) avg += od * diag_of(y);
#endif
#if GQ_TRACE
        // observable trace after steps offset + k * stride (gqis/solver.py:2177-2193)
        if (step >= a.trace_offset && (step - a.trace_offset) % a.trace_stride == 0) {
            const int kk = (step - a.trace_offset) / a.trace_stride;
            if (kk < a.num_trace) {
                const double s = point_sum(od * diag_of(y));
                if (leader && valid) {
                    bad_trace |= !gq_finite(s);
                    gq_store_complex(a.trace, (long long)kk * a.trace_point_stride + q, s, 0.0);
                }
            }
        }
#endif
    }

    bool bad = false;
#if GQ_OUT_MEAN || GQ_OUT_FINAL
#if GQ_OUT_FINAL
    avg = od * diag_of(y);
#endif
    double s = point_sum(avg);
    if (leader) {
#if GQ_OUT_MEAN
        s /= (double)(num_steps - a.warmup_steps);
#endif
        if (valid) {
            bad = !gq_finite(s);
            gq_store_complex(a.result, q, s, 0.0);
        }
    }
#elif GQ_OUT_FINAL_RHO
    if (valid) {
#pragma unroll
        for (int ct = 0; ct < HTC; ++ct) {
            const int j = jc[ct];
#pragma unroll
            for (int u = 0; u < HQ; ++u) {
                const int i = g * HQ + u;
                if (i == j && i < HN - 1) {
                    ((double*)a.result)[(long long)i * a.point_count + q] = y[ct][u];
                    bad |= !gq_finite(y[ct][u]);
                } else if (i < j) {
                    const int idx = HN - 1 + 2 * (i * HN - (i * (i + 1)) / 2 + (j - i - 1));
                    ((double*)a.result)[(long long)idx * a.point_count + q] = y[ct][u];
                    ((double*)a.result)[(long long)(idx + 1) * a.point_count + q] = y[ct][u + HQ];
                    bad |= !(gq_finite(y[ct][u]) && gq_finite(y[ct][u + HQ]));
                }
            }
        }
    }
#endif
    gq_flag(a.status, bad, GQ_STATUS_NONFINITE_RESULT);
#if GQ_TRACE
    gq_flag(a.status, bad_trace, GQ_STATUS_NONFINITE_TRACE);
#endif
}
