// Fixed-step integration of the full density matrix for models whose H_eff is sparse
// (multi-qubit chains: a few nonzeros per row), one wavefront per parameter point, on the
// vector ALUs.
//
// Same model structure and equation as hilbert_rk4.hip: H(t) = H0 + diag(v(ParX, ParY, t)),
// H_eff = H0 - (i/2) sum L^+L, and per element (i, j)
//     d rho_ij = c_ij rho_ij - i sum_k h_ik rho_kj + i sum_k conj(h_jk) rho_ik + jumps,
//     c_ij = -i (h_ii - conj(h_jj)) - i (v_i - v_j) + sum_k d_k(i) d_k(j),
// where the sums run over the off-diagonal nonzeros of H_eff: -i H_eff rho + i rho H_eff^+
// without a product of dense matrices, 2 K complex multiply-adds per element instead of 2 N.
// Jumps: single-entry operators l |a><b| add |l|^2 rho_bb to rho_aa; operators with one
// nonzero per row add l_i conj(l_j) rho[col_i][col_j].  Methods: classical RK4 (GQ_SCHEME
// undefined; stage times, k1 pipelining: gqis/cuda_solvers.py:470-517) or, with GQ_SCHEME and
// the tableau constants of fixed_tableau.hip, an explicit Butcher tableau (0), the 2N-storage
// LSRK4 (1) or AB5 (2); with GQ_ADAPTIVE, adaptive Tsit5 / DOP853 under the controller of
// adaptive.hip.  Output semantics are the reference's (gqis/solver.py:2126-2193).
//
// Off-diagonal structure, two forms:
//   GQ_XOR 1: every row i has its nonzeros at columns i ^ mask_k with values v_k that do not
//     depend on i, and each general jump operator maps i to i ^ b where i & b == cond, with
//     one weight (qubit operators sx_k, s-_k on a register of qubits): masks, values and
//     weights are compile-time constants, the neighbours' addresses one XOR away;
//   GQ_XOR 0: an ELL table of GQ_HK (column, value) entries per row, padded with zero values,
//     and the jump operators' row tables, read from LDS.
//
// Layout: rho is Hermitian, so only the N (N + 1) / 2 elements i <= j are integrated: element u
// of the upper triangle in row-major order is slot m = u / 64 of lane u % 64.  The wavefront
// keeps the stage input in an LDS copy of the full rho (complex): the caller writes each element
// and, below the diagonal, its conjugate; the RHS reads each element's neighbours rho[k][j] and
// rho[i][k] from there.  Element (i, j) of the copy sits at i HNP + (j ^ i), HNP the power of
// two >= N: the column is XOR-swizzled by the row, so the conjugate writes (lanes of one row
// writing down a column) fall into distinct banks, where a plain row-major copy with a row of
// 64 or 128 words put all 16 lanes of a write group into one bank; row reads stay a
// permutation of consecutive words.  XOR neighbours remain one XOR away: (i ^ m, j) is at
// addr ^ (m HNP + m), (i, j ^ m) at addr ^ m.  The static part of c_ij is a per-element LDS
// table (row-major), the drive part v_i - v_j comes from the drive row table.

#define HN GQ_HN
#define HK GQ_HK                  // off-diagonal nonzeros per row (ELL width)
#define HT (HN * (HN + 1) / 2)    // elements i <= j
#define HE ((HT + 63) / 64)       // complex elements (slots) per lane
#define HWAVES (GQ_BLOCK / 64)
// row stride of the state copy: the power of two >= HN
#define HNP (HN <= 8 ? 8 : HN <= 16 ? 16 : HN <= 32 ? 32 : HN <= 64 ? 64 : 128)
static_assert(HN <= 128, "state copy rows up to 128 columns");
#ifndef GQ_ADAPTIVE
#define GQ_ADAPTIVE 0
#endif
#if GQ_XOR
static_assert((HN & (HN - 1)) == 0, "the XOR form addresses a register of qubits");
#endif

// At least two wavefronts per SIMD: a wavefront is one point, and with one per SIMD the LDS
// reads of the right-hand side go unhidden.  The bound caps the registers at 256; it binds for
// the FP32 kernels with the largest stage storage at 32 levels (DOP853, AB5, Tsit5A, DOP853A:
// 274-357 registers otherwise; Tsit5A on the 5-qubit chain 1.84 -> 1.19 s, DOP853 4.40 ->
// 2.96 s on 128 x 256 points, with at most 136 B of spills).  FP64 copies of rho keep such
// kernels at one work-group per compute unit anyway (LDS).
#ifndef GQ_SPARSE_MIN_WAVES
#define GQ_SPARSE_MIN_WAVES 2
#endif
extern "C" __global__ __launch_bounds__(GQ_BLOCK) __attribute__((amdgpu_waves_per_eu(GQ_SPARSE_MIN_WAVES)))
void GQ_KERNEL(gq_hilbert_sparse)(const gq_sweep_args a) {
    __shared__ real2 s_rho[HWAVES][HN * HNP];
    __shared__ real2 s_cf[HN * HN];                 // static c_ij
    __shared__ real s_vrow[GQ_NBASIS * HN];         // drive row coefficients v_i per basis term
    __shared__ real s_jw[HN * GQ_JMAX];
    __shared__ int s_jb[HN * GQ_JMAX];
#if !GQ_XOR
    __shared__ int s_ecol[HN * HK];
    __shared__ real2 s_eval[HN * HK];
#if GQ_NMONO
    __shared__ int s_mcol[GQ_NMONO * HN];
    __shared__ real2 s_ml[GQ_NMONO * HN];
#endif
#endif

    const real* __restrict__ cst = (const real*)a.consts;   // packed operator constants
    for (int e = threadIdx.x; e < HN * HN; e += GQ_BLOCK) {
        const int i = e / HN, jj = e % HN;
        s_cf[e] = real2{cst[GQ_OFF_HDI + i] + cst[GQ_OFF_HDI + jj] + cst[GQ_OFF_DEP + e],
                        cst[GQ_OFF_HDR + jj] - cst[GQ_OFF_HDR + i]};
    }
    for (int e = threadIdx.x; e < GQ_NBASIS * HN; e += GQ_BLOCK) s_vrow[e] = cst[GQ_OFF_VCOEF + e];
    for (int e = threadIdx.x; e < HN * GQ_JMAX; e += GQ_BLOCK) {
        s_jw[e] = cst[GQ_OFF_JW + e];
        s_jb[e] = (int)cst[GQ_OFF_JB + e];
    }
#if !GQ_XOR
    for (int e = threadIdx.x; e < HN * HK; e += GQ_BLOCK) {
        s_ecol[e] = (int)cst[GQ_OFF_ECOL + e];
        s_eval[e] = real2{cst[GQ_OFF_ERE + e], cst[GQ_OFF_EIM + e]};
    }
#if GQ_NMONO
    for (int e = threadIdx.x; e < GQ_NMONO * HN; e += GQ_BLOCK) {
        s_mcol[e] = (int)cst[GQ_OFF_MCOL + e];
        s_ml[e] = real2{cst[GQ_OFF_MRE + e], cst[GQ_OFF_MIM + e]};
    }
#endif
#endif
    __syncthreads();

    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    // the lane's slots: element u = lane + 64 m of the upper triangle, (row, column) (ri, rj)
    int ri[HE], rj[HE];
    bool used[HE];
#pragma unroll
    for (int m = 0; m < HE; ++m) {
        const int u = lane + 64 * m;
        used[m] = u < HT;
        int i = 0, start = 0;
        while (i < HN - 1 && start + (HN - i) <= u) {
            start += HN - i;
            ++i;
        }
        ri[m] = used[m] ? i : 0;
        rj[m] = used[m] ? i + (u - start) : 0;
    }
    const long long q = (long long)blockIdx.x * HWAVES + wave;
    const bool valid = q < a.point_count;
    const long long qc = valid ? q : 0;
    const gq_point pt = gq_locate(a, qc);
    const real px = ((const real*)a.par_x)[pt.ix];
    const real py = ((const real*)a.par_y)[pt.iy];
    real2* __restrict__ rho = s_rho[wave];

    // ---- initial state: rho0 values (reduced layout) or |0><0| ------------------------------
    real y[2 * HE];
#pragma unroll
    for (int m = 0; m < HE; ++m) {
        const int i = ri[m], j = rj[m];
        real re = R(0.0), im = R(0.0);
#if GQ_RHO0_VALUES
        const real* __restrict__ r0 = (const real*)a.rho0 + qc * (HN * HN - 1);
        if (i == j) {
            if (i < HN - 1) {
                re = r0[i];
            } else {
                real s = R(0.0);
                for (int k = 0; k < HN - 1; ++k) s += r0[k];
                re = R(1.0) - s;
            }
        } else {
            const int lo = i < j ? i : j, hi = i < j ? j : i;
            const int u = lo * HN - (lo * (lo + 1)) / 2 + (hi - lo - 1);
            re = r0[HN - 1 + 2 * u];
            im = (i < j ? R(1.0) : R(-1.0)) * r0[HN - 1 + 2 * u + 1];
        }
#else
        re = (i == 0 && j == 0) ? R(1.0) : R(0.0);
#endif
        y[m] = re;
        y[m + HE] = im;
    }

    // ---- per-lane constants ------------------------------------------------------------------
    // observable weight of each diagonal slot (zero elsewhere)
    real od[HE];
#pragma unroll
    for (int m = 0; m < HE; ++m)
        od[m] = (used[m] && ri[m] == rj[m]) ? cst[GQ_OFF_OBS + ri[m]] : R(0.0);

    const real dt = (real)a.dt;
    const real dt2 = dt * R(0.5);
    const real dt6 = dt / R(6.0);
    const int num_steps = a.num_steps;
    real g[GQ_NBASIS];         // drive basis terms at the current stage time

    auto stage_values = [&](int row) {
        const real* __restrict__ tab = (const real*)a.time_table;
        real TT[GQ_NATOMS_ALLOC];
#pragma unroll
        for (int k = 0; k < GQ_NATOMS; ++k) TT[k] = tab[row * GQ_NATOMS + k];
        gq_basis(px, py, TT, cst, g);
    };

    // element m of the stage input into the wavefront's LDS copy (and its conjugate below the
    // diagonal)
    auto store = [&](int m, real vr, real vi) {
        if (used[m]) {
            rho[ri[m] * HNP + (rj[m] ^ ri[m])] = real2{vr, vi};
            if (ri[m] != rj[m]) rho[rj[m] * HNP + (ri[m] ^ rj[m])] = real2{vr, -vi};
        }
    };
    // the stage input x + c k
    auto put = [&](const real* __restrict__ x, real c, const real* __restrict__ kk) {
#pragma unroll
        for (int m = 0; m < HE; ++m) store(m, fma(c, kk[m], x[m]), fma(c, kk[m + HE], x[m + HE]));
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    };

    // right-hand side at the state in the LDS copy
    auto rhs = [&](real* __restrict__ kout) {
#pragma unroll
        for (int m = 0; m < HE; ++m) {
            // the element's address, opaque to the compiler: every address below derives from
            // it with an XOR or a multiply-add per call instead of living in a register
            // across the step loop (which took about 200 registers at 32 levels)
            int b = ri[m] * HN + rj[m];
            asm volatile("" : "+v"(b));
            const int i = b / HN, j = b - (b / HN) * HN;
            // the element's slot in the swizzled copy (with HNP = HN: b ^ i)
            const int bs = (HN == HNP) ? (b ^ i) : i * HNP + (j ^ i);
            const real2 x = rho[bs];
            const real2 cf = s_cf[b];
            real dvi = R(0.0), dvj = R(0.0);
#pragma unroll
            for (int k = 0; k < GQ_NBASIS; ++k) {
                dvi = fma(s_vrow[k * HN + i], g[k], dvi);
                dvj = fma(s_vrow[k * HN + j], g[k], dvj);
            }
            const real ci = cf.y - (dvi - dvj);
            real dr = fma(cf.x, x.x, -ci * x.y);
            real di = fma(cf.x, x.y, ci * x.x);
            real lr = R(0.0), li = R(0.0), ur = R(0.0), ui = R(0.0);
#if GQ_XOR
            // -i sum_k v_k rho[i ^ mask_k][j] and +i sum_k conj(v_k) rho[i][j ^ mask_k]
#pragma unroll
            for (int k = 0; k < HK; ++k) {
                // HNP = HN: (i ^ mask) HN + (j ^ i ^ mask) = bs ^ (mask HN + mask) and
                // i HN + (j ^ mask ^ i) = bs ^ mask
                const real2 p = rho[bs ^ (gq_xmask[k] * HNP + gq_xmask[k])];
                lr = fma(gq_xre[k], p.x, fma(-gq_xim[k], p.y, lr));
                li = fma(gq_xre[k], p.y, fma(gq_xim[k], p.x, li));
                const real2 u = rho[bs ^ gq_xmask[k]];
                ur = fma(gq_xre[k], u.x, fma(gq_xim[k], u.y, ur));
                ui = fma(gq_xre[k], u.y, fma(-gq_xim[k], u.x, ui));
            }
#else
#pragma unroll
            for (int k = 0; k < HK; ++k) {
                const int c = s_ecol[i * HK + k];
                const real2 hv = s_eval[i * HK + k];
                const real2 p = rho[c * HNP + (j ^ c)];
                lr = fma(hv.x, p.x, fma(-hv.y, p.y, lr));
                li = fma(hv.x, p.y, fma(hv.y, p.x, li));
                const int cj = s_ecol[j * HK + k];
                const real2 hj = s_eval[j * HK + k];
                const real2 u = rho[i * HNP + (cj ^ i)];
                ur = fma(hj.x, u.x, fma(hj.y, u.y, ur));
                ui = fma(hj.x, u.y, fma(-hj.y, u.x, ui));
            }
#endif
            dr += li - ui;
            di += ur - lr;
#if GQ_NMONO
#if GQ_XOR
            // l_i conj(l_j) rho[i ^ b][j ^ b] where both i and j map (i & b == cond)
#pragma unroll
            for (int kk = 0; kk < GQ_NMONO; ++kk) {
                if ((i & gq_mmask[kk]) == gq_mcond[kk] && (j & gq_mmask[kk]) == gq_mcond[kk]) {
                    // (i ^ b, j ^ b) sits at (i ^ b) HN + (j ^ i) = bs ^ (b HN)
                    const real2 p = rho[bs ^ (gq_mmask[kk] * HNP)];
                    dr = fma(gq_mw[kk], p.x, dr);
                    di = fma(gq_mw[kk], p.y, di);
                }
            }
#else
#pragma unroll
            for (int kk = 0; kk < GQ_NMONO; ++kk) {
                const int ci2 = s_mcol[kk * HN + i], cj2 = s_mcol[kk * HN + j];
                if (ci2 >= 0 && cj2 >= 0) {
                    // w = l_i conj(l_j), times rho[col_i][col_j]
                    const real2 l = s_ml[kk * HN + i], lj = s_ml[kk * HN + j];
                    const real wr = fma(l.x, lj.x, l.y * lj.y);
                    const real wi = fma(l.y, lj.x, -l.x * lj.y);
                    const real2 p = rho[ci2 * HNP + (cj2 ^ ci2)];
                    dr = fma(wr, p.x, fma(-wi, p.y, dr));
                    di = fma(wr, p.y, fma(wi, p.x, di));
                }
            }
#endif
#endif
            if (i == j) {
                // single-entry jumps into the population i: sum_b |l|^2 rho_bb
#pragma unroll
                for (int e = 0; e < GQ_JMAX; ++e)
                    dr = fma(s_jw[i * GQ_JMAX + e], rho[s_jb[i * GQ_JMAX + e] * HNP].x, dr);
            }
            kout[m] = dr;
            kout[m + HE] = di;
        }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    };

    // sum over the wavefront (every lane takes part; the result is in every lane)
    auto wave_sum = [&](double part) -> double {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) part += __shfl_xor(part, off);
        return part;
    };
    bool bad_trace = false;
    (void)bad_trace;
    double avg = 0.0;
    bool failed = false;

#if GQ_ADAPTIVE
    // ---- adaptive Tsit5 / DOP853 between the output points of tlist ---------------------------
    // The controller of adaptive.hip (gqis/cuda_solvers.py:731-908: snap to the output point
    // when h * 1.01 >= the remaining time, h kept across intervals, FSAL derivative kept on a
    // rejection, NaN after more than max_substeps attempts in an interval, trapezoid mean over
    // accepted steps).  One point per wavefront: the error norm is a wavefront sum (a butterfly,
    // so every lane holds the same bits) and the controller is uniform.  The norm runs over the
    // reduced state's N^2 - 1 components: the populations but the last, and the real and
    // imaginary parts above the diagonal.
    (void)dt2;
    (void)dt6;
    const double dt_out = a.dt;
    const double rtol = a.solver_param[0], atol = a.solver_param[1];
    const int max_sub = (int)a.solver_param[2];
    const double inv_avg_time = 1.0 / ((double)(num_steps - a.warmup_steps) * dt_out);
    // components in the norm, as bits of one word per lane (bit m: the real part of slot m, bit
    // HE + m: its imaginary part); an array of bools would take a 64-bit lane mask of scalar
    // registers per entry
    static_assert(2 * HE <= 32, "norm mask bits");
    unsigned cmask = 0u;
#pragma unroll
    for (int m = 0; m < HE; ++m) {
        if (used[m] && (ri[m] < rj[m] || ri[m] < HN - 1)) cmask |= 1u << m;
        if (used[m] && ri[m] < rj[m]) cmask |= 1u << (HE + m);
    }
    auto uniform = [](double v) {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_readfirstlane((int)b);
        const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
        return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    };
    auto stage_at = [&](double t_abs) {
        double T64[GQ_NATOMS_ALLOC];
        gq_atoms_f64(t_abs, cst, T64);
        real TT[GQ_NATOMS_ALLOC];
#pragma unroll
        for (int k = 0; k < GQ_NATOMS; ++k) TT[k] = (real)T64[k];
        gq_basis(px, py, TT, cst, g);
    };
    auto obs_part = [&](const real* __restrict__ x) {
        real yd = R(0.0);
#pragma unroll
        for (int m = 0; m < HE; ++m) yd = fma(od[m], x[m], yd);
        return yd;
    };
    constexpr double ncomp = (double)(HN * HN - 1);
    real k[GQ_NSTAGES][2 * HE];
    real kf[2 * HE];
    stage_at(a.t_in);
    put(y, R(0.0), y);
    rhs(k[0]);
    double tau = 0.0, h = dt_out;
    int step = 0, sub = 0, n_acc = 0, n_rej = 0;
    real op = obs_part(y);
    (void)op;
    while (step < num_steps) {
        const double remaining = dt_out - tau;
        const bool last = h * 1.01 >= remaining;
        const double hs = last ? remaining : h;
        const double t0 = (double)step * dt_out + tau;
        const real hr = (real)hs;
#pragma unroll
        for (int s = 1; s < GQ_NSTAGES; ++s) {
#pragma unroll
            for (int m = 0; m < HE; ++m) {
                real ar = R(0.0), ai = R(0.0);
#pragma unroll
                for (int j = 0; j < s; ++j) {
                    if (GQ_A[s][j] != 0.0) {
                        ar = fma((real)GQ_A[s][j], k[j][m], ar);
                        ai = fma((real)GQ_A[s][j], k[j][m + HE], ai);
                    }
                }
                store(m, fma(hr, ar, y[m]), fma(hr, ai, y[m + HE]));
            }
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            stage_at(a.t_in + (t0 + GQ_C[s] * hs));
            rhs(k[s]);
        }
        real yn[2 * HE];
#pragma unroll
        for (int e = 0; e < 2 * HE; ++e) {
            real acc = R(0.0);
#pragma unroll
            for (int j = 0; j < GQ_NSTAGES; ++j)
                if (GQ_B[j] != 0.0) acc = fma((real)GQ_B[j], k[j][e], acc);
            yn[e] = fma(hr, acc, y[e]);
        }
        // the stage values now hold t0 + hs (the last stage has c = 1)
        double err;
#if GQ_METHOD_DOP853
        {
            double s5 = 0.0, s3 = 0.0;
#pragma unroll
            for (int e = 0; e < 2 * HE; ++e) {
                if (!((cmask >> e) & 1u)) continue;
                double e5 = 0.0, e3 = 0.0;
#pragma unroll
                for (int j = 0; j < GQ_NSTAGES; ++j) {
                    if (GQ_E5[j] != 0.0) e5 = fma(GQ_E5[j], (double)k[j][e], e5);
                    if (GQ_E3[j] != 0.0) e3 = fma(GQ_E3[j], (double)k[j][e], e3);
                }
                const double sc = fma(rtol, fmax(fabs((double)y[e]), fabs((double)yn[e])), atol);
                s5 = fma(e5 / sc, e5 / sc, s5);
                s3 = fma(e3 / sc, e3 / sc, s3);
            }
            s5 = wave_sum(s5);
            s3 = wave_sum(s3);
            double deno = s5 + 0.01 * s3;
            if (!(deno > 0.0)) deno = 1.0;
            err = fabs(hs) * s5 / sqrt(deno * ncomp);
        }
#else
        put(yn, R(0.0), yn);
        rhs(kf);
        {
            double ssum = 0.0;
#pragma unroll
            for (int e = 0; e < 2 * HE; ++e) {
                if (!((cmask >> e) & 1u)) continue;
                double d = GQ_E[GQ_NSTAGES] * (double)kf[e];
#pragma unroll
                for (int j = 0; j < GQ_NSTAGES; ++j) d = fma(GQ_E[j], (double)k[j][e], d);
                d *= hs;
                const double sc = fma(rtol, fmax(fabs((double)y[e]), fabs((double)yn[e])), atol);
                ssum = fma(d / sc, d / sc, ssum);
            }
            err = sqrt(wave_sum(ssum) / ncomp);
        }
#endif
        err = uniform(err);
        const double factor = gq_step_factor(err);   // step_control.hip
        if (err <= 1.0) {
            ++n_acc;
#if GQ_METHOD_DOP853
            put(yn, R(0.0), yn);
            rhs(kf);
#endif
#if GQ_OUT_MEAN
            const real o = obs_part(yn);
            if (step >= a.warmup_steps)
                avg = fma(0.5 * hs * inv_avg_time, (double)o + (double)op, avg);
            op = o;
#endif
#pragma unroll
            for (int e = 0; e < 2 * HE; ++e) {
                y[e] = yn[e];
                k[0][e] = kf[e];
            }
            tau = last ? dt_out : tau + hs;
            h = last ? fmax(h, hs * factor) : hs * factor;
        } else {
            ++n_rej;
            h = hs * factor;
        }
        // the controller is the same in every lane: scalar registers hold it
        h = uniform(h);
        tau = uniform(tau);
        if (++sub > max_sub) {
            failed = true;
#pragma unroll
            for (int e = 0; e < 2 * HE; ++e) y[e] = (real)__builtin_nan("");
        }
        if (failed || tau >= dt_out) {
#if GQ_TRACE
            if (step >= a.trace_offset && (step - a.trace_offset) % a.trace_stride == 0) {
                const int kk = (step - a.trace_offset) / a.trace_stride;
                if (kk < a.num_trace) {
                    const double s = wave_sum((double)obs_part(y));
                    if (lane == 0 && valid) {
                        const real r = (real)s;
                        bad_trace |= !gq_finite(r);
                        gq_store_complex(a.trace, (long long)kk * a.trace_point_stride + q, r,
                                         R(0.0));
                    }
                }
            }
#endif
            ++step;
            tau = 0.0;
            sub = 0;
            if (failed) break;
        }
    }
#if GQ_COUNT_STEPS
    if (lane == 0 && valid) {
        a.adaptive_steps[2 * q] = n_acc;
        a.adaptive_steps[2 * q + 1] = n_rej;
    }
#else
    (void)n_acc;
    (void)n_rej;
#endif
#else
    stage_values(0);
#ifndef GQ_SCHEME
    // classical RK4 (table rows 2 step + 1 and 2 step + 2: midpoint and end of the step)
    real k[2 * HE];
    put(y, R(0.0), y);
    rhs(k);
    for (int step = 0; step < num_steps; ++step) {
        stage_values(2 * step + 1);
        real acc[2 * HE];
#pragma unroll
        for (int e = 0; e < 2 * HE; ++e) acc[e] = k[e];
        put(y, dt2, k);
        rhs(k);
#pragma unroll
        for (int e = 0; e < 2 * HE; ++e) acc[e] = fma(R(2.0), k[e], acc[e]);
        put(y, dt2, k);
        rhs(k);
#pragma unroll
        for (int e = 0; e < 2 * HE; ++e) acc[e] = fma(R(2.0), k[e], acc[e]);
        put(y, dt, k);
        stage_values(2 * step + 2);
        rhs(k);
#pragma unroll
        for (int e = 0; e < 2 * HE; ++e) {
            acc[e] += k[e];
            y[e] = fma(dt6, acc[e], y[e]);
        }
        if (step + 1 < num_steps) {
            put(y, R(0.0), y);
            rhs(k);
        }
#else
    // the other fixed-step methods (gqis/cuda_solvers.py:48-369, 520-702): GQ_NROWS table rows
    // per step after row 0 (t = 0); the stage arithmetic of fixed_tableau.hip per element
    (void)dt2;
    (void)dt6;
    auto row_at = [&](int step, int row) { return row < 0 ? 0 : 1 + step * GQ_NROWS + row; };
#if GQ_SCHEME == 0
    // explicit Runge-Kutta from a Butcher tableau, k1 of a step from the end of the previous
    // one.  Folded stage storage (as in fixed_tableau.hip): k_0 .. k_{F-1} are kept until stage
    // F is formed, then folded into partial sums Ps of the later stages; y_new takes each k_s
    // as soon as it exists.  Every sum runs in the unfolded form's order, with about 8 instead
    // of 14 state-sized arrays live for DOP853.
    const double p0 = a.solver_param[0];
    auto nz = [](const int s, const int j) { return GQ_A[s][j] != 0.0 || GQ_A1[s][j] != 0.0; };
    // a_sj + p0 * a1_sj only where the tableau has a fitted part (anas5): elsewhere the
    // compile-time a_sj, which a run-time p0 * 0 would turn into a value held in a register
    auto coef = [&](const int s, const int j) {
        return GQ_A1[s][j] != 0.0 ? (real)(GQ_A[s][j] + GQ_A1[s][j] * p0) : (real)GQ_A[s][j];
    };
    static_assert(GQ_TFOLD >= 1 && GQ_TFOLD < GQ_NSTAGES, "fold stage inside the step");
    real k[GQ_NSTAGES][2 * HE];
    real Ps[GQ_NSTAGES][2 * HE];
    real yn[2 * HE];
    auto take_b = [&](const int s) {
        if (GQ_B[s] != 0.0) {
#pragma unroll
            for (int e = 0; e < 2 * HE; ++e) yn[e] = fma(dt * (real)GQ_B[s], k[s][e], yn[e]);
        }
    };
    auto fold = [&](const int s) {
#pragma unroll
        for (int u = s + 1; u < GQ_NSTAGES; ++u)
            if (nz(u, s)) {
#pragma unroll
                for (int e = 0; e < 2 * HE; ++e) Ps[u][e] = fma(coef(u, s), k[s][e], Ps[u][e]);
            }
    };
    // the stage input from the stored derivatives k_0 .. k_{s-1}
    auto store_raw = [&](const int s) {
#pragma unroll
        for (int m = 0; m < HE; ++m) {
            real ar = R(0.0), ai = R(0.0);
#pragma unroll
            for (int j = 0; j < s; ++j) {
                if (nz(s, j)) {
                    ar = fma(coef(s, j), k[j][m], ar);
                    ai = fma(coef(s, j), k[j][m + HE], ai);
                }
            }
            store(m, fma(dt, ar, y[m]), fma(dt, ai, y[m + HE]));
        }
    };
    put(y, R(0.0), y);
    rhs(k[0]);
    for (int step = 0; step < num_steps; ++step) {
#pragma unroll
        for (int e = 0; e < 2 * HE; ++e) yn[e] = y[e];
        take_b(0);
        // stages 1 .. F-1 from the stored derivatives
#pragma unroll
        for (int s = 1; s < GQ_TFOLD; ++s) {
            store_raw(s);
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            if (GQ_ROW_OF[s] != GQ_ROW_OF[s - 1]) stage_values(row_at(step, GQ_ROW_OF[s]));
            rhs(k[s]);
            take_b(s);
        }
        // stage F from the stored derivatives, which are then folded
        {
            constexpr int s = GQ_TFOLD;
            store_raw(s);
#pragma unroll
            for (int u = s + 1; u < GQ_NSTAGES; ++u) {
#pragma unroll
                for (int e = 0; e < 2 * HE; ++e) {
                    real p = R(0.0);
#pragma unroll
                    for (int j = 0; j < s; ++j)
                        if (nz(u, j)) p = fma(coef(u, j), k[j][e], p);
                    Ps[u][e] = p;
                }
            }
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            if (GQ_ROW_OF[s] != GQ_ROW_OF[s - 1]) stage_values(row_at(step, GQ_ROW_OF[s]));
            rhs(k[s]);
            fold(s);
            take_b(s);
        }
        // stages F+1 .. GQ_NSTAGES-1 from the partial sums
#pragma unroll
        for (int s = GQ_TFOLD + 1; s < GQ_NSTAGES; ++s) {
#pragma unroll
            for (int m = 0; m < HE; ++m)
                store(m, fma(dt, Ps[s][m], y[m]), fma(dt, Ps[s][m + HE], y[m + HE]));
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            if (GQ_ROW_OF[s] != GQ_ROW_OF[s - 1]) stage_values(row_at(step, GQ_ROW_OF[s]));
            rhs(k[s]);
            fold(s);
            take_b(s);
        }
#pragma unroll
        for (int e = 0; e < 2 * HE; ++e) y[e] = yn[e];
        if (GQ_ROW_OF[GQ_NSTAGES - 1] != GQ_END_ROW) stage_values(row_at(step, GQ_END_ROW));
        if (step + 1 < num_steps) {
            put(y, R(0.0), y);
            rhs(k[0]);
        }
#elif GQ_SCHEME == 1
    // Carpenter-Kennedy 2N-storage LSRK4
    real r[2 * HE], k[2 * HE];
    for (int step = 0; step < num_steps; ++step) {
#pragma unroll
        for (int s = 0; s < GQ_NSTAGES; ++s) {
            if (s > 0) stage_values(row_at(step, GQ_ROW_OF[s]));
            put(y, R(0.0), y);
            rhs(k);
#pragma unroll
            for (int e = 0; e < 2 * HE; ++e) {
                r[e] = (s == 0) ? dt * k[e] : fma((real)GQ_A[s][0], r[e], dt * k[e]);
                y[e] = fma((real)GQ_B[s], r[e], y[e]);
            }
        }
        stage_values(row_at(step, GQ_END_ROW));
#else
    // Adams-Bashforth 5 after four classical RK4 steps (rows 0 and 1: midpoint and end)
    real h0[2 * HE], h1[2 * HE], h2[2 * HE], h3[2 * HE], h4[2 * HE];   // f_n ... f_{n-4}
    const real dth = dt * R(0.5), dts = dt / R(6.0);
    put(y, R(0.0), y);
    rhs(h0);
    for (int step = 0; step < num_steps; ++step) {
        if (step < 4) {
            real acc[2 * HE], kk[2 * HE];
            stage_values(row_at(step, 0));
#pragma unroll
            for (int e = 0; e < 2 * HE; ++e) acc[e] = h0[e];
            put(y, dth, h0);
            rhs(kk);
#pragma unroll
            for (int e = 0; e < 2 * HE; ++e) acc[e] = fma(R(2.0), kk[e], acc[e]);
            put(y, dth, kk);
            rhs(kk);
#pragma unroll
            for (int e = 0; e < 2 * HE; ++e) acc[e] = fma(R(2.0), kk[e], acc[e]);
            put(y, dt, kk);
            stage_values(row_at(step, 1));
            rhs(kk);
#pragma unroll
            for (int e = 0; e < 2 * HE; ++e) {
                acc[e] += kk[e];
                y[e] = fma(dts, acc[e], y[e]);
            }
        } else {
#pragma unroll
            for (int e = 0; e < 2 * HE; ++e) {
                real sum = R(0.0);
                sum = fma(R(1901.0 / 720.0), h0[e], sum);
                sum = fma(R(-1387.0 / 360.0), h1[e], sum);
                sum = fma(R(109.0 / 30.0), h2[e], sum);
                sum = fma(R(-637.0 / 360.0), h3[e], sum);
                sum = fma(R(251.0 / 720.0), h4[e], sum);
                y[e] = fma(dt, sum, y[e]);
            }
            stage_values(row_at(step, 1));
        }
        if (step + 1 < num_steps) {
#pragma unroll
            for (int e = 0; e < 2 * HE; ++e) {
                h4[e] = h3[e];
                h3[e] = h2[e];
                h2[e] = h1[e];
                h1[e] = h0[e];
            }
            put(y, R(0.0), y);
            rhs(h0);
        }
#endif
#endif
#if GQ_OUT_MEAN
        if (step >= a.warmup_steps) {
            real yd = R(0.0);
#pragma unroll
            for (int m = 0; m < HE; ++m) yd = fma(od[m], y[m], yd);
            avg += (double)yd;
        }
#endif
#if GQ_TRACE
        // observable trace after steps offset + k * stride (gqis/solver.py:2177-2193)
        if (step >= a.trace_offset && (step - a.trace_offset) % a.trace_stride == 0) {
            const int kk = (step - a.trace_offset) / a.trace_stride;
            if (kk < a.num_trace) {
                real yd = R(0.0);
#pragma unroll
                for (int m = 0; m < HE; ++m) yd = fma(od[m], y[m], yd);
                const double s = wave_sum((double)yd);
                if (lane == 0 && valid) {
                    const real r = (real)s;
                    bad_trace |= !gq_finite(r);
                    gq_store_complex(a.trace, (long long)kk * a.trace_point_stride + q, r, R(0.0));
                }
            }
        }
#endif
    }
#endif  // GQ_ADAPTIVE

    bool bad = false;
#if GQ_OUT_MEAN || GQ_OUT_FINAL
#if GQ_OUT_FINAL
    {
        real yd = R(0.0);
#pragma unroll
        for (int m = 0; m < HE; ++m) yd = fma(od[m], y[m], yd);
        avg = (double)yd;
    }
#endif
    double s = wave_sum(avg);
    if (lane == 0) {
#if GQ_OUT_MEAN && !GQ_ADAPTIVE
        s /= (double)(num_steps - a.warmup_steps);
#endif
        if (failed) s = __builtin_nan("");
        if (valid) {
            const real r = (real)s;
            bad = !gq_finite(r);
            gq_store_complex(a.result, q, r, R(0.0));
        }
    }
#elif GQ_OUT_FINAL_RHO
    if (valid) {
#pragma unroll
        for (int m = 0; m < HE; ++m) {
            const int i = ri[m], j = rj[m];
            if (!used[m]) continue;
            if (i == j && i < HN - 1) {
                ((real*)a.result)[(long long)i * a.point_count + q] = y[m];
                bad |= !gq_finite(y[m]);
            } else if (i < j) {
                const int idx = HN - 1 + 2 * (i * HN - (i * (i + 1)) / 2 + (j - i - 1));
                ((real*)a.result)[(long long)idx * a.point_count + q] = y[m];
                ((real*)a.result)[(long long)(idx + 1) * a.point_count + q] = y[m + HE];
                bad |= !(gq_finite(y[m]) && gq_finite(y[m + HE]));
            }
        }
    }
#endif
    gq_flag(a.status, bad, GQ_STATUS_NONFINITE_RESULT);
    gq_flag(a.status, failed && lane == 0 && valid, GQ_STATUS_SUBSTEP_OVERFLOW);
#if GQ_TRACE
    gq_flag(a.status, bad_trace, GQ_STATUS_NONFINITE_TRACE);
#endif
}
