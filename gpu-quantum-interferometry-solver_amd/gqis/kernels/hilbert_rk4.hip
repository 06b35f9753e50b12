// RK4 on the full density matrix with the commutator as one complex GEMM on MFMA (FP32), or
// as VALU dot products over an LDS copy of the state (FP64, GQ_FP64 = 1).
//
// Model structure (checked on the host): H(t) = H0 + diag(v(ParX, ParY, t)) with H0 constant,
// collapse operators diagonal or single-entry, observable diagonal.  With
// H_eff = H0 - (i/2) sum L^+L (constant, complex) and Y = -i H_eff rho - i diag(v) rho,
//     d rho/dt = Y + Y^+ + sum_k d_k d_k^T o rho  (diagonal L = diag(d_k))
//              + sum |l|^2 rho_bb |a><a|           (single-entry L = l |a><b|).
// X = H_eff rho is a real GEMM Hbig x Rbig with Hbig = [[Re, -Im], [Im, Re]] (2N x 2N) and
// Rbig = [Re rho; Im rho] (2N x N).  One wavefront holds P = 32/N points side by side as the 32
// columns of Rbig and runs v_mfma_f32_32x32x2f32 over K = 2N in N steps.  The K order is
// permuted so that lane (c, h) feeds at step s the s-th value it holds in the accumulator
// layout (rows 32t + 8q + 4h + r of column c): the output of one RHS is the operand of the
// next with no data movement; the matching permutation of Hbig is packed on the host
// (GQ_APACK, T x N x 64 floats, staged in LDS).  Y^+ needs Y transposed: a per-wavefront LDS
// tile with padded rows (conflict-free).  RK4 stage times, k1 pipelining and output
// semantics are the reference's (gqis/cuda_solvers.py:470-517, gqis/solver.py:2126-2152).
//
// Lane (c, h): point p = c / N, column j = c % N; it owns the N/2 complex elements
// (i_m, j), i_m = 8 (m / 4) + 4 h + (m % 4), m < N/2, with Re in register m and Im in m + N/2.

typedef float gq_f32x16 __attribute__((ext_vector_type(16)));

// Split precision (GQ_MFMA_SPLIT 1: FP16, 2: BF16): the GEMM as Ahi*xhi + Ahi*xlo + Alo*xhi on
// v_mfma_f32_32x32x16_{f16,bf16} with FP32 accumulation (the lo*lo term is dropped; FP16 parts
// carry 11 + 11 significant bits, BF16 8 + 8).  K runs in blocks of 16: element j of lane
// (c, h) in block b is the state value x[8b + j] that lane holds (Rbig row kappa(8b + j, h)),
// and the A operand is packed on the host in the same order, so again no data movement.
#ifndef GQ_MFMA_SPLIT
#define GQ_MFMA_SPLIT 0
#endif
#if GQ_MFMA_SPLIT == 1
typedef _Float16 gq_half;
#define GQ_MFMA_HALF(a, b, c) __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0)
#elif GQ_MFMA_SPLIT == 2
typedef __bf16 gq_half;
#define GQ_MFMA_HALF(a, b, c) __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0)
#endif
#if GQ_MFMA_SPLIT
typedef gq_half gq_half8 __attribute__((ext_vector_type(8)));
#endif

#define HN GQ_HN                 // Hilbert dimension N (8, 16 or 32)
#define HP (32 / HN)             // points per wavefront
#define HT ((2 * HN + 31) / 32)  // 32x32 output tiles
#define HM (HN / 2)              // complex elements per lane
#define HWAVES (GQ_BLOCK / 64)
#define HLD (HN + 1)             // padded LDS row
#define HNB (HN / 8)             // K blocks of the split-precision products

__device__ __forceinline__ int gq_row_of(int m, int h) { return 8 * (m >> 2) + 4 * h + (m & 3); }
// the part of gq_row_of that does not depend on the lane: a compile-time constant per m
__device__ __forceinline__ constexpr int gq_row_off(int m) { return 8 * (m >> 2) + (m & 3); }

// Lean register layout (N = 32, FP32 and split products): the drive row coefficients, the
// dephasing weights and the Y parts between the two passes of the RHS are read from LDS
// instead of registers, and the LDS addresses derive from four bases hidden from the compiler
// (gq_opaque), recomputed per RHS as base + a constant folded into the LDS instruction instead
// of one hoisted register per address.  Two wavefronts per SIMD then fit (250 registers, no
// spills): 256^2, 2000 steps, FP32 1.393 -> 1.242 s, FP16 hi/lo 0.765 -> 0.632 s.  At N = 16
// the same layout ran 9-16% slower (the extra LDS reads cost more than a third wavefront
// gains), so smaller N keep the values in registers.
#define GQ_LEAN (HN == 32 && !GQ_FP64)
__device__ __forceinline__ int gq_opaque(int v) {
#if GQ_LEAN
    asm volatile("" : "+v"(v));
#endif
    return v;
}

// the split-precision forms carry their product type in the name (gq_hilbert_rk4_f16x3_f32)
#if GQ_MFMA_SPLIT == 1
#define GQ_HILBERT_NAME gq_hilbert_rk4_f16x3
#elif GQ_MFMA_SPLIT == 2
#define GQ_HILBERT_NAME gq_hilbert_rk4_bf16x3
#else
#define GQ_HILBERT_NAME gq_hilbert_rk4
#endif
// At N = 32 two wavefronts per SIMD fit in 250 registers (no spills) once the state is kept
// in three vectors and the LDS addresses derive from four opaque bases; without the hint the
// compiler spends 281 and runs one wavefront, with no second one to issue VALU work while
// the matrix cores run.
#if HN == 32 && !GQ_FP64
#define GQ_HILBERT_OCCUPANCY __attribute__((amdgpu_waves_per_eu(2)))
#else
#define GQ_HILBERT_OCCUPANCY
#endif
extern "C" __global__ __launch_bounds__(GQ_BLOCK) GQ_HILBERT_OCCUPANCY void GQ_KERNEL(GQ_HILBERT_NAME)(const gq_sweep_args a) {
#if GQ_FP64
    __shared__ real s_hbig[2 * HN * 2 * HN];
    __shared__ real s_x[HWAVES][HP * 2 * HN * HN];
#elif GQ_MFMA_SPLIT
    __shared__ gq_half8 s_ahi[HT * HNB * 64];
    __shared__ gq_half8 s_alo[HT * HNB * 64];
#else
    __shared__ float s_apack[HT * HN * 64];
#endif
    __shared__ real s_y[HWAVES][2][HP * HN * HLD];
#if GQ_NMONO
    // general jump operators (one nonzero per row): a per-wavefront copy of the state and the
    // operators' row tables, (L rho L^+)_ij = l_i conj(l_j) rho[col_i][col_j]
    __shared__ real s_xm[HWAVES][2][HP * HN * HLD];
    __shared__ int s_mcol[GQ_NMONO * HN];
    __shared__ real s_mre[GQ_NMONO * HN];
    __shared__ real s_mim[GQ_NMONO * HN];
#endif
    __shared__ real s_diag[HWAVES][HP * HN];
    __shared__ real s_dep[HN * HN];           // dephasing weights sum_k d_k(i) d_k(j)
    __shared__ real s_vco[GQ_NBASIS * HN];    // drive row coefficients per basis term
    __shared__ real s_jw[HN * GQ_JMAX];       // population transfer into a: weights |l|^2 ...
    __shared__ int s_jb[HN * GQ_JMAX];        // ... from the populations b (sparse rows)
    __shared__ double s_red[HWAVES][64];

    const real* __restrict__ cst = (const real*)a.consts;   // packed operator constants
#if GQ_FP64
    for (int e = threadIdx.x; e < 4 * HN * HN; e += GQ_BLOCK) s_hbig[e] = cst[GQ_OFF_HBIG + e];
#elif GQ_MFMA_SPLIT
    for (int e = threadIdx.x; e < HT * HNB * 64 * 8; e += GQ_BLOCK) {
        // the FP32 constants hold FP16 / BF16 values exactly
        ((gq_half*)s_ahi)[e] = (gq_half)cst[GQ_OFF_AHI + e];
        ((gq_half*)s_alo)[e] = (gq_half)cst[GQ_OFF_ALO + e];
    }
#else
    for (int e = threadIdx.x; e < HT * HN * 64; e += GQ_BLOCK) s_apack[e] = cst[GQ_OFF_APACK + e];
#endif
    for (int e = threadIdx.x; e < HN * HN; e += GQ_BLOCK) s_dep[e] = cst[GQ_OFF_DEP + e];
    for (int e = threadIdx.x; e < GQ_NBASIS * HN; e += GQ_BLOCK) s_vco[e] = cst[GQ_OFF_VCOEF + e];
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
    const int c = lane & 31, h = lane >> 5;
    const int pw = c / HN, j = c % HN;
    const long long q = ((long long)blockIdx.x * HWAVES + wave) * HP + pw;
    const bool valid = q < a.point_count;
    const long long qc = valid ? q : 0;
    const gq_point pt = gq_locate(a, qc);
    const real px = ((const real*)a.par_x)[pt.ix];
    const real py = ((const real*)a.par_y)[pt.iy];
    real* __restrict__ ybuf_r = s_y[wave][0] + pw * HN * HLD;
    real* __restrict__ ybuf_i = s_y[wave][1] + pw * HN * HLD;
    real* __restrict__ dbuf = s_diag[wave] + pw * HN;
#if GQ_FP64
    real* __restrict__ xcol = s_x[wave] + pw * 2 * HN * HN;   // Rbig of this point, [row][j]
#endif

    // ---- initial state: rho0 values (reduced layout) or |0><0| ------------------------------
    real y[HN];
#pragma unroll
    for (int m = 0; m < HM; ++m) {
        const int i = gq_row_of(m, h);
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
        y[m + HM] = im;
    }

#if !GQ_LEAN
    // per-row diagonal drive coefficients v_i = sum_k coef[k][i] g_k(px, py, t)
    real vco[GQ_NBASIS][HM];
#pragma unroll
    for (int k = 0; k < GQ_NBASIS; ++k)
#pragma unroll
        for (int m = 0; m < HM; ++m) vco[k][m] = s_vco[k * HN + gq_row_of(m, h)];
#endif
    real odiag[HM];
#pragma unroll
    for (int m = 0; m < HM; ++m) odiag[m] = cst[GQ_OFF_OBS + gq_row_of(m, h)];

    // Per-lane constants of the dissipator, kept in registers: for the lane holding its
    // column's diagonal element (m = mdiag), that population's jump list (zero weights
    // elsewhere), so the RHS has no per-element branch.  The drive row coefficients and the
    // dephasing weights are read from LDS when used: in registers they took 16 (1 + NBASIS)
    // of them at N = 32, enough to keep the kernel at one wavefront per SIMD.
#if GQ_NMONO
    real* __restrict__ xm_r = s_xm[wave][0] + pw * HN * HLD;
    real* __restrict__ xm_i = s_xm[wave][1] + pw * HN * HLD;
    int mcj[GQ_NMONO];         // the general jump operators at the lane's column j
    real mlr[GQ_NMONO], mli[GQ_NMONO];
#pragma unroll
    for (int kk = 0; kk < GQ_NMONO; ++kk) {
        mcj[kk] = s_mcol[kk * HN + j];
        mlr[kk] = s_mre[kk * HN + j];
        mli[kk] = s_mim[kk * HN + j];
    }
#endif
    int mdiag = -1;
#pragma unroll
    for (int m = 0; m < HM; ++m)
        if (gq_row_of(m, h) == j) mdiag = m;
#if !GQ_LEAN
    real depw[HM];
#pragma unroll
    for (int m = 0; m < HM; ++m) depw[m] = s_dep[gq_row_of(m, h) * HN + j];
#endif
    real od = R(0.0);   // observable weight of that diagonal element (0 without one)
#pragma unroll
    for (int m = 0; m < HM; ++m) od = (m == mdiag) ? odiag[m] : od;
    real jwl[GQ_JMAX];
    int jbl[GQ_JMAX];
#pragma unroll
    for (int e = 0; e < GQ_JMAX; ++e) {
        jwl[e] = (mdiag >= 0) ? s_jw[j * GQ_JMAX + e] : R(0.0);
        jbl[e] = (mdiag >= 0) ? s_jb[j * GQ_JMAX + e] : 0;
    }

    const real dt = (real)a.dt;
    const real dt2 = dt * R(0.5);
    const real dt6 = dt / R(6.0);
    const int num_steps = a.num_steps;
    real g[GQ_NBASIS];   // drive basis terms at the current stage time
#if !GQ_LEAN
    real v[HM];          // the diagonal drive of the lane's rows at that time
#endif

    auto stage_values = [&](int row) {
        const real* __restrict__ tab = (const real*)a.time_table;
        real TT[GQ_NATOMS_ALLOC];
#pragma unroll
        for (int k = 0; k < GQ_NATOMS; ++k) TT[k] = tab[row * GQ_NATOMS + k];
        gq_basis(px, py, TT, cst, g);
#if !GQ_LEAN
#pragma unroll
        for (int m = 0; m < HM; ++m) {
            real s = R(0.0);
#pragma unroll
            for (int k = 0; k < GQ_NBASIS; ++k) s = fma(vco[k][m], g[k], s);
            v[m] = s;
        }
#endif
    };

    // the RHS at the stage input x, written over x: each element is read before the same
    // element is written, so RK4 keeps three state vectors, not four
    auto rhs = [&](real* x) {
        real* k = x;
        real X[HN];   // rows row(m, h) of H_eff rho in the Rbig layout, column c
#if GQ_FP64
#pragma unroll
        for (int m = 0; m < HN; ++m) xcol[gq_row_of(m, h) * HN + j] = x[m];
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        // rows in blocks of 8 with k inside: each state value read from LDS serves 8 rows of
        // the lane (the Hbig entries are the same for the 32 lanes of a half: broadcast reads)
#pragma unroll
        for (int mb = 0; mb < HN; mb += 8) {
            real xs[8];
#pragma unroll
            for (int m = 0; m < 8; ++m) xs[m] = R(0.0);
#pragma unroll 2
            for (int kk = 0; kk < 2 * HN; ++kk) {
                const real xv = xcol[kk * HN + j];
#pragma unroll
                for (int m = 0; m < 8; ++m)
                    xs[m] = fma(s_hbig[gq_row_of(mb + m, h) * 2 * HN + kk], xv, xs[m]);
            }
#pragma unroll
            for (int m = 0; m < 8; ++m) X[mb + m] = xs[m];
        }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
#elif GQ_MFMA_SPLIT
        gq_f32x16 acc[HT];
#pragma unroll
        for (int t = 0; t < HT; ++t)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[t][e] = 0.0f;
#pragma unroll
        for (int b = 0; b < HNB; ++b) {
            gq_half8 xh, xl;
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) {
                const float v = x[8 * b + jj];
                const gq_half hv = (gq_half)v;
                xh[jj] = hv;
                xl[jj] = (gq_half)(v - (float)hv);
            }
#pragma unroll
            for (int t = 0; t < HT; ++t) {
                const gq_half8 ah = s_ahi[(t * HNB + b) * 64 + lane];
                const gq_half8 al = s_alo[(t * HNB + b) * 64 + lane];
                acc[t] = GQ_MFMA_HALF(al, xh, acc[t]);
                acc[t] = GQ_MFMA_HALF(ah, xl, acc[t]);
                acc[t] = GQ_MFMA_HALF(ah, xh, acc[t]);
            }
        }
#pragma unroll
        for (int m = 0; m < HN; ++m) X[m] = acc[m >> 4][m & 15];
#else
        gq_f32x16 acc[HT];
#pragma unroll
        for (int t = 0; t < HT; ++t)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[t][e] = 0.0f;
#pragma unroll
        for (int s = 0; s < HN; ++s) {
#pragma unroll
            for (int t = 0; t < HT; ++t) {
                const float av = s_apack[(t * HN + s) * 64 + lane];
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, x[s], acc[t], 0, 0, 0);
            }
        }
#pragma unroll
        for (int m = 0; m < HN; ++m) X[m] = acc[m >> 4][m & 15];
#endif
        // element m of the lane is (i, j) with i = gq_row_off(m) + 4 h: its tile index is
        // b_own + gq_row_off(m) HLD, its transpose's b_tr + gq_row_off(m)
        const int b_own = gq_opaque(4 * h * HLD + j);
        const int b_tr = gq_opaque(j * HLD + 4 * h);
        const int b_dep = gq_opaque(4 * h * HN + j);
        const int b_row = gq_opaque(4 * h);
        real xd = R(0.0);   // the lane's diagonal element (population), if it holds one
#if !GQ_LEAN
        real yr[HM], yi[HM];
#endif
#pragma unroll
        for (int m = 0; m < HM; ++m) {
            const real xr = X[m];
            const real xi = X[m + HM];
            const int own = b_own + gq_row_off(m) * HLD;
#if GQ_LEAN
            // v_i = sum_k coef[k][i] g_k (the diagonal drive of row i)
            real vm = R(0.0);
#pragma unroll
            for (int kb = 0; kb < GQ_NBASIS; ++kb)
                vm = fma(s_vco[kb * HN + b_row + gq_row_off(m)], g[kb], vm);
            // Y = -i X - i v rho, into the transpose tile (read back below)
            ybuf_r[own] = fma(vm, x[m + HM], xi);
            ybuf_i[own] = -fma(vm, x[m], xr);
#else
            (void)b_row;
            yr[m] = fma(v[m], x[m + HM], xi);
            yi[m] = -fma(v[m], x[m], xr);
            ybuf_r[own] = yr[m];
            ybuf_i[own] = yi[m];
#endif
#if GQ_NMONO
            xm_r[own] = x[m];
            xm_i[own] = x[m + HM];
#endif
            xd = (m == mdiag) ? x[m] : xd;
        }
        if (mdiag >= 0) dbuf[j] = xd;
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        real jsum = R(0.0);
#pragma unroll
        for (int e = 0; e < GQ_JMAX; ++e) jsum = fma(jwl[e], dbuf[jbl[e]], jsum);
#pragma unroll
        for (int m = 0; m < HM; ++m) {
            const int i = gq_row_of(m, h);
            const int tr = b_tr + gq_row_off(m);
#if GQ_LEAN
            const int own = b_own + gq_row_off(m) * HLD;
            real dr = ybuf_r[own] + ybuf_r[tr];
            real di = ybuf_i[own] - ybuf_i[tr];
            const real dw = s_dep[b_dep + gq_row_off(m) * HN];
#else
            (void)b_dep;
            real dr = yr[m] + ybuf_r[tr];
            real di = yi[m] - ybuf_i[tr];
            const real dw = depw[m];
#endif
            dr = fma(dw, x[m], dr);
            di = fma(dw, x[m + HM], di);
#if GQ_NMONO
#pragma unroll
            for (int kk = 0; kk < GQ_NMONO; ++kk) {
                const int ci = s_mcol[kk * HN + i];
                if (ci >= 0 && mcj[kk] >= 0) {
                    // w = l_i conj(l_j), times rho[col_i][col_j]
                    const real lr = s_mre[kk * HN + i], li = s_mim[kk * HN + i];
                    const real wr = fma(lr, mlr[kk], li * mli[kk]);
                    const real wi = fma(li, mlr[kk], -lr * mli[kk]);
                    const real pr = xm_r[ci * HLD + mcj[kk]], pim = xm_i[ci * HLD + mcj[kk]];
                    dr = fma(wr, pr, fma(-wi, pim, dr));
                    di = fma(wr, pim, fma(wi, pr, di));
                }
            }
#endif
            k[m] = (m == mdiag) ? dr + jsum : dr;
            k[m + HM] = di;
        }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    };

    // sum over the 2N lanes of each point (valid in the lanes with h == 0, j == 0); every lane
    // of the wavefront takes part
    auto point_sum = [&](double part) -> double {
        s_red[wave][lane] = part;
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        double s = 0.0;
        if (h == 0 && j == 0)
            for (int l = 0; l < HN; ++l) s += s_red[wave][pw * HN + l] + s_red[wave][32 + pw * HN + l];
        return s;
    };
    bool bad_trace = false;
    (void)bad_trace;

    stage_values(0);
    real kb[HN];   // k1 of a step, then each stage input and, in place, its derivative
#pragma unroll
    for (int i = 0; i < HN; ++i) kb[i] = y[i];
    rhs(kb);
    double avg = 0.0;
    for (int step = 0; step < num_steps; ++step) {
        stage_values(2 * step + 1);
        real acc[HN];
#pragma unroll
        for (int i = 0; i < HN; ++i) {
            acc[i] = kb[i];
            kb[i] = fma(dt2, kb[i], y[i]);
        }
        rhs(kb);
#pragma unroll
        for (int i = 0; i < HN; ++i) {
            acc[i] = fma(R(2.0), kb[i], acc[i]);
            kb[i] = fma(dt2, kb[i], y[i]);
        }
        rhs(kb);
#pragma unroll
        for (int i = 0; i < HN; ++i) {
            acc[i] = fma(R(2.0), kb[i], acc[i]);
            kb[i] = fma(dt, kb[i], y[i]);
        }
        stage_values(2 * step + 2);
        rhs(kb);
#pragma unroll
        for (int i = 0; i < HN; ++i) {
            acc[i] += kb[i];
            y[i] = fma(dt6, acc[i], y[i]);
        }
        if (step + 1 < num_steps) {
#pragma unroll
            for (int i = 0; i < HN; ++i) kb[i] = y[i];
            rhs(kb);
        }
#if GQ_OUT_MEAN
        if (step >= a.warmup_steps) {
            // the lane's diagonal element (if any) weighted by the observable's diagonal
            real yd = R(0.0);
#pragma unroll
            for (int m = 0; m < HM; ++m) yd = (m == mdiag) ? y[m] : yd;
            avg += (double)(od * yd);
        }
#endif
#if GQ_TRACE
        // observable trace after steps offset + k * stride (gqis/solver.py:2177-2193)
        if (step >= a.trace_offset && (step - a.trace_offset) % a.trace_stride == 0) {
            const int kk = (step - a.trace_offset) / a.trace_stride;
            if (kk < a.num_trace) {
                real yd = R(0.0);
#pragma unroll
                for (int m = 0; m < HM; ++m) yd = (m == mdiag) ? y[m] : yd;
                const double s = point_sum((double)(od * yd));
                if (h == 0 && j == 0 && valid) {
                    const real r = (real)s;
                    bad_trace |= !gq_finite(r);
                    gq_store_complex(a.trace, (long long)kk * a.trace_point_stride + q, r, R(0.0));
                }
            }
        }
#endif
    }

    bool bad = false;
#if GQ_OUT_MEAN || GQ_OUT_FINAL
#if GQ_OUT_FINAL
    {
        real yd = R(0.0);
#pragma unroll
        for (int m = 0; m < HM; ++m) yd = (m == mdiag) ? y[m] : yd;
        avg = (double)(od * yd);
    }
#endif
    // sum the partial traces of the 2N lanes of each point
    double s = point_sum(avg);
    if (h == 0 && j == 0) {
#if GQ_OUT_MEAN
        s /= (double)(num_steps - a.warmup_steps);
#endif
        if (valid) {
            const real r = (real)s;
            bad = !gq_finite(r);
            gq_store_complex(a.result, q, r, R(0.0));
        }
    }
#elif GQ_OUT_FINAL_RHO
    if (valid) {
#pragma unroll
        for (int m = 0; m < HM; ++m) {
            const int i = gq_row_of(m, h);
            int idx = -1;
            if (i == j && i < HN - 1) {
                ((real*)a.result)[(long long)i * a.point_count + q] = y[m];
                bad |= !gq_finite(y[m]);
            } else if (i < j) {
                idx = HN - 1 + 2 * (i * HN - (i * (i + 1)) / 2 + (j - i - 1));
                ((real*)a.result)[(long long)idx * a.point_count + q] = y[m];
                ((real*)a.result)[(long long)(idx + 1) * a.point_count + q] = y[m + HM];
                bad |= !(gq_finite(y[m]) && gq_finite(y[m + HM]));
            }
        }
    }
#endif
    gq_flag(a.status, bad, GQ_STATUS_NONFINITE_RESULT);
#if GQ_TRACE
    gq_flag(a.status, bad_trace, GQ_STATUS_NONFINITE_TRACE);
#endif
}
