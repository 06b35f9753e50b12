// Adaptive embedded Runge-Kutta between the uniform output points of tlist, one parameter point
// per lane, optionally persistent with work stealing.
//
// Interval semantics of the reference's tsit5a (gqis/cuda_solvers.py:731-908): the proposed
// step h starts at the output step and is kept across intervals; a step snaps to the output
// point when h*1.01 >= the remaining time; accepted steps advance, rejected steps restore the
// state (the FSAL derivative of the step start is kept, so no RHS is recomputed); more than
// max_substeps attempts in one interval make the point NaN and stop it; "mean" outputs use the
// trapezoid rule over accepted steps weighted by hs / ((num_steps - warmup) dt_out).
// Error norms:   Tsit5  sqrt(mean((h * sum e_j k_j / sc)^2)), factor clamp(0.9 err^-1/5, 0.2, 5)
//                DOP853 Hairer's 5(3) combination,            factor clamp(0.9 err^-1/8, 1/3, 6)
// with sc = rtol * max(|y|, |y_new|) + atol.  Time and step controller are FP64; the error sums
// and norm are in the RHS precision `real` (in FP32 kernels FP32, as the reference's FP32 kernel,
// gqis/cuda_solvers.py:868-877); the state and stage sums are FP64 when GQ_STATE_FP64.
//
// Persistent mode (GQ_PERSISTENT): the grid is the number of resident work-groups; a lane whose
// point finished takes the next one from a global counter (one atomic per wavefront and refill
// round), so a wavefront does not wait for its slowest point.

#if GQ_STATE_FP64
typedef double st_t;
#else
typedef real st_t;
#endif

#define GQ_NS GQ_NSTAGES

// The derivative at the step start (FSAL, kept for a rejected step's retry) lives in LDS, one
// column per lane, when the work-group's copy fits in 64 KB: it is read by the first stages and
// the fold and written once per accepted step, and its registers go to the stages.
#if GQ_BLOCK * GQ_D * (GQ_FP64 ? 8 : 4) <= 65536
#define GQ_K0_LDS 1
#else
#define GQ_K0_LDS 0
#endif

// stage values at an FP64 time (the atoms from gq_atoms_stage: FP64, or FP32 functions of
// FP64-reduced angles in FP32 kernels)
__device__ __forceinline__ void gq_stage_at(double t_abs, real px, real py,
                                            const real* __restrict__ C, const gq_static_t& S,
                                            gq_stage_t& G) {
    real TT[GQ_NATOMS_ALLOC];
    gq_atoms_stage(t_abs, C, TT);
    gq_stage(px, py, C, S, TT, G);
}

__device__ __forceinline__ void gq_eval(const st_t* __restrict__ ys, real px, real py,
                                        const real* __restrict__ C, const gq_static_t& S,
                                        const gq_stage_t& G, real* __restrict__ k) {
    real yr[GQ_D];
#pragma unroll
    for (int i = 0; i < GQ_D; ++i) yr[i] = (real)ys[i];
    gq_rhs(yr, px, py, C, S, G, k);
}

__device__ __forceinline__ void gq_observe(const st_t* __restrict__ y, real px, real py,
                                           const real* __restrict__ C, const gq_static_t& S,
                                           const gq_stage_t& G, real& re, real& im) {
    real yr[GQ_D];
#pragma unroll
    for (int i = 0; i < GQ_D; ++i) yr[i] = (real)y[i];
    gq_obs(yr, px, py, C, S, G, re, im);
}

extern "C" __global__ __launch_bounds__(GQ_BLOCK) void GQ_KERNEL(gq_adaptive)(const gq_sweep_args a) {
    const double dt_out = a.dt;
    const double rtol = a.solver_param[0], atol = a.solver_param[1];
    const int max_sub = (int)a.solver_param[2];
    const int num_steps = a.num_steps;
    const real* __restrict__ C = (const real*)a.consts;
    const int lane = (int)(threadIdx.x & 63u);
    const long long count = a.point_count;
    const double inv_avg_time = 1.0 / ((double)(num_steps - a.warmup_steps) * dt_out);

    long long q = -1;
    bool live = false, exhausted = false;
    real px = R(0.0), py = R(0.0);
    gq_static_t S;
    gq_stage_t G;
    st_t y[GQ_D];
#if GQ_K0_LDS
    __shared__ real s_k0[GQ_D][GQ_BLOCK];
#define GQ_K0(i) s_k0[i][threadIdx.x]
#else
    real k0[GQ_D];
#define GQ_K0(i) k0[i]
#endif
    double tau = 0.0, h = dt_out;
    int step = 0, sub = 0, n_acc = 0, n_rej = 0;
    bool failed = false, bad_res = false, bad_trace = false, overflow = false;
#if GQ_OUT_MEAN
    acc_t avg_re = 0, avg_im = 0;
    real op_re = R(0.0), op_im = R(0.0);
#endif
#if GQ_OUT_MEAN_STATE
    acc_t savg[GQ_D];
#endif
#if !GQ_PERSISTENT
    bool assigned = false;
#endif

    while (true) {
        // ---- take a new point -------------------------------------------------------------
        const bool want = !live && !exhausted;
#if GQ_PERSISTENT
        const unsigned long long m = __ballot(want);
        if (m != 0ull) {
            const int leader = __ffsll((long long)m) - 1;
            unsigned int base = 0u;
            if (lane == leader) base = atomicAdd(a.work_cursor, (unsigned int)__popcll(m));
            base = __shfl(base, leader);
            if (want) {
                const unsigned long long below = m & ((1ull << lane) - 1ull);
                q = (long long)base + __popcll(below);
            }
        }
#else
        if (want) {
            q = assigned ? count : (long long)blockIdx.x * GQ_BLOCK + threadIdx.x;
            assigned = true;
        }
#endif
        if (want) {
            if (q >= count) {
                exhausted = true;
            } else {
                const gq_point pt = gq_locate(a, q);
                px = ((const real*)a.par_x)[pt.ix];
                py = ((const real*)a.par_y)[pt.iy];
                gq_statics(px, py, C, S);
                real y0[GQ_D];
                gq_init_state(y0, a, pt, px, py, C);
#pragma unroll
                for (int i = 0; i < GQ_D; ++i) y[i] = (st_t)y0[i];
                tau = 0.0;
                h = dt_out;
                step = 0;
                sub = 0;
                n_acc = n_rej = 0;
                failed = false;
                gq_stage_at(a.t_in, px, py, C, S, G);
                {
                    real kstart[GQ_D];
                    gq_eval(y, px, py, C, S, G, kstart);
#pragma unroll
                    for (int i = 0; i < GQ_D; ++i) GQ_K0(i) = kstart[i];
                }
#if GQ_OUT_MEAN
                avg_re = avg_im = 0;
                gq_observe(y, px, py, C, S, G, op_re, op_im);
#endif
#if GQ_OUT_MEAN_STATE
#pragma unroll
                for (int i = 0; i < GQ_D; ++i) savg[i] = 0;
#endif
                live = true;
            }
        }
        if (!__any(live)) break;

        bool done = live && step >= num_steps;
        if (live && !done) {
            // ---- one step attempt ---------------------------------------------------------
            const double remaining = dt_out - tau;
            const bool last = h * 1.01 >= remaining;
            const double hs = last ? remaining : h;
            const double t0 = (double)step * dt_out + tau;
            // Folded stage storage: k_0 .. k_{F-1} (F = GQ_FOLD) stay as they are until stage F
            // is formed; then they are folded into the partial sums P[r] of the later stages and
            // into the b and error sums, and every later k_s is folded as soon as it exists.  Each
            // sum still runs over j in ascending order, so stage values and y_new are those of
            // the unfolded form, while at most about F + 2 derivative-sized arrays are live
            // instead of all GQ_NS derivatives, y_new and f(y_new) at once.
            real k[GQ_NS][GQ_D];
            st_t P[GQ_NS][GQ_D];
            st_t bsum[GQ_D];
            en_t esum[GQ_NERR][GQ_D];
            // fold k_s into the later stages' partial sums and the b and error sums
            auto fold = [&](const int s) {
#pragma unroll
                for (int i = 0; i < GQ_D; ++i) {
#pragma unroll
                    for (int r = s + 1; r < GQ_NS; ++r)
                        if (GQ_A[r][s] != 0.0) P[r][i] = fma((st_t)GQ_A[r][s], (st_t)k[s][i], P[r][i]);
                    if (GQ_B[s] != 0.0) bsum[i] = fma((st_t)GQ_B[s], (st_t)k[s][i], bsum[i]);
#pragma unroll
                    for (int m = 0; m < GQ_NERR; ++m)
                        if (gq_err_w(m, s) != 0.0)
                            esum[m][i] = fma((en_t)gq_err_w(m, s), (en_t)k[s][i], esum[m][i]);
                }
            };
            // k_0 is read where it is used, so in LDS it holds no registers between uses
            auto kv = [&](const int j, const int i) -> real { return j == 0 ? GQ_K0(i) : k[j][i]; };
            // stages 1 .. F-1 from the stored derivatives
#pragma unroll
            for (int s = 1; s < GQ_FOLD; ++s) {
                st_t ys[GQ_D];
#pragma unroll
                for (int i = 0; i < GQ_D; ++i) {
                    st_t acc = 0;
#pragma unroll
                    for (int j = 0; j < s; ++j)
                        if (GQ_A[s][j] != 0.0) acc = fma((st_t)GQ_A[s][j], (st_t)kv(j, i), acc);
                    ys[i] = fma((st_t)hs, acc, y[i]);
                }
                gq_stage_at(a.t_in + (t0 + GQ_C[s] * hs), px, py, C, S, G);
                gq_eval(ys, px, py, C, S, G, k[s]);
            }
            // stage F from the stored derivatives, which are then folded
            {
                st_t ys[GQ_D];
#pragma unroll
                for (int i = 0; i < GQ_D; ++i) {
                    st_t acc = 0;
#pragma unroll
                    for (int j = 0; j < GQ_FOLD; ++j)
                        if (GQ_A[GQ_FOLD][j] != 0.0) acc = fma((st_t)GQ_A[GQ_FOLD][j], (st_t)kv(j, i), acc);
                    ys[i] = fma((st_t)hs, acc, y[i]);
#pragma unroll
                    for (int r = GQ_FOLD + 1; r < GQ_NS; ++r) {
                        st_t p = 0;
#pragma unroll
                        for (int j = 0; j < GQ_FOLD; ++j)
                            if (GQ_A[r][j] != 0.0) p = fma((st_t)GQ_A[r][j], (st_t)kv(j, i), p);
                        P[r][i] = p;
                    }
                    st_t b = 0;
#pragma unroll
                    for (int j = 0; j < GQ_FOLD; ++j)
                        if (GQ_B[j] != 0.0) b = fma((st_t)GQ_B[j], (st_t)kv(j, i), b);
                    bsum[i] = b;
#pragma unroll
                    for (int m = 0; m < GQ_NERR; ++m) {
                        en_t e = 0;
#pragma unroll
                        for (int j = 0; j < GQ_FOLD; ++j)
                            if (gq_err_w(m, j) != 0.0) e = fma((en_t)gq_err_w(m, j), (en_t)kv(j, i), e);
                        esum[m][i] = e;
                    }
                }
                gq_stage_at(a.t_in + (t0 + GQ_C[GQ_FOLD] * hs), px, py, C, S, G);
                gq_eval(ys, px, py, C, S, G, k[GQ_FOLD]);
                fold(GQ_FOLD);
            }
            // stages F+1 .. GQ_NS-1 from the partial sums
#pragma unroll
            for (int s = GQ_FOLD + 1; s < GQ_NS; ++s) {
                st_t ys[GQ_D];
#pragma unroll
                for (int i = 0; i < GQ_D; ++i) ys[i] = fma((st_t)hs, P[s][i], y[i]);
                gq_stage_at(a.t_in + (t0 + GQ_C[s] * hs), px, py, C, S, G);
                gq_eval(ys, px, py, C, S, G, k[s]);
                fold(s);
            }
            st_t yn[GQ_D];
#pragma unroll
            for (int i = 0; i < GQ_D; ++i) yn[i] = fma((st_t)hs, bsum[i], y[i]);
            // G now holds the stage values at t0 + hs (the last stage has c = 1)
            double err_norm;
            const en_t rtol_e = (en_t)rtol, atol_e = (en_t)atol;
#if GQ_METHOD_DOP853
            {
                en_t s5 = 0, s3 = 0;
#pragma unroll
                for (int i = 0; i < GQ_D; ++i) {
                    const en_t sc = fma(rtol_e, fmax(fabs((en_t)y[i]), fabs((en_t)yn[i])), atol_e);
                    const en_t e5 = gq_err_ratio(esum[0][i], sc), e3 = gq_err_ratio(esum[1][i], sc);
                    s5 = fma(e5, e5, s5);
                    s3 = fma(e3, e3, s3);
                }
                double deno = (double)s5 + 0.01 * (double)s3;
                if (!(deno > 0.0)) deno = 1.0;
                err_norm = fabs(hs) * (double)s5 / sqrt(deno * (double)GQ_D);
            }
            real kf[GQ_D];   // FSAL derivative f(y_new), computed after acceptance
#else
            real kf[GQ_D];
            gq_eval(yn, px, py, C, S, G, kf);
            {
                const en_t hse = (en_t)hs;
                en_t ssum = 0;
#pragma unroll
                for (int i = 0; i < GQ_D; ++i) {
                    const en_t e = hse * fma((en_t)GQ_E[GQ_NS], (en_t)kf[i], esum[0][i]);
                    const en_t sc = fma(rtol_e, fmax(fabs((en_t)y[i]), fabs((en_t)yn[i])), atol_e);
                    const en_t er = gq_err_ratio(e, sc);
                    ssum = fma(er, er, ssum);
                }
                err_norm = sqrt((double)ssum / (double)GQ_D);
            }
#endif
            const double factor = gq_step_factor(err_norm);   // step_control.hip
            if (err_norm <= 1.0) {
                ++n_acc;
#if GQ_METHOD_DOP853
                gq_eval(yn, px, py, C, S, G, kf);
#endif
#if GQ_OUT_MEAN
                real o_re, o_im;
                gq_observe(yn, px, py, C, S, G, o_re, o_im);
                if (step >= a.warmup_steps) {
                    const acc_t wgt = (acc_t)(0.5 * hs * inv_avg_time);
                    avg_re = fma(wgt, (acc_t)o_re + (acc_t)op_re, avg_re);
                    avg_im = fma(wgt, (acc_t)o_im + (acc_t)op_im, avg_im);
                }
                op_re = o_re;
                op_im = o_im;
#endif
#if GQ_OUT_MEAN_STATE
                if (step >= a.warmup_steps) {
                    const acc_t wgt = (acc_t)(0.5 * hs * inv_avg_time);
#pragma unroll
                    for (int i = 0; i < GQ_D; ++i)
                        savg[i] = fma(wgt, (acc_t)yn[i] + (acc_t)y[i], savg[i]);
                }
#endif
#pragma unroll
                for (int i = 0; i < GQ_D; ++i) {
                    y[i] = yn[i];
                    GQ_K0(i) = kf[i];
                }
                tau = last ? dt_out : tau + hs;
                h = last ? fmax(h, hs * factor) : hs * factor;
            } else {
                ++n_rej;
                h = hs * factor;
            }
            if (++sub > max_sub) {
                failed = true;
#pragma unroll
                for (int i = 0; i < GQ_D; ++i) y[i] = (st_t)__builtin_nan("");
            }
            if (failed || tau >= dt_out) {
                // end of an output interval (or failure inside it): trace sample of this step
#if GQ_TRACE
                if (step >= a.trace_offset && (step - a.trace_offset) % a.trace_stride == 0) {
                    const int kk = (step - a.trace_offset) / a.trace_stride;
                    if (kk < a.num_trace) {
                        const long long ld = a.trace_point_stride;
#if GQ_TRACE == 1
                        real o_re, o_im;
                        gq_observe(y, px, py, C, S, G, o_re, o_im);
                        bad_trace |= !(gq_finite(o_re) && gq_finite(o_im));
                        gq_store_complex(a.trace, (long long)kk * ld + q, o_re, o_im);
#else
#pragma unroll
                        for (int i = 0; i < GQ_D; ++i) {
                            bad_trace |= !gq_finite((real)y[i]);
                            ((real*)a.trace)[((long long)kk * GQ_D + i) * ld + q] = (real)y[i];
                        }
#endif
                    }
                }
#endif
                ++step;
                tau = 0.0;
                sub = 0;
                done = failed || step >= num_steps;
            }
        }
        if (done) {
            // ---- point finished: outputs --------------------------------------------------
            overflow |= failed;
#if GQ_OUT_MEAN
            const real r_re = failed ? (real)__builtin_nan("") : (real)avg_re;
            const real r_im = failed ? (real)__builtin_nan("") : (real)avg_im;
            bad_res |= !(gq_finite(r_re) && gq_finite(r_im));
            gq_store_complex(a.result, q, r_re, r_im);
#elif GQ_OUT_FINAL
            real r_re, r_im;
            gq_observe(y, px, py, C, S, G, r_re, r_im);
            bad_res |= !(gq_finite(r_re) && gq_finite(r_im));
            gq_store_complex(a.result, q, r_re, r_im);
#elif GQ_OUT_FINAL_RHO
#pragma unroll
            for (int i = 0; i < GQ_D; ++i) {
                const real v = (real)y[i];
                bad_res |= !gq_finite(v);
                ((real*)a.result)[(long long)i * a.point_count + q] = v;
            }
#elif GQ_OUT_MEAN_STATE
#pragma unroll
            for (int i = 0; i < GQ_D; ++i) {
                const real v = failed ? (real)__builtin_nan("") : (real)savg[i];
                bad_res |= !gq_finite(v);
                ((real*)a.result)[(long long)i * a.point_count + q] = v;
            }
#endif
#if GQ_COUNT_STEPS
            a.adaptive_steps[2 * q] = n_acc;
            a.adaptive_steps[2 * q + 1] = n_rej;
#endif
            live = false;
        }
    }
    gq_flag(a.status, bad_res, GQ_STATUS_NONFINITE_RESULT);
    gq_flag(a.status, bad_trace, GQ_STATUS_NONFINITE_TRACE);
    gq_flag(a.status, overflow, GQ_STATUS_SUBSTEP_OVERFLOW);
}
