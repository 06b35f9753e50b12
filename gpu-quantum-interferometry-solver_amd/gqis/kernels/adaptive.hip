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
// with sc = rtol * max(|y|, |y_new|) + atol.  Time, error norm and step controller are FP64; the
// state and stage sums are FP64 when GQ_STATE_FP64, the RHS is evaluated in `real`.
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
    real k0[GQ_D];
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
                gq_eval(y, px, py, C, S, G, k0);
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
            real k[GQ_NS][GQ_D];
#pragma unroll
            for (int i = 0; i < GQ_D; ++i) k[0][i] = k0[i];
#pragma unroll
            for (int s = 1; s < GQ_NS; ++s) {
                st_t ys[GQ_D];
#pragma unroll
                for (int i = 0; i < GQ_D; ++i) {
                    st_t acc = 0;
#pragma unroll
                    for (int j = 0; j < s; ++j)
                        if (GQ_A[s][j] != 0.0) acc = fma((st_t)GQ_A[s][j], (st_t)k[j][i], acc);
                    ys[i] = fma((st_t)hs, acc, y[i]);
                }
                gq_stage_at(a.t_in + (t0 + GQ_C[s] * hs), px, py, C, S, G);
                gq_eval(ys, px, py, C, S, G, k[s]);
            }
            st_t yn[GQ_D];
#pragma unroll
            for (int i = 0; i < GQ_D; ++i) {
                st_t acc = 0;
#pragma unroll
                for (int j = 0; j < GQ_NS; ++j)
                    if (GQ_B[j] != 0.0) acc = fma((st_t)GQ_B[j], (st_t)k[j][i], acc);
                yn[i] = fma((st_t)hs, acc, y[i]);
            }
            // G now holds the stage values at t0 + hs (the last stage has c = 1)
            double err_norm;
#if GQ_METHOD_DOP853
            {
                double s5 = 0.0, s3 = 0.0;
#pragma unroll
                for (int i = 0; i < GQ_D; ++i) {
                    double e5 = 0.0, e3 = 0.0;
#pragma unroll
                    for (int j = 0; j < GQ_NS; ++j) {
                        if (GQ_E5[j] != 0.0) e5 = fma(GQ_E5[j], (double)k[j][i], e5);
                        if (GQ_E3[j] != 0.0) e3 = fma(GQ_E3[j], (double)k[j][i], e3);
                    }
                    const double sc = fma(rtol, fmax(fabs((double)y[i]), fabs((double)yn[i])), atol);
                    s5 = fma(e5 / sc, e5 / sc, s5);
                    s3 = fma(e3 / sc, e3 / sc, s3);
                }
                double deno = s5 + 0.01 * s3;
                if (!(deno > 0.0)) deno = 1.0;
                err_norm = fabs(hs) * s5 / sqrt(deno * (double)GQ_D);
            }
            real kf[GQ_D];   // FSAL derivative f(y_new), computed after acceptance
#else
            real kf[GQ_D];
            gq_eval(yn, px, py, C, S, G, kf);
            {
                double ssum = 0.0;
#pragma unroll
                for (int i = 0; i < GQ_D; ++i) {
                    double e = GQ_E[GQ_NS] * (double)kf[i];
#pragma unroll
                    for (int j = 0; j < GQ_NS; ++j) e = fma(GQ_E[j], (double)k[j][i], e);
                    e *= hs;
                    const double sc = fma(rtol, fmax(fabs((double)y[i]), fabs((double)yn[i])), atol);
                    ssum = fma(e / sc, e / sc, ssum);
                }
                err_norm = sqrt(ssum / (double)GQ_D);
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
                    k0[i] = kf[i];
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
