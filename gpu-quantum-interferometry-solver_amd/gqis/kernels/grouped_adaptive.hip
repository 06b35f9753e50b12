// Adaptive embedded Runge-Kutta (Tsit5 / DOP853) between the output points, with the state
// of each parameter point split over GQ_G lanes (one per wavefront of the work-group, rows
// [R0, R0 + NR) of gq_grp<g>), as in grouped_rk4.hip.  Semantics are those of the per-lane
// adaptive kernel (adaptive.hip, after gqis/cuda_solvers.py:731-908): interval snapping,
// FSAL derivative kept on rejection, FP64 time and controller, error sums in the RHS
// precision (the waves' partials added in FP64), folded stage storage, NaN on max_substeps,
// trapezoid mean.
//
// Every wave of a group keeps an identical copy of each lane's controller state (time, step
// size, counters).  Stage inputs go through LDS as in the grouped RK4 kernel; the error norm
// is a sum over rows, so each wave writes its partial sum and every wave adds the GQ_G
// partials in the same order: the norm, and with it every accept / reject / step-size
// decision, is bitwise the same in all waves.  That sameness is needed for correct results,
// not for progress: the branches that contain barriers (loop exit, trace sample, point
// completion) are taken on a work-group vote (gq_group_vote), which every wave reads
// identically, so all waves reach the same barriers even if their copies disagreed.  Points
// are assigned one per lane for the whole integration (no work stealing across groups).
//
// The time atoms of a step's stage times (gq_atoms_stage: gq_sincos64 in FP64 kernels, an FP64
// reduction and FP32 polynomials in FP32 kernels) are the same in every
// wave: wave (s - 1) % GQ_G evaluates those of stage s at the step start and writes them to
// LDS before the step's first barrier, after which every wave reads them (s_tt,
// double-buffered by attempt parity).  The values are those of the one-lane kernel.

#if GQ_STATE_FP64
typedef double st_t;
#else
typedef real st_t;
#endif

#define GQ_NS GQ_NSTAGES
#define GQ_GW (GQ_G * 64)

template <int GI>
__device__ __forceinline__ void gq_stage_at(double t_abs, real px, real py,
                                            const real* __restrict__ C,
                                            const real* __restrict__ Sg, real* __restrict__ Gg) {
    real TT[GQ_NATOMS_ALLOC];
    gq_atoms_stage(t_abs, C, TT);
    gq_grp<GI>::stage(px, py, C, Sg, TT, Gg);
}

// stage values from the atoms of one stage time shared in LDS (tt[k * 64 + lane])
template <int GI>
__device__ __forceinline__ void gq_stage_lds(const double* __restrict__ tt, int lane, real px,
                                             real py, const real* __restrict__ C,
                                             const real* __restrict__ Sg, real* __restrict__ Gg) {
    real TT[GQ_NATOMS_ALLOC];
#pragma unroll
    for (int i = 0; i < GQ_NATOMS; ++i) TT[i] = (real)tt[i * 64 + lane];
    gq_grp<GI>::stage(px, py, C, Sg, TT, Gg);
}

#define GQ_TT_STRIDE (GQ_NATOMS_ALLOC * 64)

// The wave's rows of v into the LDS copy of the state; with a single copy (GQ_NBUF 1) after a
// barrier that waits for the readers of the previous contents.
template <int GI>
__device__ __forceinline__ void gq_share(real* __restrict__ Y, const st_t* __restrict__ v,
                                         int lane) {
    using P = gq_grp<GI>;
#if GQ_NBUF == 1
    __syncthreads();
#endif
#pragma unroll
    for (int r = 0; r < P::NR; ++r) Y[(P::R0 + r) * 64 + lane] = (real)v[r];
}

// Work-group OR of each lane's vote bits.  Every wave ORs the bits over its lanes, writes that
// to its slot of vote[parity], and after one barrier ORs the GQ_G slots: all waves get the same
// word, so a branch taken on it is uniform over the work-group.  Callers alternate the parity:
// a slot row is rewritten two votes later, after every wave has passed the barrier of the vote
// in between (and with it has read this row).
#define GQ_VOTE_BITS 3
__device__ __forceinline__ unsigned gq_group_vote(unsigned (*vote)[GQ_G], int parity, int g,
                                                  unsigned bits) {
    unsigned wave = 0;
#pragma unroll
    for (int b = 0; b < GQ_VOTE_BITS; ++b)
        if (__any((bits >> b) & 1u)) wave |= 1u << b;
    vote[parity][g] = wave;   // every lane stores the same wave-uniform word
    __syncthreads();
    unsigned all = 0;
#pragma unroll
    for (int w = 0; w < GQ_G; ++w) all |= vote[parity][w];
    return all;
}
#define GQ_VOTE_SAMPLE 1u
#define GQ_VOTE_DONE 2u
#define GQ_VOTE_LIVE 4u

// Sum of the waves' partial sums for this lane, in wave order (identical in every wave).
__device__ __forceinline__ double gq_group_sum(double* __restrict__ red, double part, int g,
                                               int lane) {
    red[g * 64 + lane] = part;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < GQ_G; ++w) s += red[w * 64 + lane];
    __syncthreads();   // the partials are rewritten by the next reduction
    return s;
}

template <int GI>
__device__ __forceinline__ void gq_adaptive_body(const gq_sweep_args& a,
                                                 real (*s_y)[GQ_D * 64], double* s_red,
                                                 double* s_tt, unsigned (*s_vote)[GQ_G],
                                                 const int lane, const long long q,
                                                 const bool valid, const real px, const real py,
                                                 const gq_point& pt) {
    using P = gq_grp<GI>;
    const double dt_out = a.dt;
    const double rtol = a.solver_param[0], atol = a.solver_param[1];
    const int max_sub = (int)a.solver_param[2];
    const int num_steps = a.num_steps;
    const real* __restrict__ C = (const real*)a.consts;
    const double inv_avg_time = 1.0 / ((double)(num_steps - a.warmup_steps) * dt_out);

    real Sg[P::NS];
    P::statics(px, py, C, Sg);
    real Gg[P::NG];
    st_t y[P::NR];
    {
        real y0[P::NR];
#if GQ_RHO0_VALUES
        const real* __restrict__ r0v = (const real*)a.rho0;
#pragma unroll
        for (int r = 0; r < P::NR; ++r) y0[r] = r0v[pt.q * GQ_D + P::R0 + r];
#else
        (void)pt;
        P::init(y0, px, py, C);
#endif
#pragma unroll
        for (int r = 0; r < P::NR; ++r) y[r] = (st_t)y0[r];
    }
    double tau = 0.0, h = dt_out;
    int step = 0, sub = 0, n_acc = 0, n_rej = 0;
    bool failed = false, bad_trace = false;
    bool live = valid && num_steps > 0;
    int buf = 0;

    gq_stage_at<GI>(a.t_in, px, py, C, Sg, Gg);
    gq_share<GI>(s_y[buf], y, lane);
    __syncthreads();
    real k0[P::NR];
    P::rhs(s_y[buf], lane, px, py, C, Sg, Gg, k0);
#if GQ_OUT_MEAN
    acc_t avg_re = 0, avg_im = 0;
    real op_re = R(0.0), op_im = R(0.0);
    if constexpr (GI == 0) P::obs(s_y[buf], lane, px, py, C, Sg, Gg, op_re, op_im);
#endif
    buf = GQ_NEXT_BUF(buf);
#if GQ_OUT_MEAN_STATE
    acc_t savg[P::NR];
#pragma unroll
    for (int r = 0; r < P::NR; ++r) savg[r] = 0;
#endif

    int attempt = 0;   // the same in every wave: selects the s_tt half of this attempt
    int vote_parity = 0;
    // lanes with nothing to do (past the end, or zero steps) take part in the barriers only
    unsigned group = gq_group_vote(s_vote, vote_parity, GI, live ? GQ_VOTE_LIVE : 0u);
    vote_parity ^= 1;
    while (group & GQ_VOTE_LIVE) {
        const double remaining = dt_out - tau;
        const bool last = h * 1.01 >= remaining;
        const double hs = last ? remaining : h;
        const double t0 = (double)step * dt_out + tau;
        double* __restrict__ tt = s_tt + (attempt & 1) * GQ_NS * GQ_TT_STRIDE;
        ++attempt;
#if GQ_NATOMS > 0
#pragma unroll
        for (int s = 1; s < GQ_NS; ++s) {
            if ((s - 1) % GQ_G == GI) {
                real TTs[GQ_NATOMS_ALLOC];
                gq_atoms_stage(a.t_in + (t0 + GQ_C[s] * hs), C, TTs);
#pragma unroll
                for (int i = 0; i < GQ_NATOMS; ++i)
                    tt[s * GQ_TT_STRIDE + i * 64 + lane] = (double)TTs[i];
            }
        }
#endif
        // Folded stage storage, as in the one-lane adaptive kernel: k_0 .. k_{F-1} are kept
        // until stage F is formed, then folded into the later stages' partial sums and the b
        // and error sums; each later k_s is folded as soon as it exists, every sum in ascending j.
        real k[GQ_NS][P::NR];
        st_t Ps[GQ_NS][P::NR];
        st_t bsum[P::NR];
        en_t esum[GQ_NERR][P::NR];
        auto fold = [&](const int s) {
#pragma unroll
            for (int r = 0; r < P::NR; ++r) {
#pragma unroll
                for (int u = s + 1; u < GQ_NS; ++u)
                    if (GQ_A[u][s] != 0.0) Ps[u][r] = fma((st_t)GQ_A[u][s], (st_t)k[s][r], Ps[u][r]);
                if (GQ_B[s] != 0.0) bsum[r] = fma((st_t)GQ_B[s], (st_t)k[s][r], bsum[r]);
#pragma unroll
                for (int m = 0; m < GQ_NERR; ++m)
                    if (gq_err_w(m, s) != 0.0)
                        esum[m][r] = fma((en_t)gq_err_w(m, s), (en_t)k[s][r], esum[m][r]);
            }
        };
#pragma unroll
        for (int r = 0; r < P::NR; ++r) k[0][r] = k0[r];
        // stages 1 .. F-1 from the stored derivatives
#pragma unroll
        for (int s = 1; s < GQ_FOLD; ++s) {
            st_t ys[P::NR];
#pragma unroll
            for (int r = 0; r < P::NR; ++r) {
                st_t acc = 0;
#pragma unroll
                for (int j = 0; j < s; ++j)
                    if (GQ_A[s][j] != 0.0) acc = fma((st_t)GQ_A[s][j], (st_t)k[j][r], acc);
                ys[r] = fma((st_t)hs, acc, y[r]);
            }
            gq_share<GI>(s_y[buf], ys, lane);
            __syncthreads();   // the first one of the step also publishes tt
            gq_stage_lds<GI>(tt + s * GQ_TT_STRIDE, lane, px, py, C, Sg, Gg);
            P::rhs(s_y[buf], lane, px, py, C, Sg, Gg, k[s]);
            buf = GQ_NEXT_BUF(buf);
        }
        // stage F from the stored derivatives, which are then folded
        {
            st_t ys[P::NR];
#pragma unroll
            for (int r = 0; r < P::NR; ++r) {
                st_t acc = 0;
#pragma unroll
                for (int j = 0; j < GQ_FOLD; ++j)
                    if (GQ_A[GQ_FOLD][j] != 0.0) acc = fma((st_t)GQ_A[GQ_FOLD][j], (st_t)k[j][r], acc);
                ys[r] = fma((st_t)hs, acc, y[r]);
#pragma unroll
                for (int u = GQ_FOLD + 1; u < GQ_NS; ++u) {
                    st_t p = 0;
#pragma unroll
                    for (int j = 0; j < GQ_FOLD; ++j)
                        if (GQ_A[u][j] != 0.0) p = fma((st_t)GQ_A[u][j], (st_t)k[j][r], p);
                    Ps[u][r] = p;
                }
                st_t b = 0;
#pragma unroll
                for (int j = 0; j < GQ_FOLD; ++j)
                    if (GQ_B[j] != 0.0) b = fma((st_t)GQ_B[j], (st_t)k[j][r], b);
                bsum[r] = b;
#pragma unroll
                for (int m = 0; m < GQ_NERR; ++m) {
                    en_t e = 0;
#pragma unroll
                    for (int j = 0; j < GQ_FOLD; ++j)
                        if (gq_err_w(m, j) != 0.0) e = fma((en_t)gq_err_w(m, j), (en_t)k[j][r], e);
                    esum[m][r] = e;
                }
            }
            gq_share<GI>(s_y[buf], ys, lane);
            __syncthreads();
            gq_stage_lds<GI>(tt + GQ_FOLD * GQ_TT_STRIDE, lane, px, py, C, Sg, Gg);
            P::rhs(s_y[buf], lane, px, py, C, Sg, Gg, k[GQ_FOLD]);
            buf = GQ_NEXT_BUF(buf);
            fold(GQ_FOLD);
        }
        // stages F+1 .. GQ_NS-1 from the partial sums
#pragma unroll
        for (int s = GQ_FOLD + 1; s < GQ_NS; ++s) {
            st_t ys[P::NR];
#pragma unroll
            for (int r = 0; r < P::NR; ++r) ys[r] = fma((st_t)hs, Ps[s][r], y[r]);
            gq_share<GI>(s_y[buf], ys, lane);
            __syncthreads();
            gq_stage_lds<GI>(tt + s * GQ_TT_STRIDE, lane, px, py, C, Sg, Gg);
            P::rhs(s_y[buf], lane, px, py, C, Sg, Gg, k[s]);
            buf = GQ_NEXT_BUF(buf);
            fold(s);
        }
        st_t yn[P::NR];
#pragma unroll
        for (int r = 0; r < P::NR; ++r) yn[r] = fma((st_t)hs, bsum[r], y[r]);
        // Gg now holds the stage values at t0 + hs (the last stage has c = 1)
        const real* __restrict__ Yn = s_y[buf];
        gq_share<GI>(s_y[buf], yn, lane);
        __syncthreads();
        buf = GQ_NEXT_BUF(buf);
        real kf[P::NR];
        double err_norm;
        const en_t rtol_e = (en_t)rtol, atol_e = (en_t)atol;
#if GQ_METHOD_DOP853
        {
            en_t s5 = 0, s3 = 0;
#pragma unroll
            for (int r = 0; r < P::NR; ++r) {
                const en_t sc = fma(rtol_e, fmax(fabs((en_t)y[r]), fabs((en_t)yn[r])), atol_e);
                const en_t e5 = gq_err_ratio(esum[0][r], sc), e3 = gq_err_ratio(esum[1][r], sc);
                s5 = fma(e5, e5, s5);
                s3 = fma(e3, e3, s3);
            }
            const double s5g = gq_group_sum(s_red, (double)s5, GI, lane);
            const double s3g = gq_group_sum(s_red, (double)s3, GI, lane);
            double deno = s5g + 0.01 * s3g;
            if (!(deno > 0.0)) deno = 1.0;
            err_norm = fabs(hs) * s5g / sqrt(deno * (double)GQ_D);
        }
#else
        P::rhs(Yn, lane, px, py, C, Sg, Gg, kf);
        {
            const en_t hse = (en_t)hs;
            en_t ssum = 0;
#pragma unroll
            for (int r = 0; r < P::NR; ++r) {
                const en_t e = hse * fma((en_t)GQ_E[GQ_NS], (en_t)kf[r], esum[0][r]);
                const en_t sc = fma(rtol_e, fmax(fabs((en_t)y[r]), fabs((en_t)yn[r])), atol_e);
                const en_t er = gq_err_ratio(e, sc);
                ssum = fma(er, er, ssum);
            }
            const double ssg = gq_group_sum(s_red, (double)ssum, GI, lane);
            err_norm = sqrt(ssg / (double)GQ_D);
        }
#endif
        const double factor = gq_step_factor(err_norm);   // step_control.hip
        const bool accept = live && err_norm <= 1.0;
#if GQ_METHOD_DOP853
        P::rhs(Yn, lane, px, py, C, Sg, Gg, kf);
#endif
#if GQ_OUT_MEAN
        real o_re = R(0.0), o_im = R(0.0);
        if constexpr (GI == 0) P::obs(Yn, lane, px, py, C, Sg, Gg, o_re, o_im);
#endif
        if (accept) {
            ++n_acc;
#if GQ_OUT_MEAN
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
                for (int r = 0; r < P::NR; ++r)
                    savg[r] = fma(wgt, (acc_t)yn[r] + (acc_t)y[r], savg[r]);
            }
#endif
#pragma unroll
            for (int r = 0; r < P::NR; ++r) {
                y[r] = yn[r];
                k0[r] = kf[r];
            }
            tau = last ? dt_out : tau + hs;
            h = last ? fmax(h, hs * factor) : hs * factor;
        } else if (live) {
            ++n_rej;
            h = hs * factor;
        }
        bool interval_end = false;
        if (live) {
            if (++sub > max_sub) {
                failed = true;
#pragma unroll
                for (int r = 0; r < P::NR; ++r) y[r] = (st_t)__builtin_nan("");
            }
            interval_end = failed || tau >= dt_out;
        }
        // this attempt's output interval; the trace sample at its end (or at a failure)
        const int interval = step;
#if GQ_TRACE
        const bool sample = interval_end && interval >= a.trace_offset &&
                            (interval - a.trace_offset) % a.trace_stride == 0 &&
                            (interval - a.trace_offset) / a.trace_stride < a.num_trace;
#else
        const bool sample = false;
#endif
        if (interval_end) {
            ++step;
            tau = 0.0;
            sub = 0;
        }
        const bool done = live && (failed || step >= num_steps);
        group = gq_group_vote(s_vote, vote_parity, GI,
                              (sample ? GQ_VOTE_SAMPLE : 0u) | (done ? GQ_VOTE_DONE : 0u) |
                                  (live && !done ? GQ_VOTE_LIVE : 0u));
        vote_parity ^= 1;
#if GQ_TRACE
        if (group & GQ_VOTE_SAMPLE) {
            const long long ld = a.trace_point_stride;
            const int kk = (interval - a.trace_offset) / a.trace_stride;
#if GQ_TRACE == 1
            gq_share<GI>(s_y[buf], y, lane);
            __syncthreads();
            if constexpr (GI == 0) {
                if (sample) {
                    real t_re, t_im;
                    P::obs(s_y[buf], lane, px, py, C, Sg, Gg, t_re, t_im);
                    bad_trace |= !(gq_finite(t_re) && gq_finite(t_im));
                    gq_store_complex(a.trace, (long long)kk * ld + q, t_re, t_im);
                }
            }
            buf = GQ_NEXT_BUF(buf);
#else
            if (sample) {
#pragma unroll
                for (int r = 0; r < P::NR; ++r) {
                    bad_trace |= !gq_finite((real)y[r]);
                    ((real*)a.trace)[((long long)kk * GQ_D + P::R0 + r) * ld + q] = (real)y[r];
                }
            }
#endif
        }
#endif
        if (group & GQ_VOTE_DONE) {
            // ---- points finished: outputs -----------------------------------------------
#if GQ_OUT_FINAL
            gq_share<GI>(s_y[buf], y, lane);
            __syncthreads();
#endif
            bool bad = false;
            if (done) {
#if GQ_OUT_MEAN
                if constexpr (GI == 0) {
                    const real r_re = failed ? (real)__builtin_nan("") : (real)avg_re;
                    const real r_im = failed ? (real)__builtin_nan("") : (real)avg_im;
                    bad = !(gq_finite(r_re) && gq_finite(r_im));
                    gq_store_complex(a.result, q, r_re, r_im);
                }
#elif GQ_OUT_FINAL
                if constexpr (GI == 0) {
                    real r_re, r_im;
                    P::obs(s_y[buf], lane, px, py, C, Sg, Gg, r_re, r_im);
                    bad = !(gq_finite(r_re) && gq_finite(r_im));
                    gq_store_complex(a.result, q, r_re, r_im);
                }
#elif GQ_OUT_FINAL_RHO
#pragma unroll
                for (int r = 0; r < P::NR; ++r) {
                    const real v = (real)y[r];
                    bad |= !gq_finite(v);
                    ((real*)a.result)[(long long)(P::R0 + r) * a.point_count + q] = v;
                }
#elif GQ_OUT_MEAN_STATE
#pragma unroll
                for (int r = 0; r < P::NR; ++r) {
                    const real v = failed ? (real)__builtin_nan("") : (real)savg[r];
                    bad |= !gq_finite(v);
                    ((real*)a.result)[(long long)(P::R0 + r) * a.point_count + q] = v;
                }
#endif
#if GQ_COUNT_STEPS
                if constexpr (GI == 0) {
                    a.adaptive_steps[2 * q] = n_acc;
                    a.adaptive_steps[2 * q + 1] = n_rej;
                }
#endif
                live = false;
            }
#if GQ_OUT_FINAL
            buf = GQ_NEXT_BUF(buf);
#endif
            gq_flag(a.status, bad, GQ_STATUS_NONFINITE_RESULT);
            if constexpr (GI == 0) gq_flag(a.status, done && failed, GQ_STATUS_SUBSTEP_OVERFLOW);
        }
    }
    gq_flag(a.status, bad_trace, GQ_STATUS_NONFINITE_TRACE);
}

extern "C" __global__ __launch_bounds__(GQ_GW) void GQ_KERNEL(gq_adaptive_grouped)(const gq_sweep_args a) {
    __shared__ real s_y[GQ_NBUF][GQ_D * 64];
    __shared__ double s_red[GQ_G * 64];
    __shared__ double s_tt[2 * GQ_NS * GQ_TT_STRIDE];
    __shared__ unsigned s_vote[2][GQ_G];
    const int g = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63u);
    const long long q_raw = (long long)blockIdx.x * 64 + lane;
    const bool valid = q_raw < a.point_count;
    const long long q = valid ? q_raw : a.point_count - 1;
    const gq_point pt = gq_locate(a, q);
    const real px = ((const real*)a.par_x)[pt.ix];
    const real py = ((const real*)a.par_y)[pt.iy];
    GQ_GROUP_DISPATCH(g, gq_adaptive_body, (a, s_y, s_red, s_tt, s_vote, lane, q, valid, px, py, pt))
}
