// Classical RK4 with the state of each parameter point split over GQ_G lanes ("lanes per
// point"): a work-group of GQ_G wavefronts integrates 64 points, wavefront g holding rows
// [R0, R0 + NR) of gq_grp<g> of every vector (state, stage input, RHS, stage sum) for the
// lane's point.  The RHS of a row needs the whole stage state, so each stage writes the wave's
// rows to LDS (point-minor, conflict-free), one barrier, and every wave evaluates its rows from
// there; two LDS buffers alternate, so one barrier per stage suffices (a buffer is rewritten
// only after every wave has passed the next stage's barrier).  Each wave runs the body
// specialised to its group (GQ_GROUP_DISPATCH on the wave-uniform index), so its statics, stage
// values and rows are compile-time register arrays.  This keeps large models (63 states in
// FP64, 255 in FP32) in registers without spills, at the cost of one LDS round trip per stage.
//
// Arithmetic and output semantics are the per-lane RK4 kernel's (gqis/cuda_solvers.py:470-517,
// gqis/solver.py:2126-2193): same stage times from the FP64 time table, k1 pipelining, right-
// endpoint mean over steps >= warmup, traces after steps offset + k * stride.  Observables are
// evaluated by wavefront 0 from the LDS copy of the state.

#define GQ_GW (GQ_G * 64)

__device__ __forceinline__ void gq_time_atoms(const gq_sweep_args& a, const real* __restrict__ C,
                                              int stage_index, real t_local,
                                              real* __restrict__ TT) {
#if GQ_TIME_TABLE
    (void)t_local;
    const real* __restrict__ tab = (const real*)a.time_table;
#pragma unroll
    for (int k = 0; k < GQ_NATOMS; ++k) TT[k] = tab[stage_index * GQ_NATOMS + k];
#else
    (void)stage_index;
    gq_atoms_real(t_local + (real)a.t_in, C, TT);
#endif
}

// The wave's rows of v into the LDS copy of the full vector.
template <int GI>
__device__ __forceinline__ void gq_share(real* __restrict__ Y, const real* __restrict__ v,
                                         int lane) {
    using P = gq_grp<GI>;
#pragma unroll
    for (int r = 0; r < P::NR; ++r) Y[(P::R0 + r) * 64 + lane] = v[r];
}

template <int GI>
__device__ __forceinline__ void gq_rk4_body(const gq_sweep_args& a, real (*s_y)[GQ_D * 64],
                                            const int lane, const long long q, const bool valid,
                                            const real px, const real py,
                                            const gq_point& pt) {
    using P = gq_grp<GI>;
    const real* __restrict__ C = (const real*)a.consts;
    real Sg[P::NS];
    P::statics(px, py, C, Sg);
    real y[P::NR];
#if GQ_RHO0_VALUES
    {
        const real* __restrict__ r0v = (const real*)a.rho0;
#pragma unroll
        for (int r = 0; r < P::NR; ++r) y[r] = r0v[pt.q * GQ_D + P::R0 + r];
    }
#else
    (void)pt;
    P::init(y, px, py, C);
#endif

    const real dt = (real)a.dt;
    const real dt2 = dt * R(0.5);
    const real dt6 = dt / R(6.0);
    const int num_steps = a.num_steps;

    real TT[GQ_NATOMS_ALLOC];
    real Gg[P::NG];
    gq_time_atoms(a, C, 0, R(0.0), TT);
    P::stage(px, py, C, Sg, TT, Gg);
    int buf = 0;
    gq_share<GI>(s_y[buf], y, lane);
    __syncthreads();
    real k[P::NR];
    P::rhs(s_y[buf], lane, px, py, C, Sg, Gg, k);
    buf ^= 1;

#if GQ_OUT_MEAN
    acc_t m_re = 0, m_im = 0;
#endif
#if GQ_OUT_MEAN_STATE
    acc_t ms[P::NR];
#pragma unroll
    for (int r = 0; r < P::NR; ++r) ms[r] = 0;
#endif
    bool trace_bad = false;

    for (int step = 0; step < num_steps; ++step) {
        real t_mid = ((real)step + R(0.5)) * dt;
        gq_time_atoms(a, C, 2 * step + 1, t_mid, TT);
        P::stage(px, py, C, Sg, TT, Gg);
        real yt[P::NR], acc[P::NR];
#pragma unroll
        for (int r = 0; r < P::NR; ++r) {
            yt[r] = fma(dt2, k[r], y[r]);
            acc[r] = k[r];
        }
        gq_share<GI>(s_y[buf], yt, lane);
        __syncthreads();
        P::rhs(s_y[buf], lane, px, py, C, Sg, Gg, k);
        buf ^= 1;
#pragma unroll
        for (int r = 0; r < P::NR; ++r) {
            yt[r] = fma(dt2, k[r], y[r]);
            acc[r] = fma(R(2.0), k[r], acc[r]);
        }
        gq_share<GI>(s_y[buf], yt, lane);
        __syncthreads();
        P::rhs(s_y[buf], lane, px, py, C, Sg, Gg, k);
        buf ^= 1;
#pragma unroll
        for (int r = 0; r < P::NR; ++r) {
            yt[r] = fma(dt, k[r], y[r]);
            acc[r] = fma(R(2.0), k[r], acc[r]);
        }
        t_mid += dt2;
        gq_time_atoms(a, C, 2 * step + 2, t_mid, TT);
        P::stage(px, py, C, Sg, TT, Gg);
        gq_share<GI>(s_y[buf], yt, lane);
        __syncthreads();
        P::rhs(s_y[buf], lane, px, py, C, Sg, Gg, k);
        buf ^= 1;
#pragma unroll
        for (int r = 0; r < P::NR; ++r) {
            acc[r] += k[r];
            y[r] = fma(dt6, acc[r], y[r]);
        }
        // the new state: shared for the next step's k1 and for the observable
        const real* __restrict__ Ycur = s_y[buf];
        gq_share<GI>(s_y[buf], y, lane);
        __syncthreads();
        if (step + 1 < num_steps) P::rhs(Ycur, lane, px, py, C, Sg, Gg, k);
        buf ^= 1;

#if GQ_OUT_MEAN
        if constexpr (GI == 0) {
            if (step >= a.warmup_steps) {
                real o_re, o_im;
                P::obs(Ycur, lane, px, py, C, Sg, Gg, o_re, o_im);
                m_re += (acc_t)o_re;
                m_im += (acc_t)o_im;
            }
        }
#endif
#if GQ_OUT_MEAN_STATE
        if (step >= a.warmup_steps) {
#pragma unroll
            for (int r = 0; r < P::NR; ++r) ms[r] += (acc_t)y[r];
        }
#endif
#if GQ_TRACE
        if (step >= a.trace_offset && (step - a.trace_offset) % a.trace_stride == 0) {
            const int kk = (step - a.trace_offset) / a.trace_stride;
            if (kk < a.num_trace && valid) {
                const long long ld = a.trace_point_stride;
#if GQ_TRACE == 1
                if constexpr (GI == 0) {
                    real o_re, o_im;
                    P::obs(Ycur, lane, px, py, C, Sg, Gg, o_re, o_im);
                    trace_bad |= !(gq_finite(o_re) && gq_finite(o_im));
                    gq_store_complex(a.trace, (long long)kk * ld + q, o_re, o_im);
                }
#else
#pragma unroll
                for (int r = 0; r < P::NR; ++r) {
                    trace_bad |= !gq_finite(y[r]);
                    ((real*)a.trace)[((long long)kk * GQ_D + P::R0 + r) * ld + q] = y[r];
                }
#endif
            }
        }
#endif
    }

    bool bad = false;
    if (valid) {
#if GQ_OUT_MEAN
        if constexpr (GI == 0) {
            const acc_t inv_t = (acc_t)1 / (acc_t)(a.num_steps - a.warmup_steps);
            const real r_re = (real)(m_re * inv_t), r_im = (real)(m_im * inv_t);
            bad = !(gq_finite(r_re) && gq_finite(r_im));
            gq_store_complex(a.result, q, r_re, r_im);
        }
#elif GQ_OUT_FINAL
        if constexpr (GI == 0) {
            // the final state is in the LDS buffer written last (the loop flipped past it)
            real r_re, r_im;
            P::obs(s_y[buf ^ 1], lane, px, py, C, Sg, Gg, r_re, r_im);
            bad = !(gq_finite(r_re) && gq_finite(r_im));
            gq_store_complex(a.result, q, r_re, r_im);
        }
#elif GQ_OUT_FINAL_RHO
#pragma unroll
        for (int r = 0; r < P::NR; ++r) {
            bad |= !gq_finite(y[r]);
            ((real*)a.result)[(long long)(P::R0 + r) * a.point_count + q] = y[r];
        }
#elif GQ_OUT_MEAN_STATE
        const acc_t inv_n = (acc_t)1 / (acc_t)(a.num_steps - a.warmup_steps);
#pragma unroll
        for (int r = 0; r < P::NR; ++r) {
            const real v = (real)(ms[r] * inv_n);
            bad |= !gq_finite(v);
            ((real*)a.result)[(long long)(P::R0 + r) * a.point_count + q] = v;
        }
#endif
    }
    gq_flag(a.status, bad, GQ_STATUS_NONFINITE_RESULT);
#if GQ_TRACE
    gq_flag(a.status, trace_bad, GQ_STATUS_NONFINITE_TRACE);
#else
    (void)trace_bad;
#endif
}

extern "C" __global__ __launch_bounds__(GQ_GW) void GQ_KERNEL(gq_rk4_grouped)(const gq_sweep_args a) {
    __shared__ real s_y[2][GQ_D * 64];
    const int g = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63u);
    const long long q_raw = (long long)blockIdx.x * 64 + lane;
    const bool valid = q_raw < a.point_count;
    // lanes past the end shadow the last point: every wave reaches every barrier, none stores
    const long long q = valid ? q_raw : a.point_count - 1;
    const gq_point pt = gq_locate(a, q);
    const real px = ((const real*)a.par_x)[pt.ix];
    const real py = ((const real*)a.par_y)[pt.iy];
    GQ_GROUP_DISPATCH(g, gq_rk4_body, (a, s_y, lane, q, valid, px, py, pt))
}
