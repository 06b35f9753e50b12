// Shared pieces of the fixed-step grouped ("lanes per point") kernels: GQ_G wavefronts of a
// work-group integrate 64 points, wavefront g holding rows [R0, R0 + NR) of gq_grp<g> of every
// vector.  A stage writes each wave's rows of its input to LDS (point-minor, conflict-free), one
// barrier, and each wave evaluates its rows of the RHS from there.  The output semantics are
// the per-lane kernels' (gqis/solver.py:2126-2193): right-endpoint mean over the steps >= warmup,
// traces after steps offset + k * stride; observables are evaluated by wavefront 0 from the LDS
// copy of the state.

#define GQ_GW (GQ_G * 64)

// The wave's rows of v into the LDS copy of the full vector.  With two alternating copies
// (GQ_NBUF 2) a copy is rewritten only after every wave has passed the barrier that follows the
// write of the other one.  With one copy (models whose two copies would not fit the LDS) a
// barrier first waits until every wave has finished reading the previous contents.
template <int GI>
__device__ __forceinline__ void gq_share(real* __restrict__ Y, const real* __restrict__ v,
                                         int lane) {
    using P = gq_grp<GI>;
#if GQ_NBUF == 1
    __syncthreads();
#endif
#pragma unroll
    for (int r = 0; r < P::NR; ++r) Y[(P::R0 + r) * 64 + lane] = v[r];
}

template <int GI>
struct gq_gout {
#if GQ_OUT_MEAN
    acc_t re, im;
#endif
#if GQ_OUT_MEAN_STATE
    acc_t ms[gq_grp<GI>::NR];
#endif
    bool trace_bad;
};

template <int GI>
__device__ __forceinline__ void gq_gout_init(gq_gout<GI>& o) {
#if GQ_OUT_MEAN
    o.re = 0;
    o.im = 0;
#endif
#if GQ_OUT_MEAN_STATE
#pragma unroll
    for (int r = 0; r < gq_grp<GI>::NR; ++r) o.ms[r] = 0;
#endif
    o.trace_bad = false;
}

// After step `step`: Ycur is the LDS copy of the new state (written before a barrier), y the
// wave's rows of it, Gg the stage values at the step's end time.
template <int GI>
__device__ __forceinline__ void gq_gout_step(gq_gout<GI>& o, const gq_sweep_args& a, int step,
                                             long long q, bool valid,
                                             const real* __restrict__ Ycur,
                                             const real* __restrict__ y, int lane, real px,
                                             real py, const real* __restrict__ C,
                                             const real* __restrict__ Sg,
                                             const real* __restrict__ Gg) {
    using P = gq_grp<GI>;
    (void)Ycur; (void)lane; (void)px; (void)py; (void)C; (void)Sg; (void)Gg; (void)q; (void)valid;
#if GQ_OUT_MEAN
    if constexpr (GI == 0) {
        if (step >= a.warmup_steps) {
            real o_re, o_im;
            P::obs(Ycur, lane, px, py, C, Sg, Gg, o_re, o_im);
            o.re += (acc_t)o_re;
            o.im += (acc_t)o_im;
        }
    }
#endif
#if GQ_OUT_MEAN_STATE
    if (step >= a.warmup_steps) {
#pragma unroll
        for (int r = 0; r < P::NR; ++r) o.ms[r] += (acc_t)y[r];
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
                o.trace_bad |= !(gq_finite(o_re) && gq_finite(o_im));
                gq_store_complex(a.trace, (long long)kk * ld + q, o_re, o_im);
            }
#else
#pragma unroll
            for (int r = 0; r < P::NR; ++r) {
                o.trace_bad |= !gq_finite(y[r]);
                ((real*)a.trace)[((long long)kk * GQ_D + P::R0 + r) * ld + q] = y[r];
            }
#endif
        }
    }
#endif
}

// End of the integration: Yfin is the LDS copy of the final state (used for "final").
template <int GI>
__device__ __forceinline__ void gq_gout_final(gq_gout<GI>& o, const gq_sweep_args& a,
                                              long long q, bool valid,
                                              const real* __restrict__ Yfin,
                                              const real* __restrict__ y, int lane, real px,
                                              real py, const real* __restrict__ C,
                                              const real* __restrict__ Sg,
                                              const real* __restrict__ Gg) {
    using P = gq_grp<GI>;
    (void)Yfin; (void)lane; (void)px; (void)py; (void)C; (void)Sg; (void)Gg; (void)y;
    bool bad = false;
    if (valid) {
#if GQ_OUT_MEAN
        if constexpr (GI == 0) {
            const acc_t inv_t = (acc_t)1 / (acc_t)(a.num_steps - a.warmup_steps);
            const real r_re = (real)(o.re * inv_t), r_im = (real)(o.im * inv_t);
            bad = !(gq_finite(r_re) && gq_finite(r_im));
            gq_store_complex(a.result, q, r_re, r_im);
        }
#elif GQ_OUT_FINAL
        if constexpr (GI == 0) {
            real r_re, r_im;
            P::obs(Yfin, lane, px, py, C, Sg, Gg, r_re, r_im);
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
            const real v = (real)(o.ms[r] * inv_n);
            bad |= !gq_finite(v);
            ((real*)a.result)[(long long)(P::R0 + r) * a.point_count + q] = v;
        }
#endif
    }
    gq_flag(a.status, bad, GQ_STATUS_NONFINITE_RESULT);
#if GQ_TRACE
    gq_flag(a.status, o.trace_bad, GQ_STATUS_NONFINITE_TRACE);
#endif
}

// Initial rows of the wave: explicit per-point values or the model's expressions.
template <int GI>
__device__ __forceinline__ void gq_ginit(real* __restrict__ y, const gq_sweep_args& a,
                                         const gq_point& pt, real px, real py,
                                         const real* __restrict__ C) {
    using P = gq_grp<GI>;
#if GQ_RHO0_VALUES
    (void)px; (void)py; (void)C;
    const real* __restrict__ r0v = (const real*)a.rho0;
#pragma unroll
    for (int r = 0; r < P::NR; ++r) y[r] = r0v[pt.q * GQ_D + P::R0 + r];
#else
    (void)a; (void)pt;
    P::init(y, px, py, C);
#endif
}
