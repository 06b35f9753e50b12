// Shared device helpers of the sweep kernels (prepended to every solver template).

// Point coordinates of local index q of a launch.
struct gq_point {
    long long q, p, ix, iy;
};

__device__ __forceinline__ gq_point gq_locate(const gq_sweep_args& a, long long q) {
    gq_point pt;
    pt.q = q;
    pt.p = a.point_begin + q;
    if (a.param_mode == 1) {          // per-point parameter lists (merged request batches)
        pt.ix = q;
        pt.iy = q;
    } else {
        pt.ix = pt.p / a.num_y;
        pt.iy = pt.p - pt.ix * a.num_y;
    }
    return pt;
}

#if !GQ_HILBERT && !GQ_GROUPED
// Initial state: explicit per-point values, else the model's expressions (|0><0| by default).
__device__ __forceinline__ void gq_init_state(real* __restrict__ y, const gq_sweep_args& a,
                                              const gq_point& pt, real px, real py,
                                              const real* __restrict__ C) {
#if GQ_RHO0_VALUES
    const real* __restrict__ r0 = (const real*)a.rho0;
#pragma unroll
    for (int i = 0; i < GQ_D; ++i) y[i] = r0[pt.q * GQ_D + i];
#else
    gq_init_expr(y, px, py, C);
#endif
}

#endif

__device__ __forceinline__ bool gq_finite(real v) { return isfinite(v); }

// One atomic OR per wavefront for the lanes that saw a problem.
__device__ __forceinline__ void gq_flag(unsigned int* status, bool bad, unsigned int bit) {
    const unsigned long long m = __ballot(bad);
    if (m != 0ull) {
        const int leader = __ffsll((long long)m) - 1;
        if ((int)(threadIdx.x & 63u) == leader) atomicOr(status, bit);
    }
}

// Status word and work cursor back to zero: the first node of a replayed launch chain (a
// kernel rather than a memset, which a replayed graph must not contain).
extern "C" __global__ void GQ_KERNEL(gq_clear_flags)(const gq_sweep_args a) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (a.status) *a.status = 0u;
        if (a.work_cursor) *a.work_cursor = 0u;
    }
}

// Outputs at local index q (layouts in gqis_hip.h).
__device__ __forceinline__ void gq_store_complex(void* base, long long idx, real re, real im) {
    real2 v;
    v.x = re;
    v.y = im;
    ((real2*)base)[idx] = v;
}

#if !GQ_HILBERT && !GQ_GROUPED
// Per-point output state of the fixed-step kernels (mean, state mean, trace) and its epilogue
// (gqis/solver.py:2126-2193 output modes).
struct gq_outputs {
#if GQ_OUT_MEAN
    acc_t re, im;
#endif
#if GQ_OUT_MEAN_STATE
    acc_t s[GQ_D];
#endif
    bool trace_bad;
};

__device__ __forceinline__ void gq_out_init(gq_outputs& o) {
#if GQ_OUT_MEAN
    o.re = 0;
    o.im = 0;
#endif
#if GQ_OUT_MEAN_STATE
#pragma unroll
    for (int i = 0; i < GQ_D; ++i) o.s[i] = 0;
#endif
    o.trace_bad = false;
}

// After step `step` (state y, stage values G at the step's end time).
__device__ __forceinline__ void gq_out_step(gq_outputs& o, const gq_sweep_args& a, int step,
                                            long long q, const real* __restrict__ y, real px,
                                            real py, const real* __restrict__ C,
                                            const gq_static_t& S, const gq_stage_t& G) {
#if GQ_OUT_MEAN
    if (step >= a.warmup_steps) {
        real o_re, o_im;
        gq_obs(y, px, py, C, S, G, o_re, o_im);
        o.re += (acc_t)o_re;
        o.im += (acc_t)o_im;
    }
#endif
#if GQ_OUT_MEAN_STATE
    if (step >= a.warmup_steps) {
#pragma unroll
        for (int i = 0; i < GQ_D; ++i) o.s[i] += (acc_t)y[i];
    }
#endif
#if GQ_TRACE
    if (step >= a.trace_offset && (step - a.trace_offset) % a.trace_stride == 0) {
        const int kk = (step - a.trace_offset) / a.trace_stride;
        if (kk < a.num_trace) {
            const long long ld = a.trace_point_stride;
#if GQ_TRACE == 1
            real o_re, o_im;
            gq_obs(y, px, py, C, S, G, o_re, o_im);
            o.trace_bad |= !(gq_finite(o_re) && gq_finite(o_im));
            gq_store_complex(a.trace, (long long)kk * ld + q, o_re, o_im);
#else
#pragma unroll
            for (int i = 0; i < GQ_D; ++i) {
                o.trace_bad |= !gq_finite(y[i]);
                ((real*)a.trace)[((long long)kk * GQ_D + i) * ld + q] = y[i];
            }
#endif
        }
    }
#endif
}

__device__ __forceinline__ void gq_out_final(gq_outputs& o, const gq_sweep_args& a, long long q,
                                             const real* __restrict__ y, real px, real py,
                                             const real* __restrict__ C, const gq_static_t& S,
                                             const gq_stage_t& G) {
    bool bad = false;
#if GQ_OUT_MEAN
    const acc_t inv_t = (acc_t)1 / (acc_t)(a.num_steps - a.warmup_steps);
    const real r_re = (real)(o.re * inv_t), r_im = (real)(o.im * inv_t);
    bad = !(gq_finite(r_re) && gq_finite(r_im));
    gq_store_complex(a.result, q, r_re, r_im);
#elif GQ_OUT_FINAL
    real r_re, r_im;
    gq_obs(y, px, py, C, S, G, r_re, r_im);
    bad = !(gq_finite(r_re) && gq_finite(r_im));
    gq_store_complex(a.result, q, r_re, r_im);
#elif GQ_OUT_FINAL_RHO
#pragma unroll
    for (int i = 0; i < GQ_D; ++i) {
        bad |= !gq_finite(y[i]);
        ((real*)a.result)[(long long)i * a.point_count + q] = y[i];
    }
#elif GQ_OUT_MEAN_STATE
    const acc_t inv_n = (acc_t)1 / (acc_t)(a.num_steps - a.warmup_steps);
#pragma unroll
    for (int i = 0; i < GQ_D; ++i) {
        const real v = (real)(o.s[i] * inv_n);
        bad |= !gq_finite(v);
        ((real*)a.result)[(long long)i * a.point_count + q] = v;
    }
#endif
    gq_flag(a.status, bad, GQ_STATUS_NONFINITE_RESULT);
#if GQ_TRACE
    gq_flag(a.status, o.trace_bad, GQ_STATUS_NONFINITE_TRACE);
#endif
}
#endif  // !GQ_HILBERT && !GQ_GROUPED
