// Shared device helpers of the sweep kernels (prepended to every solver template).

// Point coordinates of local index q of a launch.
struct gq_point {
    long long q, p, ix, iy;
};

__device__ __forceinline__ gq_point gq_locate(const gq_sweep_args& a, long long q) {
    gq_point pt;
    pt.q = q;
    pt.p = a.point_begin + q;
    pt.ix = pt.p / a.num_y;
    pt.iy = pt.p - pt.ix * a.num_y;
    return pt;
}

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

__device__ __forceinline__ bool gq_finite(real v) { return isfinite(v); }

// One atomic OR per wavefront for the lanes that saw a problem.
__device__ __forceinline__ void gq_flag(unsigned int* status, bool bad, unsigned int bit) {
    const unsigned long long m = __ballot(bad);
    if (m != 0ull) {
        const int leader = __ffsll((long long)m) - 1;
        if ((int)(threadIdx.x & 63u) == leader) atomicOr(status, bit);
    }
}

// Outputs at local index q (layouts in gqis_hip.h).
__device__ __forceinline__ void gq_store_complex(void* base, long long idx, real re, real im) {
    real2 v;
    v.x = re;
    v.y = im;
    ((real2*)base)[idx] = v;
}
