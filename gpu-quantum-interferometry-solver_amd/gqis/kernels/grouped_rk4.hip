// Classical RK4 with the state of each parameter point split over GQ_G lanes ("lanes per
// point", grouped_common.hip).  Two LDS buffers alternate, so one barrier per stage suffices
// (a buffer is rewritten only after every wave has passed the next stage's barrier); a state too
// large for two copies keeps one, with a second barrier before each write.  Each wave
// runs the body specialised to its group (GQ_GROUP_DISPATCH on the wave-uniform index), so its
// statics, stage values and rows are compile-time register arrays.  This keeps large models (63
// states in FP64, 255 in FP32) in registers, at the cost of one LDS round trip per stage.
//
// Arithmetic and output semantics are the per-lane RK4 kernel's (gqis/cuda_solvers.py:470-517,
// gqis/solver.py:2126-2193): same stage times from the FP64 time table, k1 pipelining.

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
    gq_ginit<GI>(y, a, pt, px, py, C);

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
    buf = GQ_NEXT_BUF(buf);
    gq_gout<GI> out;
    gq_gout_init<GI>(out);

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
        buf = GQ_NEXT_BUF(buf);
#pragma unroll
        for (int r = 0; r < P::NR; ++r) {
            yt[r] = fma(dt2, k[r], y[r]);
            acc[r] = fma(R(2.0), k[r], acc[r]);
        }
        gq_share<GI>(s_y[buf], yt, lane);
        __syncthreads();
        P::rhs(s_y[buf], lane, px, py, C, Sg, Gg, k);
        buf = GQ_NEXT_BUF(buf);
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
        buf = GQ_NEXT_BUF(buf);
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
        buf = GQ_NEXT_BUF(buf);
        gq_gout_step<GI>(out, a, step, q, valid, Ycur, y, lane, px, py, C, Sg, Gg);
    }
    // the final state is in the LDS buffer written last (the loop flipped past it)
    gq_gout_final<GI>(out, a, q, valid, s_y[GQ_NEXT_BUF(buf)], y, lane, px, py, C, Sg, Gg);
}

extern "C" __global__ __launch_bounds__(GQ_GW) void GQ_KERNEL(gq_rk4_grouped)(const gq_sweep_args a) {
    __shared__ real s_y[GQ_NBUF][GQ_D * 64];
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
