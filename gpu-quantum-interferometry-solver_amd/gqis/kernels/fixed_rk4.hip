// Classical RK4 over a uniform time grid, one parameter point per lane, the reduced state in
// registers.  Semantics of the reference's RK4 body (gqis/cuda_solvers.py:470-517): stage times
// from the integer step index (t_mid = (step + 1/2) dt, t_end = t_mid + dt/2), k1 of a step is
// the RHS at the end of the previous one, the mean is a right-endpoint average over the steps
// with index >= warmup (gqis/solver.py:2144-2152), traces are stored after steps
// offset + k*stride (gqis/solver.py:2177-2193).
//
// Time atoms (the time-only parts of the drives) come from a table built in FP64 once per
// launch when GQ_TIME_TABLE is 1: index 0 is t = 0, 2s+1 the midpoint and 2s+2 the end of step
// s.  With GQ_TIME_TABLE 0 each lane evaluates them in its own precision from a time computed
// as the reference does.  The table rows of a step are loaded one step ahead, so the latency of
// those (scalar) loads is hidden behind the current step's arithmetic.

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

extern "C" __global__ __launch_bounds__(GQ_BLOCK) void GQ_KERNEL(gq_rk4)(const gq_sweep_args a) {
    const long long q = (long long)blockIdx.x * GQ_BLOCK + threadIdx.x;
    if (q >= a.point_count) return;
    const gq_point pt = gq_locate(a, q);
    const real px = ((const real*)a.par_x)[pt.ix];
    const real py = ((const real*)a.par_y)[pt.iy];
    const real* __restrict__ C = (const real*)a.consts;

    gq_static_t S;
    gq_statics(px, py, C, S);
    real y[GQ_D];
    gq_init_state(y, a, pt, px, py, C);

    const real dt = (real)a.dt;
    const real dt2 = dt * R(0.5);
    const real dt6 = dt / R(6.0);
    const int num_steps = a.num_steps;

    real TT[GQ_NATOMS_ALLOC];
    gq_stage_t G;
    gq_time_atoms(a, C, 0, R(0.0), TT);
    gq_stage(px, py, C, S, TT, G);
    real k[GQ_D];
    gq_rhs(y, px, py, C, S, G, k);

    gq_outputs out;
    gq_out_init(out);
#if GQ_TIME_TABLE
    const real* __restrict__ tab = (const real*)a.time_table;
    real TT_mid[GQ_NATOMS_ALLOC], TT_end[GQ_NATOMS_ALLOC];
    {
        const int first = (num_steps > 0) ? 1 : 0;   // the table has 2 * num_steps + 1 rows
#pragma unroll
        for (int k = 0; k < GQ_NATOMS; ++k) {
            TT_mid[k] = tab[first * GQ_NATOMS + k];
            TT_end[k] = tab[2 * first * GQ_NATOMS + k];
        }
    }
#endif

    for (int step = 0; step < num_steps; ++step) {
        real t_mid = ((real)step + R(0.5)) * dt;
#if GQ_TIME_TABLE
        real TT_end_now[GQ_NATOMS_ALLOC];
        {
            // rows 2s+1 (midpoint) and 2s+2 (end) of this step; next step's rows requested now
            const int next = (step + 1 < num_steps) ? 2 * step + 3 : 2 * step + 1;
#pragma unroll
            for (int k = 0; k < GQ_NATOMS; ++k) {
                TT[k] = TT_mid[k];
                TT_end_now[k] = TT_end[k];
                TT_mid[k] = tab[next * GQ_NATOMS + k];
                TT_end[k] = tab[(next + 1) * GQ_NATOMS + k];
            }
        }
#else
        gq_time_atoms(a, C, 2 * step + 1, t_mid, TT);
#endif
        gq_stage(px, py, C, S, TT, G);
        real yt[GQ_D], acc[GQ_D];
#pragma unroll
        for (int i = 0; i < GQ_D; ++i) {
            yt[i] = fma(dt2, k[i], y[i]);
            acc[i] = k[i];
        }
        gq_rhs(yt, px, py, C, S, G, k);
#pragma unroll
        for (int i = 0; i < GQ_D; ++i) {
            yt[i] = fma(dt2, k[i], y[i]);
            acc[i] = fma(R(2.0), k[i], acc[i]);
        }
        gq_rhs(yt, px, py, C, S, G, k);
#pragma unroll
        for (int i = 0; i < GQ_D; ++i) {
            yt[i] = fma(dt, k[i], y[i]);
            acc[i] = fma(R(2.0), k[i], acc[i]);
        }
        t_mid += dt2;
#if GQ_TIME_TABLE
#pragma unroll
        for (int k = 0; k < GQ_NATOMS; ++k) TT[k] = TT_end_now[k];
#else
        gq_time_atoms(a, C, 2 * step + 2, t_mid, TT);
#endif
        gq_stage(px, py, C, S, TT, G);
        gq_rhs(yt, px, py, C, S, G, k);
#pragma unroll
        for (int i = 0; i < GQ_D; ++i) {
            acc[i] += k[i];
            y[i] = fma(dt6, acc[i], y[i]);
        }
        if (step + 1 < num_steps) gq_rhs(y, px, py, C, S, G, k);
        gq_out_step(out, a, step, q, y, px, py, C, S, G);
    }

    gq_out_final(out, a, q, y, px, py, C, S, G);
}
