// FP64 table of the time atoms at every RK4 stage time: row 0 is t = 0, row 2s+1 the
// midpoint and row 2s+2 the end of step s (the reference's FP64 stage-time arithmetic).
extern "C" __global__ void GQ_KERNEL(gq_rk4_time_table)(const gq_sweep_args a) {
    const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (s > 2 * a.num_steps) return;
    const real* __restrict__ C = (const real*)a.consts;
    double t_local = 0.0;
    if (s > 0) {
        const int step = (s - 1) >> 1;
#if GQ_FP64
        // the reference's FP64 arithmetic, bit for bit
        double t_mid = ((double)step + 0.5) * a.dt;
        if ((s & 1) == 0) t_mid += a.dt * 0.5;
        t_local = t_mid;
#else
        t_local = ((double)step + 0.5) * a.dt + (((s & 1) == 0) ? 0.5 * a.dt : 0.0);
#endif
    }
    double TT[GQ_NATOMS_ALLOC];
    gq_atoms_f64(t_local + a.t_in, C, TT);
#if GQ_TABLE_F64
    double* __restrict__ tab = (double*)a.time_table;   // the mixed-precision form's FP64 table
#pragma unroll
    for (int k = 0; k < GQ_NATOMS; ++k) tab[s * GQ_NATOMS + k] = TT[k];
#else
    real* __restrict__ tab = (real*)a.time_table;
#pragma unroll
    for (int k = 0; k < GQ_NATOMS; ++k) tab[s * GQ_NATOMS + k] = (real)TT[k];
#endif
}

