// Time atoms of the fixed-step tableau methods (not RK4): GQ_NROWS stage-time rows per step
// (row 0 of the table is t = 0, then step s row r at t = s dt + c_r dt), computed in FP64 once
// per launch when GQ_TIME_TABLE is 1, otherwise per lane from the same stage times.

// Stage at which the Butcher-tableau kernels fold their stored stage derivatives into the
// partial sums of the later stages: it balances the vectors live before the fold (y, y_new,
// k_0 .. k_{F-1}, the stage value and its derivative) against those after it (y, y_new, the
// partial sums of stages F+1 .. GQ_NSTAGES-1, the stage value and its derivative).
#define GQ_TFOLD (GQ_NSTAGES / 2)

__device__ __forceinline__ void gq_fixed_atoms(const gq_sweep_args& a, const real* __restrict__ C,
                                               int step, int row, real* __restrict__ TT) {
    // row < 0: t = 0 (the initial state)
#if GQ_TIME_TABLE
    const real* __restrict__ tab = (const real*)a.time_table;
    const int r = row < 0 ? 0 : 1 + step * GQ_NROWS + row;
#pragma unroll
    for (int k = 0; k < GQ_NATOMS; ++k) TT[k] = tab[r * GQ_NATOMS + k];
#else
    const real dt = (real)a.dt;
    const real t = row < 0 ? R(0.0) : fma((real)GQ_ROW_C[row < 0 ? 0 : row], dt, (real)step * dt);
    gq_atoms_real(t + (real)a.t_in, C, TT);
#endif
}

extern "C" __global__ void GQ_KERNEL(gq_fixed_time_table)(const gq_sweep_args a) {
    const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (s > GQ_NROWS * a.num_steps) return;
    const real* __restrict__ C = (const real*)a.consts;
    double t_local = 0.0;
    if (s > 0) {
        const int step = (s - 1) / GQ_NROWS;
        const int row = (s - 1) - step * GQ_NROWS;
        t_local = fma(GQ_ROW_C[row], a.dt, (double)step * a.dt);
    }
    double TT[GQ_NATOMS_ALLOC];
    gq_atoms_f64(t_local + a.t_in, C, TT);
    real* __restrict__ tab = (real*)a.time_table;
#pragma unroll
    for (int k = 0; k < GQ_NATOMS; ++k) tab[s * GQ_NATOMS + k] = (real)TT[k];
}

