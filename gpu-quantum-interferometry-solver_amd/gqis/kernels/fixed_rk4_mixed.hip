// Mixed-precision classical RK4 (fp64="mixed"): the RK4 step of fixed_rk4.hip (reference
// gqis/cuda_solvers.py:470-517: stage times from the step index, k1 pipelined from the previous
// step, right-endpoint mean) with the state, the sweep parameters, the statics, the time table
// and the stage values in FP64 and only the right-hand side in FP32.
//
//   * stage values: computed in FP64 from the FP64 table row of the stage time and rounded to
//     FP32 once per stage time (gq_stage_d, gq_stage_narrow);
//   * stage inputs: y rounded to FP32 once per step, plus (dt/2) k in FP32;
//   * stage sum k1 + 2 k2 + 2 k3 + k4: FP32, or FP64 with -DGQ_MIXED_ACC64=1;
//   * state update y += (dt/6) sum in FP64.
// The FP32 rounding left is that of the right-hand side's values (about 6e-8 of |k| per stage).
#ifndef GQ_MIXED_ACC64
#define GQ_MIXED_ACC64 0
#endif
#if GQ_MIXED_ACC64
typedef double gq_mix_acc_t;
#else
typedef real gq_mix_acc_t;
#endif

__device__ __forceinline__ void gq_mixed_stage(const double* __restrict__ tab, int row, double px,
                                               double py, const real* __restrict__ C,
                                               const gq_static64_t& Sd, gq_stage_t& G) {
    double TT[GQ_NATOMS_ALLOC];
#pragma unroll
    for (int k = 0; k < GQ_NATOMS; ++k) TT[k] = tab[row * GQ_NATOMS + k];
    gq_stage64_t Gd;
    gq_stage_d(px, py, C, Sd, TT, Gd);
    gq_stage_narrow(Gd, G);
}

extern "C" __global__ __launch_bounds__(GQ_BLOCK) void GQ_KERNEL(gq_rk4_mixed)(const gq_sweep_args a) {
    const long long q = (long long)blockIdx.x * GQ_BLOCK + threadIdx.x;
    if (q >= a.point_count) return;
    const gq_point pt = gq_locate(a, q);
    const double pxd = ((const double*)a.par_x)[pt.ix];
    const double pyd = ((const double*)a.par_y)[pt.iy];
    const real px = (real)pxd, py = (real)pyd;
    const real* __restrict__ C = (const real*)a.consts;
    const double* __restrict__ tab = (const double*)a.time_table;

    gq_static64_t Sd;
    gq_statics_d(pxd, pyd, C, Sd);
    gq_static_t S;
    gq_statics_narrow(Sd, S);
    real yf[GQ_D];
    gq_init_state(yf, a, pt, px, py, C);
    double y[GQ_D];
#pragma unroll
    for (int i = 0; i < GQ_D; ++i) y[i] = (double)yf[i];

    const real dtf = (real)a.dt;
    const real dt2f = (real)(0.5 * a.dt);
    const double dt6 = a.dt / 6.0;
    const int num_steps = a.num_steps;

    gq_stage_t G;
    gq_mixed_stage(tab, 0, pxd, pyd, C, Sd, G);
    real k[GQ_D];
    gq_rhs(yf, px, py, C, S, G, k);

    gq_outputs out;
    gq_out_init(out);
    for (int step = 0; step < num_steps; ++step) {
        gq_mixed_stage(tab, 2 * step + 1, pxd, pyd, C, Sd, G);
        real yt[GQ_D];
        gq_mix_acc_t acc[GQ_D];
#pragma unroll
        for (int i = 0; i < GQ_D; ++i) {
            yt[i] = fma(dt2f, k[i], yf[i]);
            acc[i] = (gq_mix_acc_t)k[i];
        }
        gq_rhs(yt, px, py, C, S, G, k);
#pragma unroll
        for (int i = 0; i < GQ_D; ++i) {
            yt[i] = fma(dt2f, k[i], yf[i]);
            acc[i] = fma((gq_mix_acc_t)2, (gq_mix_acc_t)k[i], acc[i]);
        }
        gq_rhs(yt, px, py, C, S, G, k);
#pragma unroll
        for (int i = 0; i < GQ_D; ++i) {
            yt[i] = fma(dtf, k[i], yf[i]);
            acc[i] = fma((gq_mix_acc_t)2, (gq_mix_acc_t)k[i], acc[i]);
        }
        gq_mixed_stage(tab, 2 * step + 2, pxd, pyd, C, Sd, G);
        gq_rhs(yt, px, py, C, S, G, k);
#pragma unroll
        for (int i = 0; i < GQ_D; ++i) {
            acc[i] += (gq_mix_acc_t)k[i];
            y[i] = fma(dt6, (double)acc[i], y[i]);
            yf[i] = (real)y[i];
        }
        if (step + 1 < num_steps) gq_rhs(yf, px, py, C, S, G, k);
        gq_out_step(out, a, step, q, yf, px, py, C, S, G);
    }

    gq_out_final(out, a, q, yf, px, py, C, S, G);
}
