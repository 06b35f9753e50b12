// Fixed-step methods other than RK4, one parameter point per lane, state in registers.
//   GQ_SCHEME 0: explicit Runge-Kutta from a Butcher tableau (dp5, tsit5, alshina6, dop853,
//                anas5 with its fitted coefficient solver_param[0] entering A through GQ_A1);
//                k1 of a step is the RHS at the end of the previous one
//                (gqis/cuda_solvers.py:48-369 fixed-step bodies);
//   GQ_SCHEME 1: Carpenter-Kennedy 2N-storage LSRK4 (lserk4, lserk4f; gqis/cuda_solvers.py:520-607);
//   GQ_SCHEME 2: Adams-Bashforth 5 with four RK4 start-up steps (gqis/cuda_solvers.py:609-702).
// Stage times are t = step*dt + c*dt; the time atoms come from a table with GQ_NROWS rows per
// step (row 0 is t = 0) when GQ_TIME_TABLE is 1, otherwise each lane evaluates them.
// (time atoms and their table kernel: tableau_table.hip)

extern "C" __global__ __launch_bounds__(GQ_BLOCK) void GQ_KERNEL(gq_fixed)(const gq_sweep_args a) {
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
    const int num_steps = a.num_steps;
    real TT[GQ_NATOMS_ALLOC];
    gq_stage_t G;
    gq_fixed_atoms(a, C, 0, -1, TT);
    gq_stage(px, py, C, S, TT, G);
    gq_outputs out;
    gq_out_init(out);

#if GQ_SCHEME == 0
    const double p0 = a.solver_param[0];
    // a_sj of the tableau (anas5: plus its fitted part p0 * GQ_A1)
    auto nz = [](const int s, const int j) { return GQ_A[s][j] != 0.0 || GQ_A1[s][j] != 0.0; };
    // a_sj + p0 * a1_sj only where the tableau has a fitted part (anas5): elsewhere the
    // compile-time a_sj, which a run-time p0 * 0 would turn into a value held in a register
    auto coef = [&](const int s, const int j) {
        return GQ_A1[s][j] != 0.0 ? (real)(GQ_A[s][j] + GQ_A1[s][j] * p0) : (real)GQ_A[s][j];
    };
    // Folded stage storage: k_0 .. k_{F-1} are kept until stage F is formed and then folded into
    // the partial sums Ps of the later stages; each later k_s is folded as soon as it exists.
    // Every sum runs over j in ascending order, as in the unfolded form, with about half of the
    // stage vectors live at once.
    static_assert(GQ_TFOLD >= 1 && GQ_TFOLD < GQ_NSTAGES, "fold stage inside the step");
    real k[GQ_NSTAGES][GQ_D];
    real Ps[GQ_NSTAGES][GQ_D];
    gq_rhs(y, px, py, C, S, G, k[0]);
    for (int step = 0; step < num_steps; ++step) {
        real yn[GQ_D];
        // y_new takes each k_s as soon as it exists
        auto take_b = [&](const int s) {
            if (GQ_B[s] != 0.0) {
#pragma unroll
                for (int i = 0; i < GQ_D; ++i) yn[i] = fma(dt * (real)GQ_B[s], k[s][i], yn[i]);
            }
        };
        auto fold = [&](const int s) {
#pragma unroll
            for (int u = s + 1; u < GQ_NSTAGES; ++u)
                if (nz(u, s)) {
#pragma unroll
                    for (int i = 0; i < GQ_D; ++i) Ps[u][i] = fma(coef(u, s), k[s][i], Ps[u][i]);
                }
        };
        auto stage_values = [&](const int s) {
            if (GQ_ROW_OF[s] != GQ_ROW_OF[s - 1]) {
                gq_fixed_atoms(a, C, step, GQ_ROW_OF[s], TT);
                gq_stage(px, py, C, S, TT, G);
            }
        };
#pragma unroll
        for (int i = 0; i < GQ_D; ++i) yn[i] = y[i];
        take_b(0);
        // stages 1 .. F-1 from the stored derivatives
#pragma unroll
        for (int s = 1; s < GQ_TFOLD; ++s) {
            real yt[GQ_D];
#pragma unroll
            for (int i = 0; i < GQ_D; ++i) {
                real acc = R(0.0);
#pragma unroll
                for (int j = 0; j < s; ++j)
                    if (nz(s, j)) acc = fma(coef(s, j), k[j][i], acc);
                yt[i] = fma(dt, acc, y[i]);
            }
            stage_values(s);
            gq_rhs(yt, px, py, C, S, G, k[s]);
            take_b(s);
        }
        // stage F from the stored derivatives, which are then folded
        {
            constexpr int s = GQ_TFOLD;
            real yt[GQ_D];
#pragma unroll
            for (int i = 0; i < GQ_D; ++i) {
                real acc = R(0.0);
#pragma unroll
                for (int j = 0; j < s; ++j)
                    if (nz(s, j)) acc = fma(coef(s, j), k[j][i], acc);
                yt[i] = fma(dt, acc, y[i]);
#pragma unroll
                for (int u = s + 1; u < GQ_NSTAGES; ++u) {
                    real p = R(0.0);
#pragma unroll
                    for (int j = 0; j < s; ++j)
                        if (nz(u, j)) p = fma(coef(u, j), k[j][i], p);
                    Ps[u][i] = p;
                }
            }
            stage_values(s);
            gq_rhs(yt, px, py, C, S, G, k[s]);
            fold(s);
            take_b(s);
        }
        // stages F+1 .. GQ_NSTAGES-1 from the partial sums
#pragma unroll
        for (int s = GQ_TFOLD + 1; s < GQ_NSTAGES; ++s) {
            real yt[GQ_D];
#pragma unroll
            for (int i = 0; i < GQ_D; ++i) yt[i] = fma(dt, Ps[s][i], y[i]);
            stage_values(s);
            gq_rhs(yt, px, py, C, S, G, k[s]);
            fold(s);
            take_b(s);
        }
#pragma unroll
        for (int i = 0; i < GQ_D; ++i) y[i] = yn[i];
        if (GQ_ROW_OF[GQ_NSTAGES - 1] != GQ_END_ROW) {
            gq_fixed_atoms(a, C, step, GQ_END_ROW, TT);
            gq_stage(px, py, C, S, TT, G);
        }
        if (step + 1 < num_steps) gq_rhs(y, px, py, C, S, G, k[0]);
        gq_out_step(out, a, step, q, y, px, py, C, S, G);
    }
#elif GQ_SCHEME == 1
    real r[GQ_D], k[GQ_D];
    for (int step = 0; step < num_steps; ++step) {
#pragma unroll
        for (int s = 0; s < GQ_NSTAGES; ++s) {
            if (s > 0) {
                gq_fixed_atoms(a, C, step, GQ_ROW_OF[s], TT);
                gq_stage(px, py, C, S, TT, G);
            }
            gq_rhs(y, px, py, C, S, G, k);
#pragma unroll
            for (int i = 0; i < GQ_D; ++i) {
                r[i] = (s == 0) ? dt * k[i] : fma((real)GQ_A[s][0], r[i], dt * k[i]);
                y[i] = fma((real)GQ_B[s], r[i], y[i]);
            }
        }
        gq_fixed_atoms(a, C, step, GQ_END_ROW, TT);
        gq_stage(px, py, C, S, TT, G);
        gq_out_step(out, a, step, q, y, px, py, C, S, G);
    }
#else  // GQ_SCHEME == 2: AB5
    const real dt2 = dt * R(0.5);
    const real dt6 = dt / R(6.0);
    real h0[GQ_D], h1[GQ_D], h2[GQ_D], h3[GQ_D], h4[GQ_D];   // f_n, f_{n-1}, ..., f_{n-4}
    gq_rhs(y, px, py, C, S, G, h0);
    for (int step = 0; step < num_steps; ++step) {
        if (step < 4) {
            // classical RK4 start-up step; h0 holds f(y_n)
            real yt[GQ_D], acc[GQ_D], kk[GQ_D];
            gq_fixed_atoms(a, C, step, 0, TT);
            gq_stage(px, py, C, S, TT, G);
#pragma unroll
            for (int i = 0; i < GQ_D; ++i) {
                yt[i] = fma(dt2, h0[i], y[i]);
                acc[i] = h0[i];
            }
            gq_rhs(yt, px, py, C, S, G, kk);
#pragma unroll
            for (int i = 0; i < GQ_D; ++i) {
                yt[i] = fma(dt2, kk[i], y[i]);
                acc[i] = fma(R(2.0), kk[i], acc[i]);
            }
            gq_rhs(yt, px, py, C, S, G, kk);
#pragma unroll
            for (int i = 0; i < GQ_D; ++i) {
                yt[i] = fma(dt, kk[i], y[i]);
                acc[i] = fma(R(2.0), kk[i], acc[i]);
            }
            gq_fixed_atoms(a, C, step, 1, TT);
            gq_stage(px, py, C, S, TT, G);
            gq_rhs(yt, px, py, C, S, G, kk);
#pragma unroll
            for (int i = 0; i < GQ_D; ++i) {
                acc[i] += kk[i];
                y[i] = fma(dt6, acc[i], y[i]);
            }
        } else {
#pragma unroll
            for (int i = 0; i < GQ_D; ++i) {
                real sum = R(0.0);
                sum = fma(R(1901.0 / 720.0), h0[i], sum);
                sum = fma(R(-1387.0 / 360.0), h1[i], sum);
                sum = fma(R(109.0 / 30.0), h2[i], sum);
                sum = fma(R(-637.0 / 360.0), h3[i], sum);
                sum = fma(R(251.0 / 720.0), h4[i], sum);
                y[i] = fma(dt, sum, y[i]);
            }
            gq_fixed_atoms(a, C, step, 1, TT);
            gq_stage(px, py, C, S, TT, G);
        }
        if (step + 1 < num_steps) {
#pragma unroll
            for (int i = 0; i < GQ_D; ++i) {
                h4[i] = h3[i];
                h3[i] = h2[i];
                h2[i] = h1[i];
                h1[i] = h0[i];
            }
            gq_rhs(y, px, py, C, S, G, h0);
        }
        gq_out_step(out, a, step, q, y, px, py, C, S, G);
    }
#endif
    gq_out_final(out, a, q, y, px, py, C, S, G);
}
