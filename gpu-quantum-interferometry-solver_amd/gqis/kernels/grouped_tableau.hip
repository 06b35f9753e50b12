// Fixed-step methods other than RK4 with the state of each parameter point split over GQ_G
// lanes (grouped_common.hip); the schemes and their arithmetic are those of the per-lane
// fixed_tableau.hip (gqis/cuda_solvers.py:48-702):
//   GQ_SCHEME 0: explicit Runge-Kutta from a Butcher tableau, k1 of a step is the RHS at the
//                end of the previous one;
//   GQ_SCHEME 1: 2N-storage LSRK4;
//   GQ_SCHEME 2: Adams-Bashforth 5 with four RK4 start-up steps.
// Every RHS evaluation shares its input rows through one of two alternating LDS buffers and one
// barrier (or through a single buffer, with a barrier before the write, when two do not fit).  The state after a step is shared as the input of the next step's first RHS (schemes
// 0 and 2) or, for LSRK4, only when an observable needs it.

template <int GI>
__device__ __forceinline__ void gq_tab_body(const gq_sweep_args& a, real (*s_y)[GQ_D * 64],
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
    const int num_steps = a.num_steps;
    real TT[GQ_NATOMS_ALLOC];
    real Gg[P::NG];
    gq_fixed_atoms(a, C, 0, -1, TT);
    P::stage(px, py, C, Sg, TT, Gg);
    gq_gout<GI> out;
    gq_gout_init<GI>(out);
    int buf = 0;
    // RHS of the wave's rows at the shared state v (one LDS buffer)
    auto eval = [&](const real* __restrict__ v, real* __restrict__ k) -> const real* {
        real* __restrict__ Y = s_y[buf];
        gq_share<GI>(Y, v, lane);
        __syncthreads();
        P::rhs(Y, lane, px, py, C, Sg, Gg, k);
        buf = GQ_NEXT_BUF(buf);
        return Y;
    };

#if GQ_SCHEME == 0
    const double p0 = a.solver_param[0];
    auto nz = [](const int s, const int j) { return GQ_A[s][j] != 0.0 || GQ_A1[s][j] != 0.0; };
    // a_sj + p0 * a1_sj only where the tableau has a fitted part (anas5): elsewhere the
    // compile-time a_sj, which a run-time p0 * 0 would turn into a value held in a register
    auto coef = [&](const int s, const int j) {
        return GQ_A1[s][j] != 0.0 ? (real)(GQ_A[s][j] + GQ_A1[s][j] * p0) : (real)GQ_A[s][j];
    };
    // folded stage storage, as in the one-lane kernel (fixed_tableau.hip)
    static_assert(GQ_TFOLD >= 1 && GQ_TFOLD < GQ_NSTAGES, "fold stage inside the step");
    real k[GQ_NSTAGES][P::NR];
    real Ps[GQ_NSTAGES][P::NR];
    const real* Yst = eval(y, k[0]);
    for (int step = 0; step < num_steps; ++step) {
        real yn[P::NR];
        // y_new takes each k_s as soon as it exists
        auto take_b = [&](const int s) {
            if (GQ_B[s] != 0.0) {
#pragma unroll
                for (int r = 0; r < P::NR; ++r) yn[r] = fma(dt * (real)GQ_B[s], k[s][r], yn[r]);
            }
        };
        auto fold = [&](const int s) {
#pragma unroll
            for (int u = s + 1; u < GQ_NSTAGES; ++u)
                if (nz(u, s)) {
#pragma unroll
                    for (int r = 0; r < P::NR; ++r) Ps[u][r] = fma(coef(u, s), k[s][r], Ps[u][r]);
                }
        };
        auto stage_values = [&](const int s) {
            if (GQ_ROW_OF[s] != GQ_ROW_OF[s - 1]) {
                gq_fixed_atoms(a, C, step, GQ_ROW_OF[s], TT);
                P::stage(px, py, C, Sg, TT, Gg);
            }
        };
#pragma unroll
        for (int r = 0; r < P::NR; ++r) yn[r] = y[r];
        take_b(0);
        // stages 1 .. F-1 from the stored derivatives
#pragma unroll
        for (int s = 1; s < GQ_TFOLD; ++s) {
            real yt[P::NR];
#pragma unroll
            for (int r = 0; r < P::NR; ++r) {
                real acc = R(0.0);
#pragma unroll
                for (int j = 0; j < s; ++j)
                    if (nz(s, j)) acc = fma(coef(s, j), k[j][r], acc);
                yt[r] = fma(dt, acc, y[r]);
            }
            stage_values(s);
            eval(yt, k[s]);
            take_b(s);
        }
        // stage F from the stored derivatives, which are then folded
        {
            constexpr int s = GQ_TFOLD;
            real yt[P::NR];
#pragma unroll
            for (int r = 0; r < P::NR; ++r) {
                real acc = R(0.0);
#pragma unroll
                for (int j = 0; j < s; ++j)
                    if (nz(s, j)) acc = fma(coef(s, j), k[j][r], acc);
                yt[r] = fma(dt, acc, y[r]);
#pragma unroll
                for (int u = s + 1; u < GQ_NSTAGES; ++u) {
                    real p = R(0.0);
#pragma unroll
                    for (int j = 0; j < s; ++j)
                        if (nz(u, j)) p = fma(coef(u, j), k[j][r], p);
                    Ps[u][r] = p;
                }
            }
            stage_values(s);
            eval(yt, k[s]);
            fold(s);
            take_b(s);
        }
        // stages F+1 .. GQ_NSTAGES-1 from the partial sums
#pragma unroll
        for (int s = GQ_TFOLD + 1; s < GQ_NSTAGES; ++s) {
            real yt[P::NR];
#pragma unroll
            for (int r = 0; r < P::NR; ++r) yt[r] = fma(dt, Ps[s][r], y[r]);
            stage_values(s);
            eval(yt, k[s]);
            fold(s);
            take_b(s);
        }
#pragma unroll
        for (int r = 0; r < P::NR; ++r) y[r] = yn[r];
        if (GQ_ROW_OF[GQ_NSTAGES - 1] != GQ_END_ROW) {
            gq_fixed_atoms(a, C, step, GQ_END_ROW, TT);
            P::stage(px, py, C, Sg, TT, Gg);
        }
        // the new state is shared in both cases: the next k1, and the observable
        Yst = eval(y, k[0]);
        gq_gout_step<GI>(out, a, step, q, valid, Yst, y, lane, px, py, C, Sg, Gg);
    }
#elif GQ_SCHEME == 1
    real rr[P::NR], k[P::NR];
    const real* Yst = s_y[0];
#if GQ_OUT_FINAL
    {
        gq_share<GI>(s_y[buf], y, lane);
        __syncthreads();
        Yst = s_y[buf];
        buf = GQ_NEXT_BUF(buf);
    }
#endif
    for (int step = 0; step < num_steps; ++step) {
#pragma unroll
        for (int s = 0; s < GQ_NSTAGES; ++s) {
            if (s > 0) {
                gq_fixed_atoms(a, C, step, GQ_ROW_OF[s], TT);
                P::stage(px, py, C, Sg, TT, Gg);
            }
            eval(y, k);
#pragma unroll
            for (int r = 0; r < P::NR; ++r) {
                rr[r] = (s == 0) ? dt * k[r] : fma((real)GQ_A[s][0], rr[r], dt * k[r]);
                y[r] = fma((real)GQ_B[s], rr[r], y[r]);
            }
        }
        gq_fixed_atoms(a, C, step, GQ_END_ROW, TT);
        P::stage(px, py, C, Sg, TT, Gg);
#if GQ_OUT_MEAN || GQ_OUT_FINAL || GQ_TRACE == 1
        // the observable needs the whole new state
        gq_share<GI>(s_y[buf], y, lane);
        __syncthreads();
        Yst = s_y[buf];
        buf = GQ_NEXT_BUF(buf);
#endif
        gq_gout_step<GI>(out, a, step, q, valid, Yst, y, lane, px, py, C, Sg, Gg);
    }
#else  // GQ_SCHEME == 2: AB5
    const real dt2 = dt * R(0.5);
    const real dt6 = dt / R(6.0);
    real h0[P::NR], h1[P::NR], h2[P::NR], h3[P::NR], h4[P::NR];   // f_n, ..., f_{n-4}
    const real* Yst = eval(y, h0);
    for (int step = 0; step < num_steps; ++step) {
        if (step < 4) {
            // classical RK4 start-up step; h0 holds f(y_n)
            real yt[P::NR], acc[P::NR], kk[P::NR];
            gq_fixed_atoms(a, C, step, 0, TT);
            P::stage(px, py, C, Sg, TT, Gg);
#pragma unroll
            for (int r = 0; r < P::NR; ++r) {
                yt[r] = fma(dt2, h0[r], y[r]);
                acc[r] = h0[r];
            }
            eval(yt, kk);
#pragma unroll
            for (int r = 0; r < P::NR; ++r) {
                yt[r] = fma(dt2, kk[r], y[r]);
                acc[r] = fma(R(2.0), kk[r], acc[r]);
            }
            eval(yt, kk);
#pragma unroll
            for (int r = 0; r < P::NR; ++r) {
                yt[r] = fma(dt, kk[r], y[r]);
                acc[r] = fma(R(2.0), kk[r], acc[r]);
            }
            gq_fixed_atoms(a, C, step, 1, TT);
            P::stage(px, py, C, Sg, TT, Gg);
            eval(yt, kk);
#pragma unroll
            for (int r = 0; r < P::NR; ++r) {
                acc[r] += kk[r];
                y[r] = fma(dt6, acc[r], y[r]);
            }
        } else {
#pragma unroll
            for (int r = 0; r < P::NR; ++r) {
                real sum = R(0.0);
                sum = fma(R(1901.0 / 720.0), h0[r], sum);
                sum = fma(R(-1387.0 / 360.0), h1[r], sum);
                sum = fma(R(109.0 / 30.0), h2[r], sum);
                sum = fma(R(-637.0 / 360.0), h3[r], sum);
                sum = fma(R(251.0 / 720.0), h4[r], sum);
                y[r] = fma(dt, sum, y[r]);
            }
            gq_fixed_atoms(a, C, step, 1, TT);
            P::stage(px, py, C, Sg, TT, Gg);
        }
#pragma unroll
        for (int r = 0; r < P::NR; ++r) {
            h4[r] = h3[r];
            h3[r] = h2[r];
            h2[r] = h1[r];
            h1[r] = h0[r];
        }
        Yst = eval(y, h0);
        gq_gout_step<GI>(out, a, step, q, valid, Yst, y, lane, px, py, C, Sg, Gg);
    }
#endif
    gq_gout_final<GI>(out, a, q, valid, Yst, y, lane, px, py, C, Sg, Gg);
}

extern "C" __global__ __launch_bounds__(GQ_GW) void GQ_KERNEL(gq_fixed_grouped)(const gq_sweep_args a) {
    __shared__ real s_y[GQ_NBUF][GQ_D * 64];
    const int g = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63u);
    const long long q_raw = (long long)blockIdx.x * 64 + lane;
    const bool valid = q_raw < a.point_count;
    const long long q = valid ? q_raw : a.point_count - 1;
    const gq_point pt = gq_locate(a, q);
    const real px = ((const real*)a.par_x)[pt.ix];
    const real py = ((const real*)a.par_y)[pt.iy];
    GQ_GROUP_DISPATCH(g, gq_tab_body, (a, s_y, lane, q, valid, px, py, pt))
}
