/*
 * FP64 fixed-step RK4 over a parameter grid for affine models (oracle; test infrastructure and
 * the timed CPU baseline, never part of the product path).
 *
 *   dy/dt = (M0 y + b0) + px (M1 y + b1) + py cos(omega t) (M2 y + b2)
 *
 * which covers the two-level LZS, four-level chain and random N-level sweep models.  Time and
 * update semantics restate the reference RK4 body (gqis/cuda_solvers.py:470-517) and output modes
 * (gqis/solver.py:2126-2152): stage times from the step index, t_mid = (step + 0.5) dt,
 * t_end = t_mid + dt/2, drive at t_in + t, k1 reused from the end of the previous step, right-
 * endpoint mean of o0 + o.y over steps >= warmup times 1/(num_steps - warmup).
 * Points are distributed over OpenMP threads.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int D;
    const double *M0, *b0, *M1, *b1, *M2, *b2;
    double omega;
} model_t;

static void rhs(const model_t* m, double px, double drive, const double* y, double* dy) {
    const int D = m->D;
    for (int i = 0; i < D; ++i) {
        double s0 = m->b0[i], s1 = m->b1[i], s2 = m->b2[i];
        const double* r0 = m->M0 + (size_t)i * D;
        const double* r1 = m->M1 + (size_t)i * D;
        const double* r2 = m->M2 + (size_t)i * D;
        for (int j = 0; j < D; ++j) {
            s0 += r0[j] * y[j];
            s1 += r1[j] * y[j];
            s2 += r2[j] * y[j];
        }
        dy[i] = s0 + px * s1 + drive * s2;
    }
}

/* mode 0: complex mean of the observable into out[2*p], out[2*p+1]
 * mode 1: final state into out[p*D + i]
 * y0: initial state (D values, shared by all points) or NULL for y[0] = 1.
 * Returns 0, or -1 for bad arguments. */
int rk4_affine_sweep(int D, const double* M0, const double* b0, const double* M1,
                     const double* b1, const double* M2, const double* b2, double omega,
                     long long npts, const double* px, const double* py, const double* y0,
                     double dt, int num_steps, int warmup, double t_in, const double* o_re,
                     const double* o_im, double o0_re, double o0_im, int mode, double* out,
                     int threads) {
    if (D <= 0 || npts < 0 || num_steps < 0 || (mode != 0 && mode != 1)) return -1;
    model_t m = {D, M0, b0, M1, b1, M2, b2, omega};
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel
    {
        double* buf = (double*)malloc(sizeof(double) * 5 * (size_t)D);
        double *y = buf, *k = buf + D, *yt = buf + 2 * D, *acc = buf + 3 * D;
#pragma omp for schedule(dynamic, 16)
        for (long long p = 0; p < npts; ++p) {
            const double x = px[p], a = py[p];
            if (y0) memcpy(y, y0, sizeof(double) * D);
            else { memset(y, 0, sizeof(double) * D); y[0] = 1.0; }
            const double dt2 = dt * 0.5, dt6 = dt / 6.0;
            double drive = a * cos(omega * (t_in + 0.0));
            rhs(&m, x, drive, y, k);
            double avg_re = 0.0, avg_im = 0.0;
            for (int s = 0; s < num_steps; ++s) {
                double t_mid = ((double)s + 0.5) * dt;
                drive = a * cos(omega * (t_in + t_mid));
                for (int i = 0; i < D; ++i) { yt[i] = y[i] + dt2 * k[i]; acc[i] = k[i]; }
                rhs(&m, x, drive, yt, k);
                for (int i = 0; i < D; ++i) { yt[i] = y[i] + dt2 * k[i]; acc[i] += 2.0 * k[i]; }
                rhs(&m, x, drive, yt, k);
                for (int i = 0; i < D; ++i) { yt[i] = y[i] + dt * k[i]; acc[i] += 2.0 * k[i]; }
                t_mid += dt2;
                drive = a * cos(omega * (t_in + t_mid));
                rhs(&m, x, drive, yt, k);
                for (int i = 0; i < D; ++i) { acc[i] += k[i]; y[i] += dt6 * acc[i]; }
                if (s + 1 < num_steps) rhs(&m, x, drive, y, k);
                if (mode == 0 && s >= warmup) {
                    double re = o0_re, im = o0_im;
                    for (int i = 0; i < D; ++i) { re += o_re[i] * y[i]; im += o_im[i] * y[i]; }
                    avg_re += re;
                    avg_im += im;
                }
            }
            if (mode == 0) {
                const double inv = 1.0 / (double)(num_steps - warmup);
                out[2 * p] = avg_re * inv;
                out[2 * p + 1] = avg_im * inv;
            } else {
                memcpy(out + (size_t)p * D, y, sizeof(double) * D);
            }
        }
        free(buf);
    }
    return 0;
}
