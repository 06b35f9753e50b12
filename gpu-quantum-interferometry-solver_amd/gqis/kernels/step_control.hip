// Step-size controller of the adaptive kernels (gqis/cuda_solvers.py:879): the next step is
// h * clamp(0.9 * err^(-1/p), fac_min, fac_max), and a NaN factor (NaN error norm) takes fac_min.
//   Tsit5  p = 5, factor in [0.2, 5];   DOP853  p = 8, factor in [1/3, 6].
//
// err^(-1/p) is not taken with pow(), which costs about 216 FP64 instructions on gfx950, a
// quarter of a Duffing Tsit5 step: an FP32 estimate from the hardware log2 / exp2 (relative
// error ~1e-7) is refined by two Newton steps on g(y) = err * y^p - 1 in FP64,
// y <- y - y * (err * y^p - 1) / p, which converge quadratically to within an ulp or two of the
// FP64 power.  err is first clamped to [GQ_ERR_LO, GQ_ERR_HI]; factors of both ends already lie
// beyond the clamp (0.9 * 1e-5^(-1/5) = 9 > 5, 0.9 * 1e5^(-1/5) = 0.09 < 0.2; 0.9 * 1e-8^(-1/8)
// = 9 > 6, 0.9 * 1e4^(-1/8) = 0.28 < 1/3), so the clamped factor is that of the true error,
// and the FP32 estimate neither underflows nor overflows.

#if GQ_METHOD_DOP853
#define GQ_FAC_MIN (1.0 / 3.0)
#define GQ_FAC_MAX 6.0
#define GQ_ERR_ORDER 8
#define GQ_ERR_LO 1e-8
#define GQ_ERR_HI 1e4
#else
#define GQ_FAC_MIN 0.2
#define GQ_FAC_MAX 5.0
#define GQ_ERR_ORDER 5
#define GQ_ERR_LO 1e-5
#define GQ_ERR_HI 1e5
#endif

// Error sums are kept in the RHS precision; GQ_NERR embedded combinations (Tsit5: one, DOP853:
// the 5th- and 3rd-order ones) with weights gq_err_w(m, j).
typedef real en_t;
#if GQ_METHOD_DOP853
#define GQ_NERR 2
__device__ __forceinline__ constexpr double gq_err_w(int m, int j) { return m == 0 ? GQ_E5[j] : GQ_E3[j]; }
#else
#define GQ_NERR 1
__device__ __forceinline__ constexpr double gq_err_w(int, int j) { return GQ_E[j]; }
#endif
// Stage F at which the stage derivatives kept so far are folded into the partial sums of the
// later stages (folded stage storage): it balances the arrays held before the fold (y,
// k_0 .. k_{F-1}, the stage value and its derivative) against those after it (y, k_0, the
// partial sums of stages F+1 .. GQ_NSTAGES-1, the b and error sums, the stage value and its
// derivative).
#define GQ_FOLD ((GQ_NSTAGES + 2 + GQ_NERR) / 2)
static_assert(GQ_FOLD >= 1 && GQ_FOLD < GQ_NSTAGES, "fold stage inside the step");

// e / sc in the error norms: an IEEE division in FP64; in FP32 (the error sums of FP32
// kernels) e times the hardware reciprocal of sc (within 1 ulp), about a fifth of the
// instructions of an IEEE FP32 division -- the norm only steers the step size
__device__ __forceinline__ double gq_err_ratio(double e, double sc) { return e / sc; }
__device__ __forceinline__ float gq_err_ratio(float e, float sc) { return e * __builtin_amdgcn_rcpf(sc); }

__device__ __forceinline__ double gq_step_factor(double err) {
    const double x = fmin(fmax(err, GQ_ERR_LO), GQ_ERR_HI);
    const float l2 = __builtin_amdgcn_logf((float)x);                   // log2
    double y = (double)__builtin_amdgcn_exp2f(l2 * (-1.0f / GQ_ERR_ORDER));
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double y2 = y * y;
        const double y4 = y2 * y2;
#if GQ_ERR_ORDER == 8
        const double yp = y4 * y4;
#else
        const double yp = y4 * y;
#endif
        const double e = fma(x, yp, -1.0);
        y = fma(-y, e * (1.0 / GQ_ERR_ORDER), y);
    }
    const double factor = fmin(GQ_FAC_MAX, fmax(GQ_FAC_MIN, 0.9 * y));
    return (err != err) ? GQ_FAC_MIN : factor;
}
