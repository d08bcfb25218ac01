// jk_device.h -- device-side math shared by the gfx950 kernels of libjoker.
//
// Layout and arithmetic of one prior sample's marginal likelihood (the
// quantity fast_likelihood.pyx:359-425 computes with LU + Woodbury), restated
// in the O(N L^2) form:
//   r_n  = y_n - x_n . mu,  x_n = (m_n, T_n1, ..., T_n,L-1)
//   G    = Lambda^-1 + sum_n w_n x_n x_n^T = R R^T
//   u    = sum_n w_n x_n r_n,     chi2 = sum_n w_n r_n^2 - u^T G^-1 u
//   lnL  = -1/2 [chi2 + N ln 2pi - sum ln w + sum ln Lambda + ln det G]
// with m_n = cos(omega + f_n) + e cos omega the unit-K Keplerian column
// (twobody c_rv_from_elements, called at fast_likelihood.pyx:453-455).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/joker.h"

#define JK_PI 3.14159265358979323846
#define JK_TWO_PI 6.28318530717958647692
#define JK_INV_TWO_PI 0.15915494309189533577
// 2 pi = HI + LO with HI the nearest double; the reduction subtracts k HI and
// k LO with fused multiply-adds, so the reduced angle carries an error of a
// few ulp of itself plus |k| * 1e-32.
#define JK_TWO_PI_HI 6.283185307179586
#define JK_TWO_PI_LO 2.4492935982947064e-16
#define JK_LN_TWO_PI 1.83787706640934548356
#define JK_KEPLER_MAXITER 128
#define JK_KEPLER_TOL 1e-10

// Per-star constants, device resident, read through scalar loads (every lane
// of a wavefront evaluates the same epoch at the same time).
struct StarDev {
    int n;             // epochs
    int L;             // linear parameters
    int jitter_mode;   // JK_JITTER_*
    int k_prior_mode;  // JK_KPRIOR_*
    int exact;         // batched table slots: 1 evaluates the star's rows in FP64
                       // inside mixed-precision launches (error bound 0)
    int t0_one;        // 1 when the first trend column is 1.0 at every epoch (the
                       // constant offset v0) and there is no full covariance
    double sigK0sq;    // sigma_K0^2      (FixedCompanionMass)
    double P0;         // reference period (FixedCompanionMass)
    double maxKsq;     // max_K^2         (FixedCompanionMass)
    double mu0;        // prior mean of K
    double mu[JK_LMAX];        // prior means (M_T row order; mu[0] == mu0)
    double lam0;       // prior variance of K (Normal K)
    double lam_inv[JK_LMAX];   // 1 / Lambda_i (entry 0 only for Normal K)
    double ln_lam_trend;       // sum_{i>=1} ln Lambda_i
    double sum_ln_ivar;        // sum_n ln ivar_n
    // N ln 2pi - sum_n ln ivar_n + sum_{i>=1} ln Lambda_i, plus ln Lambda_0
    // for the Normal K prior: the sample-independent part of the log-constant
    // with w = ivar; ln_c_phys the same without the ivar term (physical
    // jitter, where sum ln w comes from the sample's product of sigma^2 + s^2).
    double ln_c;
    double ln_c_phys;
    // Trend-only moments with w = ivar (the reference jitter mode, where they
    // do not depend on the sample): S_tt[j][k] = sum ivar T_j T_k,
    // S_tr[j] = sum ivar T_j r'_n, S_rr = sum ivar r'^2, r' = y - T mu_trend.
    double S_tt[JK_LMAX * JK_LMAX];
    double S_tr[JK_LMAX];
    double S_rr;
    // Posterior right-hand side pieces: sum ivar T_j y (j >= 1).
    double S_ty[JK_LMAX];
    // Epoch arrays (device): dt = t - t_ref, rp = r', y, ivar, sig2 = 1/ivar,
    // T (L-1 rows of n).
    const double *dt;
    const double *rp;
    const double *y;
    const double *ivar;
    const double *sig2;
    const double *T;
    // Full-covariance stars (nullptr otherwise): Q^T, n x n row-major, whose
    // rows are the eigenvectors of the data covariance C.  rp, y, T are then
    // rotated into that eigenbasis (Q^T r', Q^T y, Q^T T), ivar and sig2 hold
    // 1 / lambda_i and lambda_i, and only dt stays in epoch order: a row's
    // design column is computed per epoch and rotated, m' = Q^T m.
    const double *Qt;
};

// ---------------------------------------------------------------------------
// Kepler's equation in FP64.
// M is reduced to x in [-pi, pi]; E(-x) = -E(x), so the solve runs on
// ax = |x| in [0, pi], where f(E) = E - e sin E - ax is increasing and convex.
// Start: E0 = min(ax + 0.85 e, pi) (Danby 1987).  From E0, Newton's iterates
// stay in [0, pi] and converge monotonically after at most one overshoot, for
// every e in [0, 1); iteration stops when the step is below 1e-10 (the
// reference's anomaly_tol, fast_likelihood.pyx:35) or after 128 steps.
// Returns sin E and cos E of the converged anomaly; the final (sin, cos) pair
// is rotated by the last step to first order (error < step^2 / 2 < 1e-20)
// instead of being re-evaluated.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double jk_reduce_2pi(double M)
{
    double k = rint(M * JK_INV_TWO_PI);
    double x = fma(-k, JK_TWO_PI_HI, M);
    return fma(-k, JK_TWO_PI_LO, x);
}

__device__ __forceinline__ void jk_kepler_f64(double x, double e, double &sinE, double &cosE)
{
    const double ax = fabs(x);
    double E = fmin(fma(0.85, e, ax), JK_PI);
    double s, c;
    sincos(E, &s, &c);
#pragma unroll 1
    for (int it = 0; it < JK_KEPLER_MAXITER; ++it) {
        const double fp = fma(-e, c, 1.0);
        const double f = fma(-e, s, E - ax);
        const double d = f / fp;
        E -= d;
        if (fabs(d) < JK_KEPLER_TOL) {
            const double s1 = fma(-d, c, s);
            c = fma(d, s, c);
            s = s1;
            break;
        }
        sincos(E, &s, &c);
    }
    sinE = x < 0.0 ? -s : s;
    cosE = c;
}

// Unit-K Keplerian design value from (sin E, cos E):
//   cos f = (cos E - e) / (1 - e cos E),  sin f = sqrt(1-e^2) sin E / (1 - e cos E)
//   m = cos(omega + f) + e cos omega = cw cos f - sw sin f + e cw
__device__ __forceinline__ double jk_design_f64(double sE, double cE, double e, double sqrt1me2,
                                                double cw, double sw, double ecw)
{
    const double den = 1.0 / fma(-e, cE, 1.0);
    const double cf = (cE - e) * den;
    const double sf = sqrt1me2 * sE * den;
    return fma(cw, cf, fma(-sw, sf, ecw));
}

// FixedCompanionMass prior variance of K (fast_likelihood.pyx:461-464):
// sigma_K0^2 / (1 - e^2) * (P / P0)^(-2/3), clipped to max_K^2 on the
// likelihood path only.
__device__ __forceinline__ double jk_k_variance(const StarDev *__restrict__ st, double P, double e,
                                                bool clip)
{
    const double q = cbrt(P / st->P0);
    double v = st->sigK0sq / ((1.0 - e * e) * (q * q));
    if (clip) v = fmin(v, st->maxKsq);
    return v;
}

// Solve the L x L system and return lnL.
//   G = [[1/lam0 + Smm, Smt], [Smt, diag(lam_inv) + Stt]],
//   u = [Smr - mu0 Smm, Str - mu0 Smt],  chi0 = Srr - 2 mu0 Smr + mu0^2 Smm,
//   lnL = -1/2 [chi0 - u^T G^-1 u + lnc + ln(scale det G)]
// where lnc + ln(scale) = N ln 2pi - sum ln w + sum ln Lambda: the sample-
// dependent factors (the jitter product, a per-sample Lambda_0) arrive as
// `scale` and share the one logarithm of the determinant.
// With POST, post[0] = (G^-1)_00, post[1] = (G^-1 u)_0 (the posterior mean of
// the K column minus its prior mean) and post[2] = chi0 - u^T G^-1 u (the
// minimised chi^2) are also returned, for the mixed-precision error bound.
// Returns +inf if the Cholesky factorisation breaks down, as the reference
// returns INF when its LU of Ainv fails (fast_likelihood.pyx:283-291,381-382).
// Natural logarithm in FP64 of a positive finite x (subnormals included:
// frexp scales them exactly; +inf, 0, negatives and NaN give what log does),
// from a table of JK_LOG_TAB_N entries: x = m 2^k
// with m in [1, 2); for the interval of m's top seven fraction bits,
// c_i = 1 + (i + 1/2) / 128 and ltab[i] = (R_i, -ln R_i), R_i = 1 / c_i
// rounded to FP64 and its log taken in long double on the host.  Then
// r = m R_i - 1 (one fma), |r| <= 1/256 + 2^-52, and
//   ln x = k ln 2 - ln R_i + ln(1 + r),  ln(1 + r) = r - r^2/2 + ... - r^6/6
// (truncation below 3e-18).  The absolute error is a few 1e-16 plus the
// rounding of k ln 2 (below 6e-14 at the largest k): far inside the 1e-10
// relative parity of the likelihood, in about 20 instructions where the
// library's log, which carries a double-double intermediate, takes 85.
#define JK_LOG_TAB_N 128
#define JK_LN2 0.69314718055994530942

__device__ __forceinline__ double jk_log_tab(double x, const double2 *__restrict__ ltab)
{
    if (!(x > 0.0 && x < INFINITY))
        return x > 0.0 ? x : (x == 0.0 ? -INFINITY : __builtin_nan(""));
    int k;
    const double m = 2.0 * frexp(x, &k);
    const double2 t = ltab[(__double2hiint(m) >> 13) & (JK_LOG_TAB_N - 1)];
    const double r = fma(m, t.x, -1.0);
    double q = fma(r, -1.0 / 6.0, 0.2);
    q = fma(r, q, -0.25);
    q = fma(r, q, 1.0 / 3.0);
    q = fma(r, q, -0.5);
    return fma((double)(k - 1), JK_LN2, t.y + fma(r * r, q, r));
}

template <int L, bool POST = false>
__device__ __forceinline__ double jk_finish_lnl(const double *G /* LxL, lower used */,
                                                const double *u, double chi0, double lnc,
                                                double scale,
                                                const double2 *__restrict__ ltab,
                                                double *post = nullptr)
{
    if constexpr (L == 1) {
        const double g = G[0];
        if (!(g > 0.0)) return INFINITY;
        const double chi2 = chi0 - u[0] * u[0] / g;
        if constexpr (POST) {
            post[0] = 1.0 / g;
            post[1] = u[0] / g;
            post[2] = chi2;
        }
        return -0.5 * (chi2 + lnc + jk_log_tab(scale * g, ltab));
    } else if constexpr (L == 2) {
        const double g00 = G[0], g10 = G[2], g11 = G[3];
        const double det = fma(g00, g11, -g10 * g10);
        if (!(g00 > 0.0) || !(det > 0.0)) return INFINITY;
        const double q = (fma(g11, u[0] * u[0], g00 * u[1] * u[1]) - 2.0 * g10 * u[0] * u[1]) / det;
        if constexpr (POST) {
            post[0] = g11 / det;
            post[1] = fma(g11, u[0], -g10 * u[1]) / det;
            post[2] = chi0 - q;
        }
        return -0.5 * (chi0 - q + lnc + jk_log_tab(scale * det, ltab));
    } else {
        // Cholesky G = R R^T without divisions: each pivot's 1/sqrt from the
        // hardware v_rsq_f64 refined by two Newton steps; R[j][j] itself is
        // never needed, only its inverse (for the column below it and for
        // the forward substitution z = R^-1 u, chi2 correction = z^T z).
        double R[L * L];
        double rinv[L];
        double z[L];
        double zz = 0.0;
        double detm = scale;   // product of the pivots, renormalised
        int dexp = 0;
#pragma unroll
        for (int j = 0; j < L; ++j) {
            double d = G[j * L + j];
#pragma unroll
            for (int k = 0; k < j; ++k) d = fma(-R[j * L + k], R[j * L + k], d);
            if (!(d > 0.0)) return INFINITY;
            int ex;
            detm = frexp(detm * d, &ex);
            dexp += ex;
            double y = __builtin_amdgcn_rsq(d);
            y = y * fma(-0.5 * d * y, y, 1.5);
            y = y * fma(-0.5 * d * y, y, 1.5);
            rinv[j] = y;
#pragma unroll
            for (int i = j + 1; i < L; ++i) {
                double v = G[i * L + j];
#pragma unroll
                for (int k = 0; k < j; ++k) v = fma(-R[i * L + k], R[j * L + k], v);
                R[i * L + j] = v * y;
            }
        }
#pragma unroll
        for (int i = 0; i < L; ++i) {
            double v = u[i];
#pragma unroll
            for (int k = 0; k < i; ++k) v = fma(-R[i * L + k], z[k], v);
            z[i] = v * rinv[i];
            zz = fma(z[i], z[i], zz);
        }
        if constexpr (POST) {
            // ze = R^-1 e_0: (G^-1)_00 = |ze|^2 and (G^-1 u)_0 = ze . z
            double ze[L];
            double g00 = 0.0, a0 = 0.0;
#pragma unroll
            for (int i = 0; i < L; ++i) {
                double v = i == 0 ? 1.0 : 0.0;
#pragma unroll
                for (int k = 0; k < i; ++k) v = fma(-R[i * L + k], ze[k], v);
                ze[i] = v * rinv[i];
                g00 = fma(ze[i], ze[i], g00);
                a0 = fma(ze[i], z[i], a0);
            }
            post[0] = g00;
            post[1] = a0;
            post[2] = chi0 - zz;
        }
        return -0.5 * (chi0 - zz + lnc + (jk_log_tab(detm, ltab) + (double)dexp * 0.69314718055994530942));
    }
}

// ---------------------------------------------------------------------------
// 128-bit arithmetic for numpy's PCG64 (XSL-RR 128/64).
//   state <- state * A + inc   (A = 0x2360ED051FC65DA44385DF649FCCF645)
//   out    = rotr64(hi ^ lo, state >> 122) of the new state
//   uniform = (out >> 11) * 2^-53          (Generator.uniform, low=0, high=1)
// A jump by k steps is state <- A^k state + inc * D_k with
// D_k = sum_{j<k} A^j; (A^k, D_k) for k = 2^i are precomputed on the host.
// ---------------------------------------------------------------------------
struct u128 {
    uint64_t lo, hi;
};

__device__ __forceinline__ u128 u128_mul(u128 a, u128 b)
{
    u128 r;
    r.lo = a.lo * b.lo;
    r.hi = __umul64hi(a.lo, b.lo) + a.lo * b.hi + a.hi * b.lo;
    return r;
}

__device__ __forceinline__ u128 u128_add(u128 a, u128 b)
{
    u128 r;
    r.lo = a.lo + b.lo;
    r.hi = a.hi + b.hi + (r.lo < a.lo ? 1ull : 0ull);
    return r;
}

__device__ __forceinline__ u128 pcg_affine(u128 s, u128 A, u128 D, u128 inc)
{
    return u128_add(u128_mul(A, s), u128_mul(D, inc));
}

__device__ __forceinline__ uint64_t pcg_output(u128 s)
{
    const uint64_t x = s.hi ^ s.lo;
    const unsigned rot = (unsigned)(s.hi >> 58);
    return (x >> rot) | (x << ((64u - rot) & 63u));
}

__device__ __forceinline__ double pcg_to_double(uint64_t x)
{
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

// NaN-propagating maximum (np.max semantics).
__device__ __forceinline__ double jk_nanmax(double a, double b)
{
    if (a != a || b != b) return __builtin_nan("");
    return a > b ? a : b;
}
