// jk_kernels.hip -- gfx950 kernels of The Joker's likelihood + rejection path.
//
//  k_lnl<L, PHYS, MIXED>   fused per-sample marginal lnL: phase, Kepler solve,
//                          design column, moments, L x L Cholesky, lnL.  One
//                          lane per prior sample; the star's epoch arrays are
//                          wave-uniform and read through scalar loads.
//  k_max_partial / final   NaN-propagating max (np.max).
//  k_reject_words          numpy-PCG64 uniforms generated in registers for the
//                          rows that can pass, exp(ll - max) > u, one 64-bit
//                          mask word per lane.
//  k_reject_bsum/compact   order-preserving compaction of the mask words.
//  k_uniform               the raw uniforms (tests / callers).
//  k_linear_posterior      a = G^-1 rhs and G^-1 for accepted rows (one wave
//                          per row for diagonal-covariance stars).
#include <cmath>
#include <type_traits>

#include "jk_internal.h"

#define JK_PI_HALF 1.57079632679489661923
#define JK_PI_HI 3.141592653589793
#define JK_PI_LO 1.2246467991473532e-16
#define JK_PI_F_HI 3.1415927410125732f
#define JK_PI_F_LO -8.742277657347586e-08f
#define JK_LN2 0.69314718055994530942

__constant__ PcgJump c_jump;
// (A^t, D_t) for t = 0 .. 255: the jump of thread t of a 256-thread workgroup
// from its workgroup's first state.
__constant__ u128 c_tid_A[256];
__constant__ u128 c_tid_D[256];
// (A^64t, D_64t) for t = 0 .. 255: the same for lanes that own 64 rows each.
__constant__ u128 c_tid64_A[256];
__constant__ u128 c_tid64_D[256];

// (sin g_i, cos g_i) on the grid g_i = i * JK_TAB_H, i = 0 .. JK_TAB_N, over
// [0, pi].  JK_TAB_H is pi / 1024 rounded up to 42 significant bits, so i * H
// (i < 2^11) is exact in FP64.  Uploaded once per device by jk_init; the
// kernels that evaluate the likelihood stage it in LDS (26 KB with the FP32
// grid and the log table).  At this spacing cos d needs only the terms to d^4.
#define JK_TAB_N 1024
#define JK_TAB_H 0x1.921fb54443000p-9
#define JK_TAB_INV_H_F 325.9493f
__device__ double2 g_sincos_tab[JK_TAB_N + 1];
// The FP32 grid for the mixed-precision Kepler step: g_i = i * JK_TAB_H32
// (pi / 1024 rounded to FP32), (sin g_i, cos g_i) rounded once to FP32.
#define JK_TAB_H32 0x1.921fb6p-9f
__device__ float2 g_sincos_tab32[JK_TAB_N + 1];
// (R_i, -ln R_i) of jk_log_tab.
__device__ double2 g_log_tab[JK_LOG_TAB_N];
// LDS image: the FP64 table, then the FP32 table packed two per double2 slot,
// then the log table.
#define JK_TAB_LOG_OFF (JK_TAB_N + 1 + (JK_TAB_N + 2) / 2)
#define JK_TAB_LDS (JK_TAB_LOG_OFF + JK_LOG_TAB_N)

// Minimum waves per SIMD asked of the row kernels that evaluate many stars
// or scattered rows: 4 (at most 128 VGPRs) for L <= 3, where the allocator
// otherwise settled some variants at 129-132 VGPRs and 3 waves without
// spilling at 128; none for larger L, which would spill.
#define JK_ROW_WAVES(L) ((L) <= 3 ? 4 : 1)

__device__ __forceinline__ void jk_stage_tab(double2 *s_tab)
{
    float2 *t32 = reinterpret_cast<float2 *>(s_tab + JK_TAB_N + 1);
    for (int i = threadIdx.x; i <= JK_TAB_N; i += blockDim.x) {
        s_tab[i] = g_sincos_tab[i];
        t32[i] = g_sincos_tab32[i];
    }
    for (int i = threadIdx.x; i < JK_LOG_TAB_N; i += blockDim.x)
        s_tab[JK_TAB_LOG_OFF + i] = g_log_tab[i];
    __syncthreads();
}

// sin and cos in FP64 of any finite FP64 angle x: Cody-Waite reduction to
// y in [-pi, pi], then the grid as in jk_sincos_tab on |y| (exact d = |y| - g_i
// by Sterbenz's lemma), then the sign of y.
__device__ __forceinline__ void jk_sincos_tab_f64(double x, const double2 *__restrict__ tab,
                                                  double &s, double &c)
{
    const double y = jk_reduce_2pi(x);
    const double ay = fabs(y);
    int idx = (int)(ay * (JK_TAB_N / JK_PI));
    idx = min(max(idx, 0), JK_TAB_N);
    const double2 t = tab[idx];
    const double d = ay - (double)idx * JK_TAB_H;
    const double d2 = d * d;
    const double sd = fma(d * d2, fma(d2, 1.0 / 120.0, -1.0 / 6.0), d);
    const double cd = fma(d2, fma(d2, 1.0 / 24.0, -0.5), 1.0);
    const double sa = fma(t.x, cd, t.y * sd);
    s = y < 0.0 ? -sa : sa;
    c = fma(t.y, cd, -(t.x * sd));
}

// sin and cos in FP64 of an FP32 angle Ef in [0, pi] (a little outside is
// fine): Ef = g_i + d with g_i the grid point below, |d| <= 1.01 H, then
//   sin Ef = sin g_i cos d + cos g_i sin d,  cos Ef = cos g_i cos d - sin g_i sin d
// with sin d to d^5 and cos d to d^4 (truncation below 2e-18).  d is exact:
// Ef has 24 significant bits and g_i 42 on a grid of the same scale, so their
// difference needs fewer than 53.
// The index is clamped into the table whatever Ef is.
__device__ __forceinline__ void jk_sincos_tab(float Ef, const double2 *__restrict__ tab,
                                              double &s, double &c)
{
    int idx = (int)(Ef * JK_TAB_INV_H_F);
    idx = min(max(idx, 0), JK_TAB_N);
    const double2 t = tab[idx];
    const double d = (double)Ef - (double)idx * JK_TAB_H;
    const double d2 = d * d;
    const double sd = fma(d * d2, fma(d2, 1.0 / 120.0, -1.0 / 6.0), d);
    const double cd = fma(d2, fma(d2, 1.0 / 24.0, -0.5), 1.0);
    s = fma(t.x, cd, t.y * sd);
    c = fma(t.y, cd, -(t.x * sd));
}

// sin and cos in FP32 of E in [0, pi] (a little outside is fine) from the FP32
// grid staged after the FP64 table: E = g_i + d, d = E - i H32 rounded once
// (fma), |d| <= 1.01 H; sin d to d^3 and cos d to d^2 (truncation below
// 4e-12, under FP32 rounding).  The index is clamped into the table.
__device__ __forceinline__ void jk_sincos_tab_f32(float E, const double2 *__restrict__ tab,
                                                  float &s, float &c)
{
    const float2 *__restrict__ t32 = reinterpret_cast<const float2 *>(tab + JK_TAB_N + 1);
    int idx = (int)(E * JK_TAB_INV_H_F);
    idx = min(max(idx, 0), JK_TAB_N);
    const float2 t = t32[idx];
    const float d = fmaf(-(float)idx, JK_TAB_H32, E);
    const float d2 = d * d;
    const float sd = fmaf(d * d2, -1.0f / 6.0f, d);
    const float cd = fmaf(d2, -0.5f, 1.0f);
    s = fmaf(t.x, cd, t.y * sd);
    c = fmaf(t.y, cd, -(t.x * sd));
}

// Kepler in FP64 on ax = |x| in [0, pi] (see jk_kepler_f64 in jk_device.h for
// the convergence argument); this variant evaluates sin/cos with the LDS
// table (jk_sincos_tab_f64).
__device__ __forceinline__ void jk_kepler_poly_f64(double x, double e,
                                                   const double2 *__restrict__ tab,
                                                   double &sinE, double &cosE)
{
    const double ax = fabs(x);
    double E = fmin(fma(0.85, e, ax), JK_PI);
    double s, c;
    jk_sincos_tab_f64(E, tab, s, c);
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
        jk_sincos_tab_f64(E, tab, s, c);
    }
    sinE = x < 0.0 ? -s : s;
    cosE = c;
}

// ---------------------------------------------------------------------------
// Non-iterative Kepler solve.
// Starter: Markley (1995, "Kepler equation solver", Celest. Mech. Dyn.
// Astron. 63, 101), a Pade-based cubic accurate to 4.4e-4 rad for every e in
// [0, 1) and asymptotically exact as e -> 1, M -> 0.  It needs only that
// accuracy, so it runs in FP32 with the hardware sqrt / log2 / exp2 / rcp.
// Correction: with f(E) = E - e sin E - M expanded about the starter E1,
//   f(E1 + d) = f0 + f1 d + f2 d^2/2 + f3 d^3/6 + f4 d^4/24 + ...,
//   f0 = E1 - e sin E1 - M, f1 = 1 - e cos E1, f2 = e sin E1, f3 = e cos E1,
//   f4 = -f2,
// the root of the quartic is taken by series reversion in t = -f0 / f1:
//   d = t - b2 t^2 + (2 b2^2 - b3) t^3 + (-5 b2^3 + 5 b2 b3 - b4) t^4
// with b_k = f_k / (k! f1).  This is the fifth-order step of Markley's
// scheme computed with one reciprocal instead of three; from the starter's
// error it leaves an error below the conditioning of the problem
// (ulp(M) / (1 - e cos E)), the answer a converged Newton iteration gives.
// (sin E, cos E) at E = E1 + d come from rotating (sin E1, cos E1) by d with
// fourth-order series (|d| <= 4.4e-4, truncation < 2e-19).
// M is taken in [0, pi].  Per-sample constants: ome = 1 - e and
// a1 = 1.6 pi / ((1 + e) (pi^2 - 6)).
// ---------------------------------------------------------------------------
#define JK_MARKLEY_A0 7.6516382901912925f   // 3 pi^2 / (pi^2 - 6)
#define JK_MARKLEY_A1N 1.29898246041084f    // 1.6 pi / (pi^2 - 6)

__device__ __forceinline__ float jk_markley_starter(float M, float e, float ome, float a1)
{
    const float alpha = fmaf(a1, JK_PI_F_HI - M, JK_MARKLEY_A0);
    const float d = fmaf(alpha, e, 3.0f * ome);
    const float M2 = M * M;
    const float ad = alpha * d;
    const float q = fmaf(2.0f * ad, ome, -M2);
    const float r = M * fmaf(3.0f * ad, d - ome, M2);
    const float q2 = q * q;
    const float x = fabsf(r) + __builtin_amdgcn_sqrtf(fmaxf(fmaf(q2, q, r * r), 0.0f));
    const float w = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * 0.6666666667f);
    const float den = fmaf(w, w + q, q2);
    // (2 r w / den + M) / d with a single reciprocal
    return fmaf(2.0f * r, w, M * den) * __builtin_amdgcn_rcpf(den * d);
}

// 1/x in FP64 from the hardware reciprocal refined by Newton steps.
__device__ __forceinline__ double jk_rcp1(double x)
{
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

__device__ __forceinline__ double jk_rcp2(double x)
{
    const double r = jk_rcp1(x);
    return fma(r, fma(-x, r, 1.0), r);
}

// Returns false when the correction is outside the starter's error bound
// (|d| >= dmax), or always when dmax < 0 (e so close to 1 that its FP32 copy
// misleads the starter): the caller then re-evaluates the whole sample with
// the bracketed Newton iteration, outside the epoch loop, so that rare path
// adds no registers to the loop.
// With WV the epoch's variance v (sigma^2 + s^2) comes in and its reciprocal
// goes out in *w: both reciprocals come from one, R = 1 / ((1 - e cos E1) v),
// as r = R v and w = R (1 - e cos E1) (two roundings more than separate
// reciprocals, one hardware reciprocal fewer per epoch).
// sin E takes the sign of x by copysign: sin E1 and its rotation are >= 0 on
// [0, pi] up to rounding at the ends of the interval.
template <bool WV = false>
__device__ __forceinline__ bool jk_kepler_markley_f64(double x, double e, float ef, float omef,
                                                      float a1f, double dmax,
                                                      const double2 *__restrict__ tab,
                                                      double &sinE, double &cosE,
                                                      double v = 0.0, double *w = nullptr)
{
    const double ax = fabs(x);
    const float E1f = jk_markley_starter((float)ax, ef, omef, a1f);
    const double E1 = (double)E1f;
    double s1, c1;
    jk_sincos_tab(E1f, tab, s1, c1);
    const double f0 = fma(-e, s1, E1 - ax);
    // One Newton step makes r good to the square of the hardware estimate's
    // error, far below what d (<= 4.4e-4) needs.
    const double D1 = fma(-e, c1, 1.0);
    double r;
    if constexpr (WV) {
        const double R = jk_rcp1(D1 * v);
        r = R * v;
        *w = R * D1;
    } else {
        r = jk_rcp1(D1);
    }
    const double t = -f0 * r;
    const double b2 = (0.5 * e) * s1 * r;
    // e cos E1 r = r - 1 (r (1 - e cos E1) = 1 to a few ulp; the absolute
    // error of r - 1 moves d by t^3 5e-16, far below ulp(E))
    const double b3 = (r - 1.0) * (1.0 / 6.0);
    const double b22 = b2 * b2;
    const double c3 = fma(2.0, b22, -b3);
    const double c4 = b2 * fma(5.0, b3 - b22, 1.0 / 12.0);   // -b4 = b2 / 12
    const double d = t * fma(t, fma(t, fma(t, c4, c3), -b2), 1.0);
    const double d2 = d * d;
    const double cd = fma(d2, fma(d2, 1.0 / 24.0, -0.5), 1.0);
    const double sd = d * fma(d2, -1.0 / 6.0, 1.0);
    const double s = fma(s1, cd, c1 * sd);
    cosE = fma(c1, cd, -s1 * sd);
    sinE = copysign(s, x);
    return fabs(d) < dmax;
}

// FP32 Kepler for the mixed-precision row: the starter, its sin and cos, and
// one Newton step t = -f0 / f1 applied to (sin, cos) to first order.  In
// FP32 the higher orders of the FP64 correction buy nothing: emulated over
// the Beta prior and e up to 1 - 1e-6 (M down to 1e-9), the design-column
// error of this chain and of the fourth-order one are the same fraction
// (0.2) of the per-epoch error model of the bound, which FP32 rounding of
// sin E, cos E and 1 / (1 - e cos E) sets, not the truncation (t^2 ~ 2e-7).
__device__ __forceinline__ void jk_kepler_markley_f32(float x, float e, float ome, float a1,
                                                      const double2 *__restrict__ tab,
                                                      float &sinE, float &cosE)
{
    const float ax = fabsf(x);
    const float E1 = jk_markley_starter(ax, e, ome, a1);
    float s1, c1;
    jk_sincos_tab_f32(E1, tab, s1, c1);
    const float f0 = fmaf(-e, s1, E1 - ax);
    const float t = -f0 * __builtin_amdgcn_rcpf(fmaf(-e, c1, 1.0f));
    const float s = fmaf(c1, t, s1);
    cosE = fmaf(-s1, t, c1);
    sinE = x < 0.0f ? -s : s;
}

// ---------------------------------------------------------------------------
// Fused marginal-likelihood kernel.
// ---------------------------------------------------------------------------
// With BOUND (mixed precision only), *bnd receives a bound on
// |lnL(mixed) - lnL(FP64)| for the row.  First order in the design-column
// error dm_n (the only FP32 quantity that reaches the FP64 moments):
//   dlnL = a_K sum_n w_n r_n dm_n - sum_n w_n dm_n (G^-1 x_n)_0,
// with a_K the posterior mean of K and r_n the residuals at the optimum, so by
// Cauchy-Schwarz, with |dm_n| <= eps_n,
//   |dlnL| <= sqrt(sum_n w_n eps_n^2) (|a_K| sqrt(chi2_min) + sqrt((G^-1)_00)).
// eps_n = 2^-22 (1 + d_n)^2 (1 + d_n / 1000), d_n = 1 / (1 - e cos E_n), covers
// the FP32 solve and design column with a factor of two to spare (emulated
// over the Beta prior and e up to 1 - 1e-6).  eps_n grows with d_n, so
// sum_n w_n eps_n^2 <= eps(max_n d_n)^2 sum_n w_n: the epoch loop keeps only
// sum w and the largest d_n (two instructions per epoch).  The bound is
// doubled again for the terms of second order and given a relative floor of
// 1e-9.
#define JK_BOUND_EPS 2.384185791015625e-07   // 2^-22
// Error-bound value of a row whose lnL entry is only an upper bound on its
// lnL (a row pruned by k_lnl_prune_batch in mixed precision); 0 marks an FP64
// value, a positive bound a mixed-precision value.
#define JK_BOUND_UPPER (-1.0f)

// With ONE the first trend column is the constant 1 (StarDev::t0_one): its
// products fold away, with the same roundings (w * 1 = w exactly).
template <int L, bool PHYS, bool MIXED, bool SAFE = false, bool BOUND = false, bool ONE = false>
__device__ __forceinline__ double jk_lnl_row(const StarDev &st, int64_t r,
                                             const double *__restrict__ P,
                                             const void *__restrict__ e_,
                                             const void *__restrict__ om_,
                                             const void *__restrict__ M0_,
                                             const void *__restrict__ s_,
                                             const double2 *__restrict__ tab,
                                             float *bnd = nullptr)
{
    static_assert(!BOUND || MIXED, "the error bound belongs to the mixed-precision row");
    if constexpr (BOUND) *bnd = INFINITY;
    // Both precisions read the FP64 columns: the kernels are bound by
    // arithmetic, not by the prior stream, and an FP32 copy of e would carry
    // a relative error of up to 2^-24 / (1 - e) into sqrt(1 - e^2), the scale
    // of the whole design column as e -> 1.
    const double *__restrict__ ecol = static_cast<const double *>(e_);
    const double *__restrict__ omcol = static_cast<const double *>(om_);
    const double *__restrict__ M0col = static_cast<const double *>(M0_);
    const double *__restrict__ scol = static_cast<const double *>(s_);

    const double Pv = P[r];
    const double ev = ecol[r];
    const double omv = omcol[r];
    const double M0v = M0col[r];
    const double sv = PHYS ? scol[r] : 0.0;

    bool valid = isfinite(Pv) && Pv != 0.0 && ev >= 0.0 && ev < 1.0 && isfinite(omv) &&
                 isfinite(M0v);
    if (PHYS) valid = valid && isfinite(sv);
    if (!valid) return __builtin_nan("");

    double lam0;
    if (st.k_prior_mode == JK_KPRIOR_FIXED_COMPANION_MASS)
        lam0 = jk_k_variance(&st, Pv, ev, true);
    else
        lam0 = st.lam0;

    const double invP = jk_rcp2(Pv);
    const double nP = JK_TWO_PI * invP;
    const double M0t = M0v * JK_INV_TWO_PI;
    const double s2 = sv * sv;

    // Per-sample orbit constants.
    double cw, ecw, swsq;
    // FP32 copies for the Kepler starter (e capped below 1 in FP32; 1 - e
    // rounded from its FP64 value, so it keeps its relative accuracy).
    const float ef = fminf((float)ev, 0.99999994f);
    const float omef = fmaxf((float)(1.0 - ev), 5.9604645e-08f);
    const float a1f = JK_MARKLEY_A1N * __builtin_amdgcn_rcpf(1.0f + ef);
    double sw;
    jk_sincos_tab_f64(omv, tab, sw, cw);
    ecw = ev * cw;
    swsq = sw * sqrt((1.0 - ev) * (1.0 + ev));
    // FP32 copies for the mixed row's design value, which is combined in FP32
    const float cwf = (float)cw, swsqf = (float)swsq, ecwf = (float)ecw;
    // The FP32 starter sees e rounded to FP32; within 1e-7 of 1 that can put
    // it outside the correction's reach, so such samples take the fallback.
    const double dmax = ev > 0.9999999 ? -1.0 : 1e-3;

    constexpr int LT = L > 1 ? L - 1 : 1;
    double Smm = 0.0, Smr = 0.0;
    double Smt[LT];
    double Stt[LT * LT], Str[LT], Srr = 0.0;
    double prod = 1.0;
    int pexp = 0;
#pragma unroll
    for (int j = 0; j < LT; ++j) {
        Smt[j] = 0.0;
        Str[j] = 0.0;
#pragma unroll
        for (int k = 0; k < LT; ++k) Stt[j * LT + k] = 0.0;
    }

    const int N = st.n;
    bool bad = false;
    double Sw = 0.0;       // sum w (BOUND only)
    float dmx = 0.0f;      // largest 1 / (1 - e cos E) (BOUND only)
#pragma unroll 2
    for (int k = 0; k < N; ++k) {
        const double dtk = st.dt[k];
        double x;
        if constexpr (MIXED) {
            // the FP32 Kepler step reads x rounded to FP32: the low part of
            // 2 pi (k * 2.4e-16 rad) is far below that rounding
            const double ph = fma(nP, dtk, -M0v);
            x = fma(-rint(ph * JK_INV_TWO_PI), JK_TWO_PI_HI, ph);
        } else {
            // the phase in turns, dt / P - M0 / (2 pi), one rounding; its
            // fractional part is exact (Sterbenz) and scales to radians with
            // one more
            const double pt = fma(invP, dtk, -M0t);
            x = (pt - rint(pt)) * JK_TWO_PI;
        }
        double w = 0.0;
        const double vk = PHYS ? st.sig2[k] + s2 : 0.0;
        double m;
        float denb = 0.0f;
        if constexpr (MIXED) {
            float sE, cE;
            jk_kepler_markley_f32((float)x, ef, omef, a1f, tab, sE, cE);
            const float den = __builtin_amdgcn_rcpf(fmaf(-ef, cE, 1.0f));
            // the design value combined in FP32 too (an FP32 instruction
            // issues in half the cycles of an FP64 one): its rounding stays
            // inside the per-epoch error model of the bound
            m = (double)fmaf(fmaf(cwf, cE - ef, -swsqf * sE), den, ecwf);
            denb = den;
        } else {
            double sE, cE;
            if constexpr (SAFE) {
                jk_kepler_poly_f64(x, ev, tab, sE, cE);
                if constexpr (PHYS) w = jk_rcp1(vk);
            } else if constexpr (PHYS) {
                bad |= !jk_kepler_markley_f64<true>(x, ev, ef, omef, a1f, dmax, tab, sE, cE,
                                                    vk, &w);
            } else {
                bad |= !jk_kepler_markley_f64(x, ev, ef, omef, a1f, dmax, tab, sE, cE);
            }
            // m = [cos w (cos E - e) - sin w sqrt(1-e^2) sin E] / (1 - e cos E) + e cos w
            // (one Newton step on the hardware reciprocal: relative 2e-15)
            const double den = jk_rcp1(fma(-ev, cE, 1.0));
            m = fma(fma(cw, cE - ev, -swsq * sE), den, ecw);
        }
        const double rp = st.rp[k];
        if constexpr (PHYS) {
            // w = 1 / (sigma^2 + s^2): FP64 to a relative 2.5e-15 (one Newton
            // step on the hardware reciprocal, shared with the Kepler step's),
            // or in the mixed row the FP32 reciprocal (relative error below
            // 2^-22, which the bound covers); sum ln w = -ln of the running
            // product of the FP64 variances, renormalised every fourth epoch
            // (frexp scales by a power of two, exactly, so the product's
            // rounding is that of renormalising every epoch; four factors
            // stay far inside the exponent range)
            if constexpr (MIXED) w = (double)__builtin_amdgcn_rcpf((float)vk);
            prod *= vk;
            if ((k & 3) == 3) {
                int ex;
                prod = frexp(prod, &ex);
                pexp += ex;
            }
        } else {
            w = st.ivar[k];
        }
        if constexpr (BOUND) {
            Sw += w;
            dmx = fmaxf(dmx, denb);
        }
        const double wm = w * m;
        Smm = fma(wm, m, Smm);
        Smr = fma(wm, rp, Smr);
        if constexpr (L > 1) {
#pragma unroll
            for (int j = 0; j < L - 1; ++j) {
                const double Tj = ONE && j == 0 ? 1.0 : st.T[j * N + k];
                Smt[j] = fma(wm, Tj, Smt[j]);
                if constexpr (PHYS) {
                    const double wT = w * Tj;
                    Str[j] = fma(wT, rp, Str[j]);
#pragma unroll
                    for (int q = 0; q <= j; ++q)
                        Stt[j * LT + q] =
                            fma(wT, ONE && q == 0 ? 1.0 : st.T[q * N + k], Stt[j * LT + q]);
                }
            }
        }
        if constexpr (PHYS) Srr = fma(w * rp, rp, Srr);
    }

    if constexpr (!MIXED && !SAFE) {
        if (bad) return jk_lnl_row<L, PHYS, false, true>(st, r, P, e_, om_, M0_, s_, tab);
    }

    // Assemble G (lower triangle used), u, chi0, log-constant.
    double G[L * L], u[L];
    const double mu0 = st.mu0;
    const bool fcm = st.k_prior_mode == JK_KPRIOR_FIXED_COMPANION_MASS;
    G[0] = (fcm ? jk_rcp2(lam0) : st.lam_inv[0]) + Smm;
    u[0] = fma(-mu0, Smm, Smr);
    // log-constant: st.ln_c / ln_c_phys hold everything but the sample's
    // product prod(sigma_n^2 + s^2) = prod * 2^pexp and, for FixedCompanionMass,
    // ln lam0
    double chi0, lnc, scale = 1.0;
    if constexpr (PHYS) {
        chi0 = Srr;
        lnc = st.ln_c_phys + (double)pexp * JK_LN2;
        scale = prod;
    } else {
        chi0 = st.S_rr;
        lnc = st.ln_c;
    }
    if (fcm) scale *= lam0;
    chi0 = fma(mu0 * mu0, Smm, fma(-2.0 * mu0, Smr, chi0));
    if constexpr (L > 1) {
#pragma unroll
        for (int j = 1; j < L; ++j) {
            G[j * L] = Smt[j - 1];
            G[j] = Smt[j - 1];
            const double trj = PHYS ? Str[j - 1] : st.S_tr[j - 1];
            u[j] = fma(-mu0, Smt[j - 1], trj);
#pragma unroll
            for (int q = 1; q <= j; ++q) {
                const double tt = PHYS ? Stt[(j - 1) * LT + (q - 1)] : st.S_tt[(j - 1) * JK_LMAX + (q - 1)];
                const double g = (q == j ? st.lam_inv[j] : 0.0) + tt;
                G[j * L + q] = g;
                G[q * L + j] = g;
            }
        }
    }
    if constexpr (BOUND) {
        double post[3];
        const double ll = jk_finish_lnl<L, true>(G, u, chi0, lnc, scale, tab + JK_TAB_LOG_OFF, post);
        const double aK = fabs(mu0 + post[1]);
        const double t1 = 1.0 + (double)dmx;
        const double ek = t1 * t1 * fma(1e-3, (double)dmx, 1.0);
        double b = 2.0 * JK_BOUND_EPS * ek * sqrt(Sw) *
                   fma(aK, sqrt(fmax(post[2], 0.0)), sqrt(fmax(post[0], 0.0)));
        // the weights' relative error e_w moves lnL by at most
        // e_w (chi2_min + L) / 2 (envelope theorem on chi2_min; the log-det
        // term is tr(G^-1 X^T W X) e_w <= L e_w); e_w < 2^-22 for the FP32
        // reciprocal of the physical jitter mode, 2.4e-15 otherwise
        b += (PHYS ? JK_BOUND_EPS : 2e-15) * (fmax(post[2], 0.0) + (double)L);
        b = fma(1e-9, 1.0 + fabs(ll), b) * 1.000001;
        *bnd = (isfinite(b) && isfinite(ll)) ? (float)b : INFINITY;
        return ll;
    } else {
        return jk_finish_lnl<L>(G, u, chi0, lnc, scale, tab + JK_TAB_LOG_OFF);
    }
}

// ---------------------------------------------------------------------------
// Full-covariance row, FP64.  With C = Q diag(lambda) Q^T, C + s^2 I is
// diagonal in the same eigenbasis (lambda + s^2), and the marginal
// likelihood ln N(y | M mu, C + s^2 I + M Lambda M^T) does not change under
// the orthogonal change of basis (|det Q| = 1).  The star's data are stored
// rotated (StarDev::Qt); the row's design column is computed per epoch as in
// jk_lnl_row, kept in this lane's column of an LDS buffer (epoch k at
// s_m[k * JK_COV_LANES]), rotated, m'_i = sum_k Qt[i][k] m_k, and fed to the
// same moments and finish with the eigenvalues as the variances.  Each lane
// reads back only its own writes, so no barrier is needed.
// ---------------------------------------------------------------------------
#define JK_COV_LANES 64

template <int L, bool PHYS, bool SAFE = false>
__device__ double jk_lnl_row_cov(const StarDev &st, int64_t r, const double *__restrict__ P,
                                 const void *__restrict__ e_, const void *__restrict__ om_,
                                 const void *__restrict__ M0_, const void *__restrict__ s_,
                                 const double2 *__restrict__ tab, double *s_m)
{
    const double *__restrict__ ecol = static_cast<const double *>(e_);
    const double *__restrict__ omcol = static_cast<const double *>(om_);
    const double *__restrict__ M0col = static_cast<const double *>(M0_);
    const double *__restrict__ scol = static_cast<const double *>(s_);
    const double Pv = P[r], ev = ecol[r], omv = omcol[r], M0v = M0col[r];
    const double sv = PHYS ? scol[r] : 0.0;
    bool valid = isfinite(Pv) && Pv != 0.0 && ev >= 0.0 && ev < 1.0 && isfinite(omv) &&
                 isfinite(M0v);
    if (PHYS) valid = valid && isfinite(sv);
    if (!valid) return __builtin_nan("");
    const double lam0 = st.k_prior_mode == JK_KPRIOR_FIXED_COMPANION_MASS
                            ? jk_k_variance(&st, Pv, ev, true)
                            : st.lam0;
    const double nP = JK_TWO_PI * jk_rcp2(Pv);
    const double s2 = sv * sv;
    const float ef = fminf((float)ev, 0.99999994f);
    const float omef = fmaxf((float)(1.0 - ev), 5.9604645e-08f);
    const float a1f = JK_MARKLEY_A1N * __builtin_amdgcn_rcpf(1.0f + ef);
    double sw, cw;
    jk_sincos_tab_f64(omv, tab, sw, cw);
    const double ecw = ev * cw;
    const double swsq = sw * sqrt((1.0 - ev) * (1.0 + ev));
    const double dmax = ev > 0.9999999 ? -1.0 : 1e-3;
    const int N = st.n;

    // the design column in epoch order
    bool bad = false;
    for (int k = 0; k < N; ++k) {
        const double x = jk_reduce_2pi(fma(nP, st.dt[k], -M0v));
        double sE, cE;
        if constexpr (SAFE)
            jk_kepler_poly_f64(x, ev, tab, sE, cE);
        else
            bad |= !jk_kepler_markley_f64(x, ev, ef, omef, a1f, dmax, tab, sE, cE);
        const double den = jk_rcp2(fma(-ev, cE, 1.0));
        s_m[k * JK_COV_LANES] = fma(fma(cw, cE - ev, -swsq * sE), den, ecw);
    }
    if constexpr (!SAFE) {
        if (bad) return jk_lnl_row_cov<L, PHYS, true>(st, r, P, e_, om_, M0_, s_, tab, s_m);
    }

    // moments in the eigenbasis
    constexpr int LT = L > 1 ? L - 1 : 1;
    double Smm = 0.0, Smr = 0.0, Srr = 0.0;
    double Smt[LT], Stt[LT * LT], Str[LT];
    double prod = 1.0;
    int pexp = 0;
#pragma unroll
    for (int j = 0; j < LT; ++j) {
        Smt[j] = 0.0;
        Str[j] = 0.0;
#pragma unroll
        for (int q = 0; q < LT; ++q) Stt[j * LT + q] = 0.0;
    }
    for (int i = 0; i < N; ++i) {
        const double *__restrict__ q = st.Qt + (size_t)i * N;
        double m = 0.0;
        for (int k = 0; k < N; ++k) m = fma(q[k], s_m[k * JK_COV_LANES], m);
        const double rp = st.rp[i];
        double w;
        if constexpr (PHYS) {
            const double d = st.sig2[i] + s2;
            w = jk_rcp2(d);
            int ex;
            prod = frexp(prod * d, &ex);
            pexp += ex;
        } else {
            w = st.ivar[i];
        }
        const double wm = w * m;
        Smm = fma(wm, m, Smm);
        Smr = fma(wm, rp, Smr);
        if constexpr (L > 1) {
#pragma unroll
            for (int j = 0; j < L - 1; ++j) {
                const double Tj = st.T[j * N + i];
                Smt[j] = fma(wm, Tj, Smt[j]);
                if constexpr (PHYS) {
                    const double wT = w * Tj;
                    Str[j] = fma(wT, rp, Str[j]);
#pragma unroll
                    for (int qq = 0; qq <= j; ++qq)
                        Stt[j * LT + qq] = fma(wT, st.T[qq * N + i], Stt[j * LT + qq]);
                }
            }
        }
        if constexpr (PHYS) Srr = fma(w * rp, rp, Srr);
    }

    // G, u, chi0 and the log-constant as in jk_lnl_row
    double G[L * L], u[L];
    const double mu0 = st.mu0;
    const bool fcm = st.k_prior_mode == JK_KPRIOR_FIXED_COMPANION_MASS;
    G[0] = (fcm ? jk_rcp2(lam0) : st.lam_inv[0]) + Smm;
    u[0] = fma(-mu0, Smm, Smr);
    double chi0, lnc, scale = 1.0;
    if constexpr (PHYS) {
        chi0 = Srr;
        lnc = st.ln_c_phys + (double)pexp * JK_LN2;
        scale = prod;
    } else {
        chi0 = st.S_rr;
        lnc = st.ln_c;
    }
    if (fcm) scale *= lam0;
    chi0 = fma(mu0 * mu0, Smm, fma(-2.0 * mu0, Smr, chi0));
    if constexpr (L > 1) {
#pragma unroll
        for (int j = 1; j < L; ++j) {
            G[j * L] = Smt[j - 1];
            G[j] = Smt[j - 1];
            const double trj = PHYS ? Str[j - 1] : st.S_tr[j - 1];
            u[j] = fma(-mu0, Smt[j - 1], trj);
#pragma unroll
            for (int q = 1; q <= j; ++q) {
                const double tt = PHYS ? Stt[(j - 1) * LT + (q - 1)] : st.S_tt[(j - 1) * JK_LMAX + (q - 1)];
                const double g = (q == j ? st.lam_inv[j] : 0.0) + tt;
                G[j * L + q] = g;
                G[q * L + j] = g;
            }
        }
    }
    return jk_finish_lnl<L>(G, u, chi0, lnc, scale, tab + JK_TAB_LOG_OFF);
}

// ---------------------------------------------------------------------------
// Block-level helpers (256 threads = 4 wavefronts of 64).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_nanmax(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = jk_nanmax(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ int64_t wave_sum_i64(int64_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Inclusive scan across the wavefront.
__device__ __forceinline__ int64_t wave_incl_scan_i64(int64_t v, int lane)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int64_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// Dense launch: ll[i] = lnL(row rows[i] or i), one lane per row, the grid
// (jk_dense_blocks) striding over the rows so that each workgroup stages the
// sin/cos table once for several rows.  With blockmax != nullptr each wave
// also writes the NaN-propagating maximum of its rows to blockmax[wave]
// (4 jk_dense_blocks(n) entries), so the rejection step need not read the
// lnL array a second time to find the maximum.  With BOUND (mixed
// precision), bound[i] receives the row's error bound (jk_lnl_row).
template <int L, bool PHYS, bool MIXED, bool BOUND, bool ONE = false>
__global__ __launch_bounds__(256) void k_lnl(const StarDev st, int64_t n,
                                             const double *__restrict__ P,
                                             const void *__restrict__ e_,
                                             const void *__restrict__ om_,
                                             const void *__restrict__ M0_,
                                             const void *__restrict__ s_,
                                             const int64_t *__restrict__ rows,
                                             double *__restrict__ ll,
                                             double *__restrict__ blockmax,
                                             float *__restrict__ bound)
{
    __shared__ double2 s_tab[JK_TAB_LDS];
    jk_stage_tab(s_tab);
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    double vmax = -INFINITY;
    for (int64_t i = t; i < n; i += stride) {
        const int64_t r = rows ? rows[i] : i;
        double v;
        if constexpr (BOUND) {
            float b;
            v = jk_lnl_row<L, PHYS, MIXED, false, true>(st, r, P, e_, om_, M0_, s_, s_tab, &b);
            bound[i] = b;
        } else {
            v = jk_lnl_row<L, PHYS, MIXED, false, false, ONE>(st, r, P, e_, om_, M0_, s_, s_tab);
        }
        ll[i] = v;
        vmax = jk_nanmax(vmax, v);
    }
    if (blockmax) {
        // one partial per wave: no barrier, so a finished wave retires at once
        vmax = wave_nanmax(vmax);
        if ((threadIdx.x & 63) == 0) blockmax[t >> 6] = vmax;
    }
}

// Full-covariance stars, FP64: one lane per row, 64 lanes per workgroup, the
// design columns in dynamic LDS (N x 64 doubles).
template <int L, bool PHYS>
__global__ __launch_bounds__(JK_COV_LANES) void k_lnl_cov(const StarDev st, int64_t n,
                                                          const double *__restrict__ P,
                                                          const void *__restrict__ e_,
                                                          const void *__restrict__ om_,
                                                          const void *__restrict__ M0_,
                                                          const void *__restrict__ s_,
                                                          const int64_t *__restrict__ rows,
                                                          double *__restrict__ ll)
{
    __shared__ double2 s_tab[JK_TAB_LDS];
    extern __shared__ double s_m[];
    jk_stage_tab(s_tab);
    const int64_t i = (int64_t)blockIdx.x * JK_COV_LANES + threadIdx.x;
    if (i < n) {
        const int64_t r = rows ? rows[i] : i;
        ll[i] = jk_lnl_row_cov<L, PHYS>(st, r, P, e_, om_, M0_, s_, s_tab, s_m + threadIdx.x);
    }
}

template <int L, bool PHYS>
static hipError_t launch_lnl_cov_t(const jk_star *st, int64_t n, const double *P, const void *e,
                                   const void *om, const void *M0, const void *s,
                                   const int64_t *rows, double *ll, hipStream_t stream)
{
    const int64_t blocks = (n + JK_COV_LANES - 1) / JK_COV_LANES;
    const size_t lds = sizeof(double) * (size_t)st->dev.n * JK_COV_LANES;
    hipLaunchKernelGGL((k_lnl_cov<L, PHYS>), dim3((unsigned)blocks), dim3(JK_COV_LANES), lds,
                       stream, st->dev, n, P, e, om, M0, s, rows, ll);
    return hipGetLastError();
}

static hipError_t launch_lnl_cov(const jk_star *st, int64_t n, const double *P, const void *e,
                                 const void *om, const void *M0, const void *s,
                                 const int64_t *rows, double *ll, hipStream_t stream)
{
    if (st->dev.n > JK_COV_NMAX || (n + JK_COV_LANES - 1) / JK_COV_LANES > 0x7fffffff)
        return hipErrorInvalidValue;
    const bool phys = st->dev.jitter_mode == JK_JITTER_PHYSICAL;
#define JK_CCASE(LL)                                                                           \
    case LL:                                                                                   \
        return phys ? launch_lnl_cov_t<LL, true>(st, n, P, e, om, M0, s, rows, ll, stream)     \
                    : launch_lnl_cov_t<LL, false>(st, n, P, e, om, M0, s, rows, ll, stream);
    switch (st->dev.L) {
        JK_CCASE(1)
        JK_CCASE(2)
        JK_CCASE(3)
        JK_CCASE(4)
        JK_CCASE(5)
        JK_CCASE(6)
        JK_CCASE(7)
        JK_CCASE(8)
    default:
        return hipErrorInvalidValue;
    }
#undef JK_CCASE
}

// Scatter launch over a device-resident index list: for j < *count,// Scatter launch over a device-resident index list: for j < *count,
// i = idx[j], ll[i] = lnL(row rows[i] or i).  Grid-stride, so the host does
// not need the count.
template <int L, bool PHYS>
__global__ __launch_bounds__(256, JK_ROW_WAVES(L)) void k_lnl_scatter_f64(const StarDev st,
                                                         const int64_t *__restrict__ count,
                                                         const int64_t *__restrict__ idx,
                                                         const double *__restrict__ P,
                                                         const void *__restrict__ e_,
                                                         const void *__restrict__ om_,
                                                         const void *__restrict__ M0_,
                                                         const void *__restrict__ s_,
                                                         const int64_t *__restrict__ rows,
                                                         double *__restrict__ ll)
{
    __shared__ double2 s_tab[JK_TAB_LDS];
    jk_stage_tab(s_tab);
    const int64_t cnt = *count;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += stride) {
        const int64_t i = idx[j];
        const int64_t r = rows ? rows[i] : i;
        ll[i] = jk_lnl_row<L, PHYS, false>(st, r, P, e_, om_, M0_, s_, s_tab);
    }
}

// Workgroups of a dense launch over n rows: one row per lane up to 16384
// workgroups (2^22 rows), then a stride.
static int64_t jk_dense_blocks(int64_t n)
{
    const int64_t b = (n + 255) / 256;
    return b < 16384 ? (b < 1 ? 1 : b) : 16384;
}

template <int L, bool PHYS, bool MIXED>
static hipError_t launch_lnl_t(const jk_star *st, int64_t n, const double *P, const void *e,
                               const void *om, const void *M0, const void *s,
                               const int64_t *rows, double *ll, double *blockmax, float *bound,
                               hipStream_t stream)
{
    const int64_t blocks = jk_dense_blocks(n);
    if constexpr (MIXED) {
        if (bound) {
            hipLaunchKernelGGL((k_lnl<L, PHYS, true, true>), dim3((unsigned)blocks), dim3(256), 0,
                               stream, st->dev, n, P, e, om, M0, s, rows, ll, blockmax, bound);
            return hipGetLastError();
        }
    }
    if constexpr (!MIXED && (L == 2 || L == 3)) {
        // the constant offset as the first trend column (the usual v0)
        if (st->dev.t0_one) {
            hipLaunchKernelGGL((k_lnl<L, PHYS, false, false, true>), dim3((unsigned)blocks),
                               dim3(256), 0, stream, st->dev, n, P, e, om, M0, s, rows, ll,
                               blockmax, nullptr);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((k_lnl<L, PHYS, MIXED, false>), dim3((unsigned)blocks), dim3(256), 0,
                       stream, st->dev, n, P, e, om, M0, s, rows, ll, blockmax, nullptr);
    return hipGetLastError();
}

template <bool MIXED>
static hipError_t launch_lnl_dispatch(const jk_star *st, int64_t n, const double *P, const void *e,
                                      const void *om, const void *M0, const void *s,
                                      const int64_t *rows, double *ll, double *bm, float *bd,
                                      hipStream_t stream)
{
    const bool phys = st->dev.jitter_mode == JK_JITTER_PHYSICAL;
#define JK_CASE(LL)                                                                                \
    case LL:                                                                                       \
        return phys                                                                                \
                   ? launch_lnl_t<LL, true, MIXED>(st, n, P, e, om, M0, s, rows, ll, bm, bd, stream) \
                   : launch_lnl_t<LL, false, MIXED>(st, n, P, e, om, M0, s, rows, ll, bm, bd,         \
                                                    stream);
    switch (st->dev.L) {
        JK_CASE(1)
        JK_CASE(2)
        JK_CASE(3)
        JK_CASE(4)
        JK_CASE(5)
        JK_CASE(6)
        JK_CASE(7)
        JK_CASE(8)
    default:
        return hipErrorInvalidValue;
    }
#undef JK_CASE
}

hipError_t jk_launch_lnl_f64(const jk_star *st, int64_t n, const double *P, const double *e,
                             const double *om, const double *M0, const double *s,
                             const int64_t *rows, double *ll, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    if (st->dev.Qt) return launch_lnl_cov(st, n, P, e, om, M0, s, rows, ll, stream);
    return launch_lnl_dispatch<false>(st, n, P, e, om, M0, s, rows, ll, nullptr, nullptr, stream);
}

hipError_t jk_launch_lnl_mixed(const jk_star *st, int64_t n, const double *P, const double *e,
                               const double *om, const double *M0, const double *s,
                               const int64_t *rows, double *ll, float *bound, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    return launch_lnl_dispatch<true>(st, n, P, e, om, M0, s, rows, ll, nullptr, bound, stream);
}


__global__ __launch_bounds__(256) void k_max_partial(const double *__restrict__ ll, int64_t n,
                                                     double *__restrict__ partials)
{
    __shared__ double red[4];
    double m = -INFINITY;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
        m = jk_nanmax(m, ll[i]);
    m = wave_nanmax(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        partials[blockIdx.x] = jk_nanmax(jk_nanmax(red[0], red[1]), jk_nanmax(red[2], red[3]));
}

__global__ __launch_bounds__(256) void k_max_final(const double *__restrict__ partials, int np,
                                                   double *__restrict__ out)
{
    __shared__ double red[4];
    double m = -INFINITY;
    for (int i = threadIdx.x; i < np; i += 256) m = jk_nanmax(m, partials[i]);
    m = wave_nanmax(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) *out = jk_nanmax(jk_nanmax(red[0], red[1]), jk_nanmax(red[2], red[3]));
}

int jk_max_partials_count(int64_t n)
{
    int64_t b = (n + 255) / 256;
    if (b > 1024) b = 1024;
    if (b < 1) b = 1;
    return (int)b;
}

hipError_t jk_launch_max(const double *ll, int64_t n, double *partials, int n_partials,
                         double *max_out, hipStream_t stream)
{
    hipLaunchKernelGGL(k_max_partial, dim3(n_partials), dim3(256), 0, stream, ll, n, partials);
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, stream, partials, n_partials, max_out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// PCG64 uniforms and rejection.
// Thread g of a grid of T threads owns indices g, g + T, g + 2T, ...; its
// state starts at s_base (the state whose output is index 0) jumped by g
// with the 2^i table, then advances by the (A_T, D_T) jump per iteration.
// Waves step through 64-aligned index ranges, so one ballot word covers
// indices [64 w, 64 w + 64).
// ---------------------------------------------------------------------------
__device__ __forceinline__ u128 pcg_jump_table(u128 s, uint64_t k, u128 inc)
{
#pragma unroll 1
    for (int b = 0; k != 0; ++b, k >>= 1)
        if (k & 1) s = pcg_affine(s, c_jump.A[b], c_jump.D[b], inc);
    return s;
}

// State of thread g = blockIdx.x * 256 + threadIdx.x (256-thread workgroups):
// the workgroup's jump by blockIdx.x * 256 is the same for every lane, so its
// loop does not diverge; the lane's own jump by threadIdx.x is one step from
// the per-lane table.  Equal to pcg_jump_table(s, g, inc).
__device__ __forceinline__ u128 pcg_thread_start(u128 s, u128 inc)
{
    uint64_t k = blockIdx.x;
#pragma unroll 1
    for (int b = 8; k != 0; ++b, k >>= 1)
        if (k & 1) s = pcg_affine(s, c_jump.A[b], c_jump.D[b], inc);
    return pcg_affine(s, c_tid_A[threadIdx.x], c_tid_D[threadIdx.x], inc);
}

static int jk_rng_grid(int64_t n)
{
    int64_t b = (n + 256 * 16 - 1) / (256 * 16);
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    return (int)b;
}

__global__ __launch_bounds__(256) void k_uniform(int64_t n, u128 s_base, u128 inc, u128 AT, u128 DT,
                                                 double *__restrict__ u)
{
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t T = (int64_t)gridDim.x * 256;
    u128 s = pcg_thread_start(s_base, inc);
    for (int64_t i = g; i < n; i += T) {
        u[i] = pcg_to_double(pcg_output(s));
        s = pcg_affine(s, AT, DT, inc);
    }
}

// Strides of the row loop whose loads the flag kernels issue together
// (k_seg_decide on 2^24 rows of ll + bound: 66 -> 51 us).
#define JK_FLAG_UNROLL 4

// The rejection test exp(x) > u of likelihood_helpers.py:108 for a uniform u
// on numpy's grid {k 2^-53} in [0, 1), evaluated exactly as written, but
// without the exp where it cannot decide anything: for x < -40 and u > 0,
// exp(x) <= exp(-40) < 4.3e-18 < 2^-53 <= u.  Most rows of a peaked
// likelihood lie that far below the maximum, so most waves skip the exp.
__device__ __forceinline__ bool jk_accept(double x, double u)
{
    bool keep = false;
    if (x >= -40.0 || u == 0.0) keep = exp(x) > u;
    return keep;
}

// Below this x = lnL - max no uniform, not even 0, passes the test: exp(x)
// is 0 in FP64 from x < -745.14 on.
#define JK_REJECT_FLOOR (-746.0)
// Marked rows among a wave's first 1024 above which k_reject_words walks the
// wave's words row by row instead of lane by lane.
#define JK_REJECT_DENSE 128

// The rejection flags with one lane per 64-row mask word: lane g owns rows
// [64 g, 64 g + 64), so a wave owns 4096 contiguous rows.  First the wave
// marks, with coalesced loads and one ballot per 64 rows, the rows with
// x = lnL - max >= JK_REJECT_FLOOR, the only ones a uniform can accept:
// iteration k reads rows 64 k + lane of the wave's range, and its ballot is
// the candidate word of the wave's lane k.  Then each lane draws uniforms for
// the marked rows of its own word alone: the state of its first row is the
// workgroup's jump (the same for every lane, by the 2^i table from bit 14
// on: 256 lanes x 64 rows per workgroup) followed by the lane's own
// (A^64t, D_64t), and each marked row's state is reached from the previous
// one by one affine step per set bit of the gap.  Every row gets the uniform
// of its own index in the stream and the same jk_accept as a row-by-row pass,
// so the words are the same; a lane whose rows all lie far below the maximum
// (most, for a peaked likelihood) does no 128-bit arithmetic, where a
// row-by-row pass steps the LCG once per row (2^24 rows of a peaked
// likelihood: 46 us row by row, 33 us this way, 38 us with each lane reading
// its own 64 rows).
__global__ __launch_bounds__(256) void k_reject_words(const double *__restrict__ ll, int64_t n,
                                                      const double *__restrict__ llmax,
                                                      u128 s_base, u128 inc,
                                                      uint64_t *__restrict__ mask)
{
    const int t = threadIdx.x, lane = t & 63;
    const int64_t g = (int64_t)blockIdx.x * 256 + t;
    const int64_t w0 = (g - lane) * 64;          // the wave's first row
    if (w0 >= n) return;                         // the whole wave
    const double mx = *llmax;
    const int64_t r0 = g * 64;
    // the workgroup's jump, the same for every lane: 256 lanes x 64 rows
    u128 s = s_base;
    uint64_t kb = blockIdx.x;
#pragma unroll 1
    for (int b = 14; kb != 0; ++b, kb >>= 1)
        if (kb & 1) s = pcg_affine(s, c_jump.A[b], c_jump.D[b], inc);
    // marks of the wave's words, sixteen loads in flight per lane before their
    // ballots; after the first sixteen words (1024 rows), a wave with many
    // rows in reach takes the row-by-row walk instead
    uint64_t cand = 0;
    bool dense = false;
    for (int k0 = 0; k0 < 64; k0 += 16) {
        double v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int64_t r = w0 + 64 * (k0 + q) + lane;
            v[q] = r < n ? ll[r] : -INFINITY;
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint64_t bits = __ballot(v[q] - mx >= JK_REJECT_FLOOR);
            cand = lane == k0 + q ? bits : cand;
        }
        if (k0 == 0 && wave_sum_i64((int64_t)__popcll(cand)) > JK_REJECT_DENSE) {
            dense = true;
            break;
        }
    }
    uint64_t word = 0;
    if (dense) {
        // Many rows in reach (a flat likelihood): the wave walks its words in
        // order, lane L at row 64 k + L of word k, as a row-by-row pass does
        // with coalesced reads and one 64-row step of every state per word.
        // State of lane L at row L: the wave's jump by 4096 w (bits 12-13)
        // and the lane's (A^L, D_L).
        for (int b = 12, kw = t >> 6; kw != 0; ++b, kw >>= 1)
            if (kw & 1) s = pcg_affine(s, c_jump.A[b], c_jump.D[b], inc);
        s = pcg_affine(s, c_tid_A[lane], c_tid_D[lane], inc);
        const u128 A64 = c_jump.A[6], D64 = c_jump.D[6];
        for (int k0 = 0; k0 < 64; k0 += JK_FLAG_UNROLL) {
            double xv[JK_FLAG_UNROLL];
#pragma unroll
            for (int q = 0; q < JK_FLAG_UNROLL; ++q) {
                const int64_t r = w0 + 64 * (k0 + q) + lane;
                xv[q] = r < n ? ll[r] - mx : -INFINITY;
            }
#pragma unroll
            for (int q = 0; q < JK_FLAG_UNROLL; ++q) {
                bool keep = false;
                if (xv[q] >= JK_REJECT_FLOOR)
                    keep = jk_accept(xv[q], pcg_to_double(pcg_output(s)));
                const uint64_t wk = __ballot(keep);
                word = lane == k0 + q ? wk : word;
                s = pcg_affine(s, A64, D64, inc);
            }
        }
    } else if (cand != 0) {
        // Few: each lane draws the uniforms of its own marked rows.
        s = pcg_affine(s, c_tid64_A[t], c_tid64_D[t], inc);
        int p = 0;
        while (cand != 0) {
            const int j = __builtin_ctzll(cand);
            cand &= cand - 1;
            for (int d = j - p, b = 0; d != 0; ++b, d >>= 1)
                if (d & 1) s = pcg_affine(s, c_jump.A[b], c_jump.D[b], inc);
            p = j;
            if (jk_accept(ll[r0 + j] - mx, pcg_to_double(pcg_output(s)))) word |= 1ull << j;
        }
    }
    if (r0 < n) mask[g] = word;
}

// Rows to re-evaluate in FP64 after a mixed-precision pass: every row not
// below max - window, and every row when the maximum itself is not finite.
__global__ __launch_bounds__(256) void k_window_flags(const double *__restrict__ ll, int64_t n,
                                                      int64_t first,
                                                      const double *__restrict__ llmax,
                                                      double window,
                                                      uint64_t *__restrict__ mask)
{
    const int lane = threadIdx.x & 63;
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t T = (int64_t)gridDim.x * 256;
    const double mx = *llmax;
    const bool all = !isfinite(mx);
    const double thr = mx - window;
    for (int64_t base = g - lane; base < n; base += T) {
        const int64_t i = base + lane;
        const bool keep = i < n && i >= first && (all || !(ll[i] < thr));
        const uint64_t word = __ballot(keep);
        if (lane == 0) mask[base >> 6] = word;
    }
}

#define JK_WORDS_PER_BLOCK 4096   // 16 words per thread

__global__ __launch_bounds__(256) void k_reject_bsum(const uint64_t *__restrict__ mask,
                                                     int64_t nwords, int64_t *__restrict__ bsum)
{
    __shared__ int64_t red[4];
    const int64_t w0 = (int64_t)blockIdx.x * JK_WORDS_PER_BLOCK;
    int64_t c = 0;
    for (int k = threadIdx.x; k < JK_WORDS_PER_BLOCK; k += 256) {
        const int64_t w = w0 + k;
        if (w < nwords) c += __popcll(mask[w]);
    }
    c = wave_sum_i64(c);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) bsum[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void k_reject_compact(const uint64_t *__restrict__ mask,
                                                        int64_t nwords,
                                                        const int64_t *__restrict__ bsum,
                                                        int nblocks, int64_t max_keep,
                                                        int64_t *__restrict__ idx_out,
                                                        int64_t *__restrict__ n_pass_out)
{
    __shared__ int64_t red[4];
    __shared__ int64_t s_off;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;

    // Offset of this block = sum of the counts of the blocks before it; block
    // 0 also publishes the total.
    int64_t c = 0;
    const int lim = blockIdx.x == 0 ? nblocks : (int)blockIdx.x;
    for (int b = threadIdx.x; b < lim; b += 256) c += bsum[b];
    c = wave_sum_i64(c);
    if (lane == 0) red[wid] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int64_t tot = red[0] + red[1] + red[2] + red[3];
        if (blockIdx.x == 0) {
            *n_pass_out = tot;
            s_off = 0;
        } else {
            s_off = tot;
        }
    }
    __syncthreads();
    const int64_t off = s_off;
    if (off >= max_keep) return;   // uniform across the block

    const int64_t w0 = (int64_t)blockIdx.x * JK_WORDS_PER_BLOCK + (int64_t)threadIdx.x * 16;
    uint64_t words[16];
    int64_t cnt = 0;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int64_t w = w0 + q;
        words[q] = w < nwords ? mask[w] : 0ull;
        cnt += __popcll(words[q]);
    }
    // Exclusive block scan of cnt.
    const int64_t incl = wave_incl_scan_i64(cnt, lane);
    __syncthreads();
    if (lane == 63) red[wid] = incl;
    __syncthreads();
    int64_t pre = incl - cnt;
    for (int k = 0; k < wid; ++k) pre += red[k];
    int64_t rank = off + pre;
#pragma unroll 1
    for (int q = 0; q < 16 && rank < max_keep; ++q) {
        uint64_t word = words[q];
        while (word != 0ull && rank < max_keep) {
            const int bit = __builtin_ctzll(word);
            idx_out[rank++] = ((w0 + q) << 6) + bit;
            word &= word - 1ull;
        }
    }
}

int64_t jk_reject_mask_words(int64_t n) { return (n + 63) / 64; }
int64_t jk_reject_bsum_count(int64_t n)
{
    return (jk_reject_mask_words(n) + JK_WORDS_PER_BLOCK - 1) / JK_WORDS_PER_BLOCK;
}

// ---------------------------------------------------------------------------
// Host-side PCG64 jump arithmetic.
// ---------------------------------------------------------------------------
static u128 h_mul(u128 a, u128 b)
{
    unsigned __int128 x = ((unsigned __int128)a.hi << 64) | a.lo;
    unsigned __int128 y = ((unsigned __int128)b.hi << 64) | b.lo;
    unsigned __int128 z = x * y;
    u128 r;
    r.lo = (uint64_t)z;
    r.hi = (uint64_t)(z >> 64);
    return r;
}

static u128 h_add(u128 a, u128 b)
{
    unsigned __int128 x = ((unsigned __int128)a.hi << 64) | a.lo;
    unsigned __int128 y = ((unsigned __int128)b.hi << 64) | b.lo;
    unsigned __int128 z = x + y;
    u128 r;
    r.lo = (uint64_t)z;
    r.hi = (uint64_t)(z >> 64);
    return r;
}

static const u128 kPcgMult = {0x4385DF649FCCF645ULL, 0x2360ED051FC65DA4ULL};

// A^k and D_k = sum_{j<k} A^j (mod 2^128).
void jk_pcg_jump_coeffs(uint64_t k, u128 *A, u128 *D)
{
    u128 accA = {1, 0}, accD = {0, 0};
    u128 curA = kPcgMult, curD = {1, 0};
    const u128 one = {1, 0};
    while (k) {
        if (k & 1) {
            // acc <- cur o acc:  A = curA accA,  D = curA accD + curD
            accD = h_add(h_mul(curA, accD), curD);
            accA = h_mul(curA, accA);
        }
        // cur <- cur o cur:  D = curA curD + curD = (curA + 1) curD
        curD = h_mul(h_add(curA, one), curD);
        curA = h_mul(curA, curA);
        k >>= 1;
    }
    *A = accA;
    *D = accD;
}

static u128 h_jump(u128 s, u128 inc, uint64_t k)
{
    u128 A, D;
    jk_pcg_jump_coeffs(k, &A, &D);
    return h_add(h_mul(A, s), h_mul(D, inc));
}

hipError_t jk_upload_pcg_table(int device)
{
    PcgJump tab;
    for (int b = 0; b < 64; ++b) jk_pcg_jump_coeffs(1ull << b, &tab.A[b], &tab.D[b]);
    hipError_t err =
        hipMemcpyToSymbol(HIP_SYMBOL(c_jump), &tab, sizeof(tab), 0, hipMemcpyHostToDevice);
    if (err != hipSuccess) return err;
    u128 tA[256], tD[256];
    for (int t = 0; t < 256; ++t) jk_pcg_jump_coeffs((uint64_t)t, &tA[t], &tD[t]);
    err = hipMemcpyToSymbol(HIP_SYMBOL(c_tid_A), tA, sizeof(tA), 0, hipMemcpyHostToDevice);
    if (err != hipSuccess) return err;
    err = hipMemcpyToSymbol(HIP_SYMBOL(c_tid_D), tD, sizeof(tD), 0, hipMemcpyHostToDevice);
    if (err != hipSuccess) return err;
    for (int t = 0; t < 256; ++t) jk_pcg_jump_coeffs(64 * (uint64_t)t, &tA[t], &tD[t]);
    err = hipMemcpyToSymbol(HIP_SYMBOL(c_tid64_A), tA, sizeof(tA), 0, hipMemcpyHostToDevice);
    if (err != hipSuccess) return err;
    return hipMemcpyToSymbol(HIP_SYMBOL(c_tid64_D), tD, sizeof(tD), 0, hipMemcpyHostToDevice);
}

// The sin / cos grids of jk_sincos_tab and jk_sincos_tab_f32, and the log
// table of jk_log_tab, from 80-bit long double on the host rounded once to
// FP64 / FP32.
hipError_t jk_upload_sincos_table(int device)
{
    double2 tab[JK_TAB_N + 1];
    float2 tab32[JK_TAB_N + 1];
    for (int i = 0; i <= JK_TAB_N; ++i) {
        const long double g = (long double)i * (long double)JK_TAB_H;
        tab[i] = make_double2((double)sinl(g), (double)cosl(g));
        const long double g32 = (long double)i * (long double)JK_TAB_H32;
        tab32[i] = make_float2((float)sinl(g32), (float)cosl(g32));
    }
    double2 ltab[JK_LOG_TAB_N];
    for (int i = 0; i < JK_LOG_TAB_N; ++i) {
        const double R = (double)(1.0L / (1.0L + (i + 0.5L) / JK_LOG_TAB_N));
        ltab[i] = make_double2(R, (double)(-logl((long double)R)));
    }
    hipError_t err = hipMemcpyToSymbol(HIP_SYMBOL(g_sincos_tab), tab, sizeof(tab), 0,
                                       hipMemcpyHostToDevice);
    if (err != hipSuccess) return err;
    err = hipMemcpyToSymbol(HIP_SYMBOL(g_log_tab), ltab, sizeof(ltab), 0, hipMemcpyHostToDevice);
    if (err != hipSuccess) return err;
    return hipMemcpyToSymbol(HIP_SYMBOL(g_sincos_tab32), tab32, sizeof(tab32), 0,
                             hipMemcpyHostToDevice);
}

hipError_t jk_launch_uniform(u128 state, u128 inc, uint64_t offset, int64_t n, double *u,
                             hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const int grid = jk_rng_grid(n);
    u128 AT, DT;
    jk_pcg_jump_coeffs((uint64_t)grid * 256, &AT, &DT);
    const u128 s_base = h_jump(state, inc, offset + 1);
    hipLaunchKernelGGL(k_uniform, dim3(grid), dim3(256), 0, stream, n, s_base, inc, AT, DT, u);
    return hipGetLastError();
}

hipError_t jk_launch_reject(const double *ll, int64_t n, const double *llmax, u128 state, u128 inc,
                            uint64_t offset, int64_t max_keep, int64_t *idx_out,
                            int64_t *n_pass_out, uint64_t *mask, int64_t *bsum,
                            hipStream_t stream)
{
    const u128 s_base = h_jump(state, inc, offset + 1);
    const int64_t nwords = jk_reject_mask_words(n);
    const int64_t nb = jk_reject_bsum_count(n);
    hipLaunchKernelGGL(k_reject_words, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, stream,
                       ll, n, llmax, s_base, inc, mask);
    hipLaunchKernelGGL(k_reject_bsum, dim3((unsigned)nb), dim3(256), 0, stream, mask, nwords, bsum);
    hipLaunchKernelGGL(k_reject_compact, dim3((unsigned)nb), dim3(256), 0, stream, mask, nwords, bsum,
                       (int)nb, max_keep, idx_out, n_pass_out);
    return hipGetLastError();
}

// FP64 lnL plus the rejection step of the same rows, with the maximum reduced
// inside the likelihood launch (one partial per wave).
int64_t jk_lnl_reject_scratch_bytes(int64_t n)
{
    const int64_t waves = 4 * jk_dense_blocks(n);
    return (int64_t)sizeof(double) * (waves + 1024) +
           (int64_t)sizeof(uint64_t) * jk_reject_mask_words(n) +
           (int64_t)sizeof(int64_t) * jk_reject_bsum_count(n);
}

hipError_t jk_launch_lnl_reject_f64(const jk_star *st, int64_t n, const double *P, const double *e,
                                    const double *om, const double *M0, const double *s,
                                    const int64_t *rows, double *ll, u128 state, u128 inc,
                                    uint64_t offset, int64_t max_keep, int64_t *idx_out,
                                    int64_t *n_pass_out, double *llmax_out, void *scratch,
                                    hipStream_t stream)
{
    const int64_t waves = 4 * jk_dense_blocks(n);
    double *partials = (double *)scratch;       // one per likelihood wave
    double *partials2 = partials + waves;       // second level, at most 1024
    uint64_t *mask = (uint64_t *)(partials2 + 1024);
    int64_t *bsum = (int64_t *)(mask + jk_reject_mask_words(n));
    hipError_t err;
    if (st->dev.Qt) {
        // full covariance: the maximum in its own two-level pass over ll
        err = launch_lnl_cov(st, n, P, e, om, M0, s, rows, ll, stream);
        if (err != hipSuccess) return err;
        err = jk_launch_max(ll, n, partials2, jk_max_partials_count(n), llmax_out, stream);
    } else {
        err = launch_lnl_dispatch<false>(st, n, P, e, om, M0, s, rows, ll, partials, nullptr,
                                         stream);
        if (err != hipSuccess) return err;
        // the per-wave maxima reduced like any array (two parallel levels)
        err = jk_launch_max(partials, waves, partials2, jk_max_partials_count(waves), llmax_out,
                            stream);
    }
    if (err != hipSuccess) return err;
    return jk_launch_reject(ll, n, llmax_out, state, inc, offset, max_keep, idx_out, n_pass_out,
                            mask, bsum, stream);
}

template <int L, bool PHYS>
static void launch_scatter_t(const jk_star *st, const int64_t *count, const int64_t *idx, int grid,
                             const double *P, const void *e, const void *om, const void *M0,
                             const void *s, const int64_t *rows, double *ll, hipStream_t stream)
{
    hipLaunchKernelGGL((k_lnl_scatter_f64<L, PHYS>), dim3(grid), dim3(256), 0, stream, st->dev,
                       count, idx, P, e, om, M0, s, rows, ll);
}

int64_t jk_polish_scratch_bytes(int64_t n)
{
    return (int64_t)sizeof(double) * (jk_max_partials_count(n) + 1) +
           (int64_t)sizeof(uint64_t) * jk_reject_mask_words(n) +
           (int64_t)sizeof(int64_t) * (jk_reject_bsum_count(n) + n + 1);
}

hipError_t jk_launch_polish(const jk_star *st, int64_t n, const double *P, const double *e,
                            const double *om, const double *M0, const double *s,
                            const int64_t *rows, double *ll, double window, int64_t first,
                            void *scratch, int64_t *n_polished, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const int np = jk_max_partials_count(n);
    const int64_t nwords = jk_reject_mask_words(n);
    const int64_t nb = jk_reject_bsum_count(n);
    char *p = (char *)scratch;
    double *partials = (double *)p;
    double *mx = partials + np;
    uint64_t *mask = (uint64_t *)(mx + 1);
    int64_t *bsum = (int64_t *)(mask + nwords);
    int64_t *idx = bsum + nb;
    int64_t *count = n_polished;
    hipError_t err = jk_launch_max(ll, n, partials, np, mx, stream);
    if (err != hipSuccess) return err;
    const int grid = jk_rng_grid(n);
    hipLaunchKernelGGL(k_window_flags, dim3(grid), dim3(256), 0, stream, ll, n, first, mx, window,
                       mask);
    hipLaunchKernelGGL(k_reject_bsum, dim3((unsigned)nb), dim3(256), 0, stream, mask, nwords, bsum);
    hipLaunchKernelGGL(k_reject_compact, dim3((unsigned)nb), dim3(256), 0, stream, mask, nwords, bsum,
                       (int)nb, n, idx, count);
    int64_t sgrid = (n + 255) / 256;
    if (sgrid > 2048) sgrid = 2048;
    const bool phys = st->dev.jitter_mode == JK_JITTER_PHYSICAL;
#define JK_SCASE(LL)                                                                                 \
    case LL:                                                                                         \
        if (phys)                                                                                    \
            launch_scatter_t<LL, true>(st, count, idx, (int)sgrid, P, e, om, M0, s, rows, ll, stream); \
        else                                                                                         \
            launch_scatter_t<LL, false>(st, count, idx, (int)sgrid, P, e, om, M0, s, rows, ll, stream); \
        break;
    switch (st->dev.L) {
        JK_SCASE(1)
        JK_SCASE(2)
        JK_SCASE(3)
        JK_SCASE(4)
        JK_SCASE(5)
        JK_SCASE(6)
        JK_SCASE(7)
        JK_SCASE(8)
    default:
        return hipErrorInvalidValue;
    }
#undef JK_SCASE
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Many stars per launch: a persistent work queue.  Each workgroup takes the
// next (star, row tile) item from a device counter, reads the star's
// constants through scalar loads (the item index is made wave-uniform with
// readfirstlane) and evaluates the tile one lane per row.  Every workgroup
// exits once the counter passes n_items.
// ---------------------------------------------------------------------------
__global__ void k_zero_u32(unsigned *p) { *p = 0u; }

// One row of a mixed-precision batched launch: in FP64 with error bound 0
// for a star whose table slot is marked exact (jk_batch_set_exact; a star
// whose rows nearly all land within the settling margin of its maximum, so
// that mixed values would be re-evaluated anyway), else in mixed precision
// with its bound.  The branch is uniform over a work item.
template <int L, bool PHYS>
__device__ __forceinline__ double jk_lnl_row_sel(const StarDev &st, int64_t r,
                                                 const double *__restrict__ P,
                                                 const void *__restrict__ e_,
                                                 const void *__restrict__ om_,
                                                 const void *__restrict__ M0_,
                                                 const void *__restrict__ s_,
                                                 const double2 *__restrict__ tab, float *bnd)
{
    if (st.exact) {
        *bnd = 0.0f;
        return jk_lnl_row<L, PHYS, false>(st, r, P, e_, om_, M0_, s_, tab);
    }
    return jk_lnl_row<L, PHYS, true, false, true>(st, r, P, e_, om_, M0_, s_, tab, bnd);
}

// Slot `slot` of a star table evaluated in FP64 (exact != 0) or as the launch
// asks (jk_batch_set_exact); a kernel, so that it is ordered on the stream and
// reads no host memory after it returns.
__global__ void k_set_exact(StarDev *stars, int slot, int exact) { stars[slot].exact = exact; }

template <int L, bool PHYS, bool MIXED, bool BOUND>
__global__ __launch_bounds__(256, JK_ROW_WAVES(L)) void k_lnl_batch(const StarDev *__restrict__ stars,
                                                   const jk_work_item *__restrict__ items,
                                                   int64_t n_items, unsigned *__restrict__ counter,
                                                   const double *__restrict__ P,
                                                   const void *__restrict__ e_,
                                                   const void *__restrict__ om_,
                                                   const void *__restrict__ M0_,
                                                   const void *__restrict__ s_,
                                                   const int64_t *__restrict__ rows,
                                                   double *__restrict__ ll,
                                                   float *__restrict__ bound)
{
    __shared__ unsigned s_item;
    __shared__ double2 s_tab[JK_TAB_LDS];
    jk_stage_tab(s_tab);
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
        __syncthreads();
        const int64_t it = (int64_t)__builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (it >= n_items) break;
        const jk_work_item w = items[it];
        const StarDev &st = stars[w.star];
        for (int j = threadIdx.x; j < w.n_rows; j += blockDim.x) {
            const int64_t q = w.row_begin + j;
            const int64_t r = rows ? rows[q] : q;
            if constexpr (BOUND) {
                float b;
                ll[w.out_offset + j] = jk_lnl_row_sel<L, PHYS>(st, r, P, e_, om_, M0_, s_, s_tab, &b);
                bound[w.out_offset + j] = b;
            } else if constexpr (MIXED) {
                ll[w.out_offset + j] =
                    st.exact ? jk_lnl_row<L, PHYS, false>(st, r, P, e_, om_, M0_, s_, s_tab)
                             : jk_lnl_row<L, PHYS, true>(st, r, P, e_, om_, M0_, s_, s_tab);
            } else {
                ll[w.out_offset + j] = jk_lnl_row<L, PHYS, false>(st, r, P, e_, om_, M0_, s_, s_tab);
            }
        }
    }
}

hipError_t jk_launch_set_exact(const jk_batch *b, int slot, int exact, hipStream_t stream)
{
    hipLaunchKernelGGL(k_set_exact, dim3(1), dim3(1), 0, stream, b->d_stars, slot, exact ? 1 : 0);
    return hipGetLastError();
}

template <int L, bool PHYS, bool MIXED>
static void launch_batch_t(const jk_batch *b, int64_t n_items, const jk_work_item *items,
                           const double *P, const void *e, const void *om, const void *M0,
                           const void *s, const int64_t *rows, double *ll, float *bound,
                           int block, int grid, unsigned *counter, hipStream_t stream)
{
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, stream, counter);
    if constexpr (MIXED) {
        if (bound) {
            hipLaunchKernelGGL((k_lnl_batch<L, PHYS, true, true>), dim3(grid), dim3(block), 0,
                               stream, b->d_stars, items, n_items, counter, P, e, om, M0, s, rows,
                               ll, bound);
            return;
        }
    }
    hipLaunchKernelGGL((k_lnl_batch<L, PHYS, MIXED, false>), dim3(grid), dim3(block), 0, stream,
                       b->d_stars, items, n_items, counter, P, e, om, M0, s, rows, ll, nullptr);
}

template <bool MIXED>
static hipError_t launch_batch_dispatch(const jk_batch *b, int64_t n_items,
                                        const jk_work_item *items, const double *P,
                                        const void *e, const void *om, const void *M0,
                                        const void *s, const int64_t *rows, double *ll,
                                        float *bd, int block, int grid, unsigned *counter,
                                        hipStream_t stream)
{
    const bool phys = b->jitter_mode == JK_JITTER_PHYSICAL;
#define JK_BCASE(LL)                                                                              \
    case LL:                                                                                      \
        if (phys)                                                                                 \
            launch_batch_t<LL, true, MIXED>(b, n_items, items, P, e, om, M0, s, rows, ll, bd,      \
                                            block, grid, counter, stream);                        \
        else                                                                                      \
            launch_batch_t<LL, false, MIXED>(b, n_items, items, P, e, om, M0, s, rows, ll, bd,     \
                                             block, grid, counter, stream);                       \
        break;
    switch (b->L) {
        JK_BCASE(1)
        JK_BCASE(2)
        JK_BCASE(3)
        JK_BCASE(4)
        JK_BCASE(5)
        JK_BCASE(6)
        JK_BCASE(7)
        JK_BCASE(8)
    default:
        return hipErrorInvalidValue;
    }
#undef JK_BCASE
    return hipGetLastError();
}

hipError_t jk_launch_lnl_batch(const jk_batch *b, int precision, int64_t n_items,
                               const jk_work_item *items, const double *P, const void *e,
                               const void *om, const void *M0, const void *s,
                               const int64_t *rows, double *ll, float *bound, int block, int grid,
                               unsigned *counter, hipStream_t stream)
{
    if (n_items <= 0) return hipSuccess;
    if (precision == 0)
        return launch_batch_dispatch<false>(b, n_items, items, P, e, om, M0, s, rows, ll, nullptr,
                                            block, grid, counter, stream);
    return launch_batch_dispatch<true>(b, n_items, items, P, e, om, M0, s, rows, ll, bound, block,
                                       grid, counter, stream);
}

// ---------------------------------------------------------------------------
// Many stars per launch with pruning by an upper bound.  probes[k] holds star
// k's data on a subset S of its epochs.  The marginal likelihood factors as
// p(y) = p(y_S) p(y_rest | y_S), and the conditional covariance of y_rest is
// at least the noise covariance C_rest, so
//   lnL <= lnL_S - 1/2 sum_{n not in S} ln(2 pi sigma_n^2) = lnL_S + add,
// for every prior row (with jitter the variances only grow).  prune[k] =
// (thr, add): a row whose bound lnL_S (+ its mixed error bound) + add is
// finite and below thr gets that bound as its value and is not evaluated
// further (bound JK_BOUND_UPPER in mixed precision: the value is an upper
// bound on the row's lnL, nothing more, and the settling kernels treat it so); every other row is evaluated as in k_lnl_batch.  Rows
// are probed a chunk of blockDim at a time and the survivors are gathered in
// LDS over several chunks, so the full evaluations run on dense waves.  thr = -inf
// turns the probe off for the star.  n_full (optional, one entry per star)
// accumulates each star's rows evaluated in full.
// ---------------------------------------------------------------------------
#define JK_PRUNE_LIST 2048   // survivors held in LDS before their evaluation

template <int L, bool PHYS, bool MIXED>
__global__ __launch_bounds__(256, JK_ROW_WAVES(L)) void k_lnl_prune_batch(const StarDev *__restrict__ stars,
                                                         const StarDev *__restrict__ probes,
                                                         const double2 *__restrict__ prune,
                                                         const jk_work_item *__restrict__ items,
                                                         int64_t n_items,
                                                         unsigned *__restrict__ counter,
                                                         const double *__restrict__ P,
                                                         const void *__restrict__ e_,
                                                         const void *__restrict__ om_,
                                                         const void *__restrict__ M0_,
                                                         const void *__restrict__ s_,
                                                         double *__restrict__ ll,
                                                         float *__restrict__ bound,
                                                         unsigned long long *__restrict__ n_full)
{
    __shared__ unsigned s_item;
    __shared__ int s_n;
    __shared__ int s_list[JK_PRUNE_LIST];
    __shared__ double2 s_tab[JK_TAB_LDS];
    jk_stage_tab(s_tab);
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
        __syncthreads();
        const int64_t it = (int64_t)__builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (it >= n_items) break;
        const jk_work_item w = items[it];
        const StarDev &st = stars[w.star];
        const double2 pr = prune[w.star];
        if (!(pr.x > -INFINITY)) {
            for (int j = threadIdx.x; j < w.n_rows; j += blockDim.x) {
                const int64_t r = w.row_begin + j;
                if constexpr (MIXED) {
                    float b;
                    ll[w.out_offset + j] =
                        jk_lnl_row_sel<L, PHYS>(st, r, P, e_, om_, M0_, s_, s_tab, &b);
                    bound[w.out_offset + j] = b;
                } else {
                    ll[w.out_offset + j] = jk_lnl_row<L, PHYS, false>(st, r, P, e_, om_, M0_, s_, s_tab);
                    if (bound) bound[w.out_offset + j] = 0.0f;
                }
            }
            if (n_full && threadIdx.x == 0)
                atomicAdd(n_full + w.star, (unsigned long long)w.n_rows);
            continue;
        }
        const StarDev &pb = probes[w.star];
        unsigned n_surv = 0;
        if (threadIdx.x == 0) s_n = 0;
        __syncthreads();
        for (int base = 0; base < w.n_rows; base += blockDim.x) {
            const int j = base + (int)threadIdx.x;
            if (j < w.n_rows) {
                const int64_t r = w.row_begin + j;
                double v;
                if constexpr (MIXED) {
                    float b;
                    v = jk_lnl_row<L, PHYS, true, false, true>(pb, r, P, e_, om_, M0_, s_, s_tab, &b);
                    v += (double)b + pr.y;
                } else {
                    v = jk_lnl_row<L, PHYS, false>(pb, r, P, e_, om_, M0_, s_, s_tab);
                    // FP64 rounding of both evaluations, far below this slack
                    v += pr.y + 1e-9 * (1.0 + fabs(v));
                }
                if (isfinite(v) && v < pr.x) {
                    ll[w.out_offset + j] = v;
                    if (bound) bound[w.out_offset + j] = JK_BOUND_UPPER;
                } else {
                    s_list[atomicAdd(&s_n, 1)] = j;
                }
            }
            __syncthreads();
            // Survivors gather over chunks and are evaluated once another
            // chunk might overflow the list, or at the item's end: the full
            // evaluations then run on whole waves.
            const int ns = s_n;
            __syncthreads();   // every thread holds ns before the next append
            if (ns > JK_PRUNE_LIST - (int)blockDim.x || base + (int)blockDim.x >= w.n_rows) {
                n_surv += (unsigned)ns;
                for (int k = threadIdx.x; k < ns; k += blockDim.x) {
                    const int jj = s_list[k];
                    const int64_t r = w.row_begin + jj;
                    if constexpr (MIXED) {
                        float b;
                        ll[w.out_offset + jj] =
                            jk_lnl_row_sel<L, PHYS>(st, r, P, e_, om_, M0_, s_, s_tab, &b);
                        bound[w.out_offset + jj] = b;
                    } else {
                        ll[w.out_offset + jj] =
                            jk_lnl_row<L, PHYS, false>(st, r, P, e_, om_, M0_, s_, s_tab);
                        if (bound) bound[w.out_offset + jj] = 0.0f;
                    }
                }
                __syncthreads();   // every read of s_list before it is reset
                if (threadIdx.x == 0) s_n = 0;
                __syncthreads();
            }
        }
        if (n_full && threadIdx.x == 0 && n_surv)
            atomicAdd(n_full + w.star, (unsigned long long)n_surv);
    }
}

template <int L, bool PHYS, bool MIXED>
static void launch_prune_t(const jk_batch *b, const jk_batch *pb, const double2 *prune,
                           int64_t n_items, const jk_work_item *items, const double *P,
                           const void *e, const void *om, const void *M0, const void *s,
                           double *ll, float *bound, unsigned long long *n_full, int block,
                           int grid, unsigned *counter, hipStream_t stream)
{
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, stream, counter);
    hipLaunchKernelGGL((k_lnl_prune_batch<L, PHYS, MIXED>), dim3(grid), dim3(block), 0, stream,
                       b->d_stars, pb->d_stars, prune, items, n_items, counter, P, e, om, M0, s,
                       ll, bound, n_full);
}

hipError_t jk_launch_lnl_prune_batch(const jk_batch *b, const jk_batch *pb, const double *prune,
                                     int precision, int64_t n_items, const jk_work_item *items,
                                     const double *P, const void *e, const void *om,
                                     const void *M0, const void *s, double *ll, float *bound,
                                     unsigned long long *n_full, int block, int grid,
                                     unsigned *counter, hipStream_t stream)
{
    if (n_items <= 0) return hipSuccess;
    const bool phys = b->jitter_mode == JK_JITTER_PHYSICAL;
    const double2 *pr = reinterpret_cast<const double2 *>(prune);
#define JK_PCASE(LL)                                                                              \
    case LL:                                                                                      \
        if (precision == 1) {                                                                     \
            if (phys)                                                                             \
                launch_prune_t<LL, true, true>(b, pb, pr, n_items, items, P, e, om, M0, s, ll,     \
                                               bound, n_full, block, grid, counter, stream);      \
            else                                                                                  \
                launch_prune_t<LL, false, true>(b, pb, pr, n_items, items, P, e, om, M0, s, ll,    \
                                                bound, n_full, block, grid, counter, stream);     \
        } else {                                                                                  \
            if (phys)                                                                             \
                launch_prune_t<LL, true, false>(b, pb, pr, n_items, items, P, e, om, M0, s, ll,    \
                                                bound, n_full, block, grid, counter, stream);     \
            else                                                                                  \
                launch_prune_t<LL, false, false>(b, pb, pr, n_items, items, P, e, om, M0, s, ll,   \
                                                 bound, n_full, block, grid, counter, stream);    \
        }                                                                                         \
        break;
    switch (b->L) {
        JK_PCASE(1)
        JK_PCASE(2)
        JK_PCASE(3)
        JK_PCASE(4)
        JK_PCASE(5)
        JK_PCASE(6)
        JK_PCASE(7)
        JK_PCASE(8)
    default:
        return hipErrorInvalidValue;
    }
#undef JK_PCASE
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Segmented rejection: the steps of jk_launch_reject with blockIdx.y picking
// the segment.  Blocks past a segment's extent exit at once.
// ---------------------------------------------------------------------------
#define JK_SEG_MAX_BLOCKS 1024

__global__ __launch_bounds__(256) void k_seg_max_partial(const double *__restrict__ ll,
                                                         const SegDev *__restrict__ segs,
                                                         double *__restrict__ partials, int pb)
{
    __shared__ double red[4];
    const SegDev sg = segs[blockIdx.y];
    double m = -INFINITY;
    const int64_t stride = (int64_t)pb * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < sg.n; i += stride)
        m = jk_nanmax(m, ll[sg.off + i]);
    m = wave_nanmax(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        partials[(int64_t)blockIdx.y * pb + blockIdx.x] =
            jk_nanmax(jk_nanmax(red[0], red[1]), jk_nanmax(red[2], red[3]));
}

__global__ __launch_bounds__(64) void k_seg_max_final(const double *__restrict__ partials, int pb,
                                                      double *__restrict__ llmax)
{
    double m = -INFINITY;
    for (int i = threadIdx.x; i < pb; i += 64) m = jk_nanmax(m, partials[(int64_t)blockIdx.x * pb + i]);
    m = wave_nanmax(m);
    if (threadIdx.x == 0) llmax[blockIdx.x] = m;
}

// zero_u (optional): zero_u[segment] is set to 1 when one of the segment's
// uniforms is exactly 0 (probability 2^-53 per row), the one case settling
// with a margin below 746 leaves open.
__global__ __launch_bounds__(256) void k_seg_flags(const double *__restrict__ ll,
                                                   const SegDev *__restrict__ segs,
                                                   const double *__restrict__ llmax, u128 AT,
                                                   u128 DT, uint64_t *__restrict__ mask,
                                                   int64_t *__restrict__ zero_u)
{
    const SegDev sg = segs[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t T = (int64_t)gridDim.x * 256;
    if (g - lane >= sg.n) return;
    const double mx = llmax[blockIdx.y];
    u128 s = pcg_thread_start(sg.s_base, sg.inc);
    for (int64_t base0 = g - lane; base0 < sg.n; base0 += JK_FLAG_UNROLL * T) {
        // the loads of JK_FLAG_UNROLL strides first
        double xs[JK_FLAG_UNROLL];
#pragma unroll
        for (int u = 0; u < JK_FLAG_UNROLL; ++u) {
            const int64_t i = base0 + u * T + lane;
            xs[u] = i < sg.n ? ll[sg.off + i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < JK_FLAG_UNROLL; ++u) {
            const int64_t base = base0 + u * T;     // the same for every lane of the wave
            if (base >= sg.n) break;
            bool keep = false;
            if (base + lane < sg.n) {
                const double uu = pcg_to_double(pcg_output(s));
                keep = jk_accept(xs[u] - mx, uu);
                if (zero_u != nullptr && uu == 0.0) zero_u[blockIdx.y] = 1;
            }
            const uint64_t word = __ballot(keep);
            if (lane == 0) mask[sg.mask_off + (base >> 6)] = word;
            s = pcg_affine(s, AT, DT, sg.inc);
        }
    }
}

__global__ __launch_bounds__(256) void k_seg_bsum(const SegDev *__restrict__ segs,
                                                  const uint64_t *__restrict__ mask,
                                                  int64_t *__restrict__ bsum)
{
    __shared__ int64_t red[4];
    const SegDev sg = segs[blockIdx.y];
    if ((int64_t)blockIdx.x >= sg.nb) return;
    const int64_t nwords = (sg.n + 63) / 64;
    const int64_t w0 = (int64_t)blockIdx.x * JK_WORDS_PER_BLOCK;
    int64_t c = 0;
    for (int k = threadIdx.x; k < JK_WORDS_PER_BLOCK; k += 256) {
        const int64_t w = w0 + k;
        if (w < nwords) c += __popcll(mask[sg.mask_off + w]);
    }
    c = wave_sum_i64(c);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) bsum[sg.bsum_off + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void k_seg_compact(const SegDev *__restrict__ segs,
                                                     const uint64_t *__restrict__ mask,
                                                     const int64_t *__restrict__ bsum,
                                                     int64_t *__restrict__ idx_out,
                                                     int64_t *__restrict__ n_pass_out)
{
    __shared__ int64_t red[4];
    __shared__ int64_t s_off;
    const SegDev sg = segs[blockIdx.y];
    if ((int64_t)blockIdx.x >= sg.nb) return;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int64_t nwords = (sg.n + 63) / 64;
    int64_t c = 0;
    const int64_t lim = blockIdx.x == 0 ? sg.nb : (int64_t)blockIdx.x;
    for (int64_t b = threadIdx.x; b < lim; b += 256) c += bsum[sg.bsum_off + b];
    c = wave_sum_i64(c);
    if (lane == 0) red[wid] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int64_t tot = red[0] + red[1] + red[2] + red[3];
        if (blockIdx.x == 0) {
            n_pass_out[blockIdx.y] = tot;
            s_off = 0;
        } else {
            s_off = tot;
        }
    }
    __syncthreads();
    const int64_t off = s_off;
    if (off >= sg.max_keep) return;
    const int64_t w0 = (int64_t)blockIdx.x * JK_WORDS_PER_BLOCK + (int64_t)threadIdx.x * 16;
    uint64_t words[16];
    int64_t cnt = 0;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int64_t w = w0 + q;
        words[q] = w < nwords ? mask[sg.mask_off + w] : 0ull;
        cnt += __popcll(words[q]);
    }
    const int64_t incl = wave_incl_scan_i64(cnt, lane);
    __syncthreads();
    if (lane == 63) red[wid] = incl;
    __syncthreads();
    int64_t pre = incl - cnt;
    for (int k = 0; k < wid; ++k) pre += red[k];
    int64_t rank = off + pre;
#pragma unroll 1
    for (int q = 0; q < 16 && rank < sg.max_keep; ++q) {
        uint64_t word = words[q];
        while (word != 0ull && rank < sg.max_keep) {
            const int bit = __builtin_ctzll(word);
            idx_out[sg.idx_off + rank++] = ((w0 + q) << 6) + bit;
            word &= word - 1ull;
        }
    }
}

int jk_seg_flag_blocks(int64_t max_n)
{
    int64_t b = (max_n + 256 * 16 - 1) / (256 * 16);
    if (b > JK_SEG_MAX_BLOCKS) b = JK_SEG_MAX_BLOCKS;
    if (b < 1) b = 1;
    return (int)b;
}

__global__ __launch_bounds__(256) void k_fill_i64(int64_t *__restrict__ p, int n, int64_t v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t jk_launch_reject_batch(const double *ll, int n_seg, const SegDev *d_segs,
                                  int64_t max_n, int64_t max_nb, double *partials, int pb,
                                  uint64_t *mask, int64_t *bsum, int64_t *idx_out,
                                  int64_t *n_pass_out, double *llmax_out, hipStream_t stream,
                                  int64_t *zero_u)
{
    if (n_seg <= 0) return hipSuccess;
    const int fb = jk_seg_flag_blocks(max_n);
    u128 AT, DT;
    jk_pcg_jump_coeffs((uint64_t)fb * 256, &AT, &DT);
    if (zero_u)
        hipLaunchKernelGGL(k_fill_i64, dim3((n_seg + 255) / 256), dim3(256), 0, stream, zero_u,
                           n_seg, (int64_t)0);
    hipLaunchKernelGGL(k_seg_max_partial, dim3(pb, n_seg), dim3(256), 0, stream, ll, d_segs,
                       partials, pb);
    hipLaunchKernelGGL(k_seg_max_final, dim3(n_seg), dim3(64), 0, stream, partials, pb, llmax_out);
    hipLaunchKernelGGL(k_seg_flags, dim3(fb, n_seg), dim3(256), 0, stream, ll, d_segs, llmax_out,
                       AT, DT, mask, zero_u);
    hipLaunchKernelGGL(k_seg_bsum, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream, d_segs,
                       mask, bsum);
    hipLaunchKernelGGL(k_seg_compact, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream, d_segs,
                       mask, bsum, idx_out, n_pass_out);
    return hipGetLastError();
}

// Segmented FP64 re-evaluation window (jk_polish_f64 per segment): flags,
// the segmented compaction above, then a scatter of FP64 lnL over each
// segment's selected entries.
__global__ __launch_bounds__(256) void k_seg_window_flags(const double *__restrict__ ll,
                                                          const SegDev *__restrict__ segs,
                                                          const double *__restrict__ llmax,
                                                          double window,
                                                          uint64_t *__restrict__ mask)
{
    const SegDev sg = segs[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t T = (int64_t)gridDim.x * 256;
    if (g - lane >= sg.n) return;
    const double mx = llmax[blockIdx.y];
    const bool all = !isfinite(mx);
    const double thr = mx - window;
    for (int64_t base = g - lane; base < sg.n; base += T) {
        const int64_t i = base + lane;
        const bool keep = i < sg.n && i >= sg.first && (all || !(ll[sg.off + i] < thr));
        const uint64_t word = __ballot(keep);
        if (lane == 0) mask[sg.mask_off + (base >> 6)] = word;
    }
}

// Padded exclusive prefix of the per-segment counts: cum[s] = sum over t < s
// of count[t] rounded up to a multiple of 64, so that a wavefront of the
// scatter below never straddles two segments.  One workgroup.
__global__ __launch_bounds__(256) void k_seg_count_scan(const int64_t *__restrict__ count,
                                                        int n_seg, int64_t *__restrict__ cum)
{
    __shared__ int64_t wsum[4];
    const int per = (n_seg + 255) / 256;
    const int lo = threadIdx.x * per;
    const int hi = min(lo + per, n_seg);
    int64_t mine = 0;
    for (int i = lo; i < hi; ++i) mine += (count[i] + 63) & ~(int64_t)63;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int64_t incl = wave_incl_scan_i64(mine, lane);
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int64_t run = incl - mine;
    for (int k = 0; k < wid; ++k) run += wsum[k];
    for (int i = lo; i < hi; ++i) {
        cum[i] = run;
        run += (count[i] + 63) & ~(int64_t)63;
    }
    if (threadIdx.x == 255) cum[n_seg] = run;
}

// FP64 re-evaluation of every segment's selected entries as one flat list
// spread over the whole grid: wavefront w takes padded positions
// [64 w, 64 w + 64), all in one segment (found by binary search in cum and
// made wave-uniform, so the star's constants stay in scalar loads).
template <int L, bool PHYS>
__global__ __launch_bounds__(256, JK_ROW_WAVES(L)) void k_seg_scatter_f64(const StarDev *__restrict__ stars,
                                                         const SegDev *__restrict__ segs,
                                                         int n_seg,
                                                         const int64_t *__restrict__ count,
                                                         const int64_t *__restrict__ cum,
                                                         const int64_t *__restrict__ idx,
                                                         const double *__restrict__ P,
                                                         const void *__restrict__ e_,
                                                         const void *__restrict__ om_,
                                                         const void *__restrict__ M0_,
                                                         const void *__restrict__ s_,
                                                         const int64_t *__restrict__ rows,
                                                         double *__restrict__ ll,
                                                         float *__restrict__ bound)
{
    __shared__ double2 s_tab[JK_TAB_LDS];
    jk_stage_tab(s_tab);
    const int lane = threadIdx.x & 63;
    const int64_t total_waves = cum[n_seg] >> 6;
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
         w < total_waves; w += wstride) {
        const int64_t j0 = w << 6;
        int lo = 0, hi = n_seg;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (cum[mid] <= j0) lo = mid; else hi = mid;
        }
        const int sidx = __builtin_amdgcn_readfirstlane(lo);
        const SegDev &sg = segs[sidx];
        const int64_t local = j0 - cum[sidx] + lane;
        if (local < count[sidx]) {
            const int64_t p = idx[sg.idx_off + local];
            const int64_t q = sg.row0 + p;
            const int64_t r = rows ? rows[q] : q;
            ll[sg.off + p] = jk_lnl_row<L, PHYS, false>(stars[sg.star], r, P, e_, om_, M0_, s_, s_tab);
            if (bound) bound[sg.off + p] = 0.0f;
        }
    }
}

template <int L, bool PHYS>
static void launch_seg_scatter_t(const jk_batch *b, int n_seg, const SegDev *d_segs,
                                 const int64_t *count, const int64_t *cum, const int64_t *idx,
                                 int gx, const double *P, const void *e, const void *om,
                                 const void *M0, const void *s, const int64_t *rows, double *ll,
                                 float *bound, hipStream_t stream)
{
    hipLaunchKernelGGL((k_seg_scatter_f64<L, PHYS>), dim3(gx), dim3(256), 0, stream, b->d_stars,
                       d_segs, n_seg, count, cum, idx, P, e, om, M0, s, rows, ll, bound);
}

// FP64 re-evaluation of every segment's compacted list idx (counts in count)
// as one flat launch; bound entries of re-evaluated rows are set to zero.
static hipError_t launch_seg_scatter(const jk_batch *b, int n_seg, const SegDev *d_segs,
                                     const int64_t *count, int64_t *cum, const int64_t *idx,
                                     const double *P, const double *e, const double *om,
                                     const double *M0, const double *s, const int64_t *rows,
                                     double *ll, float *bound, hipStream_t stream)
{
    hipLaunchKernelGGL(k_seg_count_scan, dim3(1), dim3(256), 0, stream, count, n_seg, cum);
    const int gx = 2048;
    const bool phys = b->jitter_mode == JK_JITTER_PHYSICAL;
#define JK_SCASE2(LL)                                                                             \
    case LL:                                                                                      \
        if (phys)                                                                                 \
            launch_seg_scatter_t<LL, true>(b, n_seg, d_segs, count, cum, idx, gx, P, e, om, M0, s, \
                                           rows, ll, bound, stream);                              \
        else                                                                                      \
            launch_seg_scatter_t<LL, false>(b, n_seg, d_segs, count, cum, idx, gx, P, e, om, M0,   \
                                            s, rows, ll, bound, stream);                          \
        break;
    switch (b->L) {
        JK_SCASE2(1)
        JK_SCASE2(2)
        JK_SCASE2(3)
        JK_SCASE2(4)
        JK_SCASE2(5)
        JK_SCASE2(6)
        JK_SCASE2(7)
        JK_SCASE2(8)
    default:
        return hipErrorInvalidValue;
    }
#undef JK_SCASE2
    return hipGetLastError();
}

hipError_t jk_launch_polish_batch(const jk_batch *b, int n_seg, const SegDev *d_segs,
                                  int64_t max_n, int64_t max_nb, double *partials, int pb,
                                  uint64_t *mask, int64_t *bsum, int64_t *idx, double *llmax,
                                  int64_t *cum, const double *P, const double *e, const double *om,
                                  const double *M0, const double *s, const int64_t *rows,
                                  double *ll, double window, int64_t *n_polished,
                                  hipStream_t stream)
{
    if (n_seg <= 0) return hipSuccess;
    const int fb = jk_seg_flag_blocks(max_n);
    hipLaunchKernelGGL(k_seg_max_partial, dim3(pb, n_seg), dim3(256), 0, stream, ll, d_segs,
                       partials, pb);
    hipLaunchKernelGGL(k_seg_max_final, dim3(n_seg), dim3(64), 0, stream, partials, pb, llmax);
    hipLaunchKernelGGL(k_seg_window_flags, dim3(fb, n_seg), dim3(256), 0, stream, ll, d_segs, llmax,
                       window, mask);
    hipLaunchKernelGGL(k_seg_bsum, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream, d_segs,
                       mask, bsum);
    hipLaunchKernelGGL(k_seg_compact, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream, d_segs,
                       mask, bsum, idx, n_polished);
    return launch_seg_scatter(b, n_seg, d_segs, n_polished, cum, idx, P, e, om, M0, s, rows, ll,
                              nullptr, stream);
}

// ---------------------------------------------------------------------------
// Settling a mixed-precision lnL buffer for the rejection step
// (jk_settle_batch).  Per segment, with B the per-row bounds on
// |lnL(mixed) - lnL(FP64)| (0 for rows already FP64):
//   1. lo = max(ll - B) is a lower bound on the FP64 maximum; a row with
//      ll + B >= lo could be the maximum: re-evaluate it in FP64 (B := 0).
//      So is a row with ll + B >= lo - 40 and B > 1e-3 (k_seg_shell_flags).
//      A NaN anywhere makes lo NaN and every unsettled row is re-evaluated.
//   2. mx = max(ll) is then the FP64 maximum.
//   3. u = the row's numpy uniform (the stream the rejection step draws); a
//      row whose decision exp(x - mx) > u differs between x = ll - B and
//      x = ll + B is re-evaluated in FP64 (B := 0).
// Afterwards exp(ll - mx) > u holds for exactly the rows for which it holds
// with every row's FP64 likelihood, and every row within 40 of mx is within
// 1e-3 of its FP64 lnL.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_seg_lo_partial(const double *__restrict__ ll,
                                                        const float *__restrict__ bound,
                                                        const SegDev *__restrict__ segs,
                                                        double *__restrict__ partials, int pb)
{
    __shared__ double red[4];
    const SegDev sg = segs[blockIdx.y];
    double m = -INFINITY;
    const int64_t stride = (int64_t)pb * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < sg.n; i += stride) {
        const float b = bound[sg.off + i];
        // an upper-bound row gives no lower bound on the maximum
        if (!(b < 0.0f)) m = jk_nanmax(m, ll[sg.off + i] - (double)b);
    }
    m = wave_nanmax(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        partials[(int64_t)blockIdx.y * pb + blockIdx.x] =
            jk_nanmax(jk_nanmax(red[0], red[1]), jk_nanmax(red[2], red[3]));
}

// Rows the settling re-evaluates besides the shell of the maximum: those that
// may lie within JK_LIVE_WINDOW of it (rows with a non-negligible chance of
// acceptance, exp(-40) < 4.3e-18) and whose bound exceeds JK_LIVE_TOL, so
// that after settling every such row is within JK_LIVE_TOL of its FP64 lnL.
#define JK_LIVE_WINDOW 40.0
#define JK_LIVE_TOL 1e-3f

__global__ __launch_bounds__(256) void k_seg_shell_flags(const double *__restrict__ ll,
                                                         const float *__restrict__ bound,
                                                         const SegDev *__restrict__ segs,
                                                         const double *__restrict__ lo,
                                                         uint64_t *__restrict__ mask)
{
    const SegDev sg = segs[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t T = (int64_t)gridDim.x * 256;
    if (g - lane >= sg.n) return;
    const double l0 = lo[blockIdx.y];
    for (int64_t base = g - lane; base < sg.n; base += T) {
        const int64_t i = base + lane;
        bool keep = false;
        if (i < sg.n) {
            const float b = bound[sg.off + i];
            const double x = ll[sg.off + i];
            const double hi = x + (double)b;
            // could be the maximum; or within JK_LIVE_WINDOW of it with a
            // bound above JK_LIVE_TOL (an upper-bound row: x bounds it from
            // above only)
            keep = b > 0.0f ? (!(hi < l0) || (b > JK_LIVE_TOL && !(hi < l0 - JK_LIVE_WINDOW)))
                            : (b < 0.0f && !(x < l0 - JK_LIVE_WINDOW));
        }
        const uint64_t word = __ballot(keep);
        if (lane == 0) mask[sg.mask_off + (base >> 6)] = word;
    }
}

__global__ __launch_bounds__(256) void k_seg_uncertain_flags(const double *__restrict__ ll,
                                                             const float *__restrict__ bound,
                                                             const SegDev *__restrict__ segs,
                                                             const double *__restrict__ llmax,
                                                             u128 AT, u128 DT,
                                                             uint64_t *__restrict__ mask)
{
    const SegDev sg = segs[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t T = (int64_t)gridDim.x * 256;
    if (g - lane >= sg.n) return;
    const double mx = llmax[blockIdx.y];
    const bool live = !(mx != mx);     // a NaN maximum rejects every row anyway
    u128 s = pcg_thread_start(sg.s_base, sg.inc);
    for (int64_t base = g - lane; base < sg.n; base += T) {
        const int64_t i = base + lane;
        bool keep = false;
        if (i < sg.n) {
            const double uu = pcg_to_double(pcg_output(s));
            const float bf = bound[sg.off + i];
            if (live && bf > 0.0f) {
                const double x = ll[sg.off + i];
                const double b = fma((double)bf, 1e-6, (double)bf) + 1e-12;
                keep = jk_accept(x - b - mx, uu) != jk_accept(x + b - mx, uu);
            } else if (live && bf < 0.0f) {
                // an upper bound decides a rejection only
                keep = jk_accept(ll[sg.off + i] - mx, uu);
            }
        }
        const uint64_t word = __ballot(keep);
        if (lane == 0) mask[sg.mask_off + (base >> 6)] = word;
        s = pcg_affine(s, AT, DT, sg.inc);
    }
}

__global__ void k_seg_add_counts(const int64_t *__restrict__ a, int64_t *__restrict__ b, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] += a[i];
}

hipError_t jk_launch_settle_batch(const jk_batch *b, int n_seg, const SegDev *d_segs,
                                  int64_t max_n, int64_t max_nb, double *partials, int pb,
                                  uint64_t *mask, int64_t *bsum, int64_t *idx, double *lo,
                                  double *llmax, int64_t *cum, int64_t *count1, const double *P,
                                  const double *e, const double *om, const double *M0,
                                  const double *s, const int64_t *rows, double *ll, float *bound,
                                  int64_t *n_settled, hipStream_t stream)
{
    if (n_seg <= 0) return hipSuccess;
    const int fb = jk_seg_flag_blocks(max_n);
    // 1. the rows that could hold the FP64 maximum
    hipLaunchKernelGGL(k_seg_lo_partial, dim3(pb, n_seg), dim3(256), 0, stream, ll, bound, d_segs,
                       partials, pb);
    hipLaunchKernelGGL(k_seg_max_final, dim3(n_seg), dim3(64), 0, stream, partials, pb, lo);
    hipLaunchKernelGGL(k_seg_shell_flags, dim3(fb, n_seg), dim3(256), 0, stream, ll, bound, d_segs,
                       lo, mask);
    hipLaunchKernelGGL(k_seg_bsum, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream, d_segs,
                       mask, bsum);
    hipLaunchKernelGGL(k_seg_compact, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream, d_segs,
                       mask, bsum, idx, count1);
    hipError_t err = launch_seg_scatter(b, n_seg, d_segs, count1, cum, idx, P, e, om, M0, s, rows,
                                        ll, bound, stream);
    if (err != hipSuccess) return err;
    // 2. the FP64 maximum
    hipLaunchKernelGGL(k_seg_max_partial, dim3(pb, n_seg), dim3(256), 0, stream, ll, d_segs,
                       partials, pb);
    hipLaunchKernelGGL(k_seg_max_final, dim3(n_seg), dim3(64), 0, stream, partials, pb, llmax);
    // 3. the rows whose decision the bound leaves open
    u128 AT, DT;
    jk_pcg_jump_coeffs((uint64_t)fb * 256, &AT, &DT);
    hipLaunchKernelGGL(k_seg_uncertain_flags, dim3(fb, n_seg), dim3(256), 0, stream, ll, bound,
                       d_segs, llmax, AT, DT, mask);
    hipLaunchKernelGGL(k_seg_bsum, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream, d_segs,
                       mask, bsum);
    hipLaunchKernelGGL(k_seg_compact, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream, d_segs,
                       mask, bsum, idx, n_settled);
    err = launch_seg_scatter(b, n_seg, d_segs, n_settled, cum, idx, P, e, om, M0, s, rows, ll,
                             bound, stream);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_seg_add_counts, dim3((n_seg + 255) / 256), dim3(256), 0, stream, count1,
                       n_settled, n_seg);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Settling for every future uniform (the iterative rejection loop redraws
// the uniforms of all rows each round).  For each segment's new rows:
//   1. lo = max(floor, max(ll - B)) is a lower bound on the FP64 maximum of
//      the segment (floor: the FP64 maximum of the rows settled before);
//   2. every new row with ll + B >= lo - margin is re-evaluated in FP64.
// A row left in mixed precision then has lnL(FP64) <= ll + B < max - margin,
// and so does ll itself: with margin >= 746, exp(x - max) is 0 for both, so
// the row is rejected whatever its uniform, and it cannot be the maximum.
// The maximum row itself is always re-evaluated.  A NaN in lo re-evaluates
// every new row still in mixed precision.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_seg_live_flags(const double *__restrict__ ll,
                                                        const float *__restrict__ bound,
                                                        const SegDev *__restrict__ segs,
                                                        const double *__restrict__ lo,
                                                        double margin,
                                                        uint64_t *__restrict__ mask)
{
    const SegDev sg = segs[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t T = (int64_t)gridDim.x * 256;
    if (g - lane >= sg.n) return;
    const double thr = jk_nanmax(lo[blockIdx.y], sg.floor) - margin;
    for (int64_t base = g - lane; base < sg.n; base += T) {
        const int64_t i = base + lane;
        bool keep = false;
        if (i < sg.n) {
            const float b = bound[sg.off + i];
            // an upper-bound row (b < 0): its value bounds lnL from above
            keep = b != 0.0f && !(ll[sg.off + i] + (double)fmaxf(b, 0.0f) < thr);
        }
        const uint64_t word = __ballot(keep);
        if (lane == 0) mask[sg.mask_off + (base >> 6)] = word;
    }
}

// ---------------------------------------------------------------------------
// Settling and rejection in one sequence (jk_settle_reject_batch): after the
// rows that could hold the maximum are settled and the FP64 maximum found
// (steps 1-2 of jk_settle_batch), one pass draws every row's uniform once
// and decides it -- directly where the bound leaves no doubt, else flags it
// open.  The open rows (a few per million) are re-evaluated in FP64 and
// their decisions set from a uniform jumped to their index.  The accepted
// set is the one jk_settle_batch + jk_reject_batch give, with one pass over
// the rows and one PCG64 stream fewer.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_seg_decide(const double *__restrict__ ll,
                                                    const float *__restrict__ bound,
                                                    const SegDev *__restrict__ segs,
                                                    const double *__restrict__ llmax,
                                                    u128 AT, u128 DT,
                                                    uint64_t *__restrict__ mask,
                                                    uint64_t *__restrict__ open_mask)
{
    const SegDev sg = segs[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t T = (int64_t)gridDim.x * 256;
    if (g - lane >= sg.n) return;
    const double mx = llmax[blockIdx.y];
    u128 s = pcg_thread_start(sg.s_base, sg.inc);
    for (int64_t base0 = g - lane; base0 < sg.n; base0 += JK_FLAG_UNROLL * T) {
        // the loads of JK_FLAG_UNROLL strides first
        double xs[JK_FLAG_UNROLL];
        float bs[JK_FLAG_UNROLL];
#pragma unroll
        for (int u = 0; u < JK_FLAG_UNROLL; ++u) {
            const int64_t i = base0 + u * T + lane;
            xs[u] = i < sg.n ? ll[sg.off + i] : 0.0;
            bs[u] = i < sg.n ? bound[sg.off + i] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < JK_FLAG_UNROLL; ++u) {
            const int64_t base = base0 + u * T;     // the same for every lane of the wave
            if (base >= sg.n) break;
            const int64_t i = base + lane;
            bool acc = false, open = false;
            if (i < sg.n) {
                const double uu = pcg_to_double(pcg_output(s));
                const double x = xs[u];
                const float bf = bs[u];
                if (bf > 0.0f) {
                    // a NaN maximum makes both sides false: nothing accepted, nothing open
                    const double b = fma((double)bf, 1e-6, (double)bf) + 1e-12;
                    const bool lo = jk_accept(x - b - mx, uu), hi = jk_accept(x + b - mx, uu);
                    open = lo != hi;
                    acc = lo && hi;
                } else if (bf < 0.0f) {
                    // an upper bound decides a rejection only
                    open = jk_accept(x - mx, uu);
                } else {
                    acc = jk_accept(x - mx, uu);
                }
            }
            const uint64_t wa = __ballot(acc), wo = __ballot(open);
            if (lane == 0) {
                mask[sg.mask_off + (base >> 6)] = wa;
                open_mask[sg.mask_off + (base >> 6)] = wo;
            }
            s = pcg_affine(s, AT, DT, sg.inc);
        }
    }
}

// The decisions of the open rows, now in FP64: each row's uniform from the
// segment's stream jumped to its index; accepted rows set their bit.
__global__ __launch_bounds__(256) void k_seg_patch(const SegDev *__restrict__ segs,
                                                   const int64_t *__restrict__ count,
                                                   const int64_t *__restrict__ idx,
                                                   const double *__restrict__ ll,
                                                   const double *__restrict__ llmax,
                                                   uint64_t *__restrict__ mask)
{
    const SegDev sg = segs[blockIdx.y];
    const int64_t cnt = count[blockIdx.y];
    const double mx = llmax[blockIdx.y];
    for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < cnt;
         j += (int64_t)gridDim.x * 256) {
        const int64_t i = idx[sg.idx_off + j];
        const u128 st = pcg_jump_table(sg.s_base, (uint64_t)i, sg.inc);
        const double uu = pcg_to_double(pcg_output(st));
        if (jk_accept(ll[sg.off + i] - mx, uu))
            atomicOr((unsigned long long *)&mask[sg.mask_off + (i >> 6)], 1ull << (i & 63));
    }
}

hipError_t jk_launch_settle_reject_batch(const jk_batch *b, int n_seg, const SegDev *d_settle,
                                         const SegDev *d_reject, int64_t max_n, int64_t max_nb,
                                         double *partials, int pb, uint64_t *mask,
                                         uint64_t *open_mask, int64_t *bsum, int64_t *idx,
                                         double *lo, int64_t *cum, int64_t *count1,
                                         const double *P, const double *e, const double *om,
                                         const double *M0, const double *s, const int64_t *rows,
                                         double *ll, float *bound, int64_t *idx_out,
                                         int64_t *n_pass_out, double *llmax_out,
                                         int64_t *n_settled, hipStream_t stream)
{
    if (n_seg <= 0) return hipSuccess;
    const int fb = jk_seg_flag_blocks(max_n);
    // 1. the rows that could hold the FP64 maximum, settled
    hipLaunchKernelGGL(k_seg_lo_partial, dim3(pb, n_seg), dim3(256), 0, stream, ll, bound,
                       d_settle, partials, pb);
    hipLaunchKernelGGL(k_seg_max_final, dim3(n_seg), dim3(64), 0, stream, partials, pb, lo);
    hipLaunchKernelGGL(k_seg_shell_flags, dim3(fb, n_seg), dim3(256), 0, stream, ll, bound,
                       d_settle, lo, mask);
    hipLaunchKernelGGL(k_seg_bsum, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream, d_settle,
                       mask, bsum);
    hipLaunchKernelGGL(k_seg_compact, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream,
                       d_settle, mask, bsum, idx, count1);
    hipError_t err = launch_seg_scatter(b, n_seg, d_settle, count1, cum, idx, P, e, om, M0, s,
                                        rows, ll, bound, stream);
    if (err != hipSuccess) return err;
    // 2. the FP64 maximum
    hipLaunchKernelGGL(k_seg_max_partial, dim3(pb, n_seg), dim3(256), 0, stream, ll, d_settle,
                       partials, pb);
    hipLaunchKernelGGL(k_seg_max_final, dim3(n_seg), dim3(64), 0, stream, partials, pb,
                       llmax_out);
    // 3. every row decided or flagged open, one uniform each
    u128 AT, DT;
    jk_pcg_jump_coeffs((uint64_t)fb * 256, &AT, &DT);
    hipLaunchKernelGGL(k_seg_decide, dim3(fb, n_seg), dim3(256), 0, stream, ll, bound, d_settle,
                       llmax_out, AT, DT, mask, open_mask);
    // 4. the open rows in FP64, then their decisions
    hipLaunchKernelGGL(k_seg_bsum, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream, d_settle,
                       open_mask, bsum);
    hipLaunchKernelGGL(k_seg_compact, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream,
                       d_settle, open_mask, bsum, idx, n_settled);
    err = launch_seg_scatter(b, n_seg, d_settle, n_settled, cum, idx, P, e, om, M0, s, rows, ll,
                             bound, stream);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_seg_patch, dim3(4, n_seg), dim3(256), 0, stream, d_settle, n_settled,
                       idx, ll, llmax_out, mask);
    // 5. the ordered compaction of the accepted rows
    hipLaunchKernelGGL(k_seg_bsum, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream, d_reject,
                       mask, bsum);
    hipLaunchKernelGGL(k_seg_compact, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream,
                       d_reject, mask, bsum, idx_out, n_pass_out);
    hipLaunchKernelGGL(k_seg_add_counts, dim3((n_seg + 255) / 256), dim3(256), 0, stream, count1,
                       n_settled, n_seg);
    return hipGetLastError();
}

hipError_t jk_launch_settle_live_batch(const jk_batch *b, int n_seg, const SegDev *d_segs,
                                       int64_t max_n, int64_t max_nb, double *partials, int pb,
                                       uint64_t *mask, int64_t *bsum, int64_t *idx, double *lo,
                                       int64_t *cum, const double *P, const double *e,
                                       const double *om, const double *M0, const double *s,
                                       const int64_t *rows, double *ll, float *bound,
                                       double margin, int64_t *n_settled, hipStream_t stream)
{
    if (n_seg <= 0) return hipSuccess;
    const int fb = jk_seg_flag_blocks(max_n);
    hipLaunchKernelGGL(k_seg_lo_partial, dim3(pb, n_seg), dim3(256), 0, stream, ll, bound, d_segs,
                       partials, pb);
    hipLaunchKernelGGL(k_seg_max_final, dim3(n_seg), dim3(64), 0, stream, partials, pb, lo);
    hipLaunchKernelGGL(k_seg_live_flags, dim3(fb, n_seg), dim3(256), 0, stream, ll, bound, d_segs,
                       lo, margin, mask);
    hipLaunchKernelGGL(k_seg_bsum, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream, d_segs,
                       mask, bsum);
    hipLaunchKernelGGL(k_seg_compact, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream, d_segs,
                       mask, bsum, idx, n_settled);
    return launch_seg_scatter(b, n_seg, d_segs, n_settled, cum, idx, P, e, om, M0, s, rows, ll,
                              bound, stream);
}

// ---------------------------------------------------------------------------
// Settling to a tolerance (jk_settle_tol_batch): every entry of a segment
// whose error bound exceeds tol, or whose value is only an upper bound
// (negative bound), is re-evaluated in FP64.  Afterwards every entry is
// within tol of its FP64 lnL, whatever its distance from the maximum: the
// form for lnL arrays handed to a caller.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_seg_tol_flags(const float *__restrict__ bound,
                                                       const SegDev *__restrict__ segs, float tol,
                                                       uint64_t *__restrict__ mask)
{
    const SegDev sg = segs[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t T = (int64_t)gridDim.x * 256;
    if (g - lane >= sg.n) return;
    for (int64_t base = g - lane; base < sg.n; base += T) {
        const int64_t i = base + lane;
        bool keep = false;
        if (i < sg.n) {
            const float b = bound[sg.off + i];
            keep = b < 0.0f || b > tol || b != b;
        }
        const uint64_t word = __ballot(keep);
        if (lane == 0) mask[sg.mask_off + (base >> 6)] = word;
    }
}

hipError_t jk_launch_settle_tol_batch(const jk_batch *b, int n_seg, const SegDev *d_segs,
                                      int64_t max_n, int64_t max_nb, uint64_t *mask,
                                      int64_t *bsum, int64_t *idx, int64_t *cum, const double *P,
                                      const double *e, const double *om, const double *M0,
                                      const double *s, const int64_t *rows, double *ll,
                                      float *bound, float tol, int64_t *n_settled,
                                      hipStream_t stream)
{
    if (n_seg <= 0) return hipSuccess;
    const int fb = jk_seg_flag_blocks(max_n);
    hipLaunchKernelGGL(k_seg_tol_flags, dim3(fb, n_seg), dim3(256), 0, stream, bound, d_segs, tol,
                       mask);
    hipLaunchKernelGGL(k_seg_bsum, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream, d_segs,
                       mask, bsum);
    hipLaunchKernelGGL(k_seg_compact, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream, d_segs,
                       mask, bsum, idx, n_settled);
    return launch_seg_scatter(b, n_seg, d_segs, n_settled, cum, idx, P, e, om, M0, s, rows, ll,
                              bound, stream);
}

// ---------------------------------------------------------------------------
// Settling a buffer whose entries are values or upper bounds
// (jk_settle_pruned_batch): a probe launch leaves every row's upper bound
// (bound < 0); the rows that can matter are then evaluated in FP64 around the
// maximum, found in two steps.
//   1. hi = max over the segment of (ll + max(b, 0)) bounds the FP64 maximum
//      from above; the upper-bound rows with ll >= hi - window are evaluated.
//   2. lo = max over the value rows (b >= 0) of ll - b bounds it from below;
//      the upper-bound rows with ll >= lo - margin are evaluated.
// Afterwards every upper-bound row lies more than margin below the maximum.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_seg_hi_partial(const double *__restrict__ ll,
                                                        const float *__restrict__ bound,
                                                        const SegDev *__restrict__ segs,
                                                        double *__restrict__ partials, int pb)
{
    __shared__ double red[4];
    const SegDev sg = segs[blockIdx.y];
    double m = -INFINITY;
    const int64_t stride = (int64_t)pb * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < sg.n; i += stride)
        m = jk_nanmax(m, ll[sg.off + i] + (double)fmaxf(bound[sg.off + i], 0.0f));
    m = wave_nanmax(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        partials[(int64_t)blockIdx.y * pb + blockIdx.x] =
            jk_nanmax(jk_nanmax(red[0], red[1]), jk_nanmax(red[2], red[3]));
}

// Upper-bound rows (b < 0) not below level[segment] - margin; every
// upper-bound row when the level is NaN.
__global__ __launch_bounds__(256) void k_seg_upper_flags(const double *__restrict__ ll,
                                                         const float *__restrict__ bound,
                                                         const SegDev *__restrict__ segs,
                                                         const double *__restrict__ level,
                                                         double margin,
                                                         uint64_t *__restrict__ mask)
{
    const SegDev sg = segs[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t T = (int64_t)gridDim.x * 256;
    if (g - lane >= sg.n) return;
    const double thr = level[blockIdx.y] - margin;
    for (int64_t base = g - lane; base < sg.n; base += T) {
        const int64_t i = base + lane;
        bool keep = false;
        if (i < sg.n) keep = bound[sg.off + i] < 0.0f && !(ll[sg.off + i] < thr);
        const uint64_t word = __ballot(keep);
        if (lane == 0) mask[sg.mask_off + (base >> 6)] = word;
    }
}

hipError_t jk_launch_settle_pruned_batch(const jk_batch *b, int n_seg, const SegDev *d_segs,
                                         int64_t max_n, int64_t max_nb, double *partials, int pb,
                                         uint64_t *mask, int64_t *bsum, int64_t *idx,
                                         double *level, int64_t *cum, int64_t *count1,
                                         const double *P, const double *e, const double *om,
                                         const double *M0, const double *s, const int64_t *rows,
                                         double *ll, float *bound, double window, double margin,
                                         int64_t *n_settled, hipStream_t stream)
{
    if (n_seg <= 0) return hipSuccess;
    const int fb = jk_seg_flag_blocks(max_n);
    // 1. the rows whose bound reaches the window below the upper bound on the maximum
    hipLaunchKernelGGL(k_seg_hi_partial, dim3(pb, n_seg), dim3(256), 0, stream, ll, bound, d_segs,
                       partials, pb);
    hipLaunchKernelGGL(k_seg_max_final, dim3(n_seg), dim3(64), 0, stream, partials, pb, level);
    hipLaunchKernelGGL(k_seg_upper_flags, dim3(fb, n_seg), dim3(256), 0, stream, ll, bound, d_segs,
                       level, window, mask);
    hipLaunchKernelGGL(k_seg_bsum, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream, d_segs,
                       mask, bsum);
    hipLaunchKernelGGL(k_seg_compact, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream, d_segs,
                       mask, bsum, idx, count1);
    hipError_t err = launch_seg_scatter(b, n_seg, d_segs, count1, cum, idx, P, e, om, M0, s, rows,
                                        ll, bound, stream);
    if (err != hipSuccess) return err;
    // 2. the rows whose bound reaches the margin below the lower bound on the maximum
    hipLaunchKernelGGL(k_seg_lo_partial, dim3(pb, n_seg), dim3(256), 0, stream, ll, bound, d_segs,
                       partials, pb);
    hipLaunchKernelGGL(k_seg_max_final, dim3(n_seg), dim3(64), 0, stream, partials, pb, level);
    hipLaunchKernelGGL(k_seg_upper_flags, dim3(fb, n_seg), dim3(256), 0, stream, ll, bound, d_segs,
                       level, margin, mask);
    hipLaunchKernelGGL(k_seg_bsum, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream, d_segs,
                       mask, bsum);
    hipLaunchKernelGGL(k_seg_compact, dim3((unsigned)max_nb, n_seg), dim3(256), 0, stream, d_segs,
                       mask, bsum, idx, n_settled);
    err = launch_seg_scatter(b, n_seg, d_segs, n_settled, cum, idx, P, e, om, M0, s, rows, ll,
                             bound, stream);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_seg_add_counts, dim3((n_seg + 255) / 256), dim3(256), 0, stream, count1,
                       n_settled, n_seg);
    return hipGetLastError();
}

void jk_host_pcg_jump(u128 state, u128 inc, uint64_t k, u128 *out) { *out = h_jump(state, inc, k); }

// ---------------------------------------------------------------------------
// On-device prior cache fill.  Philox4x32-10 (Salmon, Moraes, Dror & Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC 2011): counter = (row,
// block, 0, 0), key = seed; three blocks per row give six 53-bit uniforms.
//   P     = exp(u0 (ln P_max - ln P_min) + ln P_min)    (log-uniform)
//   e     = inverse Beta CDF, linear interpolation in a table of n_tab + 1
//           equally spaced quantiles built on the host
//   omega = 2 pi u2,  M0 = 2 pi u3
//   s     = exp(mu + sigma z), z = sqrt(-2 ln(1 - u4)) cos(2 pi u5), or 0
// FP64 master columns and their FP32 copy are written in one pass.
// ---------------------------------------------------------------------------
struct Philox4 {
    uint32_t v[4];
};

__device__ __forceinline__ Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                 uint32_t k0, uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(M0, c0), lo0 = M0 * c0;
        const uint32_t hi1 = __umulhi(M1, c2), lo1 = M1 * c2;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += W0;
        k1 += W1;
    }
    Philox4 o;
    o.v[0] = c0;
    o.v[1] = c1;
    o.v[2] = c2;
    o.v[3] = c3;
    return o;
}

__device__ __forceinline__ double u53(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

__global__ __launch_bounds__(256) void k_prior_fill(int64_t n, int64_t row0, uint32_t k0,
                                                    uint32_t k1, double lnPmin, double lnPspan,
                                                    const double *__restrict__ e_tab, int n_tab,
                                                    double s_mu, double s_sigma,
                                                    double *__restrict__ P, double *__restrict__ e,
                                                    double *__restrict__ om,
                                                    double *__restrict__ M0,
                                                    double *__restrict__ s, float *__restrict__ e32,
                                                    float *__restrict__ om32,
                                                    float *__restrict__ M032,
                                                    float *__restrict__ s32)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t row = (uint64_t)(row0 + i);
    const uint32_t lo = (uint32_t)row, hi = (uint32_t)(row >> 32);
    const Philox4 a = philox4x32_10(lo, hi, 0u, 0u, k0, k1);
    const Philox4 b = philox4x32_10(lo, hi, 1u, 0u, k0, k1);
    const Philox4 c = philox4x32_10(lo, hi, 2u, 0u, k0, k1);
    const double u0 = u53(a.v[0], a.v[1]), u1 = u53(a.v[2], a.v[3]);
    const double u2 = u53(b.v[0], b.v[1]), u3 = u53(b.v[2], b.v[3]);
    const double u4 = u53(c.v[0], c.v[1]), u5 = u53(c.v[2], c.v[3]);
    const double Pv = exp(fma(u0, lnPspan, lnPmin));
    const double x = u1 * n_tab;
    int k = (int)x;
    if (k >= n_tab) k = n_tab - 1;
    const double f = x - k;
    const double ev = fma(f, e_tab[k + 1] - e_tab[k], e_tab[k]);
    const double omv = JK_TWO_PI * u2, M0v = JK_TWO_PI * u3;
    double sv = 0.0;
    if (s_sigma >= 0.0) {
        const double z = sqrt(-2.0 * log1p(-u4)) * cos(JK_TWO_PI * u5);
        sv = exp(fma(s_sigma, z, s_mu));
    }
    P[i] = Pv;
    e[i] = ev;
    om[i] = omv;
    M0[i] = M0v;
    s[i] = sv;
    if (e32) {
        e32[i] = (float)ev;
        om32[i] = (float)omv;
        M032[i] = (float)M0v;
        s32[i] = (float)sv;
    }
}

hipError_t jk_launch_prior_fill(int64_t n, int64_t row0, uint64_t seed, double P_min, double P_max,
                                const double *e_tab, int n_tab, double s_mu, double s_sigma,
                                double *P, double *e, double *om, double *M0, double *s,
                                float *e32, float *om32, float *M032, float *s32,
                                hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const double lnPmin = log(P_min);
    const double span = log(P_max) - lnPmin;
    hipLaunchKernelGGL(k_prior_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, n,
                       row0, (uint32_t)seed, (uint32_t)(seed >> 32), lnPmin, span, e_tab, n_tab,
                       s_mu, s_sigma, P, e, om, M0, s, e32, om32, M032, s32);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Linear-parameter posterior for a few rows (accepted samples): FP64, one
// lane per row, runtime L <= JK_LMAX.  Lambda[0] is not clipped here
// (fast_likelihood.pyx:519-521).
// ---------------------------------------------------------------------------
// The end of a row's linear posterior from its sums over the epochs (G lower
// triangle, rhs, u, chi0, sum ln w): the prior added, Cholesky, cov = G^-1,
// a = cov rhs, and the row's lnL.
__device__ __forceinline__ void jk_posterior_finish(const StarDev &st, int64_t i, double lam0,
                                                    double *G, double *rhs, const double *u,
                                                    double chi0, double sum_ln_w,
                                                    double *__restrict__ a_out,
                                                    double *__restrict__ cov_out,
                                                    double *__restrict__ ll_out)
{
    const int L = st.L, N = st.n;
    double lnlam = log(lam0) + st.ln_lam_trend;
    G[0] += 1.0 / lam0;
    rhs[0] += st.mu0 / lam0;
    for (int a = 1; a < L; ++a) {
        G[a * JK_LMAX + a] += st.lam_inv[a];
        rhs[a] += st.mu[a] * st.lam_inv[a];
    }
    // Cholesky G = R R^T (lower, in place).
    double lndet = 0.0;
    bool ok = true;
    for (int j = 0; j < L && ok; ++j) {
        double d = G[j * JK_LMAX + j];
        for (int k = 0; k < j; ++k) d -= G[j * JK_LMAX + k] * G[j * JK_LMAX + k];
        if (!(d > 0.0)) { ok = false; break; }
        d = sqrt(d);
        G[j * JK_LMAX + j] = d;
        lndet += 2.0 * log(d);
        for (int q = j + 1; q < L; ++q) {
            double v = G[q * JK_LMAX + j];
            for (int k = 0; k < j; ++k) v -= G[q * JK_LMAX + k] * G[j * JK_LMAX + k];
            G[q * JK_LMAX + j] = v / d;
        }
    }
    if (!ok) {
        for (int q = 0; q < L; ++q) a_out[i * L + q] = __builtin_nan("");
        for (int q = 0; q < L * L; ++q) cov_out[i * L * L + q] = __builtin_nan("");
        ll_out[i] = INFINITY;
        return;
    }
    // Rinv (lower) by forward substitution, then cov = Rinv^T Rinv.
    double Ri[JK_LMAX * JK_LMAX];
    for (int q = 0; q < JK_LMAX * JK_LMAX; ++q) Ri[q] = 0.0;
    for (int c = 0; c < L; ++c) {
        for (int a = c; a < L; ++a) {
            double v = a == c ? 1.0 : 0.0;
            for (int k = c; k < a; ++k) v -= G[a * JK_LMAX + k] * Ri[k * JK_LMAX + c];
            Ri[a * JK_LMAX + c] = v / G[a * JK_LMAX + a];
        }
    }
    for (int a = 0; a < L; ++a)
        for (int b = 0; b < L; ++b) {
            double v = 0.0;
            for (int k = (a > b ? a : b); k < L; ++k) v += Ri[k * JK_LMAX + a] * Ri[k * JK_LMAX + b];
            cov_out[i * L * L + a * L + b] = v;
        }
    // z = R^-1 u for chi2; a = cov rhs.
    double zz = 0.0;
    for (int a = 0; a < L; ++a) {
        double v = 0.0;
        for (int k = 0; k <= a; ++k) v += Ri[a * JK_LMAX + k] * u[k];
        zz += v * v;
    }
    for (int a = 0; a < L; ++a) {
        double v = 0.0;
        for (int b = 0; b < L; ++b) v += cov_out[i * L * L + a * L + b] * rhs[b];
        a_out[i * L + a] = v;
    }
    ll_out[i] = -0.5 * (chi0 - zz + (double)N * JK_LN_TWO_PI - sum_ln_w + lnlam + lndet);
}

// One accepted row of star st, one lane: a, cov and lnL into slot i of the
// outputs.  m_col: this lane's LDS column for a full-covariance star.
__device__ __forceinline__ void jk_posterior_row(const StarDev &st, int64_t i, int64_t r,
                                                 const double *__restrict__ P,
                                                 const double *__restrict__ E_,
                                                 const double *__restrict__ OM,
                                                 const double *__restrict__ MM,
                                                 const double *__restrict__ S,
                                                 double *__restrict__ a_out,
                                                 double *__restrict__ cov_out,
                                                 double *__restrict__ ll_out, double *m_col)
{
    const int L = st.L, N = st.n;
    const double Pv = P[r], ev = E_[r], omv = OM[r], M0v = MM[r];
    const double sv = st.jitter_mode == JK_JITTER_PHYSICAL ? S[r] : 0.0;
    double G[JK_LMAX * JK_LMAX], rhs[JK_LMAX], u[JK_LMAX];
    for (int q = 0; q < JK_LMAX * JK_LMAX; ++q) G[q] = 0.0;
    for (int q = 0; q < JK_LMAX; ++q) { rhs[q] = 0.0; u[q] = 0.0; }

    const bool valid = isfinite(Pv) && Pv != 0.0 && ev >= 0.0 && ev < 1.0 && isfinite(omv) &&
                       isfinite(M0v) && isfinite(sv);
    if (!valid) {
        for (int q = 0; q < L; ++q) a_out[i * L + q] = __builtin_nan("");
        for (int q = 0; q < L * L; ++q) cov_out[i * L * L + q] = __builtin_nan("");
        ll_out[i] = __builtin_nan("");
        return;
    }
    const double lam0 = st.k_prior_mode == JK_KPRIOR_FIXED_COMPANION_MASS
                            ? jk_k_variance(&st, Pv, ev, false)
                            : st.lam0;
    double sw, cw;
    sincos(omv, &sw, &cw);
    const double ecw = ev * cw, sq = sqrt((1.0 - ev) * (1.0 + ev));
    const double nP = JK_TWO_PI / Pv;
    double chi0 = 0.0, sum_ln_w = 0.0;
    double x[JK_LMAX];
    // full covariance: the design column in epoch order, kept in this lane's
    // LDS column (m_col), then rotated into the eigenbasis the star's data are in
    if (st.Qt) {
        for (int k = 0; k < N; ++k) {
            double sE, cE;
            jk_kepler_f64(jk_reduce_2pi(fma(nP, st.dt[k], -M0v)), ev, sE, cE);
            m_col[k * JK_COV_LANES] = jk_design_f64(sE, cE, ev, sq, cw, sw, ecw);
        }
    }
    for (int k = 0; k < N; ++k) {
        if (st.Qt) {
            double m = 0.0;
            for (int q = 0; q < N; ++q) m = fma(st.Qt[(size_t)k * N + q], m_col[q * JK_COV_LANES], m);
            x[0] = m;
        } else {
            double sE, cE;
            jk_kepler_f64(jk_reduce_2pi(fma(nP, st.dt[k], -M0v)), ev, sE, cE);
            x[0] = jk_design_f64(sE, cE, ev, sq, cw, sw, ecw);
        }
        for (int j = 1; j < L; ++j) x[j] = st.T[(j - 1) * N + k];
        const double iv = st.ivar[k];
        const double w = st.jitter_mode == JK_JITTER_PHYSICAL ? iv / fma(sv * sv, iv, 1.0) : iv;
        double rr = st.rp[k] - st.mu0 * x[0];
        chi0 = fma(w * rr, rr, chi0);
        sum_ln_w += log(w);
        for (int a = 0; a < L; ++a) {
            const double wx = w * x[a];
            rhs[a] = fma(wx, st.y[k], rhs[a]);
            u[a] = fma(wx, rr, u[a]);
            for (int b = 0; b <= a; ++b) G[a * JK_LMAX + b] = fma(wx, x[b], G[a * JK_LMAX + b]);
        }
    }
    jk_posterior_finish(st, i, lam0, G, rhs, u, chi0, sum_ln_w, a_out, cov_out, ll_out);
}

// One accepted row of a diagonal-covariance star per wave: lane l takes the
// epochs l, l + 64, ..., the 64 partial sums of each quantity are combined by
// the same butterfly in every launch (so the one-star and the many-star
// launches give the same bits), and lane 0 finishes the row.  A row's epochs
// then cost one wave about N / 64 epoch steps, instead of one lane N.
__device__ __forceinline__ double jk_wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void jk_posterior_row_wave(const StarDev &st, int64_t i, int64_t r,
                                                      const double *__restrict__ P,
                                                      const double *__restrict__ E_,
                                                      const double *__restrict__ OM,
                                                      const double *__restrict__ MM,
                                                      const double *__restrict__ S,
                                                      double *__restrict__ a_out,
                                                      double *__restrict__ cov_out,
                                                      double *__restrict__ ll_out)
{
    const int lane = threadIdx.x & 63;
    const int L = st.L, N = st.n;
    const double Pv = P[r], ev = E_[r], omv = OM[r], M0v = MM[r];
    const double sv = st.jitter_mode == JK_JITTER_PHYSICAL ? S[r] : 0.0;
    const bool valid = isfinite(Pv) && Pv != 0.0 && ev >= 0.0 && ev < 1.0 && isfinite(omv) &&
                       isfinite(M0v) && isfinite(sv);
    if (!valid) {
        if (lane == 0) {
            for (int q = 0; q < L; ++q) a_out[i * L + q] = __builtin_nan("");
            for (int q = 0; q < L * L; ++q) cov_out[i * L * L + q] = __builtin_nan("");
            ll_out[i] = __builtin_nan("");
        }
        return;
    }
    const double lam0 = st.k_prior_mode == JK_KPRIOR_FIXED_COMPANION_MASS
                            ? jk_k_variance(&st, Pv, ev, false)
                            : st.lam0;
    double sw, cw;
    sincos(omv, &sw, &cw);
    const double ecw = ev * cw, sq = sqrt((1.0 - ev) * (1.0 + ev));
    const double nP = JK_TWO_PI / Pv;
    double G[JK_LMAX * JK_LMAX], rhs[JK_LMAX], u[JK_LMAX], x[JK_LMAX];
    for (int q = 0; q < JK_LMAX * JK_LMAX; ++q) G[q] = 0.0;
    for (int q = 0; q < JK_LMAX; ++q) { rhs[q] = 0.0; u[q] = 0.0; }
    double chi0 = 0.0, sum_ln_w = 0.0;
    for (int k = lane; k < N; k += 64) {
        double sE, cE;
        jk_kepler_f64(jk_reduce_2pi(fma(nP, st.dt[k], -M0v)), ev, sE, cE);
        x[0] = jk_design_f64(sE, cE, ev, sq, cw, sw, ecw);
        for (int j = 1; j < L; ++j) x[j] = st.T[(j - 1) * N + k];
        const double iv = st.ivar[k];
        const double w = st.jitter_mode == JK_JITTER_PHYSICAL ? iv / fma(sv * sv, iv, 1.0) : iv;
        const double rr = st.rp[k] - st.mu0 * x[0];
        chi0 = fma(w * rr, rr, chi0);
        sum_ln_w += log(w);
        for (int a = 0; a < L; ++a) {
            const double wx = w * x[a];
            rhs[a] = fma(wx, st.y[k], rhs[a]);
            u[a] = fma(wx, rr, u[a]);
            for (int b = 0; b <= a; ++b) G[a * JK_LMAX + b] = fma(wx, x[b], G[a * JK_LMAX + b]);
        }
    }
    chi0 = jk_wave_sum(chi0);
    sum_ln_w = jk_wave_sum(sum_ln_w);
    for (int a = 0; a < L; ++a) {
        rhs[a] = jk_wave_sum(rhs[a]);
        u[a] = jk_wave_sum(u[a]);
        for (int b = 0; b <= a; ++b) G[a * JK_LMAX + b] = jk_wave_sum(G[a * JK_LMAX + b]);
    }
    if (lane == 0)
        jk_posterior_finish(st, i, lam0, G, rhs, u, chi0, sum_ln_w, a_out, cov_out, ll_out);
}

// Diagonal-covariance rows of one star, one wave (a 64-thread block) per row.
__global__ __launch_bounds__(64) void k_linear_posterior_wave(const StarDev st, int64_t n,
                                                              const double *__restrict__ P,
                                                              const double *__restrict__ E_,
                                                              const double *__restrict__ OM,
                                                              const double *__restrict__ MM,
                                                              const double *__restrict__ S,
                                                              const int64_t *__restrict__ rows,
                                                              double *__restrict__ a_out,
                                                              double *__restrict__ cov_out,
                                                              double *__restrict__ ll_out,
                                                              int64_t i0)
{
    const int64_t i = i0 + blockIdx.x;
    if (i >= n) return;
    jk_posterior_row_wave(st, i, rows ? rows[i] : i, P, E_, OM, MM, S, a_out, cov_out, ll_out);
}

// Rows per launch of the one-wave-per-row posterior kernels (the grid's
// blocks x 64 threads stay far inside 2^32).
#define JK_POST_CHUNK ((int64_t)1 << 24)

__global__ __launch_bounds__(64) void k_linear_posterior(const StarDev st, int64_t n,
                                                         const double *__restrict__ P,
                                                         const double *__restrict__ E_,
                                                         const double *__restrict__ OM,
                                                         const double *__restrict__ MM,
                                                         const double *__restrict__ S,
                                                         const int64_t *__restrict__ rows,
                                                         double *__restrict__ a_out,
                                                         double *__restrict__ cov_out,
                                                         double *__restrict__ ll_out)
{
    extern __shared__ double s_m[];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    jk_posterior_row(st, i, rows ? rows[i] : i, P, E_, OM, MM, S, a_out, cov_out, ll_out,
                     s_m + threadIdx.x);
}

// Accepted rows of many stars of a table in one launch: row i belongs to
// table entry star_idx[i] (diagonal-covariance stars only).
__global__ __launch_bounds__(64) void k_linear_posterior_batch(const StarDev *__restrict__ stars,
                                                               const int32_t *__restrict__ star_idx,
                                                               int64_t n,
                                                               const double *__restrict__ P,
                                                               const double *__restrict__ E_,
                                                               const double *__restrict__ OM,
                                                               const double *__restrict__ MM,
                                                               const double *__restrict__ S,
                                                               const int64_t *__restrict__ rows,
                                                               double *__restrict__ a_out,
                                                               double *__restrict__ cov_out,
                                                               double *__restrict__ ll_out,
                                                               int64_t i0)
{
    const int64_t i = i0 + blockIdx.x;
    if (i >= n) return;
    jk_posterior_row_wave(stars[star_idx[i]], i, rows[i], P, E_, OM, MM, S, a_out, cov_out,
                          ll_out);
}

hipError_t jk_launch_linear_posterior_batch(const jk_batch *b, int64_t n, const int32_t *star_idx,
                                            const double *P, const double *e, const double *om,
                                            const double *M0, const double *s,
                                            const int64_t *rows, double *a_out, double *cov_out,
                                            double *ll_out, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    for (int64_t i0 = 0; i0 < n; i0 += JK_POST_CHUNK) {
        const int64_t nb = n - i0 < JK_POST_CHUNK ? n - i0 : JK_POST_CHUNK;
        hipLaunchKernelGGL(k_linear_posterior_batch, dim3((unsigned)nb), dim3(64), 0, stream,
                           b->d_stars, star_idx, n, P, e, om, M0, s, rows, a_out, cov_out,
                           ll_out, i0);
    }
    return hipGetLastError();
}

hipError_t jk_launch_linear_posterior(const jk_star *st, int64_t n, const double *P,
                                      const double *e, const double *om, const double *M0,
                                      const double *s, const int64_t *rows, double *a_out,
                                      double *cov_out, double *ll_out, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    if (!st->dev.Qt) {
        for (int64_t i0 = 0; i0 < n; i0 += JK_POST_CHUNK) {
            const int64_t nb = n - i0 < JK_POST_CHUNK ? n - i0 : JK_POST_CHUNK;
            hipLaunchKernelGGL(k_linear_posterior_wave, dim3((unsigned)nb), dim3(64), 0, stream,
                               st->dev, n, P, e, om, M0, s, rows, a_out, cov_out, ll_out, i0);
        }
        return hipGetLastError();
    }
    const int64_t blocks = (n + 63) / 64;
    const size_t lds = sizeof(double) * (size_t)st->dev.n * JK_COV_LANES;
    hipLaunchKernelGGL(k_linear_posterior, dim3((unsigned)blocks), dim3(64), lds, stream, st->dev,
                       n, P, e, om, M0, s, rows, a_out, cov_out, ll_out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// One prior row in full detail (jk_debug_single): the quantities
// CJokerHelper.test_likelihood_worker leaves in its public members
// (fast_likelihood.pyx:109-114, 255-425, 547-576), for tests.  One lane, FP64,
// diagonal data covariance C = diag(1/w) (w = ivar, or ivar / (1 + s^2 ivar)
// in the physical jitter mode).  Lambda[0] is not clipped to max_K^2 here, as
// in test_likelihood_worker (:567-569).
//   M_T   (L x N)  design matrix, row 0 the unit-K Keplerian column
//   Ainv  (L x L)  Lambda^-1 + M_T C^-1 M_T^T                  (:262-276)
//   A     (L x L)  inverse of Ainv (here from its Cholesky factor)
//   a     (L)      Ainv^-1 (M_T C^-1 y + Lambda^-1 mu)         (:394-420)
//   b     (N)      M mu                                        (:305-309)
//   B     (N x N)  C + M Lambda M^T (optional)                 (:315-327)
//   Binv  (N x N)  C^-1 - C^-1 M A M^T C^-1 (optional)         (:329-337)
//   sc    lnL, chi2 = (b - y)^T Binv (b - y), ln det(2 pi B), Lambda[0]
// ln det(2 pi B) comes from the determinant lemma,
// N ln 2 pi - sum ln w + sum ln Lambda + ln det Ainv, and chi2 from the same
// algebra, sum w r^2 - v^T A v with v = M_T C^-1 r; the reference takes the
// LU of B (:342-353) and sums over Binv (:387-392), the same quantities.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_debug_row(const StarDev st, const double *__restrict__ row,
                                                 double *__restrict__ MT,
                                                 double *__restrict__ Ainv,
                                                 double *__restrict__ A,
                                                 double *__restrict__ a,
                                                 double *__restrict__ b,
                                                 double *__restrict__ B,
                                                 double *__restrict__ Binv,
                                                 double *__restrict__ sc)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int L = st.L, N = st.n;
    const bool phys = st.jitter_mode == JK_JITTER_PHYSICAL;
    const double Pv = row[0], ev = row[1], omv = row[2], M0v = row[3];
    const double sv = phys ? row[4] : 0.0;
    const double nan = __builtin_nan("");
    const bool valid = isfinite(Pv) && Pv != 0.0 && ev >= 0.0 && ev < 1.0 && isfinite(omv) &&
                       isfinite(M0v) && isfinite(sv);
    const double lam0 = st.k_prior_mode == JK_KPRIOR_FIXED_COMPANION_MASS
                            ? (valid ? jk_k_variance(&st, Pv, ev, false) : nan)
                            : st.lam0;
    double lam[JK_LMAX];
    lam[0] = lam0;
    for (int i = 1; i < L; ++i) lam[i] = 1.0 / st.lam_inv[i];
    double sw = 0.0, cw = 1.0;
    if (valid) sincos(omv, &sw, &cw);
    const double ecw = ev * cw, sq = valid ? sqrt((1.0 - ev) * (1.0 + ev)) : nan;
    const double nP = JK_TWO_PI / Pv;
    double G[JK_LMAX * JK_LMAX], rhs[JK_LMAX], v[JK_LMAX];
    for (int q = 0; q < JK_LMAX * JK_LMAX; ++q) G[q] = 0.0;
    for (int q = 0; q < JK_LMAX; ++q) rhs[q] = v[q] = 0.0;
    double wrr = 0.0, sum_ln_w = 0.0;
    for (int k = 0; k < N; ++k) {
        double x0 = nan;
        if (valid) {
            double sE, cE;
            jk_kepler_f64(jk_reduce_2pi(fma(nP, st.dt[k], -M0v)), ev, sE, cE);
            x0 = jk_design_f64(sE, cE, ev, sq, cw, sw, ecw);
        }
        MT[k] = x0;
        for (int j = 1; j < L; ++j) MT[(size_t)j * N + k] = st.T[(size_t)(j - 1) * N + k];
        const double iv = st.ivar[k];
        const double w = phys ? iv / fma(sv * sv, iv, 1.0) : iv;
        double bk = 0.0;
        for (int i = 0; i < L; ++i) bk = fma(MT[(size_t)i * N + k], st.mu[i], bk);
        b[k] = bk;
        const double r = bk - st.y[k];
        wrr = fma(w * r, r, wrr);
        sum_ln_w += log(w);
        for (int i = 0; i < L; ++i) {
            const double wx = w * MT[(size_t)i * N + k];
            rhs[i] = fma(wx, st.y[k], rhs[i]);
            v[i] = fma(wx, r, v[i]);
            for (int j = 0; j < L; ++j) G[i * JK_LMAX + j] = fma(wx, MT[(size_t)j * N + k], G[i * JK_LMAX + j]);
        }
    }
    double lnlam = 0.0;
    for (int i = 0; i < L; ++i) {
        G[i * JK_LMAX + i] += 1.0 / lam[i];
        rhs[i] += st.mu[i] / lam[i];
        lnlam += log(lam[i]);
    }
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) Ainv[i * L + j] = G[i * JK_LMAX + j];
    // Cholesky Ainv = R R^T (R lower, kept in Rl)
    double Rl[JK_LMAX * JK_LMAX];
    for (int q = 0; q < JK_LMAX * JK_LMAX; ++q) Rl[q] = 0.0;
    double lndet = 0.0;
    bool ok = valid;
    for (int j = 0; j < L && ok; ++j) {
        double d = G[j * JK_LMAX + j];
        for (int k = 0; k < j; ++k) d -= Rl[j * JK_LMAX + k] * Rl[j * JK_LMAX + k];
        if (!(d > 0.0)) { ok = false; break; }
        d = sqrt(d);
        Rl[j * JK_LMAX + j] = d;
        lndet += 2.0 * log(d);
        for (int q = j + 1; q < L; ++q) {
            double t = G[q * JK_LMAX + j];
            for (int k = 0; k < j; ++k) t -= Rl[q * JK_LMAX + k] * Rl[j * JK_LMAX + k];
            Rl[q * JK_LMAX + j] = t / d;
        }
    }
    double Am[JK_LMAX * JK_LMAX];
    if (ok) {
        // A = R^-T R^-1 from the lower inverse Ri
        double Ri[JK_LMAX * JK_LMAX];
        for (int q = 0; q < JK_LMAX * JK_LMAX; ++q) Ri[q] = 0.0;
        for (int c = 0; c < L; ++c)
            for (int r = c; r < L; ++r) {
                double t = r == c ? 1.0 : 0.0;
                for (int k = c; k < r; ++k) t -= Rl[r * JK_LMAX + k] * Ri[k * JK_LMAX + c];
                Ri[r * JK_LMAX + c] = t / Rl[r * JK_LMAX + r];
            }
        for (int i = 0; i < L; ++i)
            for (int j = 0; j < L; ++j) {
                double t = 0.0;
                for (int k = (i > j ? i : j); k < L; ++k) t += Ri[k * JK_LMAX + i] * Ri[k * JK_LMAX + j];
                Am[i * JK_LMAX + j] = t;
            }
    } else {
        for (int q = 0; q < JK_LMAX * JK_LMAX; ++q) Am[q] = nan;
    }
    double chi2 = wrr;
    for (int i = 0; i < L; ++i) {
        double ai = 0.0;
        for (int j = 0; j < L; ++j) {
            A[i * L + j] = Am[i * JK_LMAX + j];
            ai = fma(Am[i * JK_LMAX + j], rhs[j], ai);
            chi2 -= v[i] * Am[i * JK_LMAX + j] * v[j];
        }
        a[i] = ai;
    }
    if (B) {
        for (int n = 0; n < N; ++n)
            for (int m = 0; m < N; ++m) {
                double t = 0.0;
                for (int i = 0; i < L; ++i)
                    t = fma(MT[(size_t)i * N + n] * lam[i], MT[(size_t)i * N + m], t);
                if (n == m) {
                    const double iv = st.ivar[n];
                    t += phys ? fma(sv, sv, 1.0 / iv) : 1.0 / iv;
                }
                B[(size_t)n * N + m] = t;
            }
    }
    if (Binv) {
        for (int n = 0; n < N; ++n) {
            const double ivn = st.ivar[n];
            const double wn = phys ? ivn / fma(sv * sv, ivn, 1.0) : ivn;
            for (int m = 0; m < N; ++m) {
                const double ivm = st.ivar[m];
                const double wm = phys ? ivm / fma(sv * sv, ivm, 1.0) : ivm;
                double t = 0.0;
                for (int i = 0; i < L; ++i)
                    for (int j = 0; j < L; ++j)
                        t = fma(MT[(size_t)i * N + n] * Am[i * JK_LMAX + j], MT[(size_t)j * N + m], t);
                Binv[(size_t)n * N + m] = (n == m ? wn : 0.0) - wn * t * wm;
            }
        }
    }
    const double logdet = (double)N * JK_LN_TWO_PI - sum_ln_w + lnlam + lndet;
    double ll = -0.5 * (chi2 + logdet);
    if (!valid) ll = nan;
    else if (!ok) ll = INFINITY;
    sc[0] = ll;
    sc[1] = ok ? chi2 : nan;
    sc[2] = ok ? logdet : nan;
    sc[3] = lam0;
}

hipError_t jk_launch_debug_row(const jk_star *st, const double *row, double *MT, double *Ainv,
                               double *A, double *a, double *b, double *B, double *Binv,
                               double *sc, hipStream_t stream)
{
    hipLaunchKernelGGL(k_debug_row, dim3(1), dim3(64), 0, stream, st->dev, row, MT, Ainv, A, a,
                       b, B, Binv, sc);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Per-sample failure counts of an lnL array (the reference writes +INF when
// its L x L or N x N factorisation fails, fast_likelihood.pyx:283-291,
// 347-348, 381-382, 422-423; invalid rows are NaN here): counts[0] += rows
// at +inf, counts[1] += rows at -inf, counts[2] += NaN rows.  counts is
// cleared by the launcher's first kernel.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_zero_counts3(unsigned long long *counts)
{
    if (threadIdx.x < 3) counts[threadIdx.x] = 0ull;
}

__global__ __launch_bounds__(256) void k_count_failures(const double *__restrict__ ll, int64_t n,
                                                        unsigned long long *__restrict__ counts)
{
    unsigned long long c[3] = {0ull, 0ull, 0ull};
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double x = ll[i];
        c[0] += x == INFINITY;
        c[1] += x == -INFINITY;
        c[2] += x != x;
    }
    for (int q = 0; q < 3; ++q) {
        unsigned long long t = c[q];
        for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
        if ((threadIdx.x & 63) == 0 && t) atomicAdd(counts + q, t);
    }
}

hipError_t jk_launch_count_failures(const double *ll, int64_t n, unsigned long long *counts,
                                    hipStream_t stream)
{
    hipLaunchKernelGGL(k_zero_counts3, dim3(1), dim3(256), 0, stream, counts);
    if (n > 0) {
        int64_t blocks = (n + 255) / 256;
        if (blocks > 2048) blocks = 2048;
        hipLaunchKernelGGL(k_count_failures, dim3((unsigned)blocks), dim3(256), 0, stream, ll, n,
                           counts);
    }
    return hipGetLastError();
}
