/*
 * joker_oracle.c -- CPU restatement of The Joker's marginal-likelihood and
 * rejection path.  TEST INFRASTRUCTURE ONLY: this file is the checker that the
 * HIP kernels are compared against, and the timed CPU baseline.  Nothing in
 * the shipped package (thejoker_amd/) links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Reference paths
 * are relative to the adrn/thejoker source tree (v1.3.0).
 *
 * Parity status of each function (how it is pinned; see tests/):
 *   jo_kepler_twobody   -- twobody>=0.9.1 (pyproject.toml:8) is not vendored in
 *                          the reference; restated as Newton from E0 = M with
 *                          |dE| < tol stop (fast_likelihood.pyx:35-36 passes
 *                          tol=1e-10, maxiter=128).  Pinned against bisection
 *                          brute force and closed-form special cases
 *                          (tests/test_oracle_kepler.py).  The iterate returned
 *                          when the 128-step cap is hit is parity unpinned.
 *   jo_lnl_reference    -- fast_likelihood.pyx:255-469 (LU + Woodbury + N x N
 *                          LU log-det).  Pinned against numpy dense slogdet /
 *                          solve vectors and 2-D quadrature
 *                          (tests/golden/, tests/test_oracle_lnl.py).
 *   jo_lnl_chol         -- the O(N L^2) Cholesky / determinant-lemma form of the
 *                          same quantity; pinned against jo_lnl_reference and
 *                          the same golden vectors.
 *   jo_pcg64_*          -- numpy PCG64 (XSL-RR 128/64) as used by
 *                          Generator.uniform at likelihood_helpers.py:107;
 *                          pinned bit-exactly against numpy's own stream.
 *   jo_reject           -- likelihood_helpers.py:107-109; pinned against
 *                          numpy np.where on the same inputs.
 *   jo_linear_posterior -- fast_likelihood.pyx:394-423 (a = Ainv^-1 rhs,
 *                          A = Ainv^-1); pinned against numpy.linalg.solve.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define JO_LMAX 8
#define JO_TWO_PI 6.283185307179586476925286766559

/* ------------------------------------------------------------------------- */
/* Kepler's equation                                                          */
/* ------------------------------------------------------------------------- */

/* twobody c_rv_from_elements' anomaly solve (called from
 * fast_likelihood.pyx:453-455 with tol=anomaly_tol=1e-10,
 * maxiter=anomaly_maxiter=128, fast_likelihood.pyx:35-36): Newton's method on
 * f(E) = E - e sin E - M started at E0 = M, stopping when the step is below
 * tol.  *n_iter receives the number of steps taken (== maxiter when the cap
 * was hit without meeting tol). */
double jo_kepler_twobody(double M, double e, double tol, int maxiter, int *n_iter)
{
    double E0 = M, E = M;
    int i;
    for (i = 0; i < maxiter; i++) {
        E = E0 - (E0 - e * sin(E0) - M) / (1.0 - e * cos(E0));
        if (fabs(E - E0) < tol) {
            if (n_iter) *n_iter = i + 1;
            return E;
        }
        E0 = E;
    }
    if (n_iter) *n_iter = maxiter;
    return E;
}

/* Bracketed solve used as brute force in tests: bisection on [M-1, M+1]
 * (|E - M| = e|sin E| <= e < 1) to machine precision. */
double jo_kepler_bisect(double M, double e)
{
    double lo = M - 1.0, hi = M + 1.0;
    int i;
    for (i = 0; i < 200; i++) {
        double mid = 0.5 * (lo + hi);
        if (mid == lo || mid == hi) break;
        if (mid - e * sin(mid) - M > 0.0) hi = mid; else lo = mid;
    }
    return 0.5 * (lo + hi);
}

/* Unit-amplitude RV design column, restating twobody c_rv_from_elements as
 * called at fast_likelihood.pyx:453-455 (K = 1, phi0 = M0, t0 = t_ref):
 *   M_n = 2 pi (t_n - t0) / P - M0,  E_n - e sin E_n = M_n,
 *   f_n = 2 atan2(sqrt(1+e) sin(E/2), sqrt(1-e) cos(E/2)),
 *   rv_n = K [cos(omega + f_n) + e cos omega].
 * The sign/phase convention is the one the reference itself states at
 * _keplerian_orbit.py:632-645 ("cos(w + f) + e * cos(w)") and
 * samples.py:228-229 (pericentre time t_ref + P M0 / 2 pi).
 * Returns the number of epochs whose Kepler solve hit the iteration cap. */
int jo_rv_from_elements(const double *t, double *rv, int n, double P, double K,
                        double e, double omega, double phi0, double t0,
                        double tol, int maxiter)
{
    int i, n_cap = 0;
    for (i = 0; i < n; i++) {
        int it = 0;
        double M = 2 * M_PI * (t[i] - t0) / P - phi0;
        double E = jo_kepler_twobody(M, e, tol, maxiter, &it);
        double f = 2.0 * atan2(sqrt(1.0 + e) * sin(0.5 * E), sqrt(1.0 - e) * cos(0.5 * E));
        if (it >= maxiter) n_cap++;
        rv[i] = K * (cos(omega + f) + e * cos(omega));
    }
    return n_cap;
}

/* ------------------------------------------------------------------------- */
/* Star description (what CJokerHelper.__init__ holds,                        */
/* fast_likelihood.pyx:125-253)                                               */
/* ------------------------------------------------------------------------- */
typedef struct {
    int n;              /* epochs */
    int L;              /* linear parameters: K, v0, [offsets], v1, ... */
    const double *t;    /* n   observation times (BMJD days) */
    const double *rv;   /* n   RVs */
    const double *ivar; /* n   1/sigma^2 */
    const double *trend;/* n x (L-1) row-major, trend_M of likelihood_helpers.py:28-37 */
    const double *mu;   /* L   prior means of the linear parameters */
    const double *lam;  /* L   prior variances (diagonal Lambda) */
    double t0;          /* t_ref (BMJD days) */
    int k_prior_mode;   /* 0: Normal K with Lambda[0] fixed; 1: FixedCompanionMass */
    double sigma_K0, P0, max_K;
    int jitter_mode;    /* 0: reference (s ignored, fast_likelihood.pyx:458);
                           1: physical (ivar/(1+s^2 ivar), py_likelihood.py:29) */
    double tol;         /* anomaly tolerance (1e-10) */
    int maxiter;        /* anomaly iteration cap (128) */
} jo_star;

/* FixedCompanionMass variance of K (fast_likelihood.pyx:461-464; the clip to
 * max_K^2 is applied on the likelihood path only -- the posterior path at
 * fast_likelihood.pyx:519-521 omits it). */
static double k_variance(const jo_star *s, double P, double e, int clip)
{
    double v = s->sigma_K0 * s->sigma_K0 / (1.0 - e * e) * pow(P / s->P0, -2.0 / 3.0);
    if (clip && s->max_K * s->max_K < v) v = s->max_K * s->max_K;
    return v;
}

/* LU with partial pivoting (the arithmetic of LAPACK dgetrf on a small dense
 * matrix, fast_likelihood.pyx:281,345).  a is n x n row-major, overwritten.
 * Returns 0 or k+1 when U[k,k] == 0 (dgetrf's info > 0). */
static int lu_factor(double *a, int n, int *piv)
{
    int i, j, k;
    for (k = 0; k < n; k++) {
        int p = k;
        double best = fabs(a[k * n + k]);
        for (i = k + 1; i < n; i++) {
            double v = fabs(a[i * n + k]);
            if (v > best) { best = v; p = i; }
        }
        piv[k] = p;
        if (p != k)
            for (j = 0; j < n; j++) {
                double tmp = a[k * n + j]; a[k * n + j] = a[p * n + j]; a[p * n + j] = tmp;
            }
        if (a[k * n + k] == 0.0) return k + 1;
        for (i = k + 1; i < n; i++) {
            double l = a[i * n + k] / a[k * n + k];
            a[i * n + k] = l;
            for (j = k + 1; j < n; j++) a[i * n + j] -= l * a[k * n + j];
        }
    }
    return 0;
}

/* Solve with an LU factorisation (in place on b). */
static void lu_solve(const double *lu, const int *piv, int n, double *b)
{
    int i, j;
    for (i = 0; i < n; i++) {
        double t = b[i]; b[i] = b[piv[i]]; b[piv[i]] = t;
    }
    for (i = 0; i < n; i++)
        for (j = 0; j < i; j++) b[i] -= lu[i * n + j] * b[j];
    for (i = n - 1; i >= 0; i--) {
        for (j = i + 1; j < n; j++) b[i] -= lu[i * n + j] * b[j];
        b[i] /= lu[i * n + i];
    }
}

/* Inverse from an LU factorisation (dgetri, fast_likelihood.pyx:288). */
static void lu_inverse(const double *lu, const int *piv, int n, double *inv)
{
    double col[JO_LMAX];
    int i, j;
    for (j = 0; j < n; j++) {
        for (i = 0; i < n; i++) col[i] = (i == j) ? 1.0 : 0.0;
        lu_solve(lu, piv, n, col);
        for (i = 0; i < n; i++) inv[i * n + j] = col[i];
    }
}

/* Per-epoch inverse variance for sample jitter s. */
static void effective_ivar(const jo_star *st, double s, double *w)
{
    int i;
    for (i = 0; i < st->n; i++)
        w[i] = st->jitter_mode ? st->ivar[i] / (1.0 + s * s * st->ivar[i]) : st->ivar[i];
}

/* Design matrix transpose M_T (L x n): row 0 the Keplerian column, rows 1..
 * the trend columns (fast_likelihood.pyx:181-184, 453-455). */
static int fill_design(const jo_star *st, const double *row, double *MT)
{
    int i, n;
    int n_cap = jo_rv_from_elements(st->t, MT, st->n, row[0], 1.0, row[1], row[2],
                                    row[3], st->t0, st->tol, st->maxiter);
    for (n = 0; n < st->n; n++)
        for (i = 1; i < st->L; i++) MT[i * st->n + n] = st->trend[n * (st->L - 1) + i - 1];
    return n_cap;
}

/* The reference algorithm, fast_likelihood.pyx:359-425 with make_AAinv
 * (:255-297) and make_bBBinv (:299-357):
 *   Ainv = Lambda^-1 + M^T C^-1 M;  A = inv(Ainv) (LU, dgetrf/dgetri)
 *   b = M mu;  B = C + M Lambda M^T
 *   Binv = C^-1 - C^-1 M A M^T C^-1   (Woodbury, :332-339)
 *   logdet = sum_i log(2 pi |U_ii|) from LU(B)   (:345-353)
 *   chi2 = (b - y)^T Binv (b - y)   (:387-392)
 *   lnL = -0.5 (chi2 + logdet)
 * With want_a: a = solve(Ainv, M^T C^-1 y + Lambda^-1 mu) (:394-420).
 * Returns +inf when the L x L LU fails (:283-291, :381-382) and -inf when the
 * N x N LU fails (:347-348).  `lam` is the (possibly per-sample) prior variance
 * vector; `w` the inverse variances; scratch must hold 3 n^2 + 2 n doubles. */
static double lnl_reference_core(const jo_star *st, const double *MT, const double *lam,
                                 const double *w, double *scratch, int want_a,
                                 double *a_out, double *Amat_out)
{
    const int N = st->n, L = st->L;
    double Ainv[JO_LMAX * JO_LMAX], Atmp[JO_LMAX * JO_LMAX], A[JO_LMAX * JO_LMAX];
    int piv[JO_LMAX];
    double *B = scratch, *Binv = scratch + N * N, *Btmp = scratch + 2 * N * N;
    double *b = scratch + 3 * N * N;
    int *ipiv = (int *)(scratch + 3 * N * N + N);
    int i, j, n, m;
    double log_det = 0.0, chi2 = 0.0;

    /* make_AAinv */
    for (i = 0; i < L; i++)
        for (j = 0; j < L; j++) Ainv[i * L + j] = 0.0;
    for (i = 0; i < L; i++) {
        Ainv[i * L + i] = 1.0 / lam[i];
        for (j = 0; j < L; j++) {
            for (n = 0; n < N; n++) Ainv[i * L + j] += MT[j * N + n] * w[n] * MT[i * N + n];
            Atmp[i * L + j] = Ainv[i * L + j];
        }
    }
    if (lu_factor(Atmp, L, piv) != 0) return INFINITY;
    lu_inverse(Atmp, piv, L, A);

    /* make_bBBinv */
    for (n = 0; n < N; n++) {
        b[n] = 0.0;
        for (i = 0; i < L; i++) b[n] += MT[i * N + n] * st->mu[i];
        for (m = 0; m < N; m++) B[n * N + m] = 0.0;
    }
    for (n = 0; n < N; n++) {
        B[n * N + n] = 1.0 / w[n];
        for (m = 0; m < N; m++) {
            Binv[n * N + m] = 0.0;
            for (i = 0; i < L; i++) B[n * N + m] += MT[i * N + n] * lam[i] * MT[i * N + m];
            Btmp[n * N + m] = B[n * N + m];
        }
    }
    for (n = 0; n < N; n++) {
        Binv[n * N + n] = w[n];
        for (i = 0; i < L; i++)
            for (m = 0; m < N; m++)
                for (j = 0; j < L; j++)
                    Binv[n * N + m] -= w[n] * MT[i * N + n] * A[i * L + j] * MT[j * N + m] * w[m];
    }
    if (lu_factor(Btmp, N, ipiv) != 0) {
        log_det = INFINITY;
    } else {
        for (i = 0; i < N; i++) log_det += log(2 * M_PI * fabs(Btmp[i * N + i]));
    }

    for (n = 0; n < N; n++)
        for (m = 0; m < N; m++)
            chi2 += (b[m] - st->rv[m]) * Binv[n * N + m] * (b[n] - st->rv[n]);

    if (want_a) {
        double rhs[JO_LMAX];
        for (i = 0; i < L; i++) rhs[i] = 0.0;
        for (n = 0; n < N; n++)
            for (i = 0; i < L; i++) rhs[i] += MT[i * N + n] * w[n] * st->rv[n];
        for (i = 0; i < L; i++) rhs[i] += st->mu[i] / lam[i];
        for (i = 0; i < L * L; i++) Atmp[i] = Ainv[i];
        if (lu_factor(Atmp, L, piv) != 0) return INFINITY;
        lu_solve(Atmp, piv, L, rhs);
        if (a_out) for (i = 0; i < L; i++) a_out[i] = rhs[i];
        if (Amat_out) for (i = 0; i < L * L; i++) Amat_out[i] = A[i];
    }
    return -0.5 * (chi2 + log_det);
}

/* The O(N L^2) form of the same marginal likelihood (determinant lemma +
 * Woodbury on the L x L system):
 *   r = y - M mu,  G = Lambda^-1 + M^T W M = R R^T,  u = M^T W r,  z = R^-1 u
 *   chi2 = r^T W r - z^T z
 *   log det(2 pi B) = N log 2pi - sum log w + sum log Lambda + 2 sum log R_ii
 * Returns +inf when the Cholesky factorisation breaks down (the analogue of
 * the reference's L x L LU failure). */
static double lnl_chol_core(const jo_star *st, const double *MT, const double *lam,
                            const double *w)
{
    const int N = st->n, L = st->L;
    double G[JO_LMAX * JO_LMAX], u[JO_LMAX], z[JO_LMAX];
    double chi0 = 0.0, lndet = 0.0;
    int i, j, k, n;
    for (i = 0; i < L; i++) {
        u[i] = 0.0;
        for (j = 0; j < L; j++) G[i * L + j] = (i == j) ? 1.0 / lam[i] : 0.0;
    }
    for (n = 0; n < N; n++) {
        double r = st->rv[n];
        for (i = 0; i < L; i++) r -= MT[i * N + n] * st->mu[i];
        chi0 += w[n] * r * r;
        lndet -= log(w[n]);
        for (i = 0; i < L; i++) {
            double wx = w[n] * MT[i * N + n];
            u[i] += wx * r;
            for (j = 0; j <= i; j++) G[i * L + j] += wx * MT[j * N + n];
        }
    }
    /* Cholesky, lower triangle */
    for (j = 0; j < L; j++) {
        double d = G[j * L + j];
        for (k = 0; k < j; k++) d -= G[j * L + k] * G[j * L + k];
        if (!(d > 0.0)) return INFINITY;
        d = sqrt(d);
        G[j * L + j] = d;
        lndet += 2.0 * log(d) + log(lam[j]);
        for (i = j + 1; i < L; i++) {
            double v = G[i * L + j];
            for (k = 0; k < j; k++) v -= G[i * L + k] * G[j * L + k];
            G[i * L + j] = v / d;
        }
    }
    for (i = 0; i < L; i++) {
        double v = u[i];
        for (k = 0; k < i; k++) v -= G[i * L + k] * z[k];
        z[i] = v / G[i * L + i];
        chi0 -= z[i] * z[i];
    }
    return -0.5 * (chi0 + lndet + N * log(2 * M_PI));
}

/* One prior row (P, e, omega, M0, s) -> lnL.  method 0: reference (LU);
 * method 1: Cholesky.  *n_cap receives the count of capped Kepler solves. */
double jo_lnl_row(const jo_star *st, const double *row, int method, int *n_cap)
{
    const int N = st->n, L = st->L;
    double lam[JO_LMAX], ll;
    double *MT = (double *)malloc(sizeof(double) * (L * N + N + 3 * N * N + 2 * N + 8));
    double *w = MT + L * N, *scratch = w + N;
    int i, cap;
    for (i = 0; i < L; i++) lam[i] = st->lam[i];
    if (st->k_prior_mode == 1) lam[0] = k_variance(st, row[0], row[1], 1);
    cap = fill_design(st, row, MT);
    effective_ivar(st, row[4], w);
    ll = method == 0 ? lnl_reference_core(st, MT, lam, w, scratch, 0, NULL, NULL)
                     : lnl_chol_core(st, MT, lam, w);
    if (n_cap) *n_cap = cap;
    free(MT);
    return ll;
}

/* Linear-parameter posterior of one row (fast_likelihood.pyx:471-531 up to
 * the multivariate_normal draw): a (L) and A = inv(Ainv) (L x L row-major).
 * Lambda[0] for FixedCompanionMass is NOT clipped on this path
 * (fast_likelihood.pyx:519-521).  Returns the row's lnL. */
double jo_linear_posterior(const jo_star *st, const double *row, double *a, double *A)
{
    const int N = st->n, L = st->L;
    double lam[JO_LMAX], ll;
    double *MT = (double *)malloc(sizeof(double) * (L * N + N + 3 * N * N + 2 * N + 8));
    double *w = MT + L * N, *scratch = w + N;
    int i;
    for (i = 0; i < L; i++) lam[i] = st->lam[i];
    if (st->k_prior_mode == 1) lam[0] = k_variance(st, row[0], row[1], 0);
    fill_design(st, row, MT);
    effective_ivar(st, row[4], w);
    ll = lnl_reference_core(st, MT, lam, w, scratch, 1, a, A);
    free(MT);
    return ll;
}

/* ------------------------------------------------------------------------- */
/* Batched evaluation, optionally over threads (one contiguous batch per      */
/* thread, as MultiPool with n_batches = pool.size does at                    */
/* multiproc_helpers.py:39-47 and utils.py:22-72).                            */
/* ------------------------------------------------------------------------- */
typedef struct {
    const jo_star *st;
    const double *chunk; /* AoS n x 5 */
    int64_t i0, i1;
    int method;
    double *ll;
    unsigned char *capped;
    int64_t n_cap;
} jo_batch_task;

static void *batch_worker(void *arg)
{
    jo_batch_task *t = (jo_batch_task *)arg;
    const int N = t->st->n, L = t->st->L;
    double lam[JO_LMAX];
    double *MT = (double *)malloc(sizeof(double) * (L * N + N + 3 * N * N + 2 * N + 8));
    double *w = MT + L * N, *scratch = w + N;
    int64_t r;
    int i;
    t->n_cap = 0;
    for (i = 0; i < L; i++) lam[i] = t->st->lam[i];
    for (r = t->i0; r < t->i1; r++) {
        const double *row = t->chunk + 5 * r;
        if (t->st->k_prior_mode == 1) lam[0] = k_variance(t->st, row[0], row[1], 1);
        int cap = fill_design(t->st, row, MT) > 0;
        t->n_cap += cap;
        if (t->capped) t->capped[r] = (unsigned char)cap;
        effective_ivar(t->st, row[4], w);
        t->ll[r] = t->method == 0 ? lnl_reference_core(t->st, MT, lam, w, scratch, 0, NULL, NULL)
                                  : lnl_chol_core(t->st, MT, lam, w);
    }
    free(MT);
    return NULL;
}

/* batch_marginal_ln_likelihood (fast_likelihood.pyx:428-469) over n rows of
 * the packed AoS chunk (P, e, omega, M0, s; fast_likelihood.pyx:41).  Returns
 * the number of rows in which at least one Kepler solve hit the cap; when
 * `capped` is not NULL, capped[r] flags those rows. */
int64_t jo_batch_lnl(const jo_star *st, const double *chunk, int64_t n, int method,
                     int n_threads, double *ll, unsigned char *capped)
{
    jo_batch_task tasks[256];
    pthread_t th[256];
    int64_t base, rem, i1 = 0, n_cap = 0;
    int k;
    if (n_threads < 1) n_threads = 1;
    if (n_threads > 256) n_threads = 256;
    if (n < n_threads) n_threads = n > 0 ? (int)n : 1;
    base = n / n_threads;
    rem = n % n_threads;
    for (k = 0; k < n_threads; k++) {
        tasks[k].st = st; tasks[k].chunk = chunk; tasks[k].method = method; tasks[k].ll = ll;
        tasks[k].capped = capped;
        tasks[k].i0 = i1;
        i1 += base + (k < rem ? 1 : 0);
        tasks[k].i1 = i1;
    }
    if (n_threads == 1) {
        batch_worker(&tasks[0]);
    } else {
        for (k = 0; k < n_threads; k++) pthread_create(&th[k], NULL, batch_worker, &tasks[k]);
        for (k = 0; k < n_threads; k++) pthread_join(th[k], NULL);
    }
    for (k = 0; k < n_threads; k++) n_cap += tasks[k].n_cap;
    return n_cap;
}

/* ------------------------------------------------------------------------- */
/* numpy PCG64 (the bit generator behind np.random.default_rng, whose          */
/* Generator.uniform draws the rejection uniforms at likelihood_helpers.py:107 */
/* and multiproc_helpers.py:256,373).                                          */
/*   state <- state * 0x2360ED051FC65DA44385DF649FCCF645 + inc  (mod 2^128)    */
/*   out    = rotr64(hi ^ lo, state >> 122)   of the new state                 */
/*   uniform = (out >> 11) * 2^-53                                             */
/* ------------------------------------------------------------------------- */
typedef unsigned __int128 u128;

static u128 pcg_mult(void)
{
    return ((u128)0x2360ED051FC65DA4ULL << 64) | (u128)0x4385DF649FCCF645ULL;
}

static uint64_t pcg_output(u128 s)
{
    uint64_t x = (uint64_t)(s >> 64) ^ (uint64_t)s;
    unsigned rot = (unsigned)(s >> 122);
    return (x >> rot) | (x << ((64 - rot) & 63));
}

/* Jump the LCG ahead by delta steps (Brown, "Random number generation with
 * arbitrary strides", 1994). */
static u128 pcg_advance(u128 state, u128 inc, uint64_t delta)
{
    u128 cur_mult = pcg_mult(), cur_plus = inc, acc_mult = 1, acc_plus = 0;
    while (delta > 0) {
        if (delta & 1) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1;
    }
    return acc_mult * state + acc_plus;
}

/* Raw 64-bit outputs offset .. offset+n-1 of the stream whose current state
 * is (state_hi:state_lo, inc_hi:inc_lo); output k is produced by the (k+1)-th
 * step. */
void jo_pcg64_raw(uint64_t state_hi, uint64_t state_lo, uint64_t inc_hi, uint64_t inc_lo,
                  uint64_t offset, int64_t n, uint64_t *out)
{
    u128 inc = ((u128)inc_hi << 64) | inc_lo;
    u128 s = pcg_advance(((u128)state_hi << 64) | state_lo, inc, offset);
    int64_t i;
    for (i = 0; i < n; i++) {
        s = s * pcg_mult() + inc;
        out[i] = pcg_output(s);
    }
}

void jo_pcg64_uniform(uint64_t state_hi, uint64_t state_lo, uint64_t inc_hi, uint64_t inc_lo,
                      uint64_t offset, int64_t n, double *out)
{
    u128 inc = ((u128)inc_hi << 64) | inc_lo;
    u128 s = pcg_advance(((u128)state_hi << 64) | state_lo, inc, offset);
    int64_t i;
    for (i = 0; i < n; i++) {
        s = s * pcg_mult() + inc;
        out[i] = (double)(pcg_output(s) >> 11) * (1.0 / 9007199254740992.0);
    }
}

/* Rejection step, likelihood_helpers.py:107-109:
 *   good = np.where(np.exp(ll - ll.max()) > u)[0][:max_keep]
 * np.max propagates NaN, so any NaN lnL makes every comparison false.
 * Writes the (ascending) kept indices, returns the full pass count (before
 * truncation) and the max through *llmax. */
int64_t jo_reject(const double *ll, int64_t n, const double *u, int64_t max_keep,
                  int64_t *idx_out, double *llmax)
{
    double mx = -INFINITY;
    int64_t i, n_pass = 0;
    for (i = 0; i < n; i++) {
        if (isnan(ll[i])) { mx = NAN; break; }
        if (ll[i] > mx) mx = ll[i];
    }
    if (llmax) *llmax = n > 0 ? mx : NAN;
    for (i = 0; i < n; i++) {
        if (exp(ll[i] - mx) > u[i]) {
            if (n_pass < max_keep) idx_out[n_pass] = i;
            n_pass++;
        }
    }
    return n_pass;
}

/* Size of the scratch a caller-built jo_star needs -- exported so Python can
 * check the struct layout it builds with ctypes. */
int jo_star_sizeof(void) { return (int)sizeof(jo_star); }
