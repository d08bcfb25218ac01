/*
 * oracle_selftest.c -- the oracle (joker_oracle.c) driven under AddressSanitizer
 * and UndefinedBehaviorSanitizer on the host: every exported function on
 * small inputs, with answers fixed by mathematics or by numpy's documented
 * PCG64 stream, so that an out-of-bounds access, a use of freed memory or
 * undefined arithmetic in the checker aborts the run.  TEST INFRASTRUCTURE
 * ONLY; built by `make -C oracle sanitize`, run by tests/test_oracle_sanitize.py.
 */
#include <stdio.h>

#include "joker_oracle.c"

static int failures = 0;

#define CHECK(cond, ...)                                   \
    do {                                                   \
        if (!(cond)) {                                     \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            fprintf(stderr, __VA_ARGS__);                  \
            fputc('\n', stderr);                           \
            failures++;                                    \
        }                                                  \
    } while (0)

/* ln N(y | M mu, C + M Lambda M^T) by a dense N x N Cholesky written out here:
 * the textbook Gaussian, independent of both oracle forms. */
static double dense_lnl(int N, int L, const double *MT, const double *lam, const double *mu,
                        const double *w, const double *y)
{
    double *B = (double *)calloc((size_t)N * N, sizeof(double));
    double *r = (double *)malloc(sizeof(double) * N);
    double logdet = 0.0, chi2 = 0.0;
    int i, j, k;
    for (i = 0; i < N; i++) {
        double b = 0.0;
        for (k = 0; k < L; k++) b += MT[k * N + i] * mu[k];
        r[i] = y[i] - b;
        for (j = 0; j < N; j++) {
            double v = i == j ? 1.0 / w[i] : 0.0;
            for (k = 0; k < L; k++) v += MT[k * N + i] * lam[k] * MT[k * N + j];
            B[i * N + j] = v;
        }
    }
    for (j = 0; j < N; j++) { /* B = R R^T, lower R in place */
        double d = B[j * N + j];
        for (k = 0; k < j; k++) d -= B[j * N + k] * B[j * N + k];
        d = sqrt(d);
        B[j * N + j] = d;
        logdet += 2.0 * log(d);
        for (i = j + 1; i < N; i++) {
            double v = B[i * N + j];
            for (k = 0; k < j; k++) v -= B[i * N + k] * B[j * N + k];
            B[i * N + j] = v / d;
        }
    }
    for (i = 0; i < N; i++) { /* z = R^-1 r */
        double v = r[i];
        for (k = 0; k < i; k++) v -= B[i * N + k] * r[k];
        r[i] = v / B[i * N + i];
        chi2 += r[i] * r[i];
    }
    free(B);
    free(r);
    return -0.5 * (chi2 + logdet + N * log(JO_TWO_PI));
}

int main(void)
{
    int i, j, it;
    /* Kepler: closed forms and the residual */
    CHECK(jo_kepler_twobody(0.7, 0.0, 1e-10, 128, &it) == 0.7, "e = 0");
    CHECK(jo_kepler_twobody(0.0, 0.6, 1e-10, 128, &it) == 0.0, "M = 0");
    CHECK(fabs(jo_kepler_twobody(M_PI, 0.9, 1e-10, 128, &it) - M_PI) < 1e-15, "M = pi");
    for (i = 0; i < 200; i++) {
        const double M = -3.0 + 0.03 * i, e = 0.001 + 0.004 * i;
        const double E = jo_kepler_twobody(M, e, 1e-10, 128, &it);
        CHECK(fabs(E - e * sin(E) - M) < 1e-12, "residual at M=%g e=%g", M, e);
        CHECK(fabs(E - jo_kepler_bisect(M, e)) < 1e-9, "bisection at M=%g e=%g", M, e);
    }

    /* a star: N = 7 epochs, L = 3 (K, v0, v1), physical jitter */
    enum { N = 7, L = 3, NROW = 1000 };
    double t[N], rv[N], ivar[N], trend[N * (L - 1)];
    const double mu[L] = {0.5, -2.0, 0.01}, lam[L] = {900.0, 1e4, 0.25};
    for (i = 0; i < N; i++) {
        t[i] = 100.0 + 37.3 * i + 3.1 * sin(i);
        rv[i] = 4.0 * cos(0.05 * t[i]) - 3.0 + 0.2 * sin(7.0 * i);
        ivar[i] = 1.0 / (0.1 + 0.02 * i);
        trend[i * 2 + 0] = 1.0;
        trend[i * 2 + 1] = t[i] - t[0];
    }
    jo_star st = {N, L, t, rv, ivar, trend, mu, lam, t[0], 0, 0.0, 1.0, INFINITY, 1, 1e-10, 128};
    double *rows = (double *)malloc(sizeof(double) * 5 * NROW);
    for (i = 0; i < NROW; i++) {
        rows[5 * i + 0] = exp(log(2.0) + (log(2048.0) - log(2.0)) * ((i * 7919) % NROW) / NROW);
        rows[5 * i + 1] = 0.9 * ((i * 104729) % NROW) / NROW;
        rows[5 * i + 2] = JO_TWO_PI * ((i * 31) % NROW) / NROW;
        rows[5 * i + 3] = JO_TWO_PI * ((i * 17) % NROW) / NROW;
        rows[5 * i + 4] = 0.01 * (i % 13);
    }
    double *ll0 = (double *)malloc(sizeof(double) * NROW);
    double *ll1 = (double *)malloc(sizeof(double) * NROW);
    unsigned char *capped = (unsigned char *)malloc(NROW);
    jo_batch_lnl(&st, rows, NROW, 0, 4, ll0, capped);
    jo_batch_lnl(&st, rows, NROW, 1, 3, ll1, NULL);
    double MT[L * N], w[N];
    for (i = 0; i < NROW; i += 37) {
        const double *row = rows + 5 * i;
        fill_design(&st, row, MT);
        effective_ivar(&st, row[4], w);
        const double want = dense_lnl(N, L, MT, lam, mu, w, rv);
        CHECK(fabs(ll0[i] - want) < 1e-8 * fmax(1.0, fabs(want)), "LU form row %d", i);
        CHECK(fabs(ll1[i] - want) < 1e-8 * fmax(1.0, fabs(want)), "Cholesky form row %d", i);
        CHECK(jo_lnl_row(&st, row, 0, &it) == ll0[i], "single row = batch row %d", i);
    }
    /* linear posterior: a solves Ainv a = M C^-1 y + Lambda^-1 mu */
    double a[L], A[L * L];
    jo_linear_posterior(&st, rows + 5 * 11, a, A);
    fill_design(&st, rows + 5 * 11, MT);
    effective_ivar(&st, rows[5 * 11 + 4], w);
    for (i = 0; i < L; i++) {
        double lhs = 0.0, rhs = mu[i] / lam[i];
        for (j = 0; j < L; j++) {
            double g = i == j ? 1.0 / lam[i] : 0.0;
            for (it = 0; it < N; it++) g += MT[i * N + it] * w[it] * MT[j * N + it];
            lhs += g * a[j];
        }
        for (it = 0; it < N; it++) rhs += MT[i * N + it] * w[it] * rv[it];
        CHECK(fabs(lhs - rhs) < 1e-9 * fmax(1.0, fabs(rhs)), "posterior mean eq %d", i);
    }

    /* numpy PCG64: default_rng(42)'s state and its first raw outputs */
    const uint64_t sh = 0xcea44f6798798f2aULL, sl = 0xacbc7c9d68860ac8ULL;
    const uint64_t ih = 0xfa505436c9a8416eULL, il = 0x66caf2e28d25abffULL;
    uint64_t raw[3];
    double u[3];
    jo_pcg64_raw(sh, sl, ih, il, 0, 3, raw);
    CHECK(raw[0] == 0xc621fbcd16d92688ULL && raw[1] == 0x705a5661a791ffc1ULL &&
              raw[2] == 0xdbcd12c26eda1624ULL, "PCG64 raw stream");
    jo_pcg64_raw(sh, sl, ih, il, 2, 1, raw);
    CHECK(raw[0] == 0xdbcd12c26eda1624ULL, "PCG64 advance");
    jo_pcg64_uniform(sh, sl, ih, il, 0, 3, u);
    CHECK(fabs(u[0] - 0.77395605) < 1e-8 && fabs(u[2] - 0.85859792) < 1e-8, "PCG64 uniform");

    /* rejection: exp(ll - max) > u, ascending, truncated, NaN rejects all */
    const double llr[6] = {-1.0, 0.0, -50.0, -0.1, -3.0, -0.5};
    const double ur[6] = {0.3, 0.99, 0.0, 0.95, 0.04, 0.7};
    int64_t idx[6];
    double mx;
    int64_t np_ = jo_reject(llr, 6, ur, 2, idx, &mx);
    CHECK(np_ == 4 && idx[0] == 0 && idx[1] == 1 && mx == 0.0, "reject");
    const double lln[3] = {0.0, NAN, -1.0};
    CHECK(jo_reject(lln, 3, ur, 3, idx, &mx) == 0 && isnan(mx), "reject NaN");

    free(rows);
    free(ll0);
    free(ll1);
    free(capped);
    if (failures) {
        fprintf(stderr, "%d failure(s)\n", failures);
        return 1;
    }
    printf("oracle selftest ok\n");
    return 0;
}
