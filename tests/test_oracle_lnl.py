"""Pins the oracle's marginal likelihood (fast_likelihood.pyx:255-469
restated) against independent numpy dense-Gaussian vectors, 2-D quadrature of
the un-marginalised likelihood, and its own O(N L^2) form."""
import numpy as np
import pytest

from oracle import oracle
from tests._golden import lnl_stars, oracle_star, rel_err

STARS = lnl_stars()

# Tolerance against the dense golden vectors.  Any FP64 evaluation of the
# phase 2 pi (t - t_ref)/P - M0 carries ~ulp(M) / (1 - e cos E) of anomaly
# noise, ~1e-10 relative in lnL at |M| ~ 1e3; the "long" star (quadratic trend
# over 3650 d, cond(B) ~ 1e10) also loses ~1e-8 relative in the N x N LU
# log-determinant of the reference algorithm itself.  A dropped term, a wrong
# sign or a transposed operand moves lnL by far more than either.
DENSE_RTOL = {"long": 1e-6}


@pytest.mark.parametrize("name", sorted(STARS))
@pytest.mark.parametrize("mode", ["reference", "physical"])
@pytest.mark.parametrize("method", ["reference", "chol"])
def test_oracle_vs_dense_golden(name, mode, method):
    st = STARS[name]
    ost = oracle_star(st, jitter_mode=0 if mode == "reference" else 1)
    ll, capped = oracle.batch_lnl(ost, st["rows"], method=method, return_capped=True)
    want = st[f"ll_{mode}_mode"]
    ok = ~capped
    assert ok.sum() >= len(ok) - 3
    err = rel_err(ll[ok], want[ok])
    assert err.max() < DENSE_RTOL.get(name, 1e-9), (name, err.max(), np.argmax(err))


def test_reference_mode_ignores_jitter():
    # fast_likelihood.pyx:458 computes s_ivar and never reads it.
    st = STARS["cfg1"]
    ost = oracle_star(st, 0)
    rows = st["rows"].copy()
    a = oracle.batch_lnl(ost, rows)
    rows[:, 4] = 0.0
    b = oracle.batch_lnl(ost, rows)
    assert np.array_equal(a, b)


def test_modes_agree_at_zero_jitter():
    st = STARS["ecsv"]
    rows = st["rows"].copy()
    rows[:, 4] = 0.0
    a = oracle.batch_lnl(oracle_star(st, 0), rows)
    b = oracle.batch_lnl(oracle_star(st, 1), rows)
    assert rel_err(a, b).max() < 1e-13


@pytest.mark.parametrize("i_row", [3, 10, 40])
def test_marginal_equals_quadrature(i_row):
    """lnL = ln of the integral over (K, v0) of prod_n N(y_n | K m_n + v0,
    sigma_n^2) N(K | 0, 10^2) N(v0 | 0, 100^2) -- the definition the
    closed form marginalises (py_likelihood.py:32-107)."""
    from scipy import integrate

    st = STARS["three"]
    row = st["rows"][i_row]
    ost = oracle_star(st, 1)
    ll = oracle.batch_lnl(ost, row[None, :])[0]
    m = oracle.rv_unit(st["t"], row[0], row[1], row[2], row[3], st["t_ref"])
    var = st["err"] ** 2 + row[4] ** 2
    y = st["rv"]
    sK, sv = np.sqrt(st["lam"][0]), np.sqrt(st["lam"][1])
    # Centre and scale the integration box on the posterior of (K, v0).
    X = np.stack([m, np.ones_like(m)], axis=1)
    G = np.diag(1 / st["lam"]) + X.T @ (X / var[:, None])
    mean = np.linalg.solve(G, X.T @ (y / var))
    sd = np.sqrt(np.diag(np.linalg.inv(G)))

    def integrand(v0, K):
        r = y - K * m - v0
        lp = -0.5 * np.sum(r * r / var + np.log(2 * np.pi * var))
        lp += -0.5 * (K / sK) ** 2 - np.log(sK) - 0.5 * np.log(2 * np.pi)
        lp += -0.5 * (v0 / sv) ** 2 - np.log(sv) - 0.5 * np.log(2 * np.pi)
        return np.exp(lp - ll)

    val, _ = integrate.dblquad(integrand, mean[0] - 12 * sd[0], mean[0] + 12 * sd[0],
                               mean[1] - 12 * sd[1], mean[1] + 12 * sd[1],
                               epsabs=1e-12, epsrel=1e-11)
    assert abs(np.log(val)) < 1e-8


@pytest.mark.parametrize("name", ["cfg1", "ecsv", "long", "doctest"])
def test_linear_posterior_vs_dense(name):
    st = STARS[name]
    ost = oracle_star(st, 1)
    _, capped = oracle.batch_lnl(ost, st["rows"], return_capped=True)
    for i in range(len(st["post_a"])):
        if capped[i]:
            continue
        ll, a, A = oracle.linear_posterior(ost, st["rows"][i])
        assert np.allclose(a, st["post_a"][i], rtol=1e-8, atol=1e-10 * np.abs(a).max())
        assert np.allclose(A, st["post_cov"][i], rtol=1e-7, atol=1e-12 * np.abs(A).max())


def test_fixed_companion_mass_clip_only_on_likelihood_path():
    # Lambda_0 = min(max_K^2, ...) for lnL (fast_likelihood.pyx:461-464) but
    # unclipped for the posterior (fast_likelihood.pyx:519-521): a tiny period
    # makes the unclipped variance exceed max_K^2, so the two paths' lnL differ.
    st = dict(STARS["doctest"])
    ost = oracle_star(st, 0)
    row = np.array([0.01, 0.5, 1.0, 2.0, 0.0])
    ll_lik = oracle.batch_lnl(ost, row[None, :])[0]
    ll_post, _, _ = oracle.linear_posterior(ost, row)
    assert np.isfinite(ll_lik) and np.isfinite(ll_post)
    assert ll_lik != ll_post
    row2 = np.array([300.0, 0.2, 1.0, 2.0, 0.0])
    assert abs(oracle.batch_lnl(ost, row2[None, :])[0] - oracle.linear_posterior(ost, row2)[0]) < 1e-9


def test_threads_match_serial():
    st = STARS["cfg1"]
    ost = oracle_star(st, 1)
    rows = np.tile(st["rows"], (8, 1))
    a = oracle.batch_lnl(ost, rows, n_threads=1)
    b = oracle.batch_lnl(ost, rows, n_threads=5)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("jitter_mode", [0, 1])
def test_subset_upper_bound(jitter_mode):
    """The bound the probe stars of the serving loop rest on: the marginal
    likelihood of a star on a subset S of its epochs, plus
    -1/2 sum over the other epochs of ln(2 pi sigma_n^2), is at least the
    whole star's, for every row (the other epochs' conditional density
    given y_S is at most that of their noise).  Checked with the oracle's
    reference-form lnL on binaries and a constant star, jitter included."""
    from thejoker_amd import synthetic as syn
    from thejoker_amd.device import probe_epochs, probe_epochs_residual

    cols = syn.prior_columns(1 << 12, seed=51, jitter=True)
    rows = oracle.pack_rows(cols["P"], cols["e"], cols["omega"], cols["M0"], cols["s"])
    mu, lam = np.zeros(3), np.array([900.0, 1e4, 0.25])
    stars = syn.many_stars(6, seed=53, M_range=(12, 40))
    fractions = []
    for st in stars:
        def ostar(idx):
            dt = st.t[idx] - st.t_ref
            return oracle.OracleStar(st.t[idx], st.rv[idx], st.ivar()[idx],
                                     np.stack([np.ones_like(dt), dt], 1), mu, lam, st.t_ref,
                                     jitter_mode=jitter_mode)
        full, capped = oracle.batch_lnl(ostar(np.arange(st.M)), rows, n_threads=8,
                                        return_capped=True)
        dt = st.t - st.t_ref
        a = dict(rv=st.rv, ivar=st.ivar(), trend=np.stack([np.ones_like(dt), dt], 1))
        for idx in (probe_epochs(st.M), probe_epochs_residual(a)):
            rest = np.setdiff1d(np.arange(st.M), idx)
            ub = oracle.batch_lnl(ostar(idx), rows, n_threads=8)
            ub = ub - 0.5 * np.sum(np.log(2 * np.pi / st.ivar()[rest]))
            ok = np.isfinite(full) & ~capped
            assert ok.mean() > 0.99
            assert (ub[ok] >= full[ok] - 1e-9 * np.abs(full[ok])).all()
        fractions.append(np.mean(ub[ok] < full[ok].max() - 40.0))
    # the bound is not vacuous: on some binary most rows fall 40 below the
    # maximum by the bound alone (with jitter; without it, 2^12 rows leave
    # these binaries' maxima far below most rows' bounds)
    if jitter_mode == 1:
        assert max(fractions[0::2]) > 0.5, fractions


def test_probe_epochs_span_the_baseline():
    from thejoker_amd.device import probe_epochs
    for n in (12, 13, 40, 97, 200):
        idx = probe_epochs(n)
        assert idx[0] == 0 and idx[-1] == n - 1
        assert len(idx) == min(n, max(6, round(n / 6))) and np.all(np.diff(idx) > 0)
    assert np.array_equal(probe_epochs(5), np.arange(5))


def test_probe_epochs_take_the_largest_trend_residuals():
    """probe_epochs_residual: probe_size(n) epochs, in time order, whose
    weighted squared residuals from the weighted least-squares trend fit are
    the largest (checked against the fit by numpy's lstsq restated here and
    against a planted outlier); probe_size is about sqrt(n) in [5, 16]."""
    from thejoker_amd.device import probe_epochs_residual, probe_size, trend_signal
    assert [probe_size(n) for n in (6, 12, 25, 40, 100, 200, 400)] == [5, 5, 5, 6, 10, 14, 16]
    rng = np.random.default_rng(3)
    for n in (12, 40, 200):
        t = np.sort(rng.uniform(0, 1000, n))
        w = rng.uniform(1, 20, n)
        y = 3.0 + 0.01 * t + 5 * np.sin(t / 37.0) + rng.normal(0, 1, n) / np.sqrt(w)
        T = np.stack([np.ones(n), t - t[0]], 1)
        a = dict(rv=y, ivar=w, trend=T)
        idx = probe_epochs_residual(a)
        assert len(idx) == probe_size(n) and np.all(np.diff(idx) > 0)
        beta = np.linalg.solve(T.T @ (w[:, None] * T), T.T @ (w * y))
        r2 = w * (y - T @ beta) ** 2
        rest = np.setdiff1d(np.arange(n), idx)
        assert r2[idx].min() >= r2[rest].max() * (1 - 1e-12)
        assert np.isclose(trend_signal(a), 0.5 * (r2.sum() - (n - 2)))
        y2 = y.copy()
        y2[n // 3] += 1e3
        assert n // 3 in probe_epochs_residual(dict(rv=y2, ivar=w, trend=T), 1)
