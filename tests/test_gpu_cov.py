"""Data with a full covariance matrix (jk_star_create_cov): the likelihood
against the dense closed form ln N(y | M mu, C + s^2 I + M Lambda M^T) (numpy
solve / slogdet, design column from the oracle's Kepler solve), the diagonal
case against the diagonal kernel, the linear posterior against numpy, and
the refusals.  The reference has no counterpart (its fast path reads a 1-D
ivar), so these are pinned by the mathematics."""
import numpy as np
import pytest
import torch

from oracle import oracle
from thejoker_amd import _lib
from thejoker_amd import device as D
from thejoker_amd import synthetic as syn

pytestmark = pytest.mark.gpu

MU = np.zeros(2)
LAM = np.array([900.0, 1e4])


def _corr_cov(err, rho=0.6):
    n = err.size
    i = np.arange(n)
    R = rho ** np.abs(i[:, None] - i[None, :])
    return err[:, None] * R * err[None, :]


def _dense_lnl(st, cov, cols, rows, jitter, form="lemma"):
    """ln N(y | M mu, B), B = C_s + M Lambda M^T, C_s = C + s^2 I, with dense
    N x N algebra.  form "naive": solve and slogdet of B itself (B's
    condition number, ~1e8 here, limits it to ~1e-8); form "lemma": the
    matrix determinant lemma and Woodbury's identity,
    ln det B = ln det C_s + ln det Lambda + ln det G and
    r^T B^-1 r = r^T C_s^-1 r - u^T G^-1 u with G = Lambda^-1 + M^T C_s^-1 M,
    u = M^T C_s^-1 r -- dense C_s^-1 (well conditioned), no eigenbasis."""
    out = np.empty(len(rows))
    trend = st.trend()
    for j, r in enumerate(rows):
        m = oracle.rv_unit(st.t, cols["P"][r], cols["e"][r], cols["omega"][r], cols["M0"][r],
                           st.t_ref)
        Mx = np.column_stack([m, trend])
        s2 = cols["s"][r] ** 2 if jitter else 0.0
        Cs = cov + s2 * np.eye(st.M)
        res = st.rv - Mx @ MU
        if form == "naive":
            B = Cs + Mx @ np.diag(LAM) @ Mx.T
            _, logdet = np.linalg.slogdet(2 * np.pi * B)
            out[j] = -0.5 * (res @ np.linalg.solve(B, res) + logdet)
            continue
        Ci = np.linalg.inv(Cs)
        G = np.diag(1 / LAM) + Mx.T @ Ci @ Mx
        u = Mx.T @ Ci @ res
        chi2 = res @ Ci @ res - u @ np.linalg.solve(G, u)
        logdet = (st.M * np.log(2 * np.pi) + np.linalg.slogdet(Cs)[1] + np.sum(np.log(LAM))
                  + np.linalg.slogdet(G)[1])
        out[j] = -0.5 * (chi2 + logdet)
    return out


@pytest.fixture(scope="module")
def case(gpu):
    st = syn.single_star(seed=3)
    cols = syn.prior_columns(1 << 12, seed=31, jitter=True)
    cache = D.PriorCache(cols["P"], cols["e"], cols["omega"], cols["M0"], cols["s"])
    return st, cols, cache


@pytest.mark.parametrize("mode", ["physical", "reference"])
def test_cov_lnl_matches_dense_closed_form(case, mode):
    st, cols, cache = case
    cov = _corr_cov(st.err)
    star = D.DeviceStar(st.t, st.rv, None, st.trend(), MU, LAM, st.t_ref, jitter_mode=mode,
                        cov=cov)
    ll = D.lnl(star, cache, precision="f64").cpu().numpy()
    rows = np.random.default_rng(0).choice(len(cache), 300, replace=False)
    jit = mode == "physical"
    want = _dense_lnl(st, cov, cols, rows, jit)
    err = np.abs(ll[rows] - want) / np.maximum(np.abs(want), 1.0)
    assert err.max() < 1e-10, err.max()
    naive = _dense_lnl(st, cov, cols, rows[:50], jit, form="naive")
    assert np.max(np.abs(ll[rows[:50]] - naive) / np.maximum(np.abs(naive), 1.0)) < 1e-8
    # a plausible mistake -- the covariance transposed into the wrong basis
    # (rotating by Q instead of Q^T) -- is caught
    wrong = D.DeviceStar(st.t, st.rv, None, st.trend(), MU, LAM, st.t_ref, jitter_mode=mode,
                         cov=_corr_cov(st.err[::-1]))
    llw = D.lnl(wrong, cache, precision="f64").cpu().numpy()
    assert np.max(np.abs(llw[rows] - want) / np.maximum(np.abs(want), 1.0)) > 1e-6


@pytest.mark.parametrize("mode", ["physical", "reference"])
def test_diagonal_cov_equals_diagonal_star(case, mode):
    st, cols, cache = case
    diag = D.DeviceStar(st.t, st.rv, st.ivar(), st.trend(), MU, LAM, st.t_ref, jitter_mode=mode)
    full = D.DeviceStar(st.t, st.rv, None, st.trend(), MU, LAM, st.t_ref, jitter_mode=mode,
                        cov=np.diag(st.err ** 2))
    a = D.lnl(diag, cache, precision="f64").cpu().numpy()
    b = D.lnl(full, cache, precision="f64").cpu().numpy()
    assert np.max(np.abs(a - b) / np.maximum(np.abs(a), 1.0)) < 1e-12


def test_cov_linear_posterior_matches_numpy(case):
    st, cols, cache = case
    cov = _corr_cov(st.err, rho=0.4)
    star = D.DeviceStar(st.t, st.rv, None, st.trend(), MU, LAM, st.t_ref, cov=cov)
    rows = np.array([3, 77, 1000, 2047])
    a, acov, ll = D.linear_posterior(star, cache, rows)
    lnl = D.lnl(star, cache, precision="f64").cpu().numpy()
    for j, r in enumerate(rows):
        m = oracle.rv_unit(st.t, cols["P"][r], cols["e"][r], cols["omega"][r], cols["M0"][r],
                           st.t_ref)
        Mx = np.column_stack([m, st.trend()])
        Ci = np.linalg.inv(cov + cols["s"][r] ** 2 * np.eye(st.M))
        G = np.diag(1 / LAM) + Mx.T @ Ci @ Mx
        cov_ref = np.linalg.inv(G)
        a_ref = cov_ref @ (Mx.T @ Ci @ st.rv + MU / LAM)
        assert np.allclose(a[j], a_ref, rtol=1e-9, atol=1e-9 * np.abs(a_ref).max())
        assert np.allclose(acov[j], cov_ref, rtol=1e-9, atol=1e-12)
        assert abs(ll[j] - lnl[r]) <= 1e-9 * max(1.0, abs(lnl[r]))


def test_cov_rejection_equals_reject_of_lnl(case):
    st, cols, cache = case
    star = D.DeviceStar(st.t, st.rv, None, st.trend(), MU, LAM, st.t_ref, cov=_corr_cov(st.err))
    ll = D.lnl(star, cache, precision="f64")
    want = D.reject(ll, D.pcg64_state(np.random.default_rng(8)), 64)
    _, idx, n_pass, mx = D.lnl_reject(star, cache, D.pcg64_state(np.random.default_rng(8)), 64)
    assert np.array_equal(idx, want[0]) and n_pass == want[1] and mx == want[2]


def test_cov_refusals(case):
    st, cols, cache = case
    bad = _corr_cov(st.err)
    bad[0, 0] = -1.0                                   # not positive definite
    with pytest.raises(_lib.JokerLibError):
        D.DeviceStar(st.t, st.rv, None, st.trend(), MU, LAM, st.t_ref, cov=bad)
    n = 65                                             # above the LDS budget
    t = np.linspace(0, 100, n)
    with pytest.raises(_lib.JokerLibError):
        D.DeviceStar(t, np.zeros(n), None, np.ones((n, 1)), MU, LAM, 0.0, cov=np.eye(n))
    star = D.DeviceStar(st.t, st.rv, None, st.trend(), MU, LAM, st.t_ref, cov=_corr_cov(st.err))
    with pytest.raises(_lib.JokerLibError):
        D.lnl(star, cache, precision="mixed", bound=torch.empty(len(cache), dtype=torch.float32,
                                                                 device="cuda"))


def test_thejoker_with_full_covariance_data(case):
    """TheJoker on an RVData with a 2-D rv_err: FP64 rejection sampling runs
    on the covariance kernel and its ln_likelihood column is that kernel's;
    mixed precision is refused."""
    import thejoker_amd as tj
    from thejoker_amd import units as u
    st, cols, cache = case
    cov = _corr_cov(st.err)
    data = tj.RVData(st.t, st.rv * u.km / u.s, cov * u.km ** 2 / u.s ** 2)
    prior = tj.JokerPrior.default(P_min=2 * u.day, P_max=2048 * u.day, sigma_v=100 * u.km / u.s,
                                  pars={"K": tj.distributions.Normal(0.0, 30.0, unit=u.km / u.s)})
    joker = tj.TheJoker(prior, rng=np.random.default_rng(5), precision="f64",
                        jitter_mode="reference")
    ll = joker.marginal_ln_likelihood(data, cache)
    star = D.DeviceStar(st.t, st.rv, None, st.trend(), MU, LAM, st.t_ref,
                        jitter_mode="reference", cov=cov)
    want = D.lnl(star, cache, precision="f64").cpu().numpy()
    assert np.allclose(ll, want, rtol=1e-12, atol=0)
    samples = joker.rejection_sample(data, cache, max_posterior_samples=32, return_logprobs=True,
                                     in_memory=True)
    assert len(samples) >= 1
    with pytest.raises(NotImplementedError):
        tj.TheJoker(prior, precision="mixed").marginal_ln_likelihood(data, cache)
