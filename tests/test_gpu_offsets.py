"""Multi-survey data with v0 offsets through the drop-in API on the GPU
(likelihood_helpers.py:8-37, data_helpers.py:56-133,
fast_likelihood.pyx:204-220): three RVData surveys, each with its own
constant offset, and JokerPrior(v0_offsets=[...]) with a Normal K prior and
with the FixedCompanionMass prior.  The oracle star is built from the same
offset-indicator trend columns, with K's variance at index 0 and each
offset's at 2 + i; lnL, the accepted rows and the linear parameters
(K, v0, dv0_1, dv0_2) are compared with the oracle and numpy's streams."""
import numpy as np
import pytest

import thejoker_amd as tj
from oracle import oracle
from thejoker_amd import synthetic as syn
from thejoker_amd import units as u
from thejoker_amd.likelihood_helpers import validate_prepare_data

pytestmark = pytest.mark.gpu

OFFSETS = (0.0, 3.0, -2.0)          # km/s, survey 1 is the reference
OFFSET_SIGMA = (5.0, 8.0)            # prior widths of dv0_1, dv0_2, km/s


def _surveys(seed=3):
    """One binary seen by three surveys over consecutive time ranges."""
    rng = np.random.default_rng(seed)
    out = []
    t0 = 0.0
    for k, (m, span) in enumerate(((10, 400.0), (8, 300.0), (6, 300.0))):
        t = np.sort(rng.uniform(t0, t0 + span, m))
        t0 += span
        err = rng.uniform(0.1, 0.4, m)
        rv = syn.kepler_rv(t, 120.0, 0.3, 1.0, 2.0, 10.0, 0.0) - 5.0 + OFFSETS[k]
        rv = rv + rng.normal(0.0, err)
        out.append(tj.RVData(t=t, rv=rv * u.km / u.s, rv_err=err * u.km / u.s))
    return out


def _prior(kprior):
    offs = [tj.distributions.Normal(0.0, s, unit=u.km / u.s) for s in OFFSET_SIGMA]
    common = dict(P_min=2 * u.day, P_max=2048 * u.day, sigma_v=100 * u.km / u.s, v0_offsets=offs)
    if kprior == "normal":
        return tj.JokerPrior.default(pars={"K": tj.distributions.Normal(0.0, 30.0,
                                                                        unit=u.km / u.s)},
                                     **common)
    return tj.JokerPrior.default(sigma_K0=30 * u.km / u.s, P0=365.25 * u.day, **common)


def _rows(n, seed=0):
    cols = syn.prior_columns(n, seed=seed)
    s = tj.JokerSamples(poly_trend=1)
    s["P"] = cols["P"] * u.day
    s["e"] = cols["e"]
    s["omega"] = cols["omega"] * u.rad
    s["M0"] = cols["M0"] * u.rad
    s["s"] = cols["s"] * u.km / u.s
    return s, oracle.pack_rows(cols["P"], cols["e"], cols["omega"], cols["M0"], cols["s"])


def _oracle_star(data, prior):
    all_data, ids, trend = validate_prepare_data(data, prior.poly_trend, prior.n_offsets)
    assert trend.shape == (len(all_data), 3)
    # offset indicator columns in survey order
    assert (trend[:10, 1:] == 0).all() and (trend[10:18, 1] == 1).all()
    assert (trend[18:, 2] == 1).all()
    mu = np.zeros(4)
    lam = np.array([900.0, 100.0 ** 2, OFFSET_SIGMA[0] ** 2, OFFSET_SIGMA[1] ** 2])
    kw = {}
    if kprior_of(prior) == "fcm":
        lam[0] = 1.0
        kw = dict(k_prior_mode=1, sigma_K0=30.0, P0=365.25,
                  max_K=prior.pars["K"].max_K.to_value(u.km / u.s))
    ost = oracle.OracleStar(all_data._t_bmjd, all_data.rv.value, all_data.ivar_value(), trend, mu,
                            lam, all_data._t_ref_bmjd, jitter_mode=1, **kw)
    return ost


def kprior_of(prior):
    return "fcm" if isinstance(prior.pars["K"], tj.distributions.FixedCompanionMass) else "normal"


@pytest.mark.parametrize("kprior", ["normal", "fcm"])
def test_offsets_lnl_matches_oracle(gpu, kprior):
    data = _surveys()
    prior = _prior(kprior)
    samples, rows = _rows(1 << 16)
    joker = tj.TheJoker(prior)
    ll = joker.marginal_ln_likelihood(data, samples)
    want, capped = oracle.batch_lnl(_oracle_star(data, prior), rows, n_threads=16,
                                    return_capped=True)
    ok = ~capped
    rel = np.abs(ll[ok] - want[ok]) / np.maximum(np.abs(want[ok]), 1.0)
    assert rel.max() < 1e-10, rel.max()


@pytest.mark.parametrize("kprior", ["normal", "fcm"])
@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_offsets_rejection_sample(gpu, kprior, precision):
    """rejection_sample (in memory) over 2^18 rows: the accepted rows are
    np.where over the oracle's lnL with numpy's uniforms, and K, v0, dv0_1,
    dv0_2 are the multivariate-normal draws from the oracle's posterior."""
    data = _surveys()
    prior = _prior(kprior)
    samples, rows = _rows(1 << 18, seed=5)
    joker = tj.TheJoker(prior, rng=np.random.default_rng(42), precision=precision)
    out = joker.rejection_sample(data, samples, max_posterior_samples=128, in_memory=True)
    ost = _oracle_star(data, prior)
    ll_o, capped = oracle.batch_lnl(ost, rows, n_threads=16, return_capped=True)
    assert not capped[ll_o >= np.nanmax(ll_o) - 50].any()
    rng = np.random.default_rng(42)
    uu = rng.uniform(size=len(rows))
    idx = np.where(np.exp(ll_o - ll_o.max()) > uu)[0][:128]
    assert len(out) == len(idx) > 0
    assert np.array_equal(out["P"].to_value(u.day), rows[idx, 0])
    assert {"K", "v0", "dv0_1", "dv0_2"} <= set(out.tbl)
    for k, i in enumerate(idx):
        _, a, A = oracle.linear_posterior(ost, rows[i])
        lin = rng.multivariate_normal(a, A, size=1)[0]
        got = [out[name].to_value(u.km / u.s)[k] for name in ("K", "v0", "dv0_1", "dv0_2")]
        assert np.allclose(got, lin, rtol=1e-7, atol=1e-7), (k, got, lin)
    # the offsets the data carry are recovered within a few sigma
    assert abs(np.median(out["dv0_1"].to_value(u.km / u.s)) - OFFSETS[1]) < 2.0
    assert abs(np.median(out["dv0_2"].to_value(u.km / u.s)) - OFFSETS[2]) < 2.0
