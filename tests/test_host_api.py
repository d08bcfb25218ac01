"""Host-side logic of the drop-in API (no device): data validation, design
matrices, prior construction and sampling, sample packing and I/O, batch
splitting, the iterative growth rule and host RNG bookkeeping."""
import numpy as np
import pytest

import thejoker_amd as tj
from thejoker_amd import distributions as dist
from thejoker_amd import units as u
from thejoker_amd.device import pcg64_advance_state, pcg64_state, star_arrays
from thejoker_amd.likelihood_helpers import (batch_tasks, get_trend_design_matrix,
                                             next_batch_size, validate_prepare_data)


def _data(n=8, seed=0, t0=0.0):
    rng = np.random.default_rng(seed)
    t = t0 + np.sort(rng.uniform(0, 100, n))
    return tj.RVData(t=t, rv=rng.normal(0, 5, n) * u.km / u.s,
                     rv_err=rng.uniform(0.1, 0.3, n) * u.km / u.s)


# -- RVData (data.py:39-122) -------------------------------------------------
def test_rvdata_sorts_cleans_and_sets_t_ref():
    t = np.array([5.0, 1.0, np.nan, 3.0])
    d = tj.RVData(t=t, rv=[1.0, 2.0, 3.0, np.inf] * u.km / u.s, rv_err=[0.1] * 4 * u.km / u.s)
    assert len(d) == 2
    assert np.array_equal(d.t, [1.0, 5.0])
    assert np.array_equal(d.rv.value, [2.0, 1.0])
    assert d.t_ref == 1.0
    d2 = tj.RVData(t=[1.0, 2.0], rv=[1.0, 2.0] * u.m / u.s, rv_err=[0.5, 0.5] * u.m / u.s,
                   t_ref=False)
    assert d2.t_ref is None and d2._t_ref_bmjd == 0.0
    with pytest.raises(ValueError):
        tj.RVData(t=[1.0, 2.0], rv=[1.0, 2.0, 3.0] * u.km / u.s, rv_err=[0.1] * 3 * u.km / u.s)
    assert np.allclose(d.ivar_value(), 100.0)


def test_rvdata_units():
    d = tj.RVData(t=[1.0, 2.0], rv=[1.0, 2.0] * u.km / u.s, rv_err=[100.0, 200.0] * u.m / u.s)
    assert np.allclose(d.ivar_value(), [100.0, 25.0])


# -- design matrix (tests/test_likelihood_helpers.py:8-36) --------------------
def test_design_matrix_offsets():
    d1, d2, d3 = _data(8, 1), _data(4, 2, 500.0), _data(3, 3, 1000.0)
    data, ids, M = validate_prepare_data([d1, d2, d3], 1, 2)
    assert np.allclose(M[:, 0], 1.0)
    idx = np.arange(len(data))
    mask = (idx >= 8) & (idx < 12)
    assert np.allclose(M[mask, 1], 1.0) and np.allclose(M[~mask, 1], 0.0)
    mask = (idx >= 12) & (idx < 15)
    assert np.allclose(M[mask, 2], 1.0) and np.allclose(M[~mask, 2], 0.0)
    with pytest.raises(ValueError):
        validate_prepare_data([d1, d2], 1, 2)
    with pytest.raises(ValueError):
        validate_prepare_data(d1, 1, 1)


def test_trend_columns():
    d = _data(10)
    M = get_trend_design_matrix(d, None, 3)
    dt = d.t - d.t_ref
    assert M.shape == (10, 3)
    assert np.allclose(M[:, 0], 1.0) and np.allclose(M[:, 1], dt) and np.allclose(M[:, 2], dt ** 2)
    # the powers round as np.vander's running products do
    M5 = get_trend_design_matrix(d, None, 5)
    assert np.array_equal(M5[:, 1:], np.vander(dt, N=5, increasing=True)[:, 1:])


# -- batch_tasks (tests/test_utils.py:21-33) ---------------------------------
def test_batch_tasks():
    N, start = 10000, 1103
    tasks = batch_tasks(N, n_batches=16, start_idx=start)
    assert tasks[0][0][0] == start and tasks[-1][0][1] == N + start
    assert all(tasks[i][0][1] == tasks[i + 1][0][0] for i in range(15))
    tasks = batch_tasks(N, n_batches=16, start_idx=start, arr=np.random.random(8 * N))
    assert sum(t[0].size for t in tasks) == N
    assert len(batch_tasks(0, 1, arr=np.arange(0))) == 1


# -- iterative growth rule (likelihood_helpers.py:196-207) --------------------
def test_next_batch_size_known_answers():
    # int(4 * 246 / 10 * 32768) = int(3224371.2)
    assert next_batch_size(256, 10, 32768, 32768, 1 << 26, 4) == 3224371
    # capped at the samples left
    assert next_batch_size(256, 10, 32768, 32768, 100000, 4) == 100000 - 32768
    # safety factor 1 (in-memory path)
    assert next_batch_size(4, 1, 512, 512, 10 ** 6, 1) == 1536
    assert next_batch_size(4, 3, 1000, 1000, 1000, 1) <= 0


# -- prior (prior.py) --------------------------------------------------------
def test_default_prior_structure():
    p = tj.JokerPrior.default(P_min=1 * u.day, P_max=1 * u.yr, sigma_K0=25 * u.km / u.s,
                              sigma_v=[10 * u.km / u.s, 0.1 * u.km / u.s / u.yr], poly_trend=2)
    assert p.par_names == ["P", "e", "omega", "M0", "s", "K", "v0", "v1"]
    assert isinstance(p.pars["K"], dist.FixedCompanionMass)
    assert isinstance(p.pars["P"], dist.UniformLog)
    s = p.sample(size=1000, generate_linear=True, rng=np.random.default_rng(1),
                 return_logprobs=True)
    assert len(s) == 1000
    P = s["P"].to_value(u.day)
    assert P.min() >= 1 and P.max() <= 365.25
    assert np.all((s["e"].value >= 0) & (s["e"].value < 1))
    assert np.all(np.isfinite(s["ln_prior"].value))
    s2 = p.sample(size=1000, generate_linear=True, rng=np.random.default_rng(1))
    assert np.array_equal(s["P"].value, s2["P"].value)


def test_prior_validation():
    with pytest.raises(ValueError):
        tj.JokerPrior.default(P_min=1 * u.day, sigma_K0=25 * u.km / u.s, sigma_v=10 * u.km / u.s)
    with pytest.raises(ValueError):
        tj.JokerPrior.default(P_min=1 * u.day, P_max=10 * u.day, sigma_v=10 * u.km / u.s)
    with pytest.raises(ValueError):  # non-Normal linear parameter
        tj.JokerPrior.default(P_min=1 * u.day, P_max=10 * u.day, sigma_v=10 * u.km / u.s,
                              pars={"K": dist.Uniform(0, 1, unit=u.km / u.s)})
    with pytest.raises(ValueError):  # wrong unit
        tj.JokerPrior.default(P_min=1 * u.day, P_max=10 * u.day, sigma_v=10 * u.km / u.s,
                              pars={"K": dist.Normal(0, 1, unit=u.day)})


def test_uniformlog_draw_formula():
    # distributions.py:24-28: exp(U (ln b - ln a) + ln a) from rng.uniform
    d = dist.UniformLog(2.0, 2048.0, unit=u.day)
    x = d.sample(np.random.default_rng(5), 100)
    uu = np.random.default_rng(5).uniform(size=100)
    assert np.array_equal(x, np.exp(uu * (np.log(2048.0) - np.log(2.0)) + np.log(2.0)))


def test_star_arrays_prior_moments():
    p = tj.JokerPrior.default(P_min=1 * u.day, P_max=100 * u.day, sigma_v=[10 * u.km / u.s,
                              0.5 * u.km / u.s / u.day],
                              pars={"K": dist.Normal(0.0, 30000.0, unit=u.m / u.s)}, poly_trend=2)
    d = _data(6)
    a = star_arrays(d, p, get_trend_design_matrix(d, None, 2))
    assert np.allclose(a["lam"], [900.0, 100.0, 0.25])
    assert a["k_prior_mode"] == 0
    p2 = tj.JokerPrior.default(P_min=1 * u.day, P_max=100 * u.day, sigma_K0=30 * u.km / u.s,
                               sigma_v=100 * u.km / u.s)
    a2 = star_arrays(d, p2, get_trend_design_matrix(d, None, 1))
    assert a2["k_prior_mode"] == 1 and a2["sigma_K0"] == 30.0
    assert np.isclose(a2["P0"], 365.25) and np.isclose(a2["max_K"], 500.0)


# -- samples (samples.py) ----------------------------------------------------
def test_samples_pack_unpack_and_io(tmp_path):
    p = tj.JokerPrior.default(P_min=1 * u.day, P_max=100 * u.day, sigma_K0=30 * u.km / u.s,
                              sigma_v=100 * u.km / u.s)
    s = p.sample(size=50, generate_linear=True, rng=np.random.default_rng(3), t_ref=123.0)
    packed, units = s.pack()
    assert packed.shape == (50, 5)
    assert list(units) == ["P", "e", "omega", "M0", "s"]
    s2 = tj.JokerSamples.unpack(packed, units, t_ref=123.0)
    assert np.allclose(s2["P"].value, s["P"].value)
    fn = str(tmp_path / "s.npz")
    s.write(fn)
    s3 = tj.JokerSamples.read(fn)
    assert s3.t_ref == 123.0 and len(s3) == 50
    assert np.array_equal(s3["K"].value, s["K"].value)
    assert s3["K"].unit == s["K"].unit
    with pytest.raises(ValueError):
        s["bogus"] = np.zeros(50)
    with pytest.raises(u.UnitsError):
        s["P"] = np.zeros(50) * u.km / u.s
    sub = s[np.arange(5)]
    assert len(sub) == 5 and len(s[3]) == 1
    t0 = s.get_t0()
    assert np.allclose(t0, 123.0 + s["P"].value * s["M0"].value / (2 * np.pi))


def test_wrap_K():
    s = tj.JokerSamples()
    s["K"] = np.array([-1.0, 2.0]) * u.km / u.s
    s["omega"] = np.array([1.0, 1.0]) * u.rad
    s.wrap_K()
    assert np.allclose(s["K"].value, [1.0, 2.0])
    assert np.allclose(s["omega"].value, [(1.0 + np.pi) % (2 * np.pi), 1.0])


# -- host RNG bookkeeping -----------------------------------------------------
@pytest.mark.parametrize("n", [0, 1, 7, 1000, 123456789])
def test_advance_matches_numpy(n):
    g1 = np.random.default_rng(1234)
    g2 = np.random.default_rng(1234)
    pcg64_advance_state(g1, n)
    g2.bit_generator.advance(n)
    assert g1.bit_generator.state == g2.bit_generator.state
    assert pcg64_state(g1)[1] == g2.bit_generator.state["state"]["inc"]


def test_advance_equals_drawing():
    g1 = np.random.default_rng(5)
    g2 = np.random.default_rng(5)
    pcg64_advance_state(g1, 1000)
    g2.uniform(size=1000)
    assert np.array_equal(g1.uniform(size=10), g2.uniform(size=10))


def test_non_pcg_generator_rejected():
    g = np.random.Generator(np.random.MT19937(0))
    with pytest.raises(TypeError):
        pcg64_state(g)


def test_thejoker_constructor_validation():
    p = tj.JokerPrior.default(P_min=1 * u.day, P_max=100 * u.day, sigma_K0=30 * u.km / u.s,
                              sigma_v=100 * u.km / u.s)
    tj.TheJoker(p)
    tj.TheJoker(p, rng=np.random.default_rng(42))
    with pytest.raises(TypeError):
        tj.TheJoker("jsdfkj")
    with pytest.raises(TypeError):
        tj.TheJoker(p, rng="sdfks")
    with pytest.raises(TypeError):
        tj.TheJoker(p, pool="sdfks")
    with pytest.raises(ValueError):
        tj.TheJoker(p, jitter_mode="bogus")


def test_units_roundtrip():
    assert (1 * u.yr).to_value(u.day) == 365.25
    assert np.isclose((1 * u.m / u.s / u.day).to_value(u.km / u.s / u.yr), 0.36525)
    assert u.as_quantity(3.0, u.km / u.s).to_value(u.m / u.s) == 3000.0
    with pytest.raises(u.UnitsError):
        (1 * u.day).to_value(u.km)


@pytest.mark.parametrize("L", [1, 2, 3, 5, 8])
@pytest.mark.parametrize("n,size", [(1, 1), (7, 3), (256, 1), (3000, 2)])
def test_mvn_draws_equal_row_by_row(L, n, size):
    """Stacked linear-parameter draws give the values and the generator state
    of rng.multivariate_normal called row by row (fast_likelihood.pyx:529-530)."""
    from thejoker_amd.thejoker import mvn_draws
    g0 = np.random.default_rng(L * 1000 + n)
    A = g0.normal(size=(n, L, L))
    covs = A @ np.swapaxes(A, 1, 2) + 1e-3 * np.eye(L)
    covs[::5] *= 1e4
    means = g0.normal(size=(n, L)) * 50
    g1, g2 = np.random.default_rng(42), np.random.default_rng(42)
    want = np.stack([g1.multivariate_normal(means[j], covs[j], size=size) for j in range(n)])
    got = mvn_draws(g2, means, covs, size)
    assert np.array_equal(got, want)
    assert g1.bit_generator.state == g2.bit_generator.state
