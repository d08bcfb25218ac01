"""Pruning by an upper bound (probe_star, jk_lnl_prune_batch): the probe's
bound lies above the star's lnL on every row; rows that are not pruned get
exactly jk_lnl_batch's values; pruned rows hold a bound below their
threshold; the server's results do not change."""
import numpy as np
import pytest
import torch

import thejoker_amd as tj
from thejoker_amd import _lib
from thejoker_amd import batch as B
from thejoker_amd import device as D
from thejoker_amd import serving as S
from thejoker_amd import synthetic as syn
from thejoker_amd import units as u

pytestmark = pytest.mark.gpu


def _arrays(st, L, fcm=False):
    lam = np.array([900.0, 1e4, 0.25, 1e-4][:L])
    return dict(t=st.t, rv=st.rv, ivar=st.ivar(), trend=st.trend(poly_trend=L - 1),
                mu=np.zeros(L), lam=lam, t_ref=st.t_ref, cov=None,
                k_prior_mode=_lib.JK_KPRIOR_FIXED_COMPANION_MASS if fcm else _lib.JK_KPRIOR_NORMAL,
                sigma_K0=30.0, P0=365.25, max_K=300.0)


def _star(a, mode):
    return D.DeviceStar(a["t"], a["rv"], a["ivar"], a["trend"], a["mu"], a["lam"], a["t_ref"],
                        a["k_prior_mode"], a["sigma_K0"], a["P0"], a["max_K"], jitter_mode=mode)


@pytest.fixture(scope="module")
def setup(gpu):
    stars = syn.many_stars(12, seed=31, M_range=(12, 64))
    cols = syn.prior_columns(1 << 15, seed=33, jitter=True)
    cache = D.PriorCache(cols["P"], cols["e"], cols["omega"], cols["M0"], cols["s"])
    return stars, cache


@pytest.mark.parametrize("mode,L,fcm", [("physical", 3, False), ("reference", 3, False),
                                        ("physical", 2, False), ("physical", 2, True)])
def test_probe_bounds_the_star(setup, mode, L, fcm):
    stars, cache = setup
    for st in stars[:6]:
        a = _arrays(st, L, fcm)
        full = D.lnl(_star(a, mode), cache, precision="f64").cpu().numpy()
        probe, add = D.probe_star(a, mode)
        assert probe.n == len(D.probe_epochs_residual(a)) < st.M
        ub = D.lnl(probe, cache, precision="f64").cpu().numpy() + add
        ok = np.isfinite(full)
        assert ok.mean() > 0.99
        assert (ub[ok] >= full[ok]).all(), (ub[ok] - full[ok]).min()


def _tables(stars, mode="physical", L=3):
    arrs = [_arrays(st, L) for st in stars]
    full = [_star(a, mode) for a in arrs]
    probes = [D.probe_star(a, mode) for a in arrs]
    return B.StarBatch(full), B.StarBatch([p for p, _ in probes]), [add for _, add in probes]


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_prune_launch(setup, precision):
    stars, cache = setup
    table, probes, adds = _tables(stars)
    n = len(cache)
    items, off, total = B.work_items([n] * len(stars), 1000)
    items = items[np.random.default_rng(1).permutation(len(items))]
    d_items = B.upload_items(items)
    mixed = precision == "mixed"
    b_ref = torch.empty(total, dtype=torch.float32, device="cuda") if mixed else None
    ref = B.lnl_batch(table, cache, d_items, len(items), total, precision,
                      bound=b_ref).cpu().numpy()
    ref64 = B.lnl_batch(table, cache, d_items, len(items), total, "f64").cpu().numpy()
    thr = np.zeros((len(stars), 2))
    for k in range(len(stars)):
        seg = ref64[off[k]:off[k] + n]
        thr[k] = (-np.inf if k % 3 == 2 else np.nanmax(seg) - 40.0, adds[k])
    prune = torch.as_tensor(thr.reshape(-1), device="cuda")
    out = torch.full((total,), np.nan, dtype=torch.float64, device="cuda")
    bnd = torch.full((total,), 7.0, dtype=torch.float32, device="cuda") if mixed else None
    n_full = torch.zeros(len(stars), dtype=torch.int64, device="cuda")
    B.lnl_prune_batch(table, probes, prune, cache, d_items, len(items), precision, out=out,
                      bound=bnd, n_full=n_full)
    out = out.cpu().numpy()
    same = (out == ref) | (np.isnan(out) & np.isnan(ref))
    pruned = ~same
    if mixed:
        bnd = bnd.cpu().numpy()
        assert np.array_equal(bnd[same], b_ref.cpu().numpy()[same])
        # a pruned row's value is an upper bound on its lnL, so marked
        assert (bnd[pruned] == -1.0).all()
    n_full = n_full.cpu().numpy()
    for k in range(len(stars)):
        assert n_full[k] == int(same[off[k]:off[k] + n].sum()), k
    n_pruned = 0
    for k in range(len(stars)):
        sl = slice(off[k], off[k] + n)
        p = pruned[sl]
        if k % 3 == 2:
            assert not p.any()
            continue
        n_pruned += int(p.sum())
        assert (out[sl][p] < thr[k, 0]).all()
        assert (out[sl][p] >= ref64[sl][p]).all()
        # every row within the margin of the maximum is evaluated in full
        assert not p[ref64[sl] >= thr[k, 0]].any()
    assert n_pruned > len(cache)


def test_exact_slots_evaluate_in_fp64(setup):
    """jk_batch_set_exact: inside mixed launches (plain, with bounds, and with
    pruning) a marked slot's rows are the FP64 launch's values bit for bit
    with bound 0; the other slots keep their mixed values and bounds;
    set_exact(slot, False) and set_star clear the mark."""
    stars, cache = setup
    table, probes, adds = _tables(stars)
    n = len(cache)
    items, off, total = B.work_items([n] * len(stars), 1000)
    items = items[np.random.default_rng(2).permutation(len(items))]
    d_items = B.upload_items(items)

    def run(precision, bound=True):
        b = torch.full((total,), 7.0, dtype=torch.float32, device="cuda") if bound else None
        ll = B.lnl_batch(table, cache, d_items, len(items), total, precision, bound=b)
        return ll.cpu().numpy(), None if b is None else b.cpu().numpy()

    ref64, _ = run("f64", bound=False)
    mix, bmix = run("mixed")
    mix_nb, _ = run("mixed", bound=False)
    exact = {1, 4, 7}
    for k in exact:
        table.set_exact(k)
    got, bgot = run("mixed")
    got_nb, _ = run("mixed", bound=False)
    for k in range(len(stars)):
        sl = slice(off[k], off[k] + n)
        if k in exact:
            assert np.array_equal(got[sl], ref64[sl], equal_nan=True), k
            assert np.array_equal(got_nb[sl], ref64[sl], equal_nan=True), k
            assert (bgot[sl] == 0.0).all(), k
        else:
            assert np.array_equal(got[sl], mix[sl], equal_nan=True), k
            assert np.array_equal(bgot[sl], bmix[sl]), k
            assert np.array_equal(got_nb[sl], mix_nb[sl], equal_nan=True), k
    # the pruned launch: unpruned rows of marked slots are FP64 with bound 0
    thr = np.zeros((len(stars), 2))
    for k in range(len(stars)):
        thr[k] = (-np.inf if k == 4 else np.nanmax(ref64[off[k]:off[k] + n]) - 40.0, adds[k])
    prune = torch.as_tensor(thr.reshape(-1), device="cuda")
    out = torch.full((total,), np.nan, dtype=torch.float64, device="cuda")
    bnd = torch.full((total,), 7.0, dtype=torch.float32, device="cuda")
    B.lnl_prune_batch(table, probes, prune, cache, d_items, len(items), "mixed", out=out,
                      bound=bnd)
    out, bnd = out.cpu().numpy(), bnd.cpu().numpy()
    for k in range(len(stars)):
        sl = slice(off[k], off[k] + n)
        kept = bnd[sl] != -1.0
        if k == 4:
            assert kept.all()
        want, wb = (ref64[sl], 0.0) if k in exact else (mix[sl], bmix[sl])
        assert np.array_equal(out[sl][kept], want[kept], equal_nan=True), k
        assert np.array_equal(bnd[sl][kept], (np.zeros(n, np.float32) + wb)[kept]), k
    table.set_exact(1, False)
    got, bgot = run("mixed")
    assert np.array_equal(got[off[1]:off[1] + n], mix[off[1]:off[1] + n], equal_nan=True)
    # a slot table: set_star clears the mark
    slots = B.StarBatch.slots(2, 3)
    slots.set(0, table.stars[0])
    slots.set(1, table.stars[1])
    slots.set_exact(0)
    slots.set(0, table.stars[0])
    it2, off2, tot2 = B.work_items([n, n], 1000)
    ll2 = B.lnl_batch(slots, cache, B.upload_items(it2), len(it2), tot2, "mixed").cpu().numpy()
    assert np.array_equal(ll2[:n], mix_nb[off[0]:off[0] + n], equal_nan=True)


def test_zero_uniform_on_a_pruned_row(setup):
    """A pruned row whose upper bound lies within 745 of the maximum but whose
    FP64 lnL does not: with a uniform of exactly 0 at that row, the FP64 path
    rejects it (exp(lnL - max) underflows to 0).  The 40-margin settling
    leaves pruned rows alone; the 750-margin settling that follows a zero
    uniform re-evaluates every pruned row within reach, so the rejection step
    then decides like FP64 (likelihood_helpers.py:107-109)."""
    from tests.test_serving_host import zero_uniform_rng
    stars, cache = setup
    n = len(cache)
    for st in stars:
        table, probes, adds = _tables([st])
        items, off, total = B.work_items([n], 1000)
        d_items = B.upload_items(items)
        ref64 = B.lnl_batch(table, cache, d_items, len(items), total, "f64").cpu().numpy()
        mx = np.nanmax(ref64)
        prune = torch.as_tensor([mx - 40.0, adds[0]], device="cuda")
        out = torch.full((total,), np.nan, dtype=torch.float64, device="cuda")
        bnd = torch.full((total,), 7.0, dtype=torch.float32, device="cuda")
        B.lnl_prune_batch(table, probes, prune, cache, d_items, len(items), "mixed", out=out,
                          bound=bnd)
        o, b = out.cpu().numpy(), bnd.cpu().numpy()
        cand = np.flatnonzero((b == -1.0) & (o > mx - 700.0) & (ref64 < mx - 800.0))
        if len(cand):
            break
    assert len(cand), "no pruned row with a loose enough bound in the test stars"
    j = int(cand[0])
    # settling with the 40 margin, as a round after the first does: the
    # pruned rows keep their marks
    cnt = B.settle_live_batch(table, cache, out, bnd, [0], [n], [0], [mx],
                              margin=B.LIVE_MARGIN_NONZERO_U)
    b40 = bnd.cpu().numpy()
    assert (b40[b == -1.0] == -1.0).all()
    # a uniform of exactly 0 at row j; before settling for it, the stored
    # bound would accept the row
    rng = zero_uniform_rng(j)
    st = D.pcg64_state(rng)
    u_host = zero_uniform_rng(j).uniform(size=n)
    assert u_host[j] == 0.0
    want = np.flatnonzero(np.exp(ref64 - mx) > u_host)
    got, _, _, zero = B.reject_batch(out, [0], [n], [st], n, zero_check=True)
    assert zero[0] == 1
    assert j in got[0] and j not in want
    # the zero-uniform path: settle every row for every uniform, decide again
    cnt = B.settle_live_batch(table, cache, out, bnd, [0], [n], [0], [-np.inf], margin=750.0)
    o2, b2 = out.cpu().numpy(), bnd.cpu().numpy()
    assert b2[j] == 0.0 and o2[j] == ref64[j]
    left = b2 == -1.0
    assert (o2[left] < mx - 750.0).all() and (ref64[left] <= o2[left]).all()
    assert int(cnt.cpu()[0]) >= int(((b == -1.0) & (o > mx - 750.0)).sum())
    got, n_pass, llmax = B.reject_batch(out, [0], [n], [st], n)
    assert llmax[0] == mx
    assert np.array_equal(got[0], want) and n_pass[0] == len(want)


def test_prune_argument_checks(setup):
    stars, cache = setup
    table, probes, adds = _tables(stars[:3])
    other, _, _ = _tables(stars[:3], L=2)
    items, off, total = B.work_items([64] * 3, 64)
    d_items = B.upload_items(items)
    prune = torch.zeros(6, dtype=torch.float64, device="cuda")
    out = torch.empty(total, dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.JokerLibError):
        B.lnl_prune_batch(table, other, prune, cache, d_items, len(items), "f64", out=out)
    with pytest.raises(ValueError):
        B.lnl_prune_batch(table, probes, prune, cache, d_items, len(items), "mixed", out=out)
    with pytest.raises(ValueError):
        B.lnl_prune_batch(table, probes, prune[:4], cache, d_items, len(items), "f64", out=out)
    # jk_batch_set_exact: a slot outside the table is refused
    for slot in (-1, 3):
        with pytest.raises(_lib.JokerLibError):
            table.set_exact(slot)


def _prior():
    return tj.JokerPrior.default(
        P_min=2 * u.day, P_max=2048 * u.day,
        sigma_v=[100 * u.km / u.s, 0.5 * u.km / u.s / u.day], poly_trend=2,
        pars={"K": tj.distributions.Normal(0.0, 30.0, unit=u.km / u.s),
              "s": tj.distributions.LogNormal(-3.0, 1.0, unit=u.km / u.s)})


@pytest.mark.parametrize("precision", ["mixed", "f64"])
def test_server_prune_changes_no_result(setup, precision):
    """The same requests with and without pruning: identical results, and
    the pruned run evaluates fewer rows in full."""
    _, cache = setup
    prior = _prior()
    _, stars = syn.arrivals(10, seed=37, M_range=(12, 120))
    datas = [s.to_rvdata() for s in stars]
    res = {}
    for prune in (False, True):
        srv = S.JokerServer(prior, cache, n_requested=16, precision=precision, max_active=4,
                            growth_factor=8, shape=B.LaunchShape(128, 0, 1024), prune=prune)
        for i, d in enumerate(datas):
            srv.submit(d, np.random.default_rng(500 + i), rid=i)
        while srv.busy():
            srv.step()
        res[prune] = sorted(srv.finished, key=lambda r: r.id)
        if prune:
            probed, pruned = srv.prune_stats()
            assert probed > 0 and 0 < pruned <= probed
    for a, b in zip(res[False], res[True]):
        assert a.error is None and b.error is None
        for k in ("P", "e", "K", "v0", "v1", "ln_likelihood"):
            if k in a.result:
                assert np.array_equal(a.result[k].value, b.result[k].value), (a.id, k)


def test_server_prune_zero_uniform_in_a_later_round(setup):
    """A uniform of exactly 0 in the second round, at a row that the probe
    pruned: the zero-uniform path evaluates it in full, and the result equals
    the FP64 run of the star alone."""
    from tests.test_serving_host import zero_uniform_rng
    _, cache = setup
    prior = _prior()
    _, stars = syn.arrivals(1, seed=41, M_range=(60, 60))
    data = stars[0].to_rvdata()
    # round 1 draws 128 uniforms; round 2 redraws one per row evaluated so far
    k = 128 + 128 + 700
    srv = S.JokerServer(prior, cache, n_requested=16, precision="mixed", max_active=1,
                        growth_factor=8, shape=B.LaunchShape(128, 0, 1024))
    srv.submit(data, zero_uniform_rng(k), rid=0)
    while srv.busy():
        srv.step()
    r = srv.finished[0]
    assert r.error is None, r.error
    assert r.it.iteration >= 2 and r.rows_evaluated > 128 + 700
    assert srv.zero_uniform_events >= 1
    probed, pruned = srv.prune_stats()
    assert 0 < pruned <= probed
    joker = tj.TheJoker(prior, rng=zero_uniform_rng(k), precision="f64")
    alone = joker.iterative_rejection_sample(data, cache, n_requested_samples=16, growth_factor=8)
    assert np.array_equal(r.result["P"].value, alone["P"].value)
    assert np.array_equal(r.result["K"].value, alone["K"].value)


@pytest.mark.parametrize("precision", ["mixed", "f64"])
@pytest.mark.parametrize("M", [16, 50])
def test_pruned_rejection_sample_equals_plain(gpu, precision, M):
    """TheJoker(prune=True).rejection_sample: every row bounded by the probe,
    the rows within reach of the maximum evaluated in FP64 -- the samples
    equal the unpruned run's, also when a uniform of exactly 0 falls on a row
    the probe pruned (that row is then evaluated)."""
    from tests.test_serving_host import zero_uniform_rng
    star = syn.single_star(seed=5, M=M, baseline=1000.0 if M == 16 else 3650.0)
    cols = syn.prior_columns(1 << 20, seed=M, jitter=True)
    cache = D.PriorCache(cols["P"], cols["e"], cols["omega"], cols["M0"], cols["s"])
    prior = tj.JokerPrior.default(P_min=2 * u.day, P_max=2048 * u.day, sigma_v=100 * u.km / u.s,
                                  pars={"K": tj.distributions.Normal(0.0, 30.0, unit=u.km / u.s),
                                        "s": tj.distributions.LogNormal(-3.0, 1.0,
                                                                        unit=u.km / u.s)})
    data = star.to_rvdata()
    # a row far below the maximum whose probe bound is below it too
    ll64 = tj.TheJoker(prior, precision="f64").marginal_ln_likelihood(data, cache)
    far = int(np.flatnonzero(ll64 < ll64.max() - 2000.0)[7])
    for rng_of in (lambda: np.random.default_rng(42), lambda: zero_uniform_rng(far)):
        ref, ll_ref = tj.TheJoker(prior, rng=rng_of(), precision="f64").rejection_sample(
            data, cache, max_posterior_samples=64, in_memory=True, return_all_logprobs=True)
        joker = tj.TheJoker(prior, rng=rng_of(), precision=precision, prune=True)
        # below the single-pass threshold the plain pass runs; lowered here
        # so that the probe path is what is compared at 16 epochs too
        joker.PRUNE_MIN_EPOCHS_SINGLE = min(M, joker.PRUNE_MIN_EPOCHS_SINGLE)
        got, ll_got = joker.rejection_sample(data, cache, max_posterior_samples=64,
                                             in_memory=True, return_all_logprobs=True)
        assert len(got) == len(ref) > 0
        for k in ("P", "e", "K", "v0"):
            assert np.array_equal(got[k].value, ref[k].value), k
        tol = 0.0 if precision == "f64" else 1e-3
        assert np.abs(ll_got - ll_ref).max() <= tol


def test_pruned_rejection_sample_many_equals_plain(gpu):
    stars = syn.many_stars(24, seed=3, M_range=(8, 64))
    cols = syn.prior_columns(1 << 18, seed=44, jitter=True)
    cache = D.PriorCache(cols["P"], cols["e"], cols["omega"], cols["M0"], cols["s"])
    prior = _prior()
    datas = [s.to_rvdata() for s in stars]
    for precision in ("mixed", "f64"):
        res = {}
        for prune in (False, True):
            rngs = [np.random.default_rng(300 + i) for i in range(len(datas))]
            res[prune] = B.rejection_sample_many(datas, prior, cache, rngs,
                                                 max_posterior_samples=32, precision=precision,
                                                 prune=prune)
        for a, b in zip(res[False], res[True]):
            for k in ("P", "K", "v1"):
                assert np.array_equal(a[k].value, b[k].value), k
