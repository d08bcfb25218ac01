"""Many-star launches (jk_lnl_batch, jk_reject_batch, jk_polish_f64) against
the single-star entry points and the oracle: same bits alone and in a mix,
independent of work-item order and launch shape."""
import numpy as np
import pytest
import torch

from oracle import oracle
from tests._golden import rel_err
from thejoker_amd import batch as B
from thejoker_amd import device as D
from thejoker_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _stars(n_stars=24, seed=3, L=3, mode="physical"):
    out = []
    for st in syn.many_stars(n_stars, seed=seed):
        trend = st.trend(poly_trend=L - 1)
        lam = np.array([900.0, 1e4, 0.25, 1e-4][:L])
        out.append((st, D.DeviceStar(st.t, st.rv, st.ivar(), trend, np.zeros(L), lam, st.t_ref,
                                     jitter_mode=mode)))
    return out


@pytest.fixture(scope="module")
def setup(gpu):
    stars = _stars()
    cols = syn.prior_columns(1 << 14, seed=9, jitter=True)
    cache = D.PriorCache(cols["P"], cols["e"], cols["omega"], cols["M0"], cols["s"])
    return stars, cache, cols


@pytest.mark.parametrize("precision", ["f64", "mixed"])
@pytest.mark.parametrize("shape", [B.LaunchShape(256, 0, 2048), B.LaunchShape(64, 37, 777),
                                   B.LaunchShape(128, 1, 4096)])
def test_batch_equals_single_star_launches(setup, precision, shape):
    stars, cache, _ = setup
    batch = B.StarBatch([d for _, d in stars])
    n = len(cache)
    items, off, total = B.work_items([n] * len(stars), shape.tile)
    rng = np.random.default_rng(0)
    items = items[rng.permutation(len(items))]
    d_items = B.upload_items(items)
    ll = B.lnl_batch(batch, cache, d_items, len(items), total, precision, shape).cpu().numpy()
    for s, (_, dstar) in enumerate(stars):
        single = D.lnl(dstar, cache, precision=precision).cpu().numpy()
        assert np.array_equal(ll[off[s]:off[s] + n], single), s


@pytest.mark.parametrize("mode", ["physical", "reference"])
@pytest.mark.parametrize("first", ["ones", "scaled"])
def test_constant_first_column_bit_identical(gpu, mode, first):
    """A single-star FP64 launch folds a first trend column of exact ones
    (the constant offset) into its sums; the batched launch reads the column.
    Both give the same lnL bit for bit, for L = 2 and 3, and so does a star
    whose first column is not all ones (the single-star launch reads it)."""
    cols = syn.prior_columns(1 << 13, seed=21, jitter=True)
    cache = D.PriorCache(cols["P"], cols["e"], cols["omega"], cols["M0"], cols["s"])
    for L in (2, 3):
        devs = []
        for st in syn.many_stars(4, seed=8):
            trend = st.trend(poly_trend=L - 1)
            assert np.all(trend[:, 0] == 1.0)
            if first == "scaled":
                trend = trend.copy()
                trend[1:, 0] = 1.5
            lam = np.array([900.0, 1e4, 0.25][:L])
            devs.append(D.DeviceStar(st.t, st.rv, st.ivar(), trend, np.zeros(L), lam, st.t_ref,
                                     jitter_mode=mode))
        batch = B.StarBatch(devs)
        n = len(cache)
        items, off, total = B.work_items([n] * len(devs), 2048)
        ll = B.lnl_batch(batch, cache, B.upload_items(items), len(items), total,
                         "f64").cpu().numpy()
        for s, dstar in enumerate(devs):
            single = D.lnl(dstar, cache).cpu().numpy()
            assert np.array_equal(ll[off[s]:off[s] + n], single, equal_nan=True), (L, s)


def test_batch_vs_oracle(setup):
    stars, cache, cols = setup
    batch = B.StarBatch([d for _, d in stars])
    n = len(cache)
    items, off, total = B.work_items([n] * len(stars), 1024)
    ll = B.lnl_batch(batch, cache, B.upload_items(items), len(items), total, "f64").cpu().numpy()
    rows = oracle.pack_rows(cols["P"], cols["e"], cols["omega"], cols["M0"], cols["s"])
    for s in (0, 5, 11):
        st, dstar = stars[s]
        ost = oracle.OracleStar(st.t, st.rv, st.ivar(), st.trend(poly_trend=2), np.zeros(3),
                                np.array([900.0, 1e4, 0.25]), st.t_ref, jitter_mode=1)
        want, capped = oracle.batch_lnl(ost, rows, method="chol", n_threads=16, return_capped=True)
        assert rel_err(ll[off[s]:off[s] + n][~capped], want[~capped]).max() < 1e-10


def test_partial_row_ranges(setup):
    """Stars evaluated on different row ranges (the serving pattern)."""
    stars, cache, _ = setup
    batch = B.StarBatch([d for _, d in stars[:5]])
    begins = np.array([0, 100, 4000, 9999, 16000])
    sizes = np.array([50, 3000, 1, 6000, 384])
    items, off, total = B.work_items(sizes, 512, row_begin=begins)
    ll = B.lnl_batch(batch, cache, B.upload_items(items), len(items), total, "f64").cpu().numpy()
    for s in range(5):
        single = D.lnl(stars[s][1], cache, start=int(begins[s]), n=int(sizes[s])).cpu().numpy()
        assert np.array_equal(ll[off[s]:off[s] + sizes[s]], single)


def test_reject_batch_equals_single(gpu):
    g = np.random.default_rng(1)
    sizes = np.array([1, 63, 64, 65, 5000, 300000, 1 << 18, 77])
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    ll_h = np.concatenate([-g.exponential(scale=sc, size=n) for n, sc in
                           zip(sizes, [1, 0.5, 2, 0.01, 3, 1, 0.2, 10])])
    ll_h[offsets[3] + 5] = np.nan
    ll = torch.as_tensor(ll_h, device="cuda")
    states = [D.pcg64_state(np.random.default_rng(100 + i)) for i in range(len(sizes))]
    stream_off = np.array([0, 5, 0, 123456, 0, 7, 1 << 30, 3], dtype=np.uint64)
    for max_keep in (1, 256, 1 << 20):
        idx, n_pass, mx = B.reject_batch(ll, offsets, sizes, states, max_keep,
                                         stream_offsets=stream_off)
        for i in range(len(sizes)):
            seg = ll[offsets[i]:offsets[i] + sizes[i]]
            want, npw, mxw = D.reject(seg, states[i], max_keep, offset=int(stream_off[i]))
            assert np.array_equal(idx[i], want), i
            assert n_pass[i] == npw
            assert (np.isnan(mx[i]) and np.isnan(mxw)) or mx[i] == mxw


def test_polish_makes_live_set_fp64(setup):
    stars, cache, _ = setup
    for _, dstar in stars[:4]:
        f64 = D.lnl(dstar, cache, precision="f64")
        mix = D.lnl(dstar, cache, precision="mixed")
        m0 = mix.cpu().numpy()
        live = m0 >= m0.max() - 40
        cnt = D.polish(dstar, cache, mix)
        f, m = f64.cpu().numpy(), mix.cpu().numpy()
        assert int(cnt.item()) == live.sum() > 0
        assert np.array_equal(m[live], f[live])
        assert m.max() == f.max()


def test_prior_fill_distribution_and_determinism(gpu):
    from scipy import stats
    c = D.PriorCache.generate(1 << 20, seed=7, P_min=1.0, P_max=16384.0, ln_s=(-3.0, 1.0),
                              fp32_copy=True)
    c2 = D.PriorCache.generate(1 << 16, seed=7, P_min=1.0, P_max=16384.0, ln_s=(-3.0, 1.0))
    P = c.P.cpu().numpy()
    e = c.e.cpu().numpy()
    assert np.array_equal(P[:1 << 16], c2.P.cpu().numpy())      # row r depends on (seed, r)
    assert P.min() >= 1.0 and P.max() <= 16384.0
    assert stats.kstest(np.log(P), stats.uniform(0, np.log(16384.0)).cdf).pvalue > 1e-3
    assert stats.kstest(e, stats.beta(0.867, 3.03).cdf).pvalue > 1e-3
    assert stats.kstest(c.omega.cpu().numpy(), stats.uniform(0, 2 * np.pi).cdf).pvalue > 1e-3
    ls = np.log(c.s.cpu().numpy())
    assert stats.kstest(ls, stats.norm(-3, 1).cdf).pvalue > 1e-3
    assert np.array_equal(c.fp32[0].cpu().numpy(), e.astype(np.float32))
    z = D.PriorCache.generate(1000, seed=1, ln_s=None)
    assert np.all(z.s.cpu().numpy() == 0.0)


def test_polish_first_skips_settled_rows(setup):
    """Polishing rows [k, n) against the maximum of all n rows, after rows
    [0, k) were polished on their own, gives the full polish."""
    stars, cache, _ = setup
    n = len(cache)
    for _, dstar in stars[:4]:
        full = D.lnl(dstar, cache, precision="mixed")
        D.polish(dstar, cache, full)
        for k in (1, 4097, n // 2, n - 1):
            inc = D.lnl(dstar, cache, precision="mixed")
            D.polish(dstar, cache, inc[:k])
            before = inc[:k].clone()
            cnt = D.polish(dstar, cache, inc, first=k)
            assert torch.equal(inc[:k], before)
            f, g = full.cpu().numpy(), inc.cpu().numpy()
            live = f >= f.max() - 40
            assert np.array_equal(g[live], f[live]), k
            assert g.max() == f.max()
            assert int(cnt.item()) <= n - k


def test_polish_batch_first(setup):
    stars, cache, _ = setup
    batch = B.StarBatch([d for _, d in stars[:6]])
    n = len(cache)
    items, off, total = B.work_items([n] * 6, 1024)
    ll = B.lnl_batch(batch, cache, B.upload_items(items), len(items), total, "mixed")
    first = np.array([0, 1, 5000, n // 3, n - 64, n - 1])
    for s in range(6):
        k = int(first[s])
        if k:
            D.polish(stars[s][1], cache, ll[off[s]:off[s] + k])
    B.polish_batch(batch, cache, ll, off, np.full(6, n), first=first)
    for s, (_, dstar) in enumerate(stars[:6]):
        ref = D.lnl(dstar, cache, precision="mixed")
        D.polish(dstar, cache, ref)
        r, g = ref.cpu().numpy(), ll[off[s]:off[s] + n].cpu().numpy()
        live = r >= r.max() - 40
        assert np.array_equal(g[live], r[live]), s


def test_slot_table_equals_single_star_launches(setup):
    """A slot table filled in scrambled order, then with one slot refilled,
    gives each star's single-star lnL."""
    stars, cache, _ = setup
    table = B.StarBatch.slots(8, 3)
    assign = {5: 0, 0: 1, 7: 2, 2: 3}
    for slot, s in assign.items():
        table.set(slot, stars[s][1])
    n = 5000
    for _ in range(2):
        slots = np.array(sorted(assign))
        items, off, total = B.work_items([n] * len(slots), 512, row_begin=[100] * len(slots),
                                         stars=slots)
        ll = B.lnl_batch(table, cache, B.upload_items(items), len(items), total, "f64")
        for k, slot in enumerate(slots):
            single = D.lnl(stars[assign[slot]][1], cache, start=100, n=n).cpu().numpy()
            assert np.array_equal(ll[off[k]:off[k] + n].cpu().numpy(), single), slot
        assign[7] = 9
        table.set(7, stars[9][1])


def test_posterior_inputs_many_equals_single(setup):
    """The one-block, one-copy posterior inputs of many stars equal the
    single-star ones, including a star with no accepted rows."""
    from thejoker_amd.thejoker import TheJoker
    stars, cache, _ = setup
    g = np.random.default_rng(4)
    sel = [stars[i][1] for i in (0, 3, 7, 8)]
    rows = [g.choice(len(cache), size=k, replace=False) for k in (5, 0, 256, 1)]
    many = D.posterior_inputs_many(sel, cache, rows)
    for st, r, got in zip(sel, rows, many):
        want = TheJoker._posterior_inputs(st, cache, r)
        for a, b in zip(got, want):
            assert a.shape == b.shape and np.array_equal(a, b)


def _bound_check(dstar, cache, n, cols=None, label=""):
    mix = torch.empty(n, dtype=torch.float64, device="cuda")
    bnd = torch.empty(n, dtype=torch.float32, device="cuda")
    D.lnl(dstar, cache, n=n, precision="mixed", out=mix, bound=bnd)
    f64 = D.lnl(dstar, cache, n=n, precision="f64").cpu().numpy()
    m, b = mix.cpu().numpy(), bnd.cpu().numpy().astype(np.float64)
    ok = np.isfinite(f64)
    d = np.abs(m - f64)
    r = np.where(ok, d / b, 0.0)
    k = int(np.argmax(r))
    info = f"{label} row {k}: ratio {r[k]:.3g} lnL64 {f64[k]!r} mixed {m[k]!r} bound {b[k]!r}"
    if cols is not None:
        info += " " + " ".join(f"{c}={cols[c][k]!r}" for c in ("P", "e", "omega", "M0", "s"))
    return d[ok], b[ok], info


def test_mixed_error_bound_holds(setup):
    """|lnL(mixed) - lnL(FP64)| <= the kernel's per-row bound, on every row of
    binaries and constant-velocity stars (L = 3, physical jitter)."""
    stars, cache, cols = setup
    ratios = []
    for i, (_, dstar) in enumerate(stars):
        d, b, info = _bound_check(dstar, cache, len(cache), cols, f"star {i}")
        assert np.all(d <= b), info
        ratios.append((d / b).max())
    assert max(ratios) < 0.5


def test_mixed_error_bound_holds_high_eccentricity(gpu):
    """The same near e -> 1, where the FP32 solve degrades the most."""
    from thejoker_amd import synthetic as syn
    st = syn.single_star()
    cols = syn.prior_columns(1 << 16, seed=77, jitter=True)
    g = np.random.default_rng(2)
    cols["e"] = 1.0 - 10.0 ** g.uniform(-6.5, -0.5, len(cols["e"]))
    cache = D.PriorCache(cols["P"], cols["e"], cols["omega"], cols["M0"], cols["s"])
    dstar = D.DeviceStar(st.t, st.rv, st.ivar(), st.trend(), np.zeros(2), np.array([900.0, 1e4]),
                         st.t_ref)
    d, b, info = _bound_check(dstar, cache, len(cache), cols, "high e")
    assert np.all(d <= b), info


@pytest.mark.parametrize("max_keep", [64, 1 << 20])
def test_settle_then_reject_equals_fp64(setup, max_keep):
    """lnl_batch (mixed, with bounds) + settle_batch + reject_batch accepts
    exactly what the FP64 buffer accepts, for every star, re-evaluating a
    small fraction of the rows."""
    stars, cache, _ = setup
    batch = B.StarBatch([d for _, d in stars])
    n = len(cache)
    items, off, total = B.work_items([n] * len(stars), 1024)
    d_items = B.upload_items(items)
    ll64 = B.lnl_batch(batch, cache, d_items, len(items), total, "f64")
    bound = torch.empty(total, dtype=torch.float32, device="cuda")
    llm = B.lnl_batch(batch, cache, d_items, len(items), total, "mixed", bound=bound)
    states = [D.pcg64_state(np.random.default_rng(900 + i)) for i in range(len(stars))]
    sizes = np.full(len(stars), n)
    cnt = B.settle_batch(batch, cache, llm, bound, off, sizes, states).cpu().numpy()
    want, npw, mxw = B.reject_batch(ll64, off, sizes, states, max_keep)
    got, npg, mxg = B.reject_batch(llm, off, sizes, states, max_keep)
    for i in range(len(stars)):
        assert np.array_equal(got[i], want[i]), i
    assert np.array_equal(npg, npw) and np.array_equal(mxg, mxw)
    assert cnt.sum() < 0.2 * total
    # settled rows carry their FP64 values and a zero bound
    b = bound.cpu().numpy()
    assert np.array_equal(llm.cpu().numpy()[b == 0], ll64.cpu().numpy()[b == 0])


def test_settle_live_grows_and_decides_like_fp64_for_any_stream(setup):
    """Segments grown in rounds of different sizes, each round's new rows
    settled with jk_settle_live_batch (floor: the maximum the previous
    round's rejection step found): after every round the rejection step
    accepts exactly the FP64 set under several unrelated streams, and the
    maximum is the FP64 maximum."""
    stars, cache, _ = setup
    batch = B.StarBatch([d for _, d in stars])
    n = len(cache)
    items, off, total = B.work_items([n] * len(stars), 1024)
    d_items = B.upload_items(items)
    ll64 = B.lnl_batch(batch, cache, d_items, len(items), total, "f64")
    bound = torch.empty(total, dtype=torch.float32, device="cuda")
    llm = B.lnl_batch(batch, cache, d_items, len(items), total, "mixed", bound=bound)
    g = np.random.default_rng(5)
    # per star: round ends 0 < c1 < c2 < n (c1 may be tiny)
    ends = np.sort(g.integers(1, n, size=(len(stars), 2)), axis=1)
    ends = np.concatenate([ends, np.full((len(stars), 1), n)], axis=1)
    first = np.zeros(len(stars), dtype=np.int64)
    floors = np.full(len(stars), -np.inf)
    settled = 0
    for k in range(3):
        sizes = ends[:, k].astype(np.int64)
        cnt = B.settle_live_batch(batch, cache, llm, bound, off, sizes, first, floors)
        settled += int(cnt.sum())
        for seed in (11 * k + 1, 11 * k + 2, 11 * k + 3):
            states = [D.pcg64_state(np.random.default_rng(seed * 100 + i))
                      for i in range(len(stars))]
            want, npw, mxw = B.reject_batch(ll64, off, sizes, states, 128)
            got, npg, mxg = B.reject_batch(llm, off, sizes, states, 128)
            for i in range(len(stars)):
                assert np.array_equal(got[i], want[i]), (k, seed, i)
            assert np.array_equal(npg, npw) and np.array_equal(mxg, mxw), (k, seed)
        first, floors = sizes, mxg
    b = bound.cpu().numpy()
    assert np.array_equal(llm.cpu().numpy()[b == 0], ll64.cpu().numpy()[b == 0])
    print(f"settled {settled} of {total} rows", flush=True)


@pytest.mark.parametrize("precision", ["mixed", "f64"])
def test_step_graph_replays_equal_eager(setup, precision):
    """The graph-captured step (persistent lnL launch, settling, rejection on a
    device-resident segment plan) replayed with new streams, with eager work
    and waits between replays, gives the eager path's results every time."""
    stars, cache, _ = setup
    batch = B.StarBatch([d for _, d in stars])
    n = len(cache)
    shape = B.LaunchShape(256, 512, 1024)
    items, off, total = B.work_items([n] * len(stars), shape.tile)
    items = items[B.heaviest_first(items, batch.M)]
    d_items = B.upload_items(items)
    sizes = np.full(len(stars), n)

    def states(k):
        return [D.pcg64_state(np.random.default_rng(1000 * k + i)) for i in range(len(stars))]

    graph = B.StepGraph(batch, cache, d_items, len(items), total, off, sizes, states(0), 128,
                        precision, shape)
    for k in range(1, 4):
        got, npg, mxg, _ = graph.run(states(k))
        bound = torch.empty(total, dtype=torch.float32, device="cuda") if precision == "mixed" else None
        ll = B.lnl_batch(batch, cache, d_items, len(items), total, precision, shape, bound=bound)
        if precision == "mixed":
            B.settle_batch(batch, cache, ll, bound, off, sizes, states(k))
        want, npw, mxw = B.reject_batch(ll, off, sizes, states(k), 128)
        torch.cuda.synchronize()
        for i in range(len(stars)):
            assert np.array_equal(got[i], want[i]), (k, i)
        assert np.array_equal(npg, npw) and np.array_equal(mxg, mxw)
    graph.plan.close()


def test_live_settle_and_zero_check_reject_bad_arguments(setup):
    """jk_settle_live_batch refuses a margin below 40 and first > n;
    jk_reject_batch_zero_check refuses a missing flag array."""
    import ctypes
    from thejoker_amd import _lib
    stars, cache, _ = setup
    batch = B.StarBatch([d for _, d in stars[:2]])
    n = 1024
    ll = torch.zeros(2 * n, dtype=torch.float64, device="cuda")
    bound = torch.ones(2 * n, dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.JokerLibError):
        B.settle_live_batch(batch, cache, ll, bound, [0, n], [n, n], [0, 0], [-np.inf] * 2,
                            margin=30.0)
    with pytest.raises(_lib.JokerLibError):
        B.settle_live_batch(batch, cache, ll, bound, [0, n], [n, n], [0, n + 1], [-np.inf] * 2)
    lib = _lib.ensure_device(0)
    segs = np.zeros(1, dtype=B.SEGMENT_DTYPE)
    segs["n"] = n
    out = torch.zeros(4, dtype=torch.int64, device="cuda")
    p = out.data_ptr()
    rc = lib.jk_reject_batch_zero_check(ctypes.c_void_p(ll.data_ptr()), 1,
                                        segs.ctypes.data_as(ctypes.c_void_p), None,
                                        ctypes.c_void_p(p), ctypes.c_void_p(p + 8), None, None)
    assert rc != 0
    # a first = n segment has no new rows: nothing settled, count 0
    cnt = B.settle_live_batch(batch, cache, ll, bound, [0, n], [n, n], [n, n], [-np.inf] * 2)
    assert cnt.cpu().tolist() == [0, 0]


@pytest.mark.parametrize("max_keep", [1, 64, 1 << 20])
def test_settle_reject_equals_settle_then_reject(setup, max_keep):
    """jk_settle_reject_batch (one uniform per row, open rows decided after
    their FP64 re-evaluation) gives the accepted sets, pass counts, maxima,
    settled counts and settled buffer of jk_settle_batch + jk_reject_batch,
    and so the FP64 sets."""
    stars, cache, _ = setup
    batch = B.StarBatch([d for _, d in stars])
    n = len(cache)
    items, off, total = B.work_items([n] * len(stars), 1024)
    d_items = B.upload_items(items)
    ll64 = B.lnl_batch(batch, cache, d_items, len(items), total, "f64")
    b1 = torch.empty(total, dtype=torch.float32, device="cuda")
    l1 = B.lnl_batch(batch, cache, d_items, len(items), total, "mixed", bound=b1)
    b2, l2 = b1.clone(), l1.clone()
    sizes = np.full(len(stars), n)
    states = [D.pcg64_state(np.random.default_rng(77 + i)) for i in range(len(stars))]
    cnt = B.settle_batch(batch, cache, l1, b1, off, sizes, states).cpu().numpy()
    want, npw, mxw = B.reject_batch(l1, off, sizes, states, max_keep)
    got, npg, mxg, sett = B.settle_reject_batch(batch, cache, l2, b2, off, sizes, states, max_keep)
    f64, np64, _ = B.reject_batch(ll64, off, sizes, states, max_keep)
    for i in range(len(stars)):
        assert np.array_equal(got[i], want[i]), i
        assert np.array_equal(got[i], f64[i]), i
    assert np.array_equal(npg, npw) and np.array_equal(npg, np64)
    assert np.array_equal(mxg, mxw) and np.array_equal(sett, cnt)
    assert torch.equal(l1, l2) and torch.equal(b1, b2)
    # every row within 40 of the maximum is within 1e-3 of FP64 after settling
    a, m = ll64.cpu().numpy(), l2.cpu().numpy()
    for i in range(len(stars)):
        sl = slice(off[i], off[i] + n)
        live = a[sl] >= a[sl].max() - 40
        assert np.abs(a[sl][live] - m[sl][live]).max() <= 1e-3, i


def test_posterior_table_equals_per_star(setup):
    """jk_linear_posterior_batch (one launch over a star table) gives the
    per-star jk_linear_posterior results bit for bit, in any slot order."""
    stars, cache, _ = setup
    batch = B.StarBatch([d for _, d in stars])
    g = np.random.default_rng(4)
    pick = [3, 0, 17, 9]
    rows_list = [np.sort(g.choice(len(cache), k, replace=False)) for k in (5, 64, 1, 200)]
    want = D.posterior_inputs_many([stars[i][1] for i in pick], cache, rows_list)
    got = D.posterior_inputs_table(batch, pick, cache, rows_list)
    for (pw, aw, cw), (pg, ag, cg) in zip(want, got):
        assert np.array_equal(pw, pg) and np.array_equal(aw, ag) and np.array_equal(cw, cg)


def test_item_uploader_equals_pageable_upload(setup):
    """ItemUploader (pinned host buffer + kept device buffer): the device bytes
    equal the items on every call, through growth and shrinking, and a
    launch on the uploaded items equals one on upload_items' copy."""
    stars, cache, _ = setup
    batch = B.StarBatch([d for _, d in stars])
    up = B.ItemUploader()
    n = len(cache)
    rng = np.random.default_rng(4)
    for k, n_rows in enumerate([1000, 40_000, n, 300, 7 * n // 8]):
        sel = rng.choice(len(stars), size=1 + k % 3 + k, replace=True)
        items, off, total = B.work_items([n_rows] * len(sel), 512, stars=sel)
        items = items[rng.permutation(len(items))]
        d = up(items)
        assert np.array_equal(d.cpu().numpy(), np.frombuffer(items.tobytes(), np.uint8))
        got = B.lnl_batch(batch, cache, d, len(items), total, "f64").cpu().numpy()
        ref = B.lnl_batch(batch, cache, B.upload_items(items), len(items), total,
                          "f64").cpu().numpy()
        assert np.array_equal(got, ref, equal_nan=True), k
