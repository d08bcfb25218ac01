"""GPU rejection step (jk_pcg64_uniform / jk_reject through the C ABI):
numpy-identical uniforms and the accepted index set of
likelihood_helpers.py:107-109."""
import numpy as np
import pytest
import torch

from oracle import oracle
from tests._golden import load
from thejoker_amd import device as D

pytestmark = pytest.mark.gpu
PCG = load("pcg64.npz")


def _state(i):
    st = PCG[f"s{i}__state"]
    inc = PCG[f"s{i}__inc"]
    return (int(st[0]) << 64) | int(st[1]), (int(inc[0]) << 64) | int(inc[1])


@pytest.mark.parametrize("i", range(5))
def test_device_uniforms_bit_exact(gpu, i):
    state, inc = _state(i)
    u = D.uniforms((state, inc), 3000).cpu().numpy()
    assert np.array_equal(u, PCG[f"s{i}__uniform"])
    u2 = D.uniforms((state, inc), 64, offset=10 ** 7 + 3).cpu().numpy()
    assert np.array_equal(u2, PCG[f"s{i}__uniform_after_1e7p3"])


@pytest.mark.parametrize("n", [1, 63, 64, 65, 4097, 1 << 20, (1 << 22) + 17])
def test_device_uniforms_long_streams(gpu, n):
    g = np.random.default_rng(n)
    st = g.bit_generator.state["state"]
    u = D.uniforms((st["state"], st["inc"]), n).cpu().numpy()
    assert np.array_equal(u, g.uniform(size=n))


@pytest.mark.parametrize("label", ["informative", "flat"])
def test_reject_golden(gpu, label):
    acc = load("accept_sets.npz")
    ll = torch.as_tensor(acc[label + "__ll"], device="cuda")
    st = np.random.default_rng(42).bit_generator.state["state"]
    idx, n_pass, mx = D.reject(ll, (st["state"], st["inc"]), len(ll))
    assert np.array_equal(idx, acc[label + "__idx"])
    assert n_pass == len(acc[label + "__idx"])
    assert mx == acc[label + "__ll"].max()
    idx3, n3, _ = D.reject(ll, (st["state"], st["inc"]), 3)
    assert np.array_equal(idx3, acc[label + "__idx"][:3]) and n3 == n_pass


@pytest.mark.parametrize("n,scale", [(1 << 20, 1.0), (1 << 22, 3.0), ((1 << 21) + 5, 0.3),
                                     (100, 0.01)])
def test_reject_large_vs_oracle(gpu, n, scale):
    """Many acceptances spread across blocks: ordered compaction, truncation
    at several max_keep values, and a nonzero stream offset."""
    g = np.random.default_rng(n)
    ll_h = -scale * g.exponential(size=n)
    ll = torch.as_tensor(ll_h, device="cuda")
    st = np.random.default_rng(7).bit_generator.state["state"]
    for offset in (0, 12345):
        u = oracle.pcg64_uniform(st["state"], st["inc"], n, offset=offset)
        want, n_want, _ = oracle.reject(ll_h, u)
        for max_keep in (n, 256, 1):
            idx, n_pass, _ = D.reject(ll, (st["state"], st["inc"]), max_keep, offset=offset)
            assert n_pass == n_want
            assert np.array_equal(idx, want[:max_keep])


def test_reject_nan_and_inf(gpu):
    st = np.random.default_rng(1).bit_generator.state["state"]
    s = (st["state"], st["inc"])
    ll = torch.as_tensor(np.array([0.0, np.nan, -1.0] * 1000), device="cuda")
    idx, n, mx = D.reject(ll, s, 10)
    assert n == 0 and np.isnan(mx)
    ll = torch.as_tensor(np.array([np.inf, 0.0, -3.0]), device="cuda")
    idx, n, mx = D.reject(ll, s, 10)
    assert n == 0 and mx == np.inf
    ll = torch.as_tensor(np.array([-np.inf, 0.0]), device="cuda")
    idx, n, _ = D.reject(ll, s, 10)
    assert list(idx) == [1]


@pytest.mark.parametrize("n,dense", [((1 << 20) + 37, False), (70_000, False),
                                     ((1 << 20) + 37, True)])
def test_reject_sparse_rows_and_zero_uniforms(gpu, n, dense):
    """Most rows far below the maximum, where no uniform is drawn, and a few
    within reach: the accepted set equals numpy's exp(ll - max) > u, also
    with a uniform of exactly 0 on a row at -700 (accepted: exp is still
    positive) and on a row at -745.5 (rejected: exp underflows to 0).  With
    dense, a block of 20000 rows all within reach puts its waves on the
    row-by-row walk, and the zero uniforms fall inside it."""
    from tests.test_serving_host import zero_uniform_rng
    g = np.random.default_rng(n)
    ll_h = np.full(n, -5000.0)
    near = g.choice(n, size=n // 200, replace=False)
    ll_h[near] = -g.uniform(0.0, 50.0, size=len(near))
    mid = g.choice(n, size=64, replace=False)
    ll_h[mid] = -g.uniform(40.0, 745.0, size=len(mid))
    lo_blk = 300_001
    if dense:
        ll_h[lo_blk:lo_blk + 20_000] = -g.uniform(0.0, 30.0, size=20_000)
    ll_h[g.integers(n)] = 0.0
    for x, accepted in ((-700.0, True), (-745.5, False)):
        j = int(g.integers(lo_blk, lo_blk + 20_000)) if dense else int(g.integers(n))
        ll_j = ll_h.copy()
        ll_j[j] = x
        rng = zero_uniform_rng(j)
        st = D.pcg64_state(rng)
        u = zero_uniform_rng(j).uniform(size=n)
        assert u[j] == 0.0
        want = np.flatnonzero(np.exp(ll_j - ll_j.max()) > u)
        assert (j in want) == accepted
        idx, n_pass, _ = D.reject(torch.as_tensor(ll_j, device="cuda"), st, n)
        assert n_pass == len(want) and np.array_equal(idx, want)


def test_host_generator_stays_in_sync(gpu):
    """After the device draws n uniforms the host generator is advanced by n:
    the next host draw equals numpy's after uniform(size=n)."""
    g1 = np.random.default_rng(99)
    g2 = np.random.default_rng(99)
    n = 123457
    ll = torch.zeros(n, dtype=torch.float64, device="cuda")
    D.reject(ll, D.pcg64_state(g1), 5)
    D.pcg64_advance_state(g1, n)
    g2.uniform(size=n)
    assert g1.uniform() == g2.uniform()
    assert np.array_equal(g1.normal(size=10), g2.normal(size=10))


@pytest.mark.parametrize("n", [1, 255, 256, 257, 70_001, 1 << 20])
def test_lnl_reject_equals_separate_calls(gpu, n):
    """jk_lnl_reject_f64 (maximum reduced inside the likelihood launch) gives
    the lnL of jk_lnl_f64 and the rejection of jk_reject, bit for bit."""
    from thejoker_amd import synthetic as syn
    st = syn.single_star()
    cols = syn.prior_columns(n + 300, seed=n, jitter=True)
    cols["e"][5 % n] = 1.5                     # one invalid row: NaN propagates
    star = D.DeviceStar(st.t, st.rv, st.ivar(), st.trend(), np.zeros(2), np.array([900.0, 1e4]),
                        st.t_ref)
    cache = D.PriorCache(cols["P"], cols["e"], cols["omega"], cols["M0"], cols["s"])
    for nan_row in (False, True):
        if nan_row and n < 6:
            continue
        start = 0 if nan_row else 300
        g1, g2 = np.random.default_rng(9), np.random.default_rng(9)
        want_ll = D.lnl(star, cache, start=start, n=n)
        want = D.reject(want_ll, D.pcg64_state(g1), 256)
        ll, idx, n_pass, mx = D.lnl_reject(star, cache, D.pcg64_state(g2), 256, start=start, n=n)
        assert np.array_equal(ll.cpu().numpy(), want_ll.cpu().numpy(), equal_nan=True)
        assert np.array_equal(idx, want[0]) and n_pass == want[1]
        assert (np.isnan(mx) and np.isnan(want[2])) or mx == want[2]
        assert np.isnan(mx) == nan_row
    rows = torch.as_tensor(np.random.default_rng(1).permutation(n + 300)[:n], device="cuda")
    want_ll = D.lnl(star, cache, rows=rows)
    want = D.reject(want_ll, D.pcg64_state(np.random.default_rng(3)), 10)
    ll, idx, n_pass, mx = D.lnl_reject(star, cache, D.pcg64_state(np.random.default_rng(3)), 10,
                                       rows=rows)
    assert np.array_equal(ll.cpu().numpy(), want_ll.cpu().numpy(), equal_nan=True)
    assert np.array_equal(idx, want[0]) and n_pass == want[1]


def test_lnl_reject_async_pipelined_equals_sequential(gpu):
    """Calls queued back to back (each collected after the next is queued,
    the stream advanced by n in between) give the sequential results."""
    from thejoker_amd import synthetic as syn
    n = 1 << 18
    st = syn.single_star()
    cols = syn.prior_columns(n, seed=4, jitter=True)
    star = D.DeviceStar(st.t, st.rv, st.ivar(), st.trend(), np.zeros(2), np.array([900.0, 1e4]),
                        st.t_ref)
    cache = D.PriorCache(cols["P"], cols["e"], cols["omega"], cols["M0"], cols["s"])
    g = np.random.default_rng(12)
    want = []
    for _ in range(4):
        _, idx, n_pass, mx = D.lnl_reject(star, cache, D.pcg64_state(g), 64)
        D.pcg64_advance_state(g, n)
        want.append((idx, n_pass, mx))
    g = np.random.default_rng(12)
    ll = torch.empty(n, dtype=torch.float64, device="cuda")
    pending, got = [], []
    for _ in range(4):
        pending.append(D.lnl_reject_async(star, cache, D.pcg64_state(g), 64, out=ll))
        D.pcg64_advance_state(g, n)
        if len(pending) > 1:
            got.append(pending.pop(0).result()[1:])
    got.append(pending.pop(0).result()[1:])
    for (a, b, c), (x, y, z) in zip(want, got):
        assert np.array_equal(a, x) and b == y and c == z


def test_scratch_reuse_across_sizes_and_streams(gpu):
    """The library keeps each stream's scratch between calls and grows it on
    demand: rejection steps of growing, then shrinking, sizes on the default
    stream and on a second stream, interleaved, all equal numpy's decisions."""
    rng = np.random.default_rng(5)
    sizes = [1000, 1 << 20, 333, 1 << 22, 64, 1 << 21]
    side = torch.cuda.Stream()
    for k, n in enumerate(sizes):
        ll = rng.normal(0.0, 3.0, n)
        seed = 100 + k
        want_u = np.random.default_rng(seed).uniform(size=n)
        want = np.where(np.exp(ll - ll.max()) > want_u)[0][:500]
        for stream in (torch.cuda.current_stream(), side):
            with torch.cuda.stream(stream):
                d_ll = torch.as_tensor(ll, device="cuda")
                idx, n_pass, mx = D.reject(d_ll, D.pcg64_state(np.random.default_rng(seed)), 500)
            stream.synchronize()
            assert np.array_equal(idx, want), (n, stream)
            assert mx == ll.max()
