"""Pins the oracle's PCG64 restatement bit-exactly against numpy's own stream,
and its rejection step against np.where (likelihood_helpers.py:107-109)."""
import numpy as np
import pytest

from oracle import oracle
from tests._golden import load

PCG = load("pcg64.npz")


def _state(i):
    st = PCG[f"s{i}__state"]
    inc = PCG[f"s{i}__inc"]
    return (int(st[0]) << 64) | int(st[1]), (int(inc[0]) << 64) | int(inc[1])


@pytest.mark.parametrize("i", range(5))
def test_raw_and_uniform_bit_exact(i):
    state, inc = _state(i)
    assert np.array_equal(oracle.pcg64_raw(state, inc, 64), PCG[f"s{i}__raw"])
    u = oracle.pcg64_uniform(state, inc, 3000)
    assert np.array_equal(u, PCG[f"s{i}__uniform"])
    # jump-ahead: outputs 10^7 + 3 .. of the same stream
    u2 = oracle.pcg64_uniform(state, inc, 64, offset=10 ** 7 + 3)
    assert np.array_equal(u2, PCG[f"s{i}__uniform_after_1e7p3"])


def test_live_numpy_stream():
    g = np.random.default_rng(2718)
    st = g.bit_generator.state["state"]
    u = oracle.pcg64_uniform(st["state"], st["inc"], 1000, offset=17)
    g.bit_generator.advance(17)
    assert np.array_equal(u, g.uniform(size=1000))


@pytest.mark.parametrize("label", ["informative", "flat"])
def test_reject_matches_np_where(label):
    acc = load("accept_sets.npz")
    ll = acc[label + "__ll"]
    u = np.random.default_rng(42).uniform(size=ll.size)
    idx, n_pass, mx = oracle.reject(ll, u)
    assert np.array_equal(idx, acc[label + "__idx"])
    assert n_pass == len(acc[label + "__idx"])
    assert mx == ll.max()
    idx2, n_pass2, _ = oracle.reject(ll, u, max_keep=3)
    assert np.array_equal(idx2, acc[label + "__idx"][:3]) and n_pass2 == n_pass


def test_reject_edge_semantics():
    # strict '>' against the uniform
    ll = np.array([0.0, np.log(0.5), np.log(0.25)])
    u = np.array([0.999, 0.5, 0.2])
    idx, n, _ = oracle.reject(ll, u)
    assert list(idx) == [0, 2] and n == 2
    # np.max propagates NaN: nothing passes
    idx, n, mx = oracle.reject(np.array([0.0, np.nan, -1.0]), np.zeros(3))
    assert n == 0 and np.isnan(mx)
    # -inf rows never pass; +inf maximum makes every comparison false
    idx, n, _ = oracle.reject(np.array([-np.inf, 0.0]), np.zeros(2))
    assert list(idx) == [1]
    idx, n, _ = oracle.reject(np.array([np.inf, 0.0]), np.zeros(2))
    assert n == 0
