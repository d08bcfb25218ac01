"""The per-stream scratch of the rejection calls (jk_init's scratch rule):
host threads calling on one stream with growing sizes each get their own
results, and jk_release_scratch returns the buffers without changing what
later calls compute."""
import threading

import numpy as np
import pytest
import torch

from oracle import oracle
from thejoker_amd import _lib
from thejoker_amd import device as D

pytestmark = pytest.mark.gpu


def _case(seed, n):
    g = np.random.default_rng(seed)
    ll_h = -0.5 * g.exponential(size=n)
    st = np.random.default_rng(seed + 1000).bit_generator.state["state"]
    u = oracle.pcg64_uniform(st["state"], st["inc"], n)
    want, n_want, _ = oracle.reject(ll_h, u)
    return ll_h, (st["state"], st["inc"]), want, n_want


def test_two_threads_one_stream_growing_sizes(gpu):
    _lib.ensure_device(0)
    sizes = [1 << 12, 1 << 16, 1 << 19, (1 << 21) + 7, 1 << 22]
    cases = {(t, k): _case(100 * t + k, n) for t in range(2) for k, n in enumerate(sizes)}
    errors = []
    start = threading.Barrier(2)

    def worker(t):
        try:
            torch.cuda.set_device(0)
            start.wait()
            for rep in range(3):
                for k in range(len(sizes)):
                    ll_h, state, want, n_want = cases[(t, k)]
                    ll = torch.as_tensor(ll_h, device="cuda")
                    idx, n_pass, mx = D.reject(ll, state, 1 << 20)
                    if n_pass != n_want or not np.array_equal(idx, want[:1 << 20]):
                        errors.append((t, rep, k, n_pass, n_want))
        except Exception as exc:  # reported below
            errors.append((t, repr(exc)))

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(2)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=100)
    assert not any(th.is_alive() for th in threads)
    assert not errors, errors


def test_release_scratch_then_reuse(gpu):
    lib = _lib.ensure_device(0)
    ll_h, state, want, n_want = _case(7, (1 << 20) + 3)
    ll = torch.as_tensor(ll_h, device="cuda")
    idx, n_pass, _ = D.reject(ll, state, 4096)
    assert n_pass == n_want and np.array_equal(idx, want[:4096])
    _lib.check(lib.jk_release_scratch(0), "jk_release_scratch")
    _lib.check(lib.jk_release_scratch(0), "jk_release_scratch")   # nothing left: still ok
    idx, n_pass, _ = D.reject(ll, state, 4096)
    assert n_pass == n_want and np.array_equal(idx, want[:4096])
    assert lib.jk_release_scratch(-1) == 1
