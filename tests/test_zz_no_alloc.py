"""jk_lnl_reject_f64_noalloc through the C ABI: the scratch lives in a
file-scope device array, so the call allocates nothing.  Its outputs must
equal jk_lnl_reject_f64's bit for bit (lnL, kept indices, pass count,
maximum) on two calls in a row, the second with a smaller input, and sizes
beyond its maximum must be refused before any launch."""
import ctypes

import numpy as np
import pytest
import torch

from thejoker_amd import _lib
from thejoker_amd import device as D
from thejoker_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _star():
    star = syn.single_star()
    mu, lam = np.zeros(2), np.array([900.0, 1e4])
    return D.DeviceStar(star.t, star.rv, star.ivar(), star.trend(), mu, lam, star.t_ref,
                        jitter_mode="physical", device=0)


def _call(lib, fn, dstar, cache, n, max_keep, rng_state, offset=0):
    """One jk_lnl_reject_f64(_noalloc) call over cache rows [0, n); returns
    (lnL, kept indices, n_pass, max) on the host."""
    ll = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
    res = torch.full((2 + max_keep,), -7, dtype=torch.int64, device="cuda")
    p = res.data_ptr()
    state, inc = rng_state
    cols = [ctypes.c_void_p(c.data_ptr()) for c in
            (cache.P, cache.e, cache.omega, cache.M0, cache.s)]
    st = ctypes.c_void_p(torch.cuda.current_stream(0).cuda_stream)
    status = getattr(lib, fn)(dstar.handle, n, *cols, ctypes.c_void_p(0),
                              ctypes.c_void_p(ll.data_ptr()), *D._split(state), *D._split(inc),
                              offset, max_keep, ctypes.c_void_p(p + 16), ctypes.c_void_p(p),
                              ctypes.c_void_p(p + 8), st)
    _lib.check(status, fn)
    torch.cuda.synchronize()
    h = res.cpu().numpy()
    n_pass = int(h[0])
    return (ll.cpu().numpy(), h[2:2 + min(n_pass, max_keep)].copy(), n_pass,
            float(h[1:2].view(np.float64)[0]))


def test_noalloc_equals_allocating_entry_point(gpu):
    lib = _lib.ensure_device(0)
    dstar = _star()
    big, small = (1 << 22) + 1000, 3000
    cols = syn.prior_columns(big, seed=11, jitter=True)
    cache = D.PriorCache(cols["P"], cols["e"], cols["omega"], cols["M0"], cols["s"], device=0)
    g = np.random.default_rng(5)
    for n, max_keep, offset in ((big, 256, 0), (small, 64, 77)):
        state = D.pcg64_state(g)
        want = _call(lib, "jk_lnl_reject_f64", dstar, cache, n, max_keep, state, offset)
        got = _call(lib, "jk_lnl_reject_f64_noalloc", dstar, cache, n, max_keep, state, offset)
        assert np.array_equal(got[0], want[0])
        assert np.array_equal(got[1], want[1])
        assert got[2] == want[2] and got[3] == want[3]
        assert want[2] > 0 and len(want[1]) == min(want[2], max_keep)
        D.pcg64_advance_state(g, n)
    # the second (smaller) call left stale data of the first in the array;
    # a third call at the large size again matches
    state = D.pcg64_state(g)
    want = _call(lib, "jk_lnl_reject_f64", dstar, cache, big, 512, state)
    got = _call(lib, "jk_lnl_reject_f64_noalloc", dstar, cache, big, 512, state)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    assert got[2:] == want[2:]


def test_noalloc_refuses_sizes_beyond_its_maximum(gpu):
    lib = _lib.ensure_device(0)
    dstar = _star()
    one = torch.zeros(1, dtype=torch.float64, device="cuda")
    res = torch.zeros(3, dtype=torch.int64, device="cuda")
    p = ctypes.c_void_p(one.data_ptr())
    r = res.data_ptr()
    # the size check comes before any launch, so the one-element buffers are
    # never indexed
    status = lib.jk_lnl_reject_f64_noalloc(dstar.handle, (1 << 27) + 1, p, p, p, p, p,
                                           ctypes.c_void_p(0), p, 0, 1, 0, 1, 0, 1,
                                           ctypes.c_void_p(r + 16), ctypes.c_void_p(r),
                                           ctypes.c_void_p(r + 8), ctypes.c_void_p(0))
    assert status == 5   # JK_ERANGE
    with pytest.raises(_lib.JokerLibError, match="exceeds the supported maximum"):
        _lib.check(status, "jk_lnl_reject_f64_noalloc")
