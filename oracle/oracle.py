"""ctypes front end of the CPU oracle (oracle/joker_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py`` -- as the checker, or as the
timed CPU baseline -- and never by the shipped package ``thejoker_amd``.

Each wrapper cites the reference function it restates (paths relative to the
adrn/thejoker tree):

* :func:`kepler_twobody` -- twobody's anomaly solve as called from
  ``thejoker/src/fast_likelihood.pyx:453-455`` (tol 1e-10, maxiter 128).
* :func:`batch_lnl` -- ``CJokerHelper.batch_marginal_ln_likelihood``
  (``fast_likelihood.pyx:428-469``) with ``method="reference"`` (LU +
  Woodbury, ``:255-425``) or ``method="chol"`` (O(N L^2) form).
* :func:`linear_posterior` -- ``likelihood_worker(1)``'s ``a`` and ``A``
  (``fast_likelihood.pyx:394-423``) on the posterior path (``:503-530``).
* :func:`pcg64_uniform` -- numpy ``Generator.uniform`` over PCG64
  (``likelihood_helpers.py:107``).
* :func:`reject` -- ``likelihood_helpers.py:107-109``.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libjoker_oracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)


class _JoStar(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("L", ctypes.c_int),
        ("t", _D),
        ("rv", _D),
        ("ivar", _D),
        ("trend", _D),
        ("mu", _D),
        ("lam", _D),
        ("t0", ctypes.c_double),
        ("k_prior_mode", ctypes.c_int),
        ("sigma_K0", ctypes.c_double),
        ("P0", ctypes.c_double),
        ("max_K", ctypes.c_double),
        ("jitter_mode", ctypes.c_int),
        ("tol", ctypes.c_double),
        ("maxiter", ctypes.c_int),
    ]


def build(force=False):
    """Compile the oracle with its Makefile (gcc)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.jo_kepler_twobody.restype = ctypes.c_double
        L.jo_kepler_twobody.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.jo_kepler_bisect.restype = ctypes.c_double
        L.jo_kepler_bisect.argtypes = [ctypes.c_double, ctypes.c_double]
        L.jo_rv_from_elements.restype = ctypes.c_int
        L.jo_rv_from_elements.argtypes = [_D, _D, ctypes.c_int] + [ctypes.c_double] * 7 + [ctypes.c_int]
        L.jo_lnl_row.restype = ctypes.c_double
        L.jo_lnl_row.argtypes = [ctypes.POINTER(_JoStar), _D, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.jo_linear_posterior.restype = ctypes.c_double
        L.jo_linear_posterior.argtypes = [ctypes.POINTER(_JoStar), _D, _D, _D]
        L.jo_batch_lnl.restype = ctypes.c_int64
        L.jo_batch_lnl.argtypes = [ctypes.POINTER(_JoStar), _D, ctypes.c_int64, ctypes.c_int,
                                   ctypes.c_int, _D, ctypes.POINTER(ctypes.c_uint8)]
        L.jo_pcg64_uniform.restype = None
        L.jo_pcg64_uniform.argtypes = [ctypes.c_uint64] * 5 + [ctypes.c_int64, _D]
        L.jo_pcg64_raw.restype = None
        L.jo_pcg64_raw.argtypes = [ctypes.c_uint64] * 5 + [ctypes.c_int64,
                                                           ctypes.POINTER(ctypes.c_uint64)]
        L.jo_reject.restype = ctypes.c_int64
        L.jo_reject.argtypes = [_D, ctypes.c_int64, _D, ctypes.c_int64,
                                ctypes.POINTER(ctypes.c_int64), _D]
        L.jo_star_sizeof.restype = ctypes.c_int
        if L.jo_star_sizeof() != ctypes.sizeof(_JoStar):
            raise RuntimeError("oracle jo_star layout mismatch")
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D)


class OracleStar:
    """The constants CJokerHelper.__init__ holds (fast_likelihood.pyx:125-253)
    in the oracle's C struct.  Units: days, the data's RV unit."""

    def __init__(self, t, rv, ivar, trend, mu, lam, t_ref, k_prior_mode=0,
                 sigma_K0=0.0, P0=1.0, max_K=np.inf, jitter_mode=0,
                 tol=1e-10, maxiter=128):
        self.t = np.ascontiguousarray(t, dtype=np.float64)
        self.rv = np.ascontiguousarray(rv, dtype=np.float64)
        self.ivar = np.ascontiguousarray(ivar, dtype=np.float64)
        n = self.t.size
        trend = np.asarray(trend, dtype=np.float64).reshape(n, -1)
        self.trend = np.ascontiguousarray(trend)
        self.mu = np.ascontiguousarray(mu, dtype=np.float64)
        self.lam = np.ascontiguousarray(lam, dtype=np.float64)
        L = self.mu.size
        if self.trend.shape[1] != L - 1 or self.lam.size != L:
            raise ValueError("inconsistent star shapes")
        if self.trend.size == 0:
            self._trend_buf = np.zeros(1)
        else:
            self._trend_buf = self.trend
        self.c = _JoStar(n, L, _dp(self.t), _dp(self.rv), _dp(self.ivar), _dp(self._trend_buf),
                         _dp(self.mu), _dp(self.lam), float(t_ref), int(k_prior_mode),
                         float(sigma_K0), float(P0), float(max_K), int(jitter_mode),
                         float(tol), int(maxiter))

    @property
    def n(self):
        return self.t.size

    @property
    def L(self):
        return self.mu.size


def kepler_twobody(M, e, tol=1e-10, maxiter=128):
    it = ctypes.c_int(0)
    E = lib().jo_kepler_twobody(float(M), float(e), float(tol), int(maxiter), ctypes.byref(it))
    return E, it.value


def kepler_bisect(M, e):
    return lib().jo_kepler_bisect(float(M), float(e))


def rv_unit(t, P, e, omega, M0, t0, tol=1e-10, maxiter=128):
    t = np.ascontiguousarray(t, dtype=np.float64)
    out = np.empty_like(t)
    lib().jo_rv_from_elements(_dp(t), _dp(out), t.size, P, 1.0, e, omega, M0, t0, tol, maxiter)
    return out


def pack_rows(P, e, omega, M0, s):
    return np.ascontiguousarray(np.stack([np.asarray(x, dtype=np.float64) for x in
                                          (P, e, omega, M0, s)], axis=1))


def batch_lnl(star, chunk, method="reference", n_threads=1, return_capped=False):
    """chunk: (n, 5) AoS float64 rows (P, e, omega, M0, s).  With
    return_capped, also a bool mask of the rows in which an anomaly solve hit
    the 128-step cap (their value is the reference iteration's last iterate:
    parity unpinned)."""
    chunk = np.ascontiguousarray(chunk, dtype=np.float64)
    n = chunk.shape[0]
    ll = np.full(n, np.nan)
    capped = np.zeros(n, dtype=np.uint8)
    m = {"reference": 0, "chol": 1}[method]
    lib().jo_batch_lnl(ctypes.byref(star.c), _dp(chunk), n, m, int(n_threads), _dp(ll),
                       capped.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return (ll, capped.astype(bool)) if return_capped else ll


def linear_posterior(star, row):
    row = np.ascontiguousarray(row, dtype=np.float64)
    a = np.zeros(star.L)
    A = np.zeros((star.L, star.L))
    ll = lib().jo_linear_posterior(ctypes.byref(star.c), _dp(row), _dp(a), _dp(A))
    return ll, a, A


def _split128(x):
    x = int(x)
    return (x >> 64) & 0xFFFFFFFFFFFFFFFF, x & 0xFFFFFFFFFFFFFFFF


def pcg64_uniform(state, inc, n, offset=0):
    sh, sl = _split128(state)
    ih, il = _split128(inc)
    out = np.empty(int(n))
    lib().jo_pcg64_uniform(sh, sl, ih, il, int(offset), int(n), _dp(out))
    return out


def pcg64_raw(state, inc, n, offset=0):
    sh, sl = _split128(state)
    ih, il = _split128(inc)
    out = np.empty(int(n), dtype=np.uint64)
    lib().jo_pcg64_raw(sh, sl, ih, il, int(offset), int(n),
                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return out


def reject(ll, u, max_keep=None):
    ll = np.ascontiguousarray(ll, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    n = ll.size
    if max_keep is None:
        max_keep = n
    idx = np.empty(max(int(max_keep), 1), dtype=np.int64)
    mx = ctypes.c_double(0.0)
    n_pass = lib().jo_reject(_dp(ll), n, _dp(u), int(max_keep),
                             idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(mx))
    return idx[:min(n_pass, max_keep)].copy(), int(n_pass), mx.value
