"""CJokerHelper: the reference's per-star likelihood operator
(thejoker/src/fast_likelihood.pyx:70-576), as a drop-in Python class over
libjoker's C ABI.

The reference constructs it as ``CJokerHelper(data, prior, trend_M)``
(thejoker.py:87-91) and calls ``batch_marginal_ln_likelihood(chunk)`` with an
array-of-structs chunk of packed prior rows (n, 5) in the order P, e, omega,
M0, s (fast_likelihood.pyx:41).  Here the chunk is split into device
structure-of-arrays columns and evaluated by ``jk_lnl_f64``; the per-row
intermediates of ``test_likelihood_worker`` come from ``jk_debug_single``;
``batch_get_posterior_samples`` takes a and inv(Ainv) from
``jk_linear_posterior`` and draws the linear parameters with the caller's
numpy generator in row order, as fast_likelihood.pyx:529-530 does.

Differences from the reference, by design:

* ``jitter_mode`` (default "reference", the reference's behaviour: the jitter
  s is read but does not enter the likelihood, fast_likelihood.pyx:458) may be
  "physical", C = diag(sigma^2 + s^2).
* ``Lambda`` puts the K variance at index 0 also when v0 offsets are present
  (the reference's loop at fast_likelihood.pyx:230-252 writes a Normal K's
  variance to index n_offsets, where v0's then overwrites it).
* The instance holds no per-sample scratch, so it is reentrant; the B, Binv,
  b, A, Ainv and a attributes are those of the last ``test_likelihood_worker``
  call only (the reference also overwrites them in every batch call).
"""
import numpy as np

from . import device as D
from .distributions import FixedCompanionMass
from .samples import NONLINEAR_PACKED_ORDER

__all__ = ["CJokerHelper"]


def _chunk_array(chunk, what="chunk"):
    """The reference's ``double[:, ::1]`` buffer contract: float64,
    C-contiguous, two-dimensional, at least the five packed columns."""
    a = np.asarray(chunk)
    if a.dtype != np.float64:
        raise ValueError(f"Buffer dtype mismatch, expected 'double' but got '{a.dtype}' ({what})")
    if a.ndim != 2:
        raise ValueError(f"Buffer has wrong number of dimensions (expected 2, got {a.ndim})")
    if not a.flags.c_contiguous:
        raise ValueError("ndarray is not C-contiguous")
    if a.shape[1] < 5:
        raise ValueError(f"{what} needs the packed columns P, e, omega, M0, s; got "
                         f"{a.shape[1]} columns")
    return a


class CJokerHelper:
    """Per-star likelihood operator (fast_likelihood.pyx:70-576).

    Parameters
    ----------
    data : RVData
    prior : JokerPrior
    trend_M : numpy.ndarray (n_times, n_linear - 1)
        The trend columns of the design matrix (likelihood_helpers.py:28-37).
    jitter_mode : {"reference", "physical"}
    device : int
    """

    def __init__(self, data, prior, trend_M, jitter_mode="reference", device=0):
        self.prior = prior
        self.data = data
        self.jitter_mode = jitter_mode
        self.device = device
        rv_unit = data.rv.unit
        self.internal_units = D.linear_internal_units(prior, rv_unit)
        self.packed_order = list(NONLINEAR_PACKED_ORDER)
        self.n_times = len(data)
        self.n_poly = prior.poly_trend
        self.n_offsets = prior.n_offsets
        self.n_linear = 1 + self.n_poly + self.n_offsets
        self.n_pars = len(prior.par_names)
        trend_M = np.asarray(trend_M, dtype=np.float64)
        want = (self.n_times, self.n_linear - 1)
        if trend_M.ndim != 2 or trend_M.shape != want:
            raise ValueError(f"Invalid design matrix shape: {trend_M.shape}, expected: {want}")
        self.trend_M = np.ascontiguousarray(trend_M)
        a = D.star_arrays(data, prior, self.trend_M)
        self.t0 = a["t_ref"]
        self.t = a["t"]
        self.rv = a["rv"]
        self.ivar = a["ivar"]
        self.mu = a["mu"]
        self.Lambda = a["lam"].copy()
        # 0: FixedCompanionMass (Lambda[0] per sample), 1: fixed K variance,
        # as the reference's flag (fast_likelihood.pyx:225-228)
        self.fixed_K_prior = 0 if isinstance(prior.pars["K"], FixedCompanionMass) else 1
        self.sigma_K0, self.P0, self.max_K = a["sigma_K0"], a["P0"], a["max_K"]
        self._star = D.DeviceStar(a["t"], a["rv"], a["ivar"], a["trend"], a["mu"], a["lam"],
                                  a["t_ref"], a["k_prior_mode"], a["sigma_K0"], a["P0"],
                                  a["max_K"], jitter_mode=jitter_mode, device=device, cov=a["cov"])
        n, L = self.n_times, self.n_linear
        self.B = np.zeros((n, n))
        self.Binv = np.zeros((n, n))
        self.b = np.zeros(n)
        self.A = np.zeros((L, L))
        self.Ainv = np.zeros((L, L))
        self.a = np.zeros(L)
        self.M_T = np.zeros((L, n))
        self.failures = (0, 0, 0)

    def __reduce__(self):
        return (CJokerHelper, (self.data, self.prior, np.array(self.trend_M), self.jitter_mode,
                               self.device))

    @property
    def star(self):
        """The device star (DeviceStar) the kernels read."""
        return self._star

    def _cache(self, chunk):
        return D.PriorCache(chunk[:, 0], chunk[:, 1], chunk[:, 2], chunk[:, 3], chunk[:, 4],
                            device=self.device)

    def batch_marginal_ln_likelihood(self, chunk):
        """fast_likelihood.pyx:428-469: the marginal lnL of each packed prior
        row of ``chunk`` (n, 5+) as a float64 array of n values (+inf where
        the L x L system is singular, NaN for invalid rows).  The counts of
        such rows are left in ``self.failures`` = (n_posinf, n_neginf, n_nan)."""
        chunk = _chunk_array(chunk)
        n = chunk.shape[0]
        if n == 0:
            self.failures = (0, 0, 0)
            return np.zeros(0)
        cache = self._cache(chunk)
        ll = D.lnl(self._star, cache, precision="f64")
        self.failures = D.count_failures(ll)
        return D.to_host(ll).copy()

    def batch_get_posterior_samples(self, chunk, n_linear_samples_per, rng):
        """fast_likelihood.pyx:471-545: for each packed row, n_linear_samples_per
        draws of the linear parameters from N(a, inv(Ainv)) with
        ``rng.multivariate_normal`` in row order.  Returns (samples
        (n * k, 5 + n_linear): the row's P, e, omega, M0, s then the linear
        parameters; ll (n * k,): the row's lnL repeated, with Lambda[0] not
        clipped (:519-521))."""
        from .thejoker import mvn_draws

        chunk = _chunk_array(chunk)
        n, k, L = chunk.shape[0], int(n_linear_samples_per), self.n_linear
        samples = np.zeros((n, k, 5 + L))
        ll = np.full((n, k), np.nan)
        if n == 0 or k == 0:
            return samples.reshape(n * k, 5 + L), ll.reshape(n * k)
        cache = self._cache(chunk)
        a, cov, lrow = D.linear_posterior(self._star, cache, np.arange(n))
        samples[:, :, :5] = chunk[:, None, :5]
        samples[:, :, 5:] = mvn_draws(rng, a, cov, k)
        ll[:] = lrow[:, None]
        return samples.reshape(n * k, 5 + L), ll.reshape(n * k)

    def test_likelihood_worker(self, chunk_row):
        """fast_likelihood.pyx:547-576: the lnL of one packed row, leaving the
        intermediates in the public attributes B, Binv, b, A, Ainv, a (and the
        design matrix in M_T).  Lambda[0] is not clipped to max_K^2 here."""
        row = np.asarray(chunk_row)
        if row.dtype != np.float64:
            raise ValueError(f"Buffer dtype mismatch, expected 'double' but got '{row.dtype}'")
        if row.ndim != 1 or row.shape[0] < 5:
            raise ValueError("chunk_row needs the packed values P, e, omega, M0, s")
        out = D.debug_single(self._star, row[:5])
        self.M_T = out["M_T"]
        self.Ainv, self.A, self.a = out["Ainv"], out["A"], out["a"]
        self.b, self.B, self.Binv = out["b"], out["B"], out["Binv"]
        self.Lambda = self.Lambda.copy()
        self.Lambda[0] = out["Lambda0"]
        self.chi2, self.log_det = out["chi2"], out["logdet"]
        return out["ll"]
