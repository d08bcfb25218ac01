"""Device-side objects of the likelihood path.

* :class:`DeviceStar` -- the per-star constants of CJokerHelper.__init__
  (fast_likelihood.pyx:125-253) held on the GPU (a ``jk_star``).
* :class:`PriorCache` -- the prior samples as device-resident
  structure-of-arrays columns: an FP64 master (P, e, omega, M0, s) read by
  both precisions, plus an optional FP32 copy of e, omega, M0, s
  (``fp32_copy=True``) that no kernel reads.  Replaces the
  reference's HDF5 temp-file transport (utils.py:248-302,
  multiproc_helpers.py:63-120).
* :func:`lnl`, :func:`reject`, :func:`linear_posterior` -- thin wrappers of
  the C ABI on torch's current HIP stream.

torch is used for device memory and streams only; every arithmetic step runs
in libjoker's HIP kernels.
"""
import ctypes

import numpy as np

from . import _lib
from . import units as u
from .samples import NONLINEAR_INTERNAL_UNITS, NONLINEAR_PACKED_ORDER

__all__ = ["DeviceStar", "PriorCache", "star_arrays", "probe_star", "probe_epochs", "probe_epochs_residual", "probe_size", "trend_signal", "lnl", "lnl_reject", "lnl_reject_async",
           "PendingReject", "polish", "reject",
           "posterior_inputs_many", "posterior_inputs_table", "to_host",
           "linear_posterior", "uniforms", "debug_single", "count_failures",
           "pcg64_state", "pcg64_advance_state", "JITTER_MODES"]

JITTER_MODES = {"reference": _lib.JK_JITTER_REFERENCE, "physical": _lib.JK_JITTER_PHYSICAL}

_MASK64 = (1 << 64) - 1
_MASK128 = (1 << 128) - 1
_PCG_MULT = 0x2360ED051FC65DA44385DF649FCCF645


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.JokerLibError("no HIP device visible to torch: the likelihood path runs only "
                                 "on a gfx950 GPU (there is no CPU fallback)")
    return torch


def _dptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _hp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _stream(device):
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


# ---------------------------------------------------------------------------
# numpy PCG64 state handling (the bit generator of np.random.default_rng).
# ---------------------------------------------------------------------------
def pcg64_state(rng):
    """(state, inc) of a numpy Generator backed by PCG64."""
    bg = rng.bit_generator
    st = bg.state
    if st.get("bit_generator") != "PCG64":
        raise TypeError("device rejection sampling reproduces numpy's PCG64 stream; got a "
                        f"{st.get('bit_generator')} generator (use np.random.default_rng)")
    return int(st["state"]["state"]), int(st["state"]["inc"])


def _lcg_jump(state, inc, k):
    acc_mult, acc_plus = 1, 0
    cur_mult, cur_plus = _PCG_MULT, inc
    while k > 0:
        if k & 1:
            acc_mult = (acc_mult * cur_mult) & _MASK128
            acc_plus = (acc_plus * cur_mult + cur_plus) & _MASK128
        cur_plus = ((cur_mult + 1) * cur_plus) & _MASK128
        cur_mult = (cur_mult * cur_mult) & _MASK128
        k >>= 1
    return (acc_mult * state + acc_plus) & _MASK128


def pcg64_advance_state(rng, n):
    """Advance `rng` by n 64-bit draws, exactly as n calls of
    rng.uniform() would (the buffered 32-bit half, if any, is left alone)."""
    bg = rng.bit_generator
    st = bg.state
    state, inc = int(st["state"]["state"]), int(st["state"]["inc"])
    st["state"]["state"] = _lcg_jump(state, inc, int(n))
    bg.state = st


def _split(x):
    return (x >> 64) & _MASK64, x & _MASK64


# ---------------------------------------------------------------------------
class DeviceStar:
    """One star's constants on the device.

    Parameters mirror CJokerHelper.__init__ (fast_likelihood.pyx:125-253):
    t, rv, ivar (n,), trend (n, L-1), mu and lam (L,) in M_T row order
    (K, v0, offsets..., v1, ...), t_ref; k_prior_mode/sigma_K0/P0/max_K for
    the FixedCompanionMass K prior; jitter_mode "physical" or "reference".
    cov: instead of ivar, a full (n, n) data covariance (jk_star_create_cov;
    n <= 64, FP64 single-star calls only).
    """

    def __init__(self, t, rv, ivar, trend, mu, lam, t_ref, k_prior_mode=_lib.JK_KPRIOR_NORMAL,
                 sigma_K0=0.0, P0=1.0, max_K=np.inf, jitter_mode="physical", device=0, cov=None):
        self.device = device
        lib = _lib.ensure_device(device)
        self._lib = lib
        self.t = np.ascontiguousarray(t, dtype=np.float64)
        self.rv = np.ascontiguousarray(rv, dtype=np.float64)
        n = self.t.size
        self.cov = None
        if cov is not None:
            self.cov = np.ascontiguousarray(cov, dtype=np.float64)
            if self.cov.shape != (n, n):
                raise ValueError(f"Invalid covariance shape {self.cov.shape}, expected {(n, n)}")
            self.ivar = 1.0 / np.diag(self.cov)
        else:
            self.ivar = np.ascontiguousarray(ivar, dtype=np.float64)
        self.mu = np.ascontiguousarray(mu, dtype=np.float64)
        self.lam = np.ascontiguousarray(lam, dtype=np.float64)
        L = self.mu.size
        trend = np.asarray(trend, dtype=np.float64)
        if trend.shape != (n, L - 1):
            raise ValueError(f"Invalid design matrix shape: {trend.shape}, expected: {(n, L - 1)}")
        self.trend = np.ascontiguousarray(trend) if trend.size else np.zeros(1)
        if jitter_mode not in JITTER_MODES:
            raise ValueError(f"jitter_mode must be one of {list(JITTER_MODES)}")
        self.jitter_mode = jitter_mode
        self.k_prior_mode = int(k_prior_mode)
        self.t_ref = float(t_ref)
        self.sigma_K0, self.P0, self.max_K = float(sigma_K0), float(P0), float(max_K)
        h = ctypes.c_void_p()
        if self.cov is None:
            _lib.check(lib.jk_star_create(n, _hp(self.t), _hp(self.rv), _hp(self.ivar), L,
                                          _hp(self.trend), _hp(self.mu), _hp(self.lam),
                                          self.t_ref, self.k_prior_mode, self.sigma_K0, self.P0,
                                          self.max_K, JITTER_MODES[jitter_mode],
                                          ctypes.byref(h)), "jk_star_create")
        else:
            _lib.check(lib.jk_star_create_cov(n, _hp(self.t), _hp(self.rv), _hp(self.cov), L,
                                              _hp(self.trend), _hp(self.mu), _hp(self.lam),
                                              self.t_ref, self.k_prior_mode, self.sigma_K0,
                                              self.P0, self.max_K, JITTER_MODES[jitter_mode],
                                              ctypes.byref(h)), "jk_star_create_cov")
        self._h = h

    @property
    def n(self):
        return self.t.size

    @property
    def L(self):
        return self.mu.size

    def arrays(self):
        """The star's host arrays and constants in the form of star_arrays()."""
        return dict(t=self.t, rv=self.rv, ivar=self.ivar,
                    trend=self.trend.reshape(self.n, -1) if self.L > 1
                    else np.zeros((self.n, 0)),
                    mu=self.mu, lam=self.lam, t_ref=self.t_ref, cov=self.cov,
                    k_prior_mode=self.k_prior_mode, sigma_K0=self.sigma_K0, P0=self.P0,
                    max_K=self.max_K)

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.jk_star_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @classmethod
    def from_data_prior(cls, data, prior, trend_M, jitter_mode="physical", device=0):
        """Star constants from RVData + JokerPrior + trend_M (see
        :func:`star_arrays`)."""
        a = star_arrays(data, prior, trend_M)
        return cls(a["t"], a["rv"], a["ivar"], a["trend"], a["mu"], a["lam"], a["t_ref"],
                   a["k_prior_mode"], a["sigma_K0"], a["P0"], a["max_K"],
                   jitter_mode=jitter_mode, device=device, cov=a["cov"])


def probe_epochs(n, k=None):
    """k of n epochs (default n / 6, at least 6) spread evenly over the
    time-sorted list: the probe of earlier versions, kept for callers that
    choose the subset by time."""
    k = max(6, int(round(n / 6.0))) if k is None else int(k)
    k = min(k, n)
    return np.unique(np.round(np.linspace(0, n - 1, k)).astype(np.int64))


def probe_size(n):
    """Epochs of a star's probe: about sqrt(n), at least 5 and at most 16,
    fewer than n."""
    return int(min(n - 1, max(5, min(16, round(float(np.sqrt(n)))))))


def _trend_residual(a):
    """(weighted squared residuals, degrees of freedom) of the weighted
    least-squares fit of the RVs on the trend columns of star_arrays()
    output `a`."""
    y = np.asarray(a["rv"], dtype=np.float64)
    w = np.asarray(a["ivar"], dtype=np.float64)
    T = np.asarray(a["trend"], dtype=np.float64).reshape(y.size, -1)
    k = T.shape[1]
    r = y
    if k:
        sw = np.sqrt(w)
        beta, *_ = np.linalg.lstsq(T * sw[:, None], y * sw, rcond=None)
        r = y - T @ beta
    return w * r * r, y.size - k


def probe_epochs_residual(a, k=None):
    """The epochs of a star's probe: the k (default probe_size(n)) epochs
    where the trend-only fit leaves the largest weighted squared residual,
    in time order.  Those are the epochs an orbit has to explain, so a prior
    row that misses them loses the most there.  Estimated with the CPU
    restatement on the probed stars of the serving stream (M log-uniform on
    [3, 200], 2^16 rows, margin 40): the probe's epochs plus the rows left
    for the full evaluation come to 0.17-0.19 of the unpruned work, against
    0.29 for n / 6 epochs spread evenly over the baseline."""
    r2, _ = _trend_residual(a)
    n = r2.size
    k = probe_size(n) if k is None else min(int(k), n)
    return np.sort(np.argsort(-r2, kind="stable")[:k]).astype(np.int64)


def probe_star(a, jitter_mode="physical", device=0, k=None):
    """(probe DeviceStar, add) for the star of star_arrays() output `a`: the
    star on the epochs probe_epochs_residual(a, k) with the same prior, and
    add = -1/2 sum over the other epochs of ln(2 pi sigma_n^2).  For every
    prior row, lnL_probe + add >= lnL of the whole star: p(y) = p(y_S)
    p(y_rest | y_S), and the conditional covariance of y_rest is at least the
    noise covariance of those epochs, so their density is at most
    prod_n (2 pi sigma_n^2)^(-1/2) (with jitter, sigma_n^2 + s^2 only makes it
    smaller).  The bound holds for any subset S.  Diagonal noise only."""
    if a.get("cov") is not None:
        raise ValueError("probe stars need a diagonal noise covariance")
    n = a["t"].size
    idx = probe_epochs_residual(a, k)
    rest = np.setdiff1d(np.arange(n), idx)
    add = -0.5 * float(np.sum(np.log(2.0 * np.pi / a["ivar"][rest])))
    trend = np.asarray(a["trend"], dtype=np.float64).reshape(n, -1)[idx]
    star = DeviceStar(a["t"][idx], a["rv"][idx], a["ivar"][idx], trend, a["mu"], a["lam"],
                      a["t_ref"], a["k_prior_mode"], a["sigma_K0"], a["P0"], a["max_K"],
                      jitter_mode=jitter_mode, device=device)
    return star, add


def trend_signal(a):
    """How far a Keplerian orbit can lift a star's likelihood above its
    trend-only fit: half the excess of the trend-only chi^2 over its degrees
    of freedom, (chi2 - (n - n_trend)) / 2, from the weighted least-squares
    fit of the RVs on the trend columns of star_arrays() output `a`.  Data a
    constant (or trend) model fits leave every prior row within a few units
    of the maximum, so a probe prunes nothing there."""
    r2, dof = _trend_residual(a)
    return 0.5 * (float(np.sum(r2)) - dof)


def star_arrays(data, prior, trend_M):
    """The arrays CJokerHelper.__init__ derives from (data, prior, trend_M)
    (fast_likelihood.pyx:134-252): times, RVs and 1/sigma^2 (or, for data with
    a full covariance, that covariance; cov is None otherwise) in the data's RV
    unit, and the prior mean / variance of the linear parameters in M_T row
    order -- K at 0, v0 at 1, the v0 offsets at 2.., v_i (i >= 1) after them
    -- in internal units.  For a FixedCompanionMass K prior, sigma_K0, P0
    [day] and max_K are returned instead of a fixed Lambda[0]."""
    from .distributions import FixedCompanionMass

    rv_unit = data.rv.unit
    n_off = prior.n_offsets
    L = 1 + prior.poly_trend + n_off
    mu = np.zeros(L)
    lam = np.zeros(L)
    internal = linear_internal_units(prior, rv_unit)
    Kp = prior.pars["K"]
    kmode, sK0, P0, maxK = _lib.JK_KPRIOR_NORMAL, 0.0, 1.0, np.inf
    if isinstance(Kp, FixedCompanionMass):
        kmode = _lib.JK_KPRIOR_FIXED_COMPANION_MASS
        sK0 = Kp.sigma_K0.to_value(rv_unit)
        P0 = Kp.P0.to_value(u.day)
        maxK = Kp.max_K.to_value(rv_unit)
        mu[0] = Kp.mu * Kp.unit.to(rv_unit)
        lam[0] = 1.0
    else:
        m0, s0 = Kp.mean_std(rv_unit)
        mu[0], lam[0] = m0, s0 ** 2
    for i, name in enumerate(prior._v_trend_names):
        m_, s_ = prior.pars[name].mean_std(internal[name])
        j = 1 if i == 0 else 1 + n_off + i
        mu[j], lam[j] = m_, s_ ** 2
    for i in range(n_off):
        m_, s_ = prior.pars[f"dv0_{i + 1}"].mean_std(rv_unit)
        mu[2 + i], lam[2 + i] = m_, s_ ** 2
    cov = data.cov_value() if data._has_cov else None
    return dict(t=np.asarray(data._t_bmjd, dtype=np.float64),
                rv=np.asarray(data.rv.value, dtype=np.float64),
                ivar=None if cov is not None else data.ivar_value(), cov=cov,
                trend=np.asarray(trend_M, dtype=np.float64), mu=mu, lam=lam,
                t_ref=float(data._t_ref_bmjd), k_prior_mode=kmode, sigma_K0=sK0, P0=P0,
                max_K=maxK)


def linear_internal_units(prior, rv_unit):
    """The internal units of CJokerHelper (fast_likelihood.pyx:134-148), in
    packed order: P, e, omega, M0, s, K, v0, dv0_1..., v1, v2, ..."""
    units = dict(NONLINEAR_INTERNAL_UNITS)
    units["s"] = rv_unit
    units["K"] = rv_unit
    units["v0"] = rv_unit
    for i in range(prior.n_offsets):
        units[f"dv0_{i + 1}"] = rv_unit
    for i, name in enumerate(prior._v_trend_names):
        units[name] = rv_unit / u.day ** i
    return units


# ---------------------------------------------------------------------------
class PriorCache:
    """Prior samples resident in HBM as SoA columns.

    FP64 columns: P [day], e, omega [rad], M0 [rad], s [rv unit] (40 B/row),
    read by both precisions.  fp32_copy adds FP32 copies of e, omega, M0, s
    (16 B/row) for callers of their own; no kernel here reads them.
    `ln_prior` (optional) stays on the host.
    """

    def __init__(self, P, e, omega, M0, s=None, ln_prior=None, device=0, fp32_copy=False):
        torch = _torch()
        _lib.ensure_device(device)
        self.device = device
        dev = torch.device("cuda", device)
        n = int(np.size(P))
        cols = [P, e, omega, M0, np.zeros(n) if s is None else s]
        self.P, self.e, self.omega, self.M0, self.s = [
            torch.as_tensor(np.ascontiguousarray(c, dtype=np.float64)).to(dev) for c in cols]
        self.fp32 = None
        if fp32_copy:
            self.fp32 = tuple(c.to(torch.float32) for c in (self.e, self.omega, self.M0, self.s))
        self.ln_prior = None if ln_prior is None else np.asarray(ln_prior, dtype=np.float64)

    def __len__(self):
        return int(self.P.shape[0])

    @classmethod
    def from_packed(cls, packed, ln_prior=None, device=0, fp32_copy=False):
        packed = np.asarray(packed, dtype=np.float64)
        return cls(packed[:, 0], packed[:, 1], packed[:, 2], packed[:, 3], packed[:, 4],
                   ln_prior=ln_prior, device=device, fp32_copy=fp32_copy)

    @classmethod
    def from_samples(cls, samples, rv_unit=u.km / u.s, device=0, fp32_copy=False):
        units = dict(NONLINEAR_INTERNAL_UNITS)
        units["s"] = rv_unit
        packed, _ = samples.pack(units=units, names=NONLINEAR_PACKED_ORDER)
        lnp = np.asarray(samples["ln_prior"].value) if "ln_prior" in samples else None
        return cls.from_packed(packed, ln_prior=lnp, device=device, fp32_copy=fp32_copy)

    @classmethod
    def from_hdf5(cls, filename, rv_unit=u.km / u.s, path="samples", device=0, fp32_copy=False,
                  block_rows=1 << 22):
        """Stream the table of an HDF5 samples file -- the reference's prior
        cache files (samples.py:480-545), read there in row batches with a
        unit conversion per column (utils.py:168-228) -- into the device
        columns, `block_rows` rows at a time, converting each column to the
        internal units (P day, e 1, omega and M0 rad, s `rv_unit`).  A missing
        s column is zero; ln_prior, when present, is kept on the host.  A file
        without its YAML header is taken to hold internal units."""
        import warnings

        from . import hdf5
        from .samples import read_hdf5_header

        torch = _torch()
        _lib.ensure_device(device)
        dev = torch.device("cuda", device)
        want = dict(NONLINEAR_INTERNAL_UNITS)
        want["s"] = rv_unit
        with hdf5.File(filename) as f:
            header, units = read_hdf5_header(f, path)
            if not units:
                warnings.warn(f"{filename}: no table header; columns taken to be in internal "
                              "units (day, rad, the data's RV unit)", RuntimeWarning)
                units = dict(want)
            ds = f[path]
            names = ds.dtype.names or ()
            missing = [k for k in ("P", "e", "omega", "M0") if k not in names]
            if missing:
                raise ValueError(f"{filename}: table '{path}' lacks the column(s) {missing}")
            n = len(ds)
            self = cls.__new__(cls)
            self.device = device
            self.P, self.e, self.omega, self.M0, self.s = [
                torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(5)]
            cols = dict(zip(NONLINEAR_PACKED_ORDER, (self.P, self.e, self.omega, self.M0, self.s)))
            scale = {k: u._as_unit(units.get(k, u.one)).to(want[k])
                     for k in NONLINEAR_PACKED_ORDER if k in names}
            lnp = [] if "ln_prior" in names else None
            buf = np.empty(min(n, int(block_rows)) * ds.dtype.itemsize, dtype=np.uint8)
            for lo in range(0, n, int(block_rows)):
                hi = min(n, lo + int(block_rows))
                block = ds.read(lo, hi, buf=buf)
                for k, fac in scale.items():
                    v = np.ascontiguousarray(block[k], dtype=np.float64)
                    if fac != 1.0:
                        v = v * fac
                    cols[k][lo:hi].copy_(torch.from_numpy(v))
                if lnp is not None:
                    lnp.append(np.array(block["ln_prior"], dtype=np.float64))
        self.ln_prior = None if lnp is None else (np.concatenate(lnp) if lnp else np.zeros(0))
        self.fp32 = None
        if fp32_copy:
            self.fp32 = tuple(c.to(torch.float32) for c in (self.e, self.omega, self.M0, self.s))
        return self

    @classmethod
    def generate(cls, n, seed=0, P_min=2.0, P_max=2048.0, e_ab=(0.867, 3.03), ln_s=None,
                 device=0, fp32_copy=False, n_tab=1 << 16):
        """n prior rows drawn on the device (jk_prior_fill): P log-uniform on
        [P_min, P_max] days, e ~ Beta(e_ab) (inverse CDF table), omega and M0
        uniform on [0, 2 pi), ln s ~ N(ln_s) km/s or s = 0.  Row r depends
        only on (seed, r)."""
        from scipy.special import betaincinv

        torch = _torch()
        lib = _lib.ensure_device(device)
        dev = torch.device("cuda", device)
        self = cls.__new__(cls)
        self.device = device
        self.ln_prior = None
        n = int(n)
        q = np.linspace(0.0, 1.0, int(n_tab) + 1)
        tab = betaincinv(e_ab[0], e_ab[1], q)
        tab[0], tab[-1] = 0.0, min(tab[-1], np.nextafter(1.0, 0.0))
        d_tab = torch.as_tensor(tab, device=dev)
        self.P, self.e, self.omega, self.M0, self.s = [
            torch.empty(n, dtype=torch.float64, device=dev) for _ in range(5)]
        self.fp32 = None
        f32 = [None] * 4
        if fp32_copy:
            self.fp32 = tuple(torch.empty(n, dtype=torch.float32, device=dev) for _ in range(4))
            f32 = list(self.fp32)
        mu, sig = (0.0, -1.0) if ln_s is None else (float(ln_s[0]), float(ln_s[1]))
        _lib.check(lib.jk_prior_fill(n, 0, int(seed), float(P_min), float(P_max), _dptr(d_tab),
                                     int(n_tab), mu, sig, _dptr(self.P), _dptr(self.e),
                                     _dptr(self.omega), _dptr(self.M0), _dptr(self.s),
                                     *[_dptr(x) for x in f32], _stream(device)), "jk_prior_fill")
        return self

    def host_rows(self, idx):
        """(len(idx), 5) float64 rows P, e, omega, M0, s copied to the host."""
        torch = _torch()
        it = torch.as_tensor(np.asarray(idx, dtype=np.int64), device=self.P.device)
        cols = [c.index_select(0, it) for c in (self.P, self.e, self.omega, self.M0, self.s)]
        return torch.stack(cols, dim=1).cpu().numpy()

    def nbytes(self):
        b = 8 * 5 * len(self)
        if self.fp32 is not None:
            b += 4 * 4 * len(self)
        return b


# ---------------------------------------------------------------------------
def lnl(star, cache, rows=None, n=None, precision="f64", out=None, start=0, bound=None):
    """Marginal lnL of cache rows (all, `rows` (device int64 tensor) or
    [start, start + n)) into a device float64 tensor.  `bound` (mixed only): a
    device float32 tensor of n entries that receives each row's bound on
    |lnL(mixed) - lnL(FP64)|."""
    torch = _torch()
    lib = _lib.ensure_device(cache.device)
    if rows is not None:
        n = int(rows.shape[0])
    elif n is None:
        n = len(cache) - start
    if out is None:
        out = torch.empty(n, dtype=torch.float64, device=cache.P.device)
    if n == 0:
        return out
    off64 = 8 * start
    if rows is not None:
        off64 = 0
    P = ctypes.c_void_p(cache.P.data_ptr() + off64)
    if precision == "f64":
        cols = [ctypes.c_void_p(c.data_ptr() + off64) for c in (cache.e, cache.omega, cache.M0, cache.s)]
        fn, what = lib.jk_lnl_f64, "jk_lnl_f64"
    elif precision == "mixed":
        # the FP64 columns too: the FP32 arithmetic starts inside the kernel
        cols = [ctypes.c_void_p(c.data_ptr() + off64) for c in (cache.e, cache.omega, cache.M0, cache.s)]
        _lib.check(lib.jk_lnl_mixed(star.handle, n, P, *cols, _dptr(rows), _dptr(out),
                                    _dptr(bound), _stream(cache.device)), "jk_lnl_mixed")
        return out
    else:
        raise ValueError("precision must be 'f64' or 'mixed'")
    if bound is not None:
        raise ValueError("the error bound belongs to the mixed-precision likelihood")
    _lib.check(fn(star.handle, n, P, *cols, _dptr(rows), _dptr(out), _stream(cache.device)), what)
    return out


class PendingReject:
    """Result of :func:`lnl_reject_async`: the copy of the accepted indices to
    pinned host memory is queued behind the launches; :meth:`result` waits
    for it alone, so work queued after it on the stream keeps running."""

    def __init__(self, out, host, event, max_keep):
        self._out, self._host, self._event, self._max_keep = out, host, event, max_keep

    def result(self):
        """(device lnL tensor, idx, n_pass, llmax)."""
        self._event.synchronize()
        h = self._host.numpy()
        n_pass = int(h[0])
        k = min(n_pass, self._max_keep)
        return self._out, h[2:2 + k].copy(), n_pass, float(h[1:2].view(np.float64)[0])


def lnl_reject(star, cache, rng_state, max_keep, rows=None, n=None, start=0, out=None, offset=0):
    """Plain FP64 lnL of cache rows (as :func:`lnl`) and the rejection step over
    them (as :func:`reject`) in one call (jk_lnl_reject_f64).  Returns
    (device lnL tensor, idx, n_pass, llmax)."""
    return lnl_reject_async(star, cache, rng_state, max_keep, rows=rows, n=n, start=start,
                            out=out, offset=offset).result()


def lnl_reject_async(star, cache, rng_state, max_keep, rows=None, n=None, start=0, out=None,
                     offset=0):
    """:func:`lnl_reject` without the wait: returns a :class:`PendingReject`.
    The host may queue the next call (with the stream advanced by n, which
    does not depend on this result) before collecting this one."""
    torch = _torch()
    lib = _lib.ensure_device(cache.device)
    if rows is not None:
        n = int(rows.shape[0])
    elif n is None:
        n = len(cache) - start
    if n <= 0:
        raise ValueError("zero-size array to reduction operation maximum which has no identity")
    if out is None:
        out = torch.empty(n, dtype=torch.float64, device=cache.P.device)
    off = 0 if rows is not None else 8 * start
    cols = [ctypes.c_void_p(c.data_ptr() + off) for c in
            (cache.P, cache.e, cache.omega, cache.M0, cache.s)]
    max_keep = int(min(max_keep, n))
    res = torch.empty(2 + max(max_keep, 1), dtype=torch.int64, device=cache.P.device)
    p = res.data_ptr()
    state, inc = rng_state
    _lib.check(lib.jk_lnl_reject_f64(star.handle, n, *cols, _dptr(rows), _dptr(out), *_split(state),
                                     *_split(inc), int(offset), max_keep, ctypes.c_void_p(p + 16),
                                     ctypes.c_void_p(p), ctypes.c_void_p(p + 8),
                                     _stream(cache.device)), "jk_lnl_reject_f64")
    host = torch.empty(res.shape, dtype=res.dtype, pin_memory=True)
    host.copy_(res, non_blocking=True)
    ev = torch.cuda.Event()
    ev.record(torch.cuda.current_stream(cache.device))
    return PendingReject(out, host, ev, max_keep)


def polish(star, cache, ll, rows=None, start=0, window=40.0, first=0):
    """FP64 re-evaluation, in place, of every entry ll[i], i >= first, of `ll`
    (the lnL of cache rows `rows` or [start, start + len(ll)), mixed precision
    from `first` on) not below max(ll) - window.  Returns a device int64
    tensor holding the count."""
    torch = _torch()
    lib = _lib.ensure_device(cache.device)
    n = int(ll.shape[0])
    count = torch.zeros(1, dtype=torch.int64, device=cache.P.device)
    if n == 0:
        return count
    off = 0 if rows is not None else 8 * start
    cols = [ctypes.c_void_p(c.data_ptr() + off) for c in
            (cache.P, cache.e, cache.omega, cache.M0, cache.s)]
    _lib.check(lib.jk_polish_f64(star.handle, n, *cols, _dptr(rows), _dptr(ll), float(window),
                                 int(first), _dptr(count), _stream(cache.device)), "jk_polish_f64")
    return count


def uniforms(rng_state, n, offset=0, device=0):
    """numpy-identical uniforms of a PCG64 state (tests and callers)."""
    torch = _torch()
    lib = _lib.ensure_device(device)
    state, inc = rng_state
    out = torch.empty(int(n), dtype=torch.float64, device=torch.device("cuda", device))
    _lib.check(lib.jk_pcg64_uniform(*_split(state), *_split(inc), int(offset), int(n), _dptr(out),
                                    _stream(device)), "jk_pcg64_uniform")
    return out


def reject(ll, rng_state, max_keep, offset=0, llmax=None, device=0):
    """Rejection step on a device lnL tensor.  Returns (idx (np.int64, ascending,
    <= max_keep), n_pass, llmax)."""
    torch = _torch()
    lib = _lib.ensure_device(device)
    n = int(ll.shape[0])
    if n == 0:
        raise ValueError("zero-size array to reduction operation maximum which has no identity")
    dev = ll.device
    max_keep = int(min(max_keep, n))
    # one device block [n_pass, max, idx...] and one copy back to pinned memory
    res = torch.empty(2 + max(max_keep, 1), dtype=torch.int64, device=dev)
    p = res.data_ptr()
    state, inc = rng_state
    if llmax is None:
        _lib.check(lib.jk_reject(_dptr(ll), n, *_split(state), *_split(inc), int(offset), max_keep,
                                 ctypes.c_void_p(p + 16), ctypes.c_void_p(p),
                                 ctypes.c_void_p(p + 8), _stream(device)), "jk_reject")
    else:
        _lib.check(lib.jk_reject_with_max(_dptr(ll), n, _dptr(llmax), *_split(state), *_split(inc),
                                          int(offset), max_keep, ctypes.c_void_p(p + 16),
                                          ctypes.c_void_p(p), _stream(device)),
                   "jk_reject_with_max")
        res[1:2].copy_(llmax.reshape(1).view(torch.int64))
    h = to_host(res)
    n_pass = int(h[0])
    k = min(n_pass, max_keep)
    return h[2:2 + k].copy(), n_pass, float(h[1:2].view(np.float64)[0])


def posterior_inputs_many(stars, cache, rows_list):
    """For each (star, accepted cache rows): (packed rows (k, 5), a (k, L),
    cov (k, L, L)) -- jk_linear_posterior launched once per star into one
    device block, the prior rows gathered into the same block, one copy back."""
    torch = _torch()
    lib = _lib.ensure_device(cache.device)
    dev = cache.P.device
    ks = [len(r) for r in rows_list]
    K = int(sum(ks))
    if K == 0:
        return [(np.zeros((0, 5)), np.zeros((0, st.L)), np.zeros((0, st.L, st.L)))
                for st in stars]
    per = [k * (st.L + st.L * st.L + 1) for k, st in zip(ks, stars)]
    base = int(sum(per))
    res = torch.empty(base + 5 * K, dtype=torch.float64, device=dev)
    rows = torch.as_tensor(np.concatenate([np.asarray(r, dtype=np.int64) for r in rows_list]),
                           device=dev)
    p0, r0 = res.data_ptr(), rows.data_ptr()
    off, roff = 0, 0
    cols = [_dptr(c) for c in (cache.P, cache.e, cache.omega, cache.M0, cache.s)]
    for st, k, n in zip(stars, ks, per):
        if k:
            L = st.L
            a_ptr = p0 + 8 * off
            _lib.check(lib.jk_linear_posterior(st.handle, k, *cols, ctypes.c_void_p(r0 + 8 * roff),
                                               ctypes.c_void_p(a_ptr),
                                               ctypes.c_void_p(a_ptr + 8 * k * L),
                                               ctypes.c_void_p(a_ptr + 8 * k * (L + L * L)),
                                               _stream(cache.device)), "jk_linear_posterior")
        off += n
        roff += k
    packed = res[base:].view(5, K)
    for j, c in enumerate((cache.P, cache.e, cache.omega, cache.M0, cache.s)):
        torch.index_select(c, 0, rows, out=packed[j])
    h = to_host(res)
    out, off, roff = [], 0, 0
    hp = h[base:].reshape(5, K)
    for st, k, n in zip(stars, ks, per):
        L = st.L
        blk = h[off:off + n]
        out.append((hp[:, roff:roff + k].T.copy(), blk[:k * L].reshape(k, L).copy(),
                    blk[k * L:k * (L + L * L)].reshape(k, L, L).copy()))
        off += n
        roff += k
    return out


def posterior_inputs_table(table, slots, cache, rows_list):
    """:func:`posterior_inputs_many` for stars held in a device star table
    (`table`, a batch.StarBatch; `slots` their entries): one
    jk_linear_posterior_batch launch for every star's accepted rows, the
    prior rows gathered into the same block, one copy back."""
    torch = _torch()
    lib = _lib.ensure_device(cache.device)
    dev = cache.P.device
    L = table.L
    ks = [len(r) for r in rows_list]
    K = int(sum(ks))
    if K == 0:
        return [(np.zeros((0, 5)), np.zeros((0, L)), np.zeros((0, L, L))) for _ in ks]
    per = L + L * L + 1
    res = torch.empty(K * per + 5 * K, dtype=torch.float64, device=dev)
    d_rows = torch.as_tensor(np.concatenate([np.asarray(r, dtype=np.int64) for r in rows_list]),
                             device=dev)
    d_sidx = torch.as_tensor(np.repeat(np.asarray(slots, dtype=np.int32), ks), device=dev)
    p0 = res.data_ptr()
    _lib.check(lib.jk_linear_posterior_batch(
        table.handle, K, _dptr(d_sidx), _dptr(cache.P), _dptr(cache.e), _dptr(cache.omega),
        _dptr(cache.M0), _dptr(cache.s), _dptr(d_rows), ctypes.c_void_p(p0),
        ctypes.c_void_p(p0 + 8 * K * L), ctypes.c_void_p(p0 + 8 * K * (L + L * L)),
        _stream(cache.device)), "jk_linear_posterior_batch")
    packed = res[K * per:].view(5, K)
    for j, c in enumerate((cache.P, cache.e, cache.omega, cache.M0, cache.s)):
        torch.index_select(c, 0, d_rows, out=packed[j])
    h = to_host(res)
    a = h[:K * L].reshape(K, L)
    cov = h[K * L:K * (L + L * L)].reshape(K, L, L)
    hp = h[K * per:].reshape(5, K)
    out, o = [], 0
    for k in ks:
        out.append((hp[:, o:o + k].T.copy(), a[o:o + k].copy(), cov[o:o + k].copy()))
        o += k
    return out


def to_host(t):
    """Device tensor -> numpy array through a pinned staging buffer (one DMA
    copy ordered on the current stream, then a wait for that stream)."""
    torch = _torch()
    h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    h.copy_(t, non_blocking=True)
    torch.cuda.current_stream(t.device).synchronize()
    return h.numpy()


def linear_posterior(star, cache, rows):
    """(a (k, L), cov (k, L, L), ll (k,)) for cache rows `rows` (host ints)."""
    torch = _torch()
    lib = _lib.ensure_device(cache.device)
    k = len(rows)
    L = star.L
    dev = cache.P.device
    if k == 0:
        return np.zeros((0, L)), np.zeros((0, L, L)), np.zeros(0)
    r = torch.as_tensor(np.asarray(rows, dtype=np.int64), device=dev)
    # a, cov and lnL in one device block, one copy back
    res = torch.empty(k * (L + L * L + 1), dtype=torch.float64, device=dev)
    p = res.data_ptr()
    _lib.check(lib.jk_linear_posterior(star.handle, k, _dptr(cache.P), _dptr(cache.e),
                                       _dptr(cache.omega), _dptr(cache.M0), _dptr(cache.s),
                                       _dptr(r), ctypes.c_void_p(p),
                                       ctypes.c_void_p(p + 8 * k * L),
                                       ctypes.c_void_p(p + 8 * k * (L + L * L)),
                                       _stream(cache.device)), "jk_linear_posterior")
    h = to_host(res)
    return (h[:k * L].reshape(k, L).copy(), h[k * L:k * (L + L * L)].reshape(k, L, L).copy(),
            h[k * (L + L * L):].copy())


def debug_single(star, row):
    """jk_debug_single: every intermediate of one prior row (P, e, omega, M0,
    s), as CJokerHelper.test_likelihood_worker leaves them
    (fast_likelihood.pyx:547-576).  Returns a dict of host arrays M_T (L, n),
    Ainv, A (L, L), a (L,), b (n,), B, Binv (n, n) and floats ll, chi2,
    logdet, Lambda0."""
    torch = _torch()
    lib = _lib.ensure_device(star.device)
    n, L = star.n, star.L
    dev = torch.device("cuda", star.device)
    sizes = [("M_T", L * n), ("Ainv", L * L), ("A", L * L), ("a", L), ("b", n), ("B", n * n),
             ("Binv", n * n), ("sc", 4)]
    total = sum(k for _, k in sizes)
    res = torch.empty(total + 5, dtype=torch.float64, device=dev)
    res[total:].copy_(torch.as_tensor(np.asarray(row, dtype=np.float64).reshape(5)))
    ptrs, off = {}, 0
    for name, k in sizes:
        ptrs[name] = ctypes.c_void_p(res.data_ptr() + 8 * off)
        off += k
    _lib.check(lib.jk_debug_single(star.handle, ctypes.c_void_p(res.data_ptr() + 8 * total),
                                   *[ptrs[k] for k, _ in sizes], _stream(star.device)),
               "jk_debug_single")
    h = to_host(res)
    out, off = {}, 0
    shapes = {"M_T": (L, n), "Ainv": (L, L), "A": (L, L), "a": (L,), "b": (n,), "B": (n, n),
              "Binv": (n, n), "sc": (4,)}
    for name, k in sizes:
        out[name] = h[off:off + k].reshape(shapes[name]).copy()
        off += k
    sc = out.pop("sc")
    out.update(ll=float(sc[0]), chi2=float(sc[1]), logdet=float(sc[2]), Lambda0=float(sc[3]))
    return out


def count_failures(ll):
    """jk_count_failures on a device lnL tensor: (n_posinf, n_neginf, n_nan)."""
    torch = _torch()
    lib = _lib.ensure_device(ll.device.index or 0)
    counts = torch.empty(3, dtype=torch.int64, device=ll.device)
    _lib.check(lib.jk_count_failures(_dptr(ll), int(ll.shape[0]), _dptr(counts),
                                     _stream(ll.device.index or 0)), "jk_count_failures")
    h = to_host(counts)
    return int(h[0]), int(h[1]), int(h[2])
