"""RVData: radial-velocity measurements of one target (thejoker/data.py).

Times are Barycentric MJD days (``t`` may also be anything with
``.tcb.mjd``, e.g. an astropy Time); ``rv`` and ``rv_err`` are speeds
(Quantities, or bare numbers taken as km/s).  Construction filters
non-finite rows (data.py:75-95), sorts by time (data.py:97-105) and sets
``t_ref`` to the earliest time unless given (data.py:107-122).
"""
import logging

import numpy as np

from . import units as u

__all__ = ["RVData"]

logger = logging.getLogger("thejoker")


def _bmjd(t):
    if hasattr(t, "tcb") and hasattr(t.tcb, "mjd"):
        return np.atleast_1d(np.asarray(t.tcb.mjd, dtype=np.float64))
    return np.atleast_1d(np.asarray(t, dtype=np.float64))


class RVData:
    def __init__(self, t, rv, rv_err, t_ref=None, clean=True):
        self._t_bmjd = _bmjd(t)
        rvq = u.as_quantity(rv, u.km / u.s)
        if not rvq.unit.is_equivalent(u.km / u.s):
            raise u.UnitsError("rv must be a speed")
        self.rv = u.Quantity(np.atleast_1d(rvq.value), rvq.unit)
        errq = u.as_quantity(rv_err, self.rv.unit)
        err_val = np.atleast_1d(np.asarray(errq.value, dtype=np.float64))
        if err_val.ndim == 1:
            if not errq.unit.is_equivalent(u.km / u.s):
                raise u.UnitsError("rv_err must be a speed")
            self._has_cov = False
        elif err_val.ndim == 2:
            if not errq.unit.is_equivalent((u.km / u.s) ** 2):
                raise u.UnitsError("a 2-D rv_err (covariance) must have units of speed^2")
            self._has_cov = True
        else:
            self._has_cov = False
        self.rv_err = u.Quantity(err_val, errq.unit)

        nrv = self.rv.size
        if err_val.shape != (nrv, nrv) and err_val.shape != (nrv,):
            raise ValueError(f"Invalid shape for input RV error {err_val.shape}. Should either be "
                             f"({nrv},) or ({nrv}, {nrv})")
        if self._t_bmjd.shape != np.shape(self.rv.value):
            raise ValueError("Shape of input times and RVs must be consistent "
                             f"({self._t_bmjd.shape} vs {np.shape(self.rv.value)})")

        if clean:
            idx = np.isfinite(self._t_bmjd) & np.isfinite(self.rv.value)
            if self._has_cov:
                idx &= np.isfinite(err_val).all(axis=0)
            else:
                idx &= np.isfinite(err_val)
            n_filter = len(idx) - int(idx.sum())
            if n_filter > 0:
                logger.info(f"Filtering {n_filter} NaN/Inf data points")
            self._apply_index(idx)

        self._apply_index(self._t_bmjd.argsort())

        if t_ref is False:
            self.t_ref = None
            self._t_ref_bmjd = 0.0
        else:
            if t_ref is None:
                t_ref = self._t_bmjd.min()
            t_ref = float(_bmjd(t_ref)[0])
            self.t_ref = t_ref
            self._t_ref_bmjd = t_ref

    def _apply_index(self, idx):
        self._t_bmjd = self._t_bmjd[idx]
        self.rv = u.Quantity(np.asarray(self.rv.value)[idx], self.rv.unit)
        ev = np.asarray(self.rv_err.value)
        ev = ev[idx][:, idx] if self._has_cov else ev[idx]
        self.rv_err = u.Quantity(ev, self.rv_err.unit)

    @property
    def t(self):
        """Observation times (BMJD days)."""
        return self._t_bmjd

    @property
    def cov(self):
        if self._has_cov:
            return self.rv_err
        return u.Quantity(np.diag(np.asarray(self.rv_err.value) ** 2), self.rv_err.unit ** 2)

    @property
    def ivar(self):
        """Inverse variance, in 1 / (rv unit)^2 (data.py:146-152)."""
        if self._has_cov:
            return u.Quantity(np.linalg.inv(np.asarray(self.rv_err.value)), u.one / self.rv_err.unit)
        return u.Quantity(1.0 / np.asarray(self.rv_err.value) ** 2, u.one / self.rv_err.unit ** 2)

    def ivar_value(self):
        """1 / sigma^2 in 1 / (rv unit)^2, as CJokerHelper reads it
        (fast_likelihood.pyx:165-166)."""
        if self._has_cov:
            raise ValueError("this RVData has a full covariance: use cov_value()")
        err = self.rv_err.to_value(self.rv.unit)
        return 1.0 / np.asarray(err) ** 2

    def cov_value(self):
        """The full covariance matrix in (rv unit)^2.  A matrix given in the
        RV unit itself is taken as covariance values in that unit squared (the
        reference's ivar reads it so, data.py:146-150)."""
        if not self._has_cov:
            return np.diag(1.0 / self.ivar_value())
        try:
            return np.asarray(self.rv_err.to_value(self.rv.unit ** 2), dtype=np.float64)
        except Exception:
            return np.asarray(self.rv_err.to_value(self.rv.unit), dtype=np.float64)

    def phase(self, P, t_ref=None):
        if t_ref is None:
            t_ref = self._t_ref_bmjd
        return ((self._t_bmjd - t_ref) / u.to_value(P, u.day)) % 1.0

    def __copy__(self):
        return self.__class__(t=self._t_bmjd.copy(), rv=u.Quantity(np.array(self.rv.value), self.rv.unit),
                              rv_err=u.Quantity(np.array(self.rv_err.value), self.rv_err.unit),
                              t_ref=self.t_ref if self.t_ref is not None else False)

    copy = __copy__

    def __getitem__(self, slc):
        ev = np.asarray(self.rv_err.value)
        ev = ev[slc][:, slc] if self._has_cov else ev[slc]
        return self.__class__(t=self._t_bmjd[slc], rv=u.Quantity(np.asarray(self.rv.value)[slc], self.rv.unit),
                              rv_err=u.Quantity(ev, self.rv_err.unit))

    def __len__(self):
        return int(np.size(self.rv.value))

    def __repr__(self):
        return f"<RVData: {len(self)} epochs>"
