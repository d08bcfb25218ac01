"""Synthetic stars and prior caches of the benchmark workloads.

numpy only.  Each workload fixes its seeds so that the bench, the tests and
the smoke check see the same inputs:

* ``single_star(seed=1, M=16, baseline=1000)`` -- one star, times uniform on
  [0, baseline] d, t_ref = min t; true orbit P = 120 d, e = 0.3, omega = 1.0,
  M0 = 2.0, K = 10 km/s, v0 = -5 km/s; sigma ~ U[0.1, 0.5] km/s; Gaussian
  noise.  (M = 50, baseline = 3650, seed 5 for the large-precision workload.)
* ``prior_columns(n, seed=0, ...)`` -- prior rows as SoA float64 columns:
  P log-uniform on [P_min, P_max] (the UniformLog draw of
  distributions.py:24-28), e ~ Beta(0.867, 3.03), omega, M0 ~ U[0, 2 pi),
  s = 0 or ln s ~ N(-3, 1) km/s.
* ``many_stars(n_stars=1024, seed=3)`` -- M ~ U{3..64}, baselines
  U[200, 3000] d, half binaries with orbits from the prior (K ~ N(0, 30)),
  half constant; v0 ~ N(0, 100); sigma ~ U[0.05, 1].
* ``arrivals(rate=200, n=..., seed=4)`` -- Poisson arrival times and
  log-uniform M in [3, 200] for the serving workload.
"""
import numpy as np

__all__ = ["kepler_rv", "single_star", "prior_columns", "many_stars", "arrivals", "Star"]


def kepler_rv(t, P, e, omega, M0, K, t_ref):
    """K [cos(omega + f) + e cos omega] with M = 2 pi (t - t_ref)/P - M0
    (vectorised Newton; used only to synthesise data)."""
    t = np.asarray(t, dtype=np.float64)
    M = 2 * np.pi * (t - t_ref) / P - M0
    M = np.mod(M + np.pi, 2 * np.pi) - np.pi
    E = M + e * np.sin(M)
    for _ in range(60):
        E = E - (E - e * np.sin(E) - M) / (1 - e * np.cos(E))
    f = 2 * np.arctan2(np.sqrt(1 + e) * np.sin(E / 2), np.sqrt(1 - e) * np.cos(E / 2))
    return K * (np.cos(omega + f) + e * np.cos(omega))


class Star:
    """Plain arrays of one synthetic star (days, km/s)."""

    def __init__(self, t, rv, err, t_ref=None, truth=None):
        self.t = np.asarray(t, dtype=np.float64)
        self.rv = np.asarray(rv, dtype=np.float64)
        self.err = np.asarray(err, dtype=np.float64)
        self.t_ref = float(self.t.min() if t_ref is None else t_ref)
        self.truth = truth or {}

    @property
    def M(self):
        return self.t.size

    def ivar(self):
        return 1.0 / self.err ** 2

    def trend(self, poly_trend=1):
        """[1, dt, dt^2, ...] (likelihood_helpers.py:28-37 with one survey)."""
        dt = self.t - self.t_ref
        return np.hstack([np.ones((self.M, 1)), np.vander(dt, N=poly_trend, increasing=True)[:, 1:]])

    def to_rvdata(self):
        from . import units as u
        from .data import RVData
        return RVData(t=self.t, rv=u.Quantity(self.rv, u.km / u.s),
                      rv_err=u.Quantity(self.err, u.km / u.s))


def single_star(seed=1, M=16, baseline=1000.0, P=120.0, e=0.3, omega=1.0, M0=2.0, K=10.0,
                v0=-5.0, err_range=(0.1, 0.5)):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0.0, baseline, M))
    err = rng.uniform(err_range[0], err_range[1], M)
    t_ref = t.min()
    rv = kepler_rv(t, P, e, omega, M0, K, t_ref) + v0 + rng.normal(0.0, err)
    return Star(t, rv, err, t_ref, truth=dict(P=P, e=e, omega=omega, M0=M0, K=K, v0=v0))


def prior_columns(n, seed=0, P_min=2.0, P_max=2048.0, jitter=False, e_ab=(0.867, 3.03),
                  dtype=np.float64):
    """Dict of SoA prior columns P, e, omega, M0, s (draw order P, e, omega,
    M0, s from default_rng(seed))."""
    rng = np.random.default_rng(seed)
    n = int(n)
    fac = np.log(P_max) - np.log(P_min)
    P = np.exp(rng.uniform(size=n) * fac + np.log(P_min))
    e = rng.beta(e_ab[0], e_ab[1], size=n)
    omega = rng.uniform(0.0, 2 * np.pi, size=n)
    M0 = rng.uniform(0.0, 2 * np.pi, size=n)
    s = np.exp(rng.normal(-3.0, 1.0, size=n)) if jitter else np.zeros(n)
    return {k: v.astype(dtype) for k, v in dict(P=P, e=e, omega=omega, M0=M0, s=s).items()}


def _one_of_many(rng, M, baseline, binary, P_min=2.0, P_max=2048.0):
    t = np.sort(rng.uniform(0.0, baseline, M))
    err = rng.uniform(0.05, 1.0, M)
    v0 = rng.normal(0.0, 100.0)
    t_ref = t.min()
    truth = dict(v0=v0, binary=bool(binary))
    rv = np.full(M, v0)
    if binary:
        P = np.exp(rng.uniform() * (np.log(P_max) - np.log(P_min)) + np.log(P_min))
        e = rng.beta(0.867, 3.03)
        omega = rng.uniform(0, 2 * np.pi)
        M0 = rng.uniform(0, 2 * np.pi)
        K = rng.normal(0.0, 30.0)
        rv = rv + kepler_rv(t, P, e, omega, M0, K, t_ref)
        truth.update(P=P, e=e, omega=omega, M0=M0, K=K)
    rv = rv + rng.normal(0.0, err)
    return Star(t, rv, err, t_ref, truth=truth)


def many_stars(n_stars=1024, seed=3, M_range=(3, 64), baseline_range=(200.0, 3000.0)):
    """Stars of the batched workload: M uniform on the integers of M_range,
    baseline uniform on baseline_range, even index binaries."""
    rng = np.random.default_rng(seed)
    stars = []
    for i in range(n_stars):
        M = int(rng.integers(M_range[0], M_range[1] + 1))
        T = rng.uniform(*baseline_range)
        stars.append(_one_of_many(rng, M, T, binary=(i % 2 == 0)))
    return stars


def arrivals(n, rate=200.0, seed=4, M_range=(3, 200), baseline_range=(200.0, 3000.0)):
    """(arrival times [s], stars) of a Poisson stream: exponential gaps of
    mean 1/rate, M log-uniform on M_range (rounded), star parameters as in
    many_stars."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=n)
    times = np.cumsum(gaps) - gaps[0]
    stars = []
    for i in range(n):
        M = int(np.round(np.exp(rng.uniform(np.log(M_range[0]), np.log(M_range[1])))))
        T = rng.uniform(*baseline_range)
        stars.append(_one_of_many(rng, M, T, binary=(i % 2 == 0)))
    return times, stars
