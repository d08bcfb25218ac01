"""numpy prior distributions with the reference's parameterisations.

The reference builds its priors as pymc random variables
(thejoker/prior.py:410-575, thejoker/distributions.py); pymc is not a
dependency here, so each distribution is a small object that knows its unit,
how to draw from a numpy Generator and its log-density.  Draw formulas follow
the reference where it defines them:

* :class:`UniformLog` -- ``exp(U * (ln b - ln a) + ln a)``
  (distributions.py:24-28 / 117-120); its log-density restates
  distributions.py:137-144 as written (``-value - ln(ln b - ln a)``).
* :class:`Kipping13Global` / ``Long`` / ``Short`` -- Beta(0.867, 3.03),
  Beta(1.12, 3.09), Beta(0.697, 3.27) (distributions.py:247-268).
* :class:`FixedCompanionMass` -- Normal(mu, sigma_K) with
  ``sigma_K = clip(sigma_K0 (P/P0)^(-1/3) / sqrt(1 - e^2), 0, max_K)``
  (distributions.py:195-244).
* :class:`Normal`, :class:`Uniform`, :class:`LogNormal`, :class:`Constant` --
  the pymc distributions of the same names (``Constant`` is the
  ``pm.Deterministic`` jitter of prior.py:458-461).
"""
import numpy as np

from . import units as u

__all__ = ["Distribution", "UniformLog", "Uniform", "Angle", "Beta", "Kipping13Global",
           "Kipping13Long", "Kipping13Short", "Normal", "LogNormal", "Constant",
           "FixedCompanionMass"]


class Distribution:
    """Base class: `unit` tags the values (thejoker/units.py:10-26)."""

    name = "Distribution"
    is_normal = False

    def __init__(self, unit=u.one):
        self.unit = u._as_unit(unit)

    def sample(self, rng, size, context=None):
        raise NotImplementedError

    def logp(self, x, context=None):
        raise NotImplementedError

    def __repr__(self):
        return f"<{self.name} [{self.unit}]>"


class UniformLog(Distribution):
    """p(x) ~ 1/x on (a, b) (distributions.py:109-144)."""

    name = "UniformLog"

    def __init__(self, a, b, unit=u.one):
        super().__init__(unit)
        self.a = float(a)
        self.b = float(b)
        if not (0 < self.a < self.b):
            raise ValueError("UniformLog requires 0 < a < b")

    def sample(self, rng, size, context=None):
        fac = np.log(self.b) - np.log(self.a)
        uu = rng.uniform(size=size)
        return np.exp(uu * fac + np.log(self.a))

    def logp(self, x, context=None):
        fac = np.log(self.b) - np.log(self.a)
        return -np.asarray(x) - np.log(fac)


class Uniform(Distribution):
    name = "Uniform"

    def __init__(self, lower, upper, unit=u.one):
        super().__init__(unit)
        self.lower = float(lower)
        self.upper = float(upper)

    def sample(self, rng, size, context=None):
        return rng.uniform(self.lower, self.upper, size=size)

    def logp(self, x, context=None):
        x = np.asarray(x)
        inside = (x >= self.lower) & (x <= self.upper)
        return np.where(inside, -np.log(self.upper - self.lower), -np.inf)


class Angle(Uniform):
    """Uniform angle on [0, 2 pi) radians (the prior of omega and M0)."""

    name = "Angle"

    def __init__(self, unit=u.rad):
        super().__init__(0.0, 2 * np.pi, unit)

    def sample(self, rng, size, context=None):
        return rng.uniform(0.0, 2 * np.pi, size=size)


class Beta(Distribution):
    name = "Beta"

    def __init__(self, alpha, beta, unit=u.one):
        super().__init__(unit)
        self.alpha = float(alpha)
        self.beta = float(beta)

    def sample(self, rng, size, context=None):
        return rng.beta(self.alpha, self.beta, size=size)

    def logp(self, x, context=None):
        from scipy.special import betaln
        x = np.asarray(x, dtype=float)
        with np.errstate(divide="ignore", invalid="ignore"):
            lp = ((self.alpha - 1) * np.log(x) + (self.beta - 1) * np.log1p(-x)
                  - betaln(self.alpha, self.beta))
        return np.where((x > 0) & (x < 1), lp, -np.inf)


class Kipping13Global(Beta):
    name = "Kipping13Global"

    def __init__(self, unit=u.one):
        super().__init__(0.867, 3.03, unit)


class Kipping13Long(Beta):
    name = "Kipping13Long"

    def __init__(self, unit=u.one):
        super().__init__(1.12, 3.09, unit)


class Kipping13Short(Beta):
    name = "Kipping13Short"

    def __init__(self, unit=u.one):
        super().__init__(0.697, 3.27, unit)


class Normal(Distribution):
    name = "Normal"
    is_normal = True

    def __init__(self, mu, sigma, unit=u.one):
        super().__init__(unit)
        self.mu = float(mu)
        self.sigma = float(sigma)
        if not self.sigma > 0:
            raise ValueError("Normal requires sigma > 0")

    def mean_std(self, to_unit):
        """(mu, sigma) converted to `to_unit` (utils.py:317-334)."""
        f = self.unit.to(to_unit)
        return self.mu * f, self.sigma * f

    def sample(self, rng, size, context=None):
        return rng.normal(self.mu, self.sigma, size=size)

    def logp(self, x, context=None):
        z = (np.asarray(x) - self.mu) / self.sigma
        return -0.5 * z * z - np.log(self.sigma) - 0.5 * np.log(2 * np.pi)


class LogNormal(Distribution):
    """ln x ~ N(mu, sigma) (the jitter prior ln s ~ N(-3, 1) of the configs)."""

    name = "LogNormal"

    def __init__(self, mu, sigma, unit=u.one):
        super().__init__(unit)
        self.mu = float(mu)
        self.sigma = float(sigma)

    def sample(self, rng, size, context=None):
        return np.exp(rng.normal(self.mu, self.sigma, size=size))

    def logp(self, x, context=None):
        x = np.asarray(x, dtype=float)
        with np.errstate(divide="ignore", invalid="ignore"):
            z = (np.log(x) - self.mu) / self.sigma
            lp = -0.5 * z * z - np.log(self.sigma * x) - 0.5 * np.log(2 * np.pi)
        return np.where(x > 0, lp, -np.inf)


class Constant(Distribution):
    """A fixed value (prior.py:458-461 wraps the default jitter in
    pm.Deterministic).  Has no log-density: the reference skips it with a
    warning when summing ln_prior (prior.py:366-376)."""

    name = "Constant"

    def __init__(self, value, unit=u.one):
        super().__init__(unit)
        self.value = float(value)

    def sample(self, rng, size, context=None):
        return np.full(size, self.value)

    def logp(self, x, context=None):
        return None


class FixedCompanionMass(Normal):
    """Normal prior on K whose width scales with P and e
    (distributions.py:195-244):
    sigma_K = clip(sigma_K0 (P/P0)^(-1/3) / sqrt(1 - e^2), 0, max_K).

    `sigma_K0`, `max_K` are speeds; `P0` a time.  Bare numbers are taken to be
    in km/s and days."""

    name = "FixedCompanionMass"

    def __init__(self, sigma_K0, P0, mu=0.0, max_K=500 * u.km / u.s, unit=None):
        sk = u.as_quantity(sigma_K0, u.km / u.s)
        if unit is None:
            unit = sk.unit
        Distribution.__init__(self, unit)
        self.mu = float(mu)
        self.sigma_K0 = sk
        self.P0 = u.as_quantity(P0, u.day)
        self.max_K = u.as_quantity(max_K, u.km / u.s)
        self.sigma = float("nan")

    def sigma_K(self, P_day, e):
        sk0 = self.sigma_K0.to_value(self.unit)
        mk = self.max_K.to_value(self.unit)
        P0 = self.P0.to_value(u.day)
        sig = sk0 * (np.asarray(P_day) / P0) ** (-1.0 / 3) / np.sqrt(1 - np.asarray(e) ** 2)
        return np.clip(sig, 0.0, mk)

    def sample(self, rng, size, context=None):
        if context is None or "P" not in context or "e" not in context:
            raise ValueError("FixedCompanionMass needs the P and e samples")
        return rng.normal(self.mu, self.sigma_K(context["P"], context["e"]), size=size)

    def logp(self, x, context=None):
        sig = self.sigma_K(context["P"], context["e"])
        z = (np.asarray(x) - self.mu) / sig
        return -0.5 * z * z - np.log(sig) - 0.5 * np.log(2 * np.pi)
