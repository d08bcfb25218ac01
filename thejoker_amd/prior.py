"""JokerPrior: the prior over The Joker's parameters (thejoker/prior.py).

Same constructor surface as the reference -- ``JokerPrior(pars, poly_trend,
v0_offsets)`` and ``JokerPrior.default(P_min, P_max, sigma_K0, P0, sigma_v, s,
poly_trend, v0_offsets, pars)`` -- with numpy distribution objects
(:mod:`thejoker_amd.distributions`) in place of pymc variables.  Every
parameter must carry a unit equivalent to the reference's
(prior.py:123-150) and the linear parameters must be Normal or
FixedCompanionMass (prior.py:152-170).
"""
import logging

import numpy as np

from . import distributions as dist
from . import units as u
from .prior_helpers import (get_linear_equiv_units, get_nonlinear_equiv_units,
                            get_v0_offsets_equiv_units, validate_poly_trend, validate_sigma_v)

__all__ = ["JokerPrior", "default_nonlinear_prior", "default_linear_prior"]

logger = logging.getLogger("thejoker")


class JokerPrior:
    def __init__(self, pars=None, poly_trend=1, v0_offsets=None, model=None):
        if pars is None:
            pars = {}
        elif isinstance(pars, dist.Distribution):
            raise ValueError("pars must be a dict mapping parameter names to distributions")
        else:
            pars = dict(pars)

        self.poly_trend, self._v_trend_names = validate_poly_trend(poly_trend)

        if v0_offsets is None:
            v0_offsets = []
        if hasattr(v0_offsets, "items"):
            v0_offsets = list(v0_offsets.values())
        try:
            v0_offsets = list(v0_offsets)
        except Exception as e:
            raise TypeError("Constant velocity offsets must be an iterable of distributions that "
                            "define the priors on each offset term.") from e
        self.v0_offsets = v0_offsets
        for i, p in enumerate(self.v0_offsets):
            pars[f"dv0_{i + 1}"] = p

        self._nonlinear_equiv_units = get_nonlinear_equiv_units()
        self._linear_equiv_units = get_linear_equiv_units(self.poly_trend)
        self._v0_offsets_equiv_units = get_v0_offsets_equiv_units(self.n_offsets)
        self._all_par_unit_equiv = {**self._nonlinear_equiv_units, **self._linear_equiv_units,
                                    **self._v0_offsets_equiv_units}

        for name in self.par_names:
            if name not in pars:
                raise ValueError(f"Missing prior for parameter '{name}': you must specify a prior "
                                 "distribution for all parameters.")
            p = pars[name]
            if not isinstance(p, dist.Distribution):
                raise ValueError(f"Parameter '{name}' must be a thejoker_amd.distributions "
                                 f"object, not {type(p)}")
            if not p.unit.is_equivalent(self._all_par_unit_equiv[name]):
                raise ValueError(f"Parameter '{name}' has an invalid unit: The units for this "
                                 "parameter must be transformable to "
                                 f"'{self._all_par_unit_equiv[name]}'")

        for name in list(self._linear_equiv_units) + list(self._v0_offsets_equiv_units):
            if not pars[name].is_normal:
                raise ValueError("Priors on the linear parameters (K, v0, etc.) must be "
                                 f"independent Normal distributions, not '{pars[name].name}' "
                                 f"(for {name})")
        self.pars = pars

    @classmethod
    def default(cls, P_min=None, P_max=None, sigma_K0=None, P0=1 * u.yr, sigma_v=None, s=None,
                poly_trend=1, v0_offsets=None, model=None, pars=None):
        """The default prior of prior.py:180-268: P log-uniform on (P_min,
        P_max), e ~ Kipping13Global, omega and M0 uniform angles, s fixed (0
        unless given), K ~ FixedCompanionMass(sigma_K0, P0), v_i Normal(0,
        sigma_v[i])."""
        nl = default_nonlinear_prior(P_min, P_max, s=s, pars=pars)
        lin = default_linear_prior(sigma_K0=sigma_K0, P0=P0, sigma_v=sigma_v,
                                   poly_trend=poly_trend, pars=pars)
        return cls(pars={**nl, **lin}, poly_trend=poly_trend, v0_offsets=v0_offsets)

    @property
    def par_names(self):
        return (list(self._nonlinear_equiv_units) + list(self._linear_equiv_units)
                + list(self._v0_offsets_equiv_units))

    @property
    def par_units(self):
        return {name: p.unit for name, p in self.pars.items()}

    @property
    def n_offsets(self):
        return len(self.v0_offsets)

    def __repr__(self):
        return f'<JokerPrior [{", ".join(self.par_names)}]>'

    def __str__(self):
        return ", ".join(self.par_names)

    def sample(self, size=1, generate_linear=False, return_logprobs=False, rng=None, dtype=None,
               **kwargs):
        """Draw prior samples (prior.py:297-407) from a numpy Generator.

        Draw order: P, e, omega, M0, s, then (with generate_linear) K, v0,
        offsets, v1, ...; pm.draw's internal order is not reproduced."""
        from .samples import JokerSamples

        if dtype is None:
            dtype = np.float64
        if rng is None:
            rng = np.random.default_rng()
        elif isinstance(rng, (int, np.integer)):
            rng = np.random.default_rng(rng)
        size = int(size)

        names = list(self._nonlinear_equiv_units)
        if generate_linear:
            names = names + list(self._linear_equiv_units) + list(self._v0_offsets_equiv_units)

        raw = {}
        context = {}
        for name in names:
            p = self.pars[name]
            vals = np.asarray(p.sample(rng, size, context=context))
            raw[name] = vals
            if name == "P":
                context["P"] = vals * p.unit.to(u.day)
            elif name == "e":
                context["e"] = vals

        samples = JokerSamples(poly_trend=self.poly_trend, n_offsets=self.n_offsets, **kwargs)
        for name in names:
            samples[name] = u.Quantity(raw[name].astype(dtype), self.pars[name].unit)

        if return_logprobs:
            logp = []
            for name in names:
                lp = self.pars[name].logp(raw[name], context=context)
                if lp is None:
                    logger.warning(f"Cannot auto-compute log-prior value for parameter {name}")
                    continue
                logp.append(lp)
            samples["ln_prior"] = np.sum(logp, axis=0)
        return samples


def default_nonlinear_prior(P_min=None, P_max=None, s=None, model=None, pars=None):
    """prior.py:410-493."""
    if pars is None:
        pars = {}
    if s is None:
        s = 0 * u.m / u.s
    if isinstance(s, dist.Distribution):
        pars = dict(pars)
        pars["s"] = pars.get("s", s)
    else:
        sq = u.as_quantity(s, u.m / u.s)
        if not sq.unit.is_equivalent(u.km / u.s):
            raise u.UnitsError("Invalid unit for s: must be equivalent to km/s")

    out = {}
    if "e" not in pars:
        out["e"] = dist.Kipping13Global()
    if "omega" not in pars:
        out["omega"] = dist.Angle()
    if "M0" not in pars:
        out["M0"] = dist.Angle()
    if "s" not in pars:
        out["s"] = dist.Constant(sq.value, sq.unit)
    if "P" not in pars:
        if P_min is None or P_max is None:
            raise ValueError("If you are using the default period prior, you must pass in both "
                             "P_min and P_max to set the period prior domain.")
        pmin = u.as_quantity(P_min, u.day)
        out["P"] = dist.UniformLog(pmin.value, u.as_quantity(P_max, u.day).to_value(pmin.unit),
                                   unit=pmin.unit)
    for k in pars:
        out[k] = pars[k]
    return out


def default_linear_prior(sigma_K0=None, P0=None, sigma_v=None, poly_trend=1, model=None,
                         pars=None):
    """prior.py:496-575."""
    if pars is None:
        pars = {}
    poly_trend, v_names = validate_poly_trend(poly_trend)
    if v_names and "v0" not in pars:
        sigma_v = validate_sigma_v(sigma_v, poly_trend, v_names)

    out = {}
    if "K" not in pars:
        if sigma_K0 is None or P0 is None:
            raise ValueError("If using the default prior form on K, you must pass in a variance "
                             "scale (sigma_K0) and a reference period (P0)")
        sk = u.as_quantity(sigma_K0, u.km / u.s)
        out["K"] = dist.FixedCompanionMass(sk, u.as_quantity(P0, u.day), unit=sk.unit)
    for name in v_names:
        if name not in pars:
            sv = u.as_quantity(sigma_v[name], u.km / u.s / u.day ** int(name[1:]))
            out[name] = dist.Normal(0.0, sv.value, unit=sv.unit)
    for k in pars.keys():
        out[k] = pars[k]
    return out
