"""Parameter names and unit equivalences of the prior (the helpers of
thejoker/prior_helpers.py:11-92).  The names, the unit table and the
exception types are the reference's interface; the messages are this
package's own."""
from . import units as u

__all__ = ["get_nonlinear_equiv_units", "validate_poly_trend", "get_linear_equiv_units",
           "validate_sigma_v", "validate_n_offsets", "get_v0_offsets_equiv_units"]


def get_nonlinear_equiv_units():
    """Units each nonlinear parameter must be convertible to
    (prior_helpers.py:12-19)."""
    return {"P": u.day, "e": u.one, "omega": u.rad, "M0": u.rad, "s": u.m / u.s}


def validate_poly_trend(poly_trend):
    """(poly_trend as an int, trend parameter names v0 .. v{poly_trend-1})
    (prior_helpers.py:23-31)."""
    try:
        n = int(poly_trend)
    except Exception as exc:
        raise ValueError(f"poly_trend={poly_trend!r}: expected an integer count of polynomial "
                         "velocity-trend terms (1: a constant v0)") from exc
    return n, [f"v{i}" for i in range(n)]


def get_linear_equiv_units(poly_trend):
    """Units of K and the trend parameters v_i, RV / day^i
    (prior_helpers.py:34-39)."""
    _, names = validate_poly_trend(poly_trend)
    units = {"K": u.m / u.s}
    for power, name in enumerate(names):
        units[name] = u.m / u.s / u.day ** power
    return units


def _is_quantity(x):
    return isinstance(x, u.Quantity) or (hasattr(x, "unit") and hasattr(x, "to_value"))


def validate_sigma_v(sigma_v, poly_trend, v_names):
    """The prior widths of the trend parameters as a dict keyed by v_names
    (prior_helpers.py:42-73): a scalar Quantity (v0 only), a dict with every
    name, or a sequence with one value per trend term."""
    if _is_quantity(sigma_v):
        q = u.as_quantity(sigma_v, u.km / u.s)
        if not q.isscalar:
            raise ValueError("a single sigma_v Quantity must be a scalar (the width of the v0 "
                             "prior); give a sequence or a dict for several trend terms")
        sigma_v = {"v0": q}
    if hasattr(sigma_v, "keys"):
        missing = [name for name in v_names if name not in sigma_v.keys()]
        if missing:
            raise ValueError(f"sigma_v has no width for {missing}; a dict must hold one for "
                             f"every trend parameter {v_names} (got keys {list(sigma_v.keys())})")
        return sigma_v
    try:
        n_given = len(sigma_v)
    except TypeError as exc:
        raise TypeError(f"sigma_v of type {type(sigma_v)} is not usable: give a scalar Quantity "
                        "(poly_trend=1), a sequence of Quantities or a dict keyed by "
                        f"{v_names}") from exc
    if n_given != poly_trend:
        raise ValueError(f"sigma_v holds {n_given} widths but poly_trend={poly_trend} needs one "
                         "per trend term")
    return dict(zip(v_names, sigma_v))


def validate_n_offsets(n_offsets):
    """(n_offsets as an int, offset names dv0_1 .. dv0_{n})
    (prior_helpers.py:77-87)."""
    try:
        n = int(n_offsets)
    except Exception as exc:
        raise ValueError(f"n_offsets={n_offsets!r}: expected an integer count of v0 offset "
                         "parameters (one per survey beyond the first)") from exc
    return n, [f"dv0_{i}" for i in range(1, n + 1)]


def get_v0_offsets_equiv_units(n_offsets):
    """Units of the v0 offsets, all RV (prior_helpers.py:90-92)."""
    _, names = validate_n_offsets(n_offsets)
    return dict.fromkeys(names, u.m / u.s)
