"""Parameter names and unit equivalences (thejoker/prior_helpers.py:11-92)."""
from . import units as u

__all__ = ["get_nonlinear_equiv_units", "validate_poly_trend", "get_linear_equiv_units",
           "validate_sigma_v", "validate_n_offsets", "get_v0_offsets_equiv_units"]


def get_nonlinear_equiv_units():
    """prior_helpers.py:12-19."""
    return {"P": u.day, "e": u.one, "omega": u.rad, "M0": u.rad, "s": u.m / u.s}


def validate_poly_trend(poly_trend):
    """prior_helpers.py:23-31."""
    try:
        poly_trend = int(poly_trend)
    except Exception as e:
        raise ValueError("poly_trend must be an integer that specifies the number of polynomial "
                         "(in time) trend terms to include in The Joker.") from e
    return poly_trend, [f"v{i}" for i in range(poly_trend)]


def get_linear_equiv_units(poly_trend):
    """prior_helpers.py:34-39."""
    poly_trend, v_names = validate_poly_trend(poly_trend)
    return {"K": u.m / u.s, **{name: u.m / u.s / u.day ** i for i, name in enumerate(v_names)}}


def validate_sigma_v(sigma_v, poly_trend, v_names):
    """prior_helpers.py:42-73."""
    if isinstance(sigma_v, u.Quantity) or (hasattr(sigma_v, "unit") and hasattr(sigma_v, "to_value")):
        q = u.as_quantity(sigma_v, u.km / u.s)
        if not q.isscalar:
            raise ValueError("You must pass in a scalar value for sigma_v if passing in a single "
                             "quantity.")
        sigma_v = {"v0": q}
    if hasattr(sigma_v, "keys"):
        for name in v_names:
            if name not in sigma_v.keys():
                raise ValueError("If specifying the standard-deviations of the polynomial trend "
                                 "parameter prior, you must pass in values for all parameter "
                                 f"names. Expected keys: {v_names}, received: {sigma_v.keys()}")
        return sigma_v
    try:
        if len(sigma_v) != poly_trend:
            raise ValueError("You must pass in a single sigma value for each velocity trend "
                             f"parameter: You passed in {len(sigma_v)} values, but "
                             f"poly_trend={poly_trend}")
        sigma_v = {name: val for name, val in zip(v_names, sigma_v)}
    except TypeError as e:
        raise TypeError("Invalid input for velocity trend prior sigma values. This must either be "
                        "a scalar Quantity (if poly_trend=1) or an iterable of Quantity objects "
                        "(if poly_trend>1)") from e
    return sigma_v


def validate_n_offsets(n_offsets):
    """prior_helpers.py:77-87."""
    try:
        n_offsets = int(n_offsets)
    except Exception as e:
        raise ValueError("n_offsets must be an integer that specifies the number of v0 offset "
                         "parameters to include in The Joker.") from e
    return n_offsets, [f"dv0_{i}" for i in range(1, n_offsets + 1)]


def get_v0_offsets_equiv_units(n_offsets):
    """prior_helpers.py:90-92."""
    n_offsets, names = validate_n_offsets(n_offsets)
    return {name: u.m / u.s for name in names}
