"""A small unit system for the drop-in API (astropy is not a dependency).

The reference tags every prior and data column with an astropy unit
(thejoker/units.py:10-26, prior_helpers.py:12-39).  This module provides the
subset the likelihood path needs -- time, speed, angle and their products --
with the same spelling (``u.day``, ``u.km / u.s``, ``u.m / u.s / u.day``):

    >>> from thejoker_amd import units as u
    >>> (30 * u.km / u.s).to_value(u.m / u.s)
    30000.0

If astropy is importable, astropy Quantities are accepted wherever a
Quantity is (they are converted through their SI value).
"""
import numbers
import re

import numpy as np

__all__ = ["Unit", "Quantity", "as_quantity", "to_value", "day", "d", "yr", "year", "s", "m",
           "km", "rad", "radian", "deg", "one", "dimensionless_unscaled", "au", "hour", "minute",
           "parse_unit", "UnitsError"]


class Unit:
    """scale x (m^a s^b rad^c); `dims` = (a, b, c)."""

    __slots__ = ("scale", "dims", "name")
    # numpy defers binary operators to the Unit (array * u.km -> Quantity).
    __array_ufunc__ = None

    def __init__(self, scale, dims, name=None):
        self.scale = float(scale)
        self.dims = tuple(int(x) for x in dims)
        self.name = name

    def _combine(self, other, sign):
        other = _as_unit(other)
        dims = tuple(a + sign * b for a, b in zip(self.dims, other.dims))
        scale = self.scale * other.scale ** sign
        name = None
        if self.name is not None and other.name is not None:
            if sign > 0:
                name = f"{self.name} {other.name}".strip()
            else:
                name = f"{self.name} / {other.name}" if self.name else f"1 / {other.name}"
        return Unit(scale, dims, name)

    def __mul__(self, other):
        if isinstance(other, (Unit,)):
            return self._combine(other, 1)
        if isinstance(other, Quantity):
            return Quantity(other.value, self * other.unit)
        return Quantity(other, self)

    __rmul__ = __mul__

    def __truediv__(self, other):
        if isinstance(other, Unit):
            return self._combine(other, -1)
        return NotImplemented

    def __rtruediv__(self, other):
        inv = Unit(1.0 / self.scale, tuple(-x for x in self.dims),
                   f"1 / {self.name}" if self.name else None)
        if isinstance(other, numbers.Number) or isinstance(other, np.ndarray):
            return Quantity(other, inv)
        return NotImplemented

    def __pow__(self, p):
        if p == 0:
            return one
        if p == 1:
            return self
        return Unit(self.scale ** p, tuple(x * p for x in self.dims),
                    f"{self.name}{p}" if self.name else None)

    def is_equivalent(self, other):
        return self.dims == _as_unit(other).dims

    def to(self, other):
        """Conversion factor from this unit to `other`."""
        other = _as_unit(other)
        if not self.is_equivalent(other):
            raise UnitsError(f"'{self}' is not convertible to '{other}'")
        return self.scale / other.scale

    def __eq__(self, other):
        try:
            other = _as_unit(other)
        except (TypeError, UnitsError):
            return False
        return self.dims == other.dims and np.isclose(self.scale, other.scale, rtol=1e-15, atol=0)

    def __hash__(self):
        return hash((round(self.scale, 12), self.dims))

    def __repr__(self):
        return f"Unit({self})"

    def __str__(self):
        return self.name if self.name is not None else f"{self.scale:g} m^{self.dims[0]} s^{self.dims[1]} rad^{self.dims[2]}"


class UnitsError(ValueError):
    pass


class Quantity:
    """A value (scalar or numpy array) with a Unit."""

    __array_priority__ = 1000
    __array_ufunc__ = None

    def __init__(self, value, unit=None):
        if isinstance(value, Quantity):
            if unit is not None:
                value = value.to_value(unit)
            else:
                value, unit = value.value, value.unit
        self.value = np.asarray(value, dtype=np.float64) if not np.isscalar(value) else float(value)
        self.unit = _as_unit(unit) if unit is not None else one

    def to_value(self, unit=None):
        if unit is None:
            return self.value
        return self.value * self.unit.to(unit)

    def to(self, unit):
        return Quantity(self.to_value(unit), _as_unit(unit))

    def __mul__(self, other):
        if isinstance(other, Unit):
            return Quantity(self.value, self.unit * other)
        if isinstance(other, Quantity):
            return Quantity(self.value * other.value, self.unit * other.unit)
        return Quantity(self.value * np.asarray(other), self.unit)

    __rmul__ = __mul__

    def __truediv__(self, other):
        if isinstance(other, Unit):
            return Quantity(self.value, self.unit / other)
        if isinstance(other, Quantity):
            return Quantity(self.value / other.value, self.unit / other.unit)
        return Quantity(self.value / np.asarray(other), self.unit)

    def __neg__(self):
        return Quantity(-self.value, self.unit)

    def _same(self, other):
        if isinstance(other, Quantity):
            return other.to_value(self.unit)
        if self.unit.dims == (0, 0, 0):
            return np.asarray(other) / self.unit.scale
        raise UnitsError("cannot combine a Quantity with a bare number")

    def __add__(self, other):
        return Quantity(self.value + self._same(other), self.unit)

    __radd__ = __add__

    def __sub__(self, other):
        return Quantity(self.value - self._same(other), self.unit)

    def __lt__(self, other):
        return self.value < self._same(other)

    def __gt__(self, other):
        return self.value > self._same(other)

    def __getitem__(self, idx):
        return Quantity(np.asarray(self.value)[idx], self.unit)

    def __len__(self):
        return len(np.atleast_1d(self.value))

    def __array__(self, dtype=None, copy=None):
        return np.asarray(self.value, dtype=dtype)

    @property
    def shape(self):
        return np.shape(self.value)

    @property
    def ndim(self):
        return np.ndim(self.value)

    @property
    def size(self):
        return np.size(self.value)

    @property
    def isscalar(self):
        return np.ndim(self.value) == 0

    def __repr__(self):
        return f"<Quantity {self.value!r} {self.unit}>"


def _as_unit(u):
    if isinstance(u, Unit):
        return u
    if isinstance(u, str):
        key = u.strip()
        if key in _NAMED:
            return _NAMED[key]
        return parse_unit(key)
    # astropy unit (duck-typed): go through SI.
    si = getattr(u, "si", None)
    if si is not None and hasattr(si, "scale") and hasattr(si, "bases"):
        dims = [0, 0, 0]
        for base, power in zip(si.bases, si.powers):
            bname = getattr(base, "name", str(base))
            idx = {"m": 0, "s": 1, "rad": 2}.get(bname)
            if idx is None:
                raise UnitsError(f"unsupported base unit {bname}")
            dims[idx] += int(power)
        return Unit(si.scale, dims, str(u))
    if u is None:
        return one
    raise TypeError(f"not a unit: {u!r}")


def as_quantity(x, default_unit):
    """Quantity from a Quantity, an astropy Quantity, or a bare number (taken
    to be in `default_unit`)."""
    if isinstance(x, Quantity):
        return x
    if hasattr(x, "unit") and hasattr(x, "value") and hasattr(x, "to_value"):
        return Quantity(np.asarray(x.value), _as_unit(x.unit))
    return Quantity(x, default_unit)


def to_value(x, unit):
    """Value of `x` in `unit`; bare numbers are taken to be in `unit`."""
    return as_quantity(x, unit).to_value(unit)


one = Unit(1.0, (0, 0, 0), "")
dimensionless_unscaled = one
m = Unit(1.0, (1, 0, 0), "m")
km = Unit(1000.0, (1, 0, 0), "km")
au = Unit(1.495978707e11, (1, 0, 0), "AU")
s = Unit(1.0, (0, 1, 0), "s")
minute = Unit(60.0, (0, 1, 0), "min")
hour = Unit(3600.0, (0, 1, 0), "h")
day = Unit(86400.0, (0, 1, 0), "d")
d = day
yr = Unit(365.25 * 86400.0, (0, 1, 0), "yr")
year = yr
rad = Unit(1.0, (0, 0, 1), "rad")
radian = rad
deg = Unit(np.pi / 180.0, (0, 0, 1), "deg")

_NAMED = {
    "": one, "m": m, "km": km, "AU": au, "s": s, "min": minute, "h": hour, "d": day,
    "day": day, "yr": yr, "year": yr, "rad": rad, "deg": deg, "au": au, "dimensionless": one,
    "km / s": km / s, "m / s": m / s,
}

_UNIT_TOKEN = re.compile(r"\s*(?:(\*\*|\^|\*|\.|/|\(|\))|([A-Za-z_]+)([-+]?\d+)?|"
                         r"([-+]?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?))")


def parse_unit(text):
    """Unit from a string in astropy's generic format, as astropy writes them
    into table headers: ``'km / s'``, ``'km / (d s)'``, ``'m / (d2 s)'``,
    ``'s-1'``, ``'m s**-2'``.  Products are juxtaposition, ``*`` or ``.``;
    ``/`` divides by the next factor; a power follows a name directly
    (``d2``) or after ``**`` / ``^``."""
    text = str(text).strip()
    if text in _NAMED:
        return _NAMED[text]
    toks, pos = [], 0
    while pos < len(text):
        mt = _UNIT_TOKEN.match(text, pos)
        if mt is None or mt.end() == pos:
            if text[pos:].strip() == "":
                break
            raise UnitsError(f"cannot parse unit '{text}'")
        toks.append(mt.groups())
        pos = mt.end()
    i = 0

    def peek():
        return toks[i] if i < len(toks) else (None, None, None, None)

    def number():
        nonlocal i
        op, name, pw, num = peek()
        if op == "(":
            i += 1
            val = number()
            if peek()[0] != ")":
                raise UnitsError(f"cannot parse unit '{text}'")
            i += 1
            return val
        if num is None:
            raise UnitsError(f"cannot parse unit '{text}'")
        i += 1
        return float(num)

    def factor():
        nonlocal i
        op, name, pw, num = peek()
        if op == "(":
            i += 1
            val = product()
            if peek()[0] != ")":
                raise UnitsError(f"cannot parse unit '{text}'")
            i += 1
        elif name is not None:
            if name not in _NAMED:
                raise UnitsError(f"unknown unit '{name}' in '{text}'")
            i += 1
            val = _NAMED[name]
            if pw is not None:
                val = val ** _int_power(float(pw))
        elif num is not None:
            i += 1
            val = Unit(float(num), (0, 0, 0), num)
        else:
            raise UnitsError(f"cannot parse unit '{text}'")
        if peek()[0] in ("**", "^"):
            i += 1
            val = val ** _int_power(number())
        return val

    def product():
        nonlocal i
        val = factor()
        while i < len(toks):
            op = peek()[0]
            if op == ")":
                break
            if op == "/":
                i += 1
                val = val / factor()
            elif op in ("*", "."):
                i += 1
                val = val * factor()
            else:
                val = val * factor()
        return val

    val = product()
    if i != len(toks):
        raise UnitsError(f"cannot parse unit '{text}'")
    return Unit(val.scale, val.dims, text)


def _int_power(p):
    if p != int(p):
        raise UnitsError(f"fractional unit power {p} is not supported")
    return int(p)
