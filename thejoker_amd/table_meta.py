"""The YAML table header stored beside a table of samples.

The reference writes its tables with astropy's serialised header
(samples_helpers.py:204-211 via ``astropy.table.meta.get_yaml_from_table``)
and reads column units back from it (utils.py:75-87) and the table meta
(samples.py:546-562).  The header is YAML with three astropy tags:
``!astropy.units.Unit {unit: km / s}``, ``!astropy.table.SerializedColumn
{name: P}`` and ``!astropy.time.Time {format: mjd, jd1: ..., jd2: ...,
scale: tcb, ...}``; the ECSV files under docs/examples hold the same header
behind ``# `` prefixes.  This module reads and writes that header with
PyYAML's safe loader and the package's own units (astropy is not needed).
"""
import math
import warnings

import yaml

from . import units as u

__all__ = ["read_header", "column_units", "table_meta", "make_header", "unit_string", "TimeTag"]

_MJD0 = 2400000.5


class TimeTag(dict):
    """An ``!astropy.time.Time`` mapping (jd1, jd2, format, scale, ...)."""

    @property
    def mjd(self):
        """The time as an MJD in its own scale: (jd1 - 2400000.5) + jd2."""
        return (float(self["jd1"]) - _MJD0) + float(self["jd2"])


class _Loader(yaml.SafeLoader):
    pass


def _unit(loader, node):
    return u.parse_unit(loader.construct_mapping(node).get("unit", ""))


def _time(loader, node):
    return TimeTag(loader.construct_mapping(node, deep=True))


def _serialized_column(loader, node):
    return dict(loader.construct_mapping(node))


def _other_tag(loader, suffix, node):
    if isinstance(node, yaml.MappingNode):
        return loader.construct_mapping(node, deep=True)
    if isinstance(node, yaml.SequenceNode):
        return loader.construct_sequence(node, deep=True)
    return loader.construct_scalar(node)


_Loader.add_constructor("!astropy.units.Unit", _unit)
_Loader.add_constructor("!astropy.time.Time", _time)
_Loader.add_constructor("!astropy.table.SerializedColumn", _serialized_column)
_Loader.add_multi_constructor("!", _other_tag)


def read_header(lines):
    """The header dict from its lines (bytes or str).  ECSV comment prefixes
    (``# ``), the ``%ECSV`` version line and the ``---`` marker are
    skipped, so an ECSV file's first lines parse as well."""
    text = []
    for ln in lines:
        if isinstance(ln, bytes):
            ln = ln.decode("utf-8")
        if ln.startswith("#"):
            ln = ln[2:] if ln.startswith("# ") else ln[1:]
        if ln.startswith("%ECSV") or ln.strip() == "---":
            continue
        text.append(ln)
    header = yaml.load("\n".join(text), Loader=_Loader) or {}
    meta = header.get("meta", {})
    if isinstance(meta, list):  # !!omap: a list of (key, value) pairs
        meta = dict(meta)
    header["meta"] = meta or {}
    return header


def column_units(header):
    """{column name: Unit} (utils.py:75-87): a column with no unit is
    dimensionless; a serialised Quantity's unit is used when the datatype
    entry carries none."""
    ser = header["meta"].get("__serialized_columns__", {}) or {}
    out = {}
    for row in header.get("datatype", []):
        name = row["name"]
        unit = row.get("unit")
        if unit is None and isinstance(ser.get(name), dict) and "unit" in ser[name]:
            unit = ser[name]["unit"]
        out[name] = u.one if unit is None else (unit if isinstance(unit, u.Unit)
                                               else u.parse_unit(unit))
    return out


def table_meta(header):
    """The table meta without serialisation entries; an ``!astropy.time.Time``
    becomes its MJD (days) in its own scale (JokerSamples keeps t_ref as a
    BMJD in days, TCB as the reference's data)."""
    out = {}
    for k, v in header["meta"].items():
        if k == "__serialized_columns__":
            continue
        if isinstance(v, TimeTag):
            scale = v.get("scale", "tcb")
            if scale != "tcb":
                warnings.warn(f"meta '{k}' is a time in scale '{scale}'; kept as an MJD in that "
                              "scale", RuntimeWarning)
            v = v.mjd
        out[k] = v
    return out


def _scalar(v):
    if hasattr(v, "item") and getattr(v, "ndim", 1) == 0:
        v = v.item()                      # numpy scalars
    if v is None:
        return "null"
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, int):
        return str(v)
    if isinstance(v, float):
        if math.isnan(v):
            return ".nan"
        if math.isinf(v):
            return ".inf" if v > 0 else "-.inf"
        return repr(v)
    return yaml.safe_dump(v, default_flow_style=True).strip().removesuffix("...").strip()


def _time_tag(mjd):
    jd1 = float(round(float(mjd) + _MJD0))
    jd2 = float(mjd) - (jd1 - _MJD0)
    return ("!astropy.time.Time {format: mjd, in_subfmt: '*', "
            f"jd1: {jd1!r}, jd2: {jd2!r}, out_subfmt: '*', precision: 3, scale: tcb}}")


def unit_string(unit):
    """A unit in astropy's generic format ('' when dimensionless)."""
    if unit.dims == (0, 0, 0) and unit.scale == 1.0:
        return ""
    if unit.name:
        return unit.name
    parts = [repr(unit.scale)] if unit.scale != 1.0 else []
    parts += [f"{b}{p}" if p != 1 else b for b, p in zip(("m", "s", "rad"), unit.dims) if p]
    return " ".join(parts)


def make_header(columns, meta, time_keys=("t_ref",)):
    """Header lines for float64 columns [(name, Unit)] and a meta dict, in
    the layout astropy gives a QTable of Quantity columns: the datatype list,
    an ordered meta mapping (times in `time_keys` as ``!astropy.time.Time``
    in MJD/TCB) with ``__serialized_columns__``, and the schema line."""
    lines = ["datatype:"]
    for name, unit in columns:
        us = unit_string(unit)
        lines.append(f"- {{name: {name}, unit: {us}, datatype: float64}}" if us
                     else f"- {{name: {name}, datatype: float64}}")
    lines.append("meta: !!omap")
    for k, v in meta.items():
        if k in time_keys and v is not None:
            lines.append(f"- {k}: {_time_tag(v)}")
        else:
            lines.append(f"- {{{k}: {_scalar(v)}}}")
    lines.append("- __serialized_columns__:")
    for name, unit in columns:
        us = unit_string(unit) or "''"
        lines += [f"    {name}:",
                  "      __class__: astropy.units.quantity.Quantity",
                  f"      unit: !astropy.units.Unit {{unit: {us}}}",
                  f"      value: !astropy.table.SerializedColumn {{name: {name}}}"]
    lines.append("schema: astropy-2.0")
    return lines
