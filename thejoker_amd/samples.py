"""JokerSamples: table of prior or posterior samples (thejoker/samples.py).

Columns are :class:`thejoker_amd.units.Quantity` arrays keyed by parameter
name; table metadata holds ``t_ref`` (BMJD days or None), ``poly_trend`` and
``n_offsets`` (samples.py:36-104).  ``pack``/``unpack`` follow
samples.py:404-478: nonlinear parameters pack in the order P, e, omega, M0, s
(fast_likelihood.pyx:41) with internal units day, 1, rad, rad
(fast_likelihood.pyx:42-45).
"""
import os
from collections import OrderedDict

import numpy as np

from . import units as u
from .prior_helpers import (get_linear_equiv_units, get_nonlinear_equiv_units,
                            get_v0_offsets_equiv_units, validate_n_offsets, validate_poly_trend)

__all__ = ["JokerSamples", "NONLINEAR_PACKED_ORDER", "NONLINEAR_INTERNAL_UNITS"]

NONLINEAR_PACKED_ORDER = ["P", "e", "omega", "M0", "s"]
NONLINEAR_INTERNAL_UNITS = {"P": u.day, "e": u.one, "omega": u.rad, "M0": u.rad}


class JokerSamples:
    _hdf5_path = "samples"

    def __init__(self, samples=None, t_ref=None, n_offsets=None, poly_trend=None, **kwargs):
        if poly_trend is None:
            poly_trend = 1
        if n_offsets is None:
            n_offsets = 0
        if isinstance(samples, JokerSamples):
            meta = dict(samples.meta)
            t_ref = meta.pop("t_ref", t_ref)
            poly_trend = meta.pop("poly_trend", poly_trend)
            n_offsets = meta.pop("n_offsets", n_offsets)
            kwargs.update(meta)
            samples = samples.tbl
        poly_trend, _ = validate_poly_trend(poly_trend)
        n_offsets, _ = validate_n_offsets(n_offsets)

        valid = {**get_nonlinear_equiv_units(), **get_linear_equiv_units(poly_trend),
                 **get_v0_offsets_equiv_units(n_offsets)}
        valid["ln_prior"] = u.one
        valid["ln_likelihood"] = u.one
        valid["ln_posterior"] = u.one
        self._valid_units = valid

        self.tbl = OrderedDict()
        self.meta = {"t_ref": t_ref, "poly_trend": poly_trend, "n_offsets": n_offsets}
        self.meta.update(kwargs)
        if samples is not None:
            for name, col in dict(samples).items():
                self[name] = col
        self._cache = {}

    # -- table access -------------------------------------------------------
    def __getitem__(self, key):
        if isinstance(key, str):
            return self.tbl[key]
        if isinstance(key, (int, np.integer)):
            key = slice(int(key), int(key) + 1 if int(key) != -1 else None)
        new = self.__class__(t_ref=self.t_ref, poly_trend=self.poly_trend,
                             n_offsets=self.n_offsets,
                             **{k: v for k, v in self.meta.items()
                                if k not in ("t_ref", "poly_trend", "n_offsets")})
        for name, col in self.tbl.items():
            new.tbl[name] = u.Quantity(np.atleast_1d(np.asarray(col.value)[key]), col.unit)
        return new

    def __setitem__(self, key, val):
        if key not in self._valid_units:
            raise ValueError(f"Invalid parameter name '{key}'. Must be one of: "
                             f"{list(self._valid_units.keys())}")
        expected = self._valid_units[key]
        if isinstance(val, u.Quantity) or (hasattr(val, "unit") and hasattr(val, "to_value")):
            q = u.as_quantity(val, expected)
        else:
            q = u.Quantity(np.asarray(val, dtype=np.float64), u.one)
        if not q.unit.is_equivalent(expected):
            raise u.UnitsError(f"Units of '{key}' must be convertable to {expected}")
        q = u.Quantity(np.atleast_1d(np.asarray(q.value, dtype=np.float64)), q.unit)
        if self.tbl and len(q) != len(self):
            raise ValueError(f"column '{key}' has {len(q)} rows, the table {len(self)}")
        self.tbl[key] = q

    def __contains__(self, key):
        return key in self.tbl

    @property
    def t_ref(self):
        return self.meta["t_ref"]

    @property
    def poly_trend(self):
        return self.meta["poly_trend"]

    @property
    def n_offsets(self):
        return self.meta["n_offsets"]

    @property
    def par_names(self):
        return list(self.tbl.keys())

    def __len__(self):
        if not self.tbl:
            return 0
        return len(next(iter(self.tbl.values())))

    def __repr__(self):
        return f'<JokerSamples [{", ".join(self.par_names)}] ({len(self)} samples)>'

    __str__ = __repr__

    @property
    def isscalar(self):
        return len(self) == 1

    # -- derived ------------------------------------------------------------
    def get_time_with_phase(self, phase=0.0, t_ref=None):
        """BMJD time at which the orbit has the given phase (samples.py:199-233):
        t_ref + P M0 / 2 pi + P phase / 2 pi."""
        if t_ref is None:
            if self.t_ref is None:
                raise ValueError("This samples object has no reference time t_ref, so you must "
                                 "pass in the reference time via t_ref")
            t_ref = self.t_ref
        elif self.t_ref is not None:
            raise ValueError("You passed in a reference time t_ref, but this samples object "
                             "already has a reference time.")
        P = self["P"].to_value(u.day)
        M0 = self["M0"].to_value(u.rad)
        ph = u.to_value(phase, u.rad)
        return np.squeeze(t_ref + P * M0 / (2 * np.pi) + P * ph / (2 * np.pi))

    def get_t0(self, t_ref=None):
        return self.get_time_with_phase(0.0, t_ref=t_ref)

    def _apply(self, func):
        new = self.__class__(**self.meta)
        for k, col in self.tbl.items():
            new.tbl[k] = u.Quantity(np.atleast_1d(func(np.asarray(col.value))), col.unit)
        return new

    def mean(self):
        return self._apply(np.mean)

    def std(self):
        return self._apply(np.std)

    def median_period(self):
        P = np.asarray(self["P"].value)
        idx = np.argpartition(P, len(P) // 2)[len(P) // 2]
        return self[int(idx)]

    def wrap_K(self):
        """samples.py:391-401."""
        K = np.asarray(self.tbl["K"].value)
        mask = K < 0
        if np.any(mask):
            om = self.tbl["omega"]
            omv = np.array(om.to_value(u.rad))
            K = K.copy()
            K[mask] = np.abs(K[mask])
            omv[mask] = (omv[mask] + np.pi) % (2 * np.pi)
            self.tbl["K"] = u.Quantity(K, self.tbl["K"].unit)
            self.tbl["omega"] = u.Quantity(omv * u.rad.to(om.unit), om.unit)
        return self

    def copy(self):
        new = self.__class__(**self.meta)
        for k, col in self.tbl.items():
            new.tbl[k] = u.Quantity(np.array(col.value, copy=True), col.unit)
        return new

    # -- packing (samples.py:404-478) ---------------------------------------
    def pack(self, units=None, names=None, nonlinear_only=True):
        if units is None:
            units = dict()
        units = dict(units)
        out_units = OrderedDict()
        for k, v in NONLINEAR_INTERNAL_UNITS.items():
            units.setdefault(k, v)
        if names is None:
            names = NONLINEAR_PACKED_ORDER if nonlinear_only else self.par_names
        arrs = []
        for name in names:
            col = self.tbl[name]
            unit = u._as_unit(units.get(name, col.unit))
            arrs.append(col.to_value(unit))
            out_units[name] = unit
        return np.stack(arrs, axis=1), out_units

    @classmethod
    def unpack(cls, packed_samples, units, **kwargs):
        packed_samples = np.array(packed_samples)
        nsamples, npars = packed_samples.shape
        samples = cls(**kwargs)
        for i, k in enumerate(list(units.keys())[:npars]):
            samples[k] = u.Quantity(packed_samples[:, i], units[k])
        return samples

    # -- I/O ---------------------------------------------------------------
    def write(self, output, overwrite=False, append=False):
        """Write to an HDF5 (``.hdf5``/``.h5``) or ``.npz`` file
        (samples.py:480-545).

        HDF5 files have the reference's layout (samples_helpers.py:35-272):
        the compound dataset ``samples`` plus ``samples.__table_column_meta__``
        holding the astropy YAML header, so the reference's readers open
        them.  As there, an existing file raises OSError unless `overwrite`
        (replace it) or `append` (extend its table; the meta and the column
        names and units must match, else ValueError).  FITS needs astropy
        and is not supported."""
        ext = os.path.splitext(output)[1]
        if ext == ".fits":
            raise NotImplementedError("FITS output needs astropy; write .hdf5, .h5 or .npz")
        if ext not in (".hdf5", ".h5", ".npz"):
            raise NotImplementedError("We currently only support writing to HDF5 files, with "
                                      "extension .hdf5 or .h5, or .npz files.")
        if os.path.exists(output) and not append:
            if not overwrite:
                raise OSError(f"File exists: {output}")
            os.remove(output)
        src = self
        if append and os.path.exists(output):
            src = JokerSamples.read(output)._appended(self)
        if ext == ".npz":
            src._write_npz(output)
            return
        from . import hdf5, table_meta
        cols = [(k, v.unit) for k, v in src.tbl.items()]
        header = table_meta.make_header(cols, src.meta)
        hdf5.write_table(output, [(k, np.asarray(v.value, dtype=np.float64).ravel())
                                  for k, v in src.tbl.items()], header, name=self._hdf5_path)

    def _appended(self, other):
        """self followed by other's rows (samples_helpers.py:159-243 checks)."""
        for key in set(self.meta) | set(other.meta):
            a, b = self.meta.get(key), other.meta.get(key)
            same = (a is None and b is None) or (a is not None and b is not None and (
                a == b or (isinstance(a, float) and isinstance(b, float) and np.isnan(a)
                           and np.isnan(b))))
            if not same:
                raise ValueError("Cannot append table to existing file because the existing file "
                                 f"table metadata and this object's metadata do not match ({key}: "
                                 f"{a!r} vs. {b!r})")
        if list(self.tbl) != list(other.tbl) or any(
                not self.tbl[k].unit.is_equivalent(other.tbl[k].unit) for k in self.tbl):
            raise ValueError("Cannot append table to existing file because the existing file "
                             "table datatype and this object's table datatype do not match")
        out = self.copy()
        for k in self.tbl:
            out.tbl[k] = u.Quantity(np.concatenate([self.tbl[k].value,
                                                    other.tbl[k].to_value(self.tbl[k].unit)]),
                                    self.tbl[k].unit)
        return out

    def _write_npz(self, output):
        arrays = {f"col__{k}": np.asarray(v.value) for k, v in self.tbl.items()}
        unit_rec = np.array([[k, repr(v.unit.scale), ",".join(map(str, v.unit.dims)),
                              v.unit.name or ""] for k, v in self.tbl.items()], dtype=object)
        meta = {k: (np.nan if v is None else v) for k, v in self.meta.items()}
        np.savez(output, **arrays, __units__=unit_rec.astype(str),
                 __meta_keys__=np.array(list(meta.keys()), dtype=str),
                 __meta_vals__=np.array([float(v) for v in meta.values()]))

    @classmethod
    def read(cls, filename, path=None):
        """Read samples written by :meth:`write` or by the reference
        (samples.py:564-601): HDF5 (``.hdf5``/``.h5``, the table at `path`,
        default ``samples``) or ``.npz``."""
        ext = os.path.splitext(filename)[1]
        if ext in (".hdf5", ".h5"):
            return cls._read_hdf5(filename, path or cls._hdf5_path)
        if ext != ".npz":
            raise NotImplementedError(f"cannot read samples from '{ext}' files")
        with np.load(filename, allow_pickle=False) as f:
            units = {row[0]: u.Unit(float(row[1]), [int(x) for x in row[2].split(",")],
                                    row[3] or None) for row in f["__units__"]}
            meta = {}
            for k, v in zip(f["__meta_keys__"], f["__meta_vals__"]):
                meta[str(k)] = None if np.isnan(v) else (int(v) if k in ("poly_trend", "n_offsets")
                                                          else float(v))
            out = cls(**meta)
            for k in units:
                out.tbl[k] = u.Quantity(f[f"col__{k}"], units[k])
        return out

    @classmethod
    def _read_hdf5(cls, filename, path):
        from . import hdf5
        with hdf5.File(filename) as f:
            header, units = read_hdf5_header(f, path)
            data = f[path].read()
        meta = table_meta_of(header)
        out = cls(**meta)
        for name in data.dtype.names:
            out[name] = u.Quantity(np.ascontiguousarray(data[name], dtype=np.float64),
                                   units.get(name, u.one))
        return out


def read_hdf5_header(f, path="samples"):
    """(header dict, {column: Unit}) of the table at `path` of an open
    :class:`thejoker_amd.hdf5.File` (utils.py:75-87).  A table written
    without its header has dimensionless columns and an empty meta."""
    from . import table_meta
    mpath = path + ".__table_column_meta__"
    if mpath not in f:
        return {"meta": {}}, {}
    header = table_meta.read_header(f[mpath].read())
    return header, table_meta.column_units(header)


def table_meta_of(header):
    from . import table_meta
    meta = table_meta.table_meta(header)
    for key in ("poly_trend", "n_offsets"):
        if meta.get(key) is not None:
            meta[key] = int(meta[key])
    return meta
