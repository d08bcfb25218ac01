"""HDF5 sample tables: the numpy-only reader against files the HDF5 library
wrote (tests/golden/make_hdf5_fixtures.py, h5py 3.3.0 on HDF5 1.10.6), the
writer read back by this reader and, when an interpreter with h5py is
present, by h5py itself (including the reference's append-by-resize), the
astropy YAML header and the unit strings it carries."""
import os
import subprocess

import numpy as np
import pytest

from thejoker_amd import JokerSamples, hdf5, table_meta
from thejoker_amd import units as u

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
H5PY_PYTHON = "/opt/conda/bin/python3.9"


def _golden(name):
    return os.path.join(GOLDEN, name)


@pytest.fixture(scope="module")
def expected():
    with np.load(_golden("hdf5_expected.npz")) as f:
        return {k: f[k] for k in f.files}


@pytest.mark.parametrize("fname,key,nchunks", [
    ("hdf5_prior_default.hdf5", "prior_default", 2),
    ("hdf5_posterior_appended.hdf5", "posterior_appended", 4),
    ("hdf5_prior_small_chunks.hdf5", "prior_small_chunks", 125),
])
def test_reader_matches_hdf5_library(expected, fname, key, nchunks):
    ref = expected[key]
    with hdf5.File(_golden(fname)) as f:
        assert f.keys() == ["samples", "samples.__table_column_meta__"]
        ds = f["samples"]
        assert ds.shape == ref.shape and ds.maxshape == (None,)
        assert ds.dtype.names == ref.dtype.names
        assert ds._layout[0] == "chunked" and len(ds._layout[2]) == nchunks
        got = ds.read()
        for name in ref.dtype.names:
            np.testing.assert_array_equal(got[name], ref[name])
        # row slices across chunk boundaries, one field; and into a reused buffer
        buf = np.full(len(ref) * ref.dtype.itemsize, 255, dtype=np.uint8)
        for lo, hi in [(0, 1), (7, 9), (123, 391), (len(ref) - 3, len(ref) + 10), (5, 5)]:
            np.testing.assert_array_equal(ds.read(lo, hi, field="omega"), ref["omega"][lo:hi])
            np.testing.assert_array_equal(ds.read(lo, hi, buf=buf)["e"], ref["e"][lo:hi])


def test_header_strings_fixed_and_variable_length():
    """The reference's writer passes a list of bytes objects, which h5py
    stores as variable-length strings; a numpy 'S' array is stored as
    fixed-length strings."""
    with hdf5.File(_golden("hdf5_prior_default.hdf5")) as f:
        vlen = f["samples.__table_column_meta__"].read()
    with hdf5.File(_golden("hdf5_prior_small_chunks.hdf5")) as f:
        fixed = f["samples.__table_column_meta__"].read()
    assert fixed.dtype.kind == "S" and vlen.dtype == object
    assert [bytes(x) for x in fixed] == list(vlen)
    assert fixed[0] == b"datatype:" and vlen[-1] == b"schema: astropy-2.0"


def test_newest_file_format(expected):
    with hdf5.File(_golden("hdf5_libver_latest.hdf5")) as f:
        assert set(f.keys()) == {"contiguous", "samples"}
        got = f["contiguous"].read()
        np.testing.assert_array_equal(got, expected["libver_latest"])
        with pytest.raises(NotImplementedError, match="extensible array"):
            f["samples"]


def test_not_hdf5(tmp_path):
    p = tmp_path / "x.hdf5"
    p.write_bytes(b"not an hdf5 file at all" * 40)
    with pytest.raises(hdf5.H5FormatError):
        hdf5.File(p)
    p.write_bytes(b"")
    with pytest.raises(hdf5.H5FormatError):
        hdf5.File(p)
    with hdf5.File(_golden("hdf5_prior_default.hdf5")) as f:
        with pytest.raises(KeyError):
            f["nope"]


def test_samples_read_units_and_meta(expected):
    js = JokerSamples.read(_golden("hdf5_posterior_appended.hdf5"))
    ref = expected["posterior_appended"]
    assert js.poly_trend == 2 and js.n_offsets == 0
    # jd1 + jd2 of the stored Time, as an MJD
    assert js.t_ref == (2458512.0 - 2400000.5) + -0.20488644757278962
    assert js["s"].unit == u.m / u.s and js["v1"].unit == u.km / u.s / u.day
    np.testing.assert_array_equal(js["K"].value, ref["K"])
    np.testing.assert_allclose(js["s"].to_value(u.km / u.s), ref["s"] / 1000.0, rtol=1e-15)
    js = JokerSamples.read(_golden("hdf5_prior_default.hdf5"))
    assert js.t_ref is None and js.poly_trend == 1
    assert js["e"].unit == u.one and js["P"].unit == u.day
    packed, units = js.pack(units={"s": u.m / u.s})
    np.testing.assert_array_equal(packed[:, 4], expected["prior_default"]["s"] * 1000.0)


def test_ecsv_header_of_reference_data():
    """The YAML header of the reference's docs/examples/data.ecsv: its t_ref
    Time tag decodes to the first epoch, 2458511.7951135524 BJD (the file's
    first data row), as an MJD."""
    with open(_golden("ecsv_header_data.txt")) as fh:
        h = table_meta.read_header(fh.read().splitlines())
    units = table_meta.column_units(h)
    assert units["rv"] == u.km / u.s and units["rv_err"] == u.km / u.s
    assert units["bjd"] == u.one
    meta = table_meta.table_meta(h)
    assert list(meta) == ["t_ref"]
    assert abs(meta["t_ref"] - (2458511.7951135524 - 2400000.5)) < 1e-9


def test_header_round_trip():
    cols = [("P", u.day), ("e", u.one), ("s", u.m / u.s), ("v1", u.km / u.s / u.day),
            ("v2", u.km / u.s / u.day ** 2)]
    meta = {"t_ref": 58511.29511355243, "poly_trend": 3, "n_offsets": 0, "tag": "x y"}
    lines = table_meta.make_header(cols, meta)
    h = table_meta.read_header(lines)
    back = table_meta.column_units(h)
    assert list(back) == [c for c, _ in cols]
    for c, unit in cols:
        assert back[c] == unit, c
    got = table_meta.table_meta(h)
    assert got == meta
    # the Time tag splits exactly as the stored reference value
    assert any("jd1: 2458512.0, jd2: -0.20488644757278962" in ln for ln in lines)


@pytest.mark.parametrize("text,scale,dims", [
    ("km / s", 1e3, (1, -1, 0)), ("km / (d s)", 1e3 / 86400, (1, -2, 0)),
    ("m / (d2 s)", 1 / 86400.0 ** 2, (1, -3, 0)), ("s-1", 1.0, (0, -1, 0)),
    ("m s**-2", 1.0, (1, -2, 0)), ("km / s / d", 1e3 / 86400, (1, -2, 0)), ("", 1.0, (0, 0, 0)),
    ("rad", 1.0, (0, 0, 1)), ("deg", np.pi / 180, (0, 0, 1)), ("AU / yr", 1.495978707e11 /
                                                                 (365.25 * 86400), (1, -1, 0)),
])
def test_parse_unit(text, scale, dims):
    unit = u.parse_unit(text)
    assert unit.dims == dims
    assert abs(unit.scale / scale - 1) < 1e-15


@pytest.mark.parametrize("bad", ["km / / s", "furlong", "km (", "m**0.5"])
def test_parse_unit_errors(bad):
    with pytest.raises(u.UnitsError):
        u.parse_unit(bad)


def _table(n, seed=0):
    rng = np.random.default_rng(seed)
    js = JokerSamples(t_ref=58000.25, poly_trend=2)
    js["P"] = np.exp(rng.uniform(0, 7, n)) * u.day
    js["e"] = rng.uniform(0, 1, n)
    js["omega"] = rng.uniform(0, 6, n) * u.rad
    js["M0"] = rng.uniform(0, 6, n) * u.rad
    js["s"] = rng.uniform(0, 1, n) * u.m / u.s
    js["K"] = rng.normal(0, 10, n) * u.km / u.s
    js["v0"] = rng.normal(0, 10, n) * u.km / u.s
    js["v1"] = rng.normal(0, 1, n) * u.km / u.s / u.day
    return js


@pytest.mark.parametrize("n", [0, 1, 1023, 1025, 70000])
def test_writer_round_trip(tmp_path, n):
    js = _table(n)
    path = str(tmp_path / "s.hdf5")
    js.write(path)
    back = JokerSamples.read(path)
    assert back.meta == js.meta
    assert back.par_names == js.par_names
    for k in js.par_names:
        assert back[k].unit == js[k].unit
        np.testing.assert_array_equal(back[k].value, js[k].value)


def test_writer_exists_append_and_conflicts(tmp_path):
    path = str(tmp_path / "s.h5")
    a, b = _table(100, 1), _table(50, 2)
    a.write(path)
    with pytest.raises(OSError, match="File exists"):
        a.write(path)
    a.write(path, overwrite=True)
    b.write(path, append=True)
    back = JokerSamples.read(path)
    np.testing.assert_array_equal(back["P"].value, np.concatenate([a["P"].value, b["P"].value]))
    c = _table(10, 3)
    c.meta["poly_trend"] = 2
    c.meta["t_ref"] = 1.0
    with pytest.raises(ValueError, match="metadata"):
        c.write(path, append=True)
    with pytest.raises(NotImplementedError):
        a.write(str(tmp_path / "s.fits"))


def test_writer_chunk_limit(tmp_path):
    """One B-tree leaf indexes at most 64 chunks; the default chunk length
    keeps any table within it."""
    cols = [("P", np.arange(100.0))]
    with pytest.raises(ValueError, match="B-tree"):
        hdf5.write_table(str(tmp_path / "a.hdf5"), cols, ["datatype:"], chunk_rows=1)
    hdf5.write_table(str(tmp_path / "b.hdf5"), cols, ["datatype:"], chunk_rows=2)
    with hdf5.File(str(tmp_path / "b.hdf5")) as f:
        assert len(f["samples"]._layout[2]) == 50
        np.testing.assert_array_equal(f["samples"].read(field="P"), cols[0][1])


@pytest.mark.skipif(not (os.path.exists(H5PY_PYTHON) and subprocess.run(
    [H5PY_PYTHON, "-c", "import h5py"], capture_output=True).returncode == 0),
    reason="no interpreter with h5py")
def test_h5py_reads_and_appends_written_file(tmp_path):
    """h5py opens a file this writer made, reads every value and appends to
    it the way the reference's writer appends (resize + write,
    samples_helpers.py:239-243); the reader then sees both parts."""
    js = _table(5000, 4)
    path = str(tmp_path / "w.hdf5")
    js.write(path)
    np.save(tmp_path / "P.npy", js["P"].value)
    script = f"""
import numpy as np, h5py
ref = np.load({str(tmp_path / 'P.npy')!r})
with h5py.File({path!r}, 'r') as f:
    s = f['samples']
    assert s.shape == (5000,) and s.maxshape == (None,), (s.shape, s.maxshape)
    assert np.array_equal(s['P'], ref)
    lines = [x.decode() for x in f['samples.__table_column_meta__'][:]]
    assert lines[0] == 'datatype:', lines[:2]
with h5py.File({path!r}, 'a') as f:
    s = f['samples']
    add = np.zeros(3000, s.dtype)
    add['P'] = np.arange(3000.0)
    s.resize((8000,))
    s[5000:] = add
print('ok')
"""
    r = subprocess.run([H5PY_PYTHON, "-c", script], capture_output=True, text=True,
                       cwd=str(tmp_path))
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr
    back = JokerSamples.read(path)
    assert len(back) == 8000
    np.testing.assert_array_equal(back["P"].value[:5000], js["P"].value)
    np.testing.assert_array_equal(back["P"].value[5000:], np.arange(3000.0))


def test_header_numpy_scalars():
    lines = table_meta.make_header([("P", u.day)], {"t_ref": np.float64(58000.5),
                                                    "poly_trend": np.int64(2),
                                                    "n_offsets": np.int32(0),
                                                    "flag": np.bool_(True)})
    meta = table_meta.table_meta(table_meta.read_header(lines))
    assert meta == {"t_ref": 58000.5, "poly_trend": 2, "n_offsets": 0, "flag": True}
