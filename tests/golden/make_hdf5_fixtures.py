"""Write the HDF5 sample-table fixtures with h5py (the HDF5 library itself).

Run with an interpreter that has h5py (this image: /opt/conda/bin/python3.9,
h5py 3.3.0 on HDF5 1.10.6):

    /opt/conda/bin/python3.9 tests/golden/make_hdf5_fixtures.py

The files follow the reference's writer (samples_helpers.py:199-211): the
table as ``create_dataset('samples', data=table.as_array(), maxshape=(None,))``
(chunked, no filters) and the YAML header as
``create_dataset('samples.__table_column_meta__', data=[line.encode(), ...])``;
an append resizes and writes the tail (samples_helpers.py:239-243).  The
header lines are astropy's layout for a QTable of Quantity columns, as in the
ECSV files of the reference's docs/examples.  The expected column values go
to hdf5_expected.npz, written from the same arrays.
"""
import os

import h5py
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PRIOR_COLS = [("P", "d"), ("e", ""), ("omega", "rad"), ("M0", "rad"), ("s", "km / s"),
              ("ln_prior", "")]
POST_COLS = [("P", "d"), ("e", ""), ("omega", "rad"), ("M0", "rad"), ("s", "m / s"),
             ("K", "km / s"), ("v0", "km / s"), ("v1", "km / (d s)")]


def header(cols, meta_lines):
    lines = ["datatype:"]
    for name, unit in cols:
        lines.append(f"- {{name: {name}, unit: {unit}, datatype: float64}}" if unit
                     else f"- {{name: {name}, datatype: float64}}")
    lines.append("meta: !!omap")
    lines += meta_lines
    lines.append("- __serialized_columns__:")
    for name, unit in cols:
        lines += [f"    {name}:",
                  "      __class__: astropy.units.quantity.Quantity",
                  f"      unit: !astropy.units.Unit {{unit: {unit or chr(39) * 2}}}",
                  f"      value: !astropy.table.SerializedColumn {{name: {name}}}"]
    lines.append("schema: astropy-2.0")
    return lines


def prior_table(rng, n, cols):
    dt = np.dtype([(c, "<f8") for c, _ in cols])
    a = np.zeros(n, dt)
    for c, _ in cols:
        a[c] = {"P": lambda: np.exp(rng.uniform(np.log(2.0), np.log(2048.0), n)),
                "e": lambda: rng.beta(0.867, 3.03, n),
                "omega": lambda: rng.uniform(0.0, 2 * np.pi, n),
                "M0": lambda: rng.uniform(0.0, 2 * np.pi, n),
                "s": lambda: np.exp(rng.normal(-3.0, 1.0, n)),
                }.get(c, lambda: rng.normal(0.0, 10.0, n))()
    return a


def main():
    rng = np.random.default_rng(20240611)
    expected = {}
    meta_prior = ["- {t_ref: null}", "- {poly_trend: 1}", "- {n_offsets: 0}"]
    meta_post = ["- t_ref: !astropy.time.Time {format: mjd, in_subfmt: '*', jd1: 2458512.0, "
                 "jd2: -0.20488644757278962, out_subfmt: '*', precision: 3,",
                 "    scale: tcb}", "- {poly_trend: 2}", "- {n_offsets: 0}"]

    # 1. a prior cache as the reference writes it (h5py's automatic chunking)
    a = prior_table(rng, 400, PRIOR_COLS)
    with h5py.File(os.path.join(HERE, "hdf5_prior_default.hdf5"), "w") as f:
        f.create_dataset("samples", data=a, maxshape=(None,))
        f.create_dataset("samples.__table_column_meta__",
                         data=[h.encode("utf-8") for h in header(PRIOR_COLS, meta_prior)])
    expected["prior_default"] = a

    # 2. posterior samples written, then appended to (resize + write)
    b1, b2 = prior_table(rng, 300, POST_COLS), prior_table(rng, 200, POST_COLS)
    with h5py.File(os.path.join(HERE, "hdf5_posterior_appended.hdf5"), "w") as f:
        f.create_dataset("samples", data=b1, maxshape=(None,))
        f.create_dataset("samples.__table_column_meta__",
                         data=[h.encode("utf-8") for h in header(POST_COLS, meta_post)])
    with h5py.File(os.path.join(HERE, "hdf5_posterior_appended.hdf5"), "a") as f:
        f["samples"].resize((500,))
        f["samples"][300:] = b2
    expected["posterior_appended"] = np.concatenate([b1, b2])

    # 3. small chunks (a B-tree of more than one level) and the header as
    #    fixed-length strings (h5py stores the list of bytes objects of
    #    cases 1 and 2 as variable-length strings, in the global heap)
    c = prior_table(rng, 1000, PRIOR_COLS)
    with h5py.File(os.path.join(HERE, "hdf5_prior_small_chunks.hdf5"), "w") as f:
        f.create_dataset("samples", data=c, maxshape=(None,), chunks=(8,))
        f.create_dataset("samples.__table_column_meta__",
                         data=np.array([h.encode("utf-8") for h in header(PRIOR_COLS, meta_prior)]))
    expected["prior_small_chunks"] = c

    # 4. the newest file format: version 2 object headers, link messages, a
    #    contiguous table (readable) and a chunked one with an extensible
    #    array index (not supported)
    d = prior_table(rng, 64, PRIOR_COLS)
    with h5py.File(os.path.join(HERE, "hdf5_libver_latest.hdf5"), "w", libver="latest") as f:
        f.create_dataset("contiguous", data=d)
        f.create_dataset("samples", data=d, maxshape=(None,))
    expected["libver_latest"] = d

    np.savez(os.path.join(HERE, "hdf5_expected.npz"), **expected)


if __name__ == "__main__":
    main()
