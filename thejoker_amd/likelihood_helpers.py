"""Host-side helpers of the likelihood path (thejoker/likelihood_helpers.py,
thejoker/data_helpers.py, thejoker/utils.py): design-matrix trend columns,
data validation / multi-survey concatenation, batch splitting and the
iterative-rejection growth rule."""
import numpy as np

from . import units as u

__all__ = ["get_constant_term_design_matrix", "get_trend_design_matrix", "validate_prepare_data",
           "batch_tasks", "next_batch_size", "ln_normal"]


def get_constant_term_design_matrix(data, ids=None):
    """likelihood_helpers.py:8-25: a column of ones plus one indicator column
    per extra data source."""
    if ids is None:
        ids = np.zeros(len(data), dtype=int)
    ids = np.array(ids)
    unq_ids = np.unique(ids)
    constant_part = np.zeros((len(data), len(unq_ids)))
    constant_part[:, 0] = 1.0
    for j, id_ in enumerate(unq_ids[1:]):
        constant_part[ids == id_, j + 1] = 1.0
    return constant_part


def get_trend_design_matrix(data, ids, poly_trend):
    """likelihood_helpers.py:28-37: [1, offsets..., dt, dt^2, ...] with
    dt = t - t_ref."""
    const_M = get_constant_term_design_matrix(data, ids)
    dt = data._t_bmjd - data._t_ref_bmjd
    trend_M = np.vander(dt, N=poly_trend, increasing=True)[:, 1:]
    return np.hstack((const_M, trend_M))


def validate_prepare_data(data, poly_trend, n_offsets):
    """data_helpers.py:56-133: one RVData, or a list/dict of them (one per
    survey), to (RVData, ids, trend_M)."""
    from .data import RVData

    if isinstance(data, RVData):
        if n_offsets != 0:
            raise ValueError("If sampling over velocity offsets between data sources, you must "
                             "pass in multiple data sources. To do this, pass in a list of RVData "
                             "instances or a dictionary with RVData instances as values.")
        return data, np.zeros(len(data), dtype=int), get_trend_design_matrix(data, None, poly_trend)

    if not hasattr(data, "keys"):
        try:
            data = {i: d for i, d in enumerate(data)}
        except Exception as e:
            raise TypeError("Failed to parse input data: data must either be an RVData instance, "
                            "an iterable of RVData instances, or a dictionary with RVData "
                            f"instances as values. Received: {type(data)}") from e

    rv_unit = None
    t, rv, err, ids = [], [], [], []
    for k in data.keys():
        d = data[k]
        if not isinstance(d, RVData):
            raise TypeError("All data must be specified as RVData instances: Object at key "
                            f"'{k}' is a '{type(d)}' instead.")
        if d._has_cov:
            raise NotImplementedError("We currently don't support multi-survey data when a full "
                                      "covariance matrix is specified.")
        if rv_unit is None:
            rv_unit = d.rv.unit
        t.append(d._t_bmjd)
        rv.append(d.rv.to_value(rv_unit))
        err.append(d.rv_err.to_value(rv_unit))
        ids.append([k] * len(d))
    t = np.concatenate(t)
    rv = np.concatenate(rv)
    err = np.concatenate(err)
    ids = np.concatenate(ids)
    if (len(np.unique(ids)) - 1) != n_offsets:
        raise ValueError("Number of data IDs + 1 must equal the number of priors on constant "
                         "offsets specified (i.e. v0_offsets)")
    # As at data_helpers.py:633-636: RVData sorts the concatenated epochs by
    # time while `ids` keeps the concatenation order, and the offset columns
    # are built from `ids` as given.  Kept as is for parity; it only matters
    # when surveys interleave in time.
    all_data = RVData(t=t, rv=u.Quantity(rv, rv_unit), rv_err=u.Quantity(err, rv_unit))
    return all_data, ids, get_trend_design_matrix(all_data, ids, poly_trend)


def batch_tasks(n_tasks, n_batches, arr=None, args=None, start_idx=0):
    """utils.py:22-72: split n_tasks into n_batches contiguous batches (the
    first n_tasks % n_batches batches one larger)."""
    if args is None:
        args = []
    args = list(args)
    tasks = []
    if n_batches > 0 and n_tasks >= n_batches:
        base = n_tasks // n_batches
        rmdr = n_tasks % n_batches
        i1 = start_idx
        for i in range(n_batches):
            i2 = i1 + base + (1 if i < rmdr else 0)
            tasks.append([(i1, i2) if arr is None else arr[i1:i2], i1] + args)
            i1 = i2
    else:
        if arr is None:
            tasks.append([(start_idx, n_tasks + start_idx), start_idx] + args)
        else:
            tasks.append([arr[start_idx:n_tasks + start_idx], start_idx] + args)
    return tasks


def next_batch_size(n_requested, n_good, n_ll_evals, start_idx, n_total, safety_factor):
    """Size of the next iterative-rejection batch
    (likelihood_helpers.py:196-207, multiproc_helpers.py:387-397):
    int(safety * n_need / n_good * n_evals), capped at the samples left.
    Returns the size (<= 0 means stop)."""
    n_need = n_requested - n_good
    n_process = int(safety_factor * n_need / n_good * n_ll_evals)
    if start_idx + n_process > n_total:
        n_process = n_total - start_idx
    return n_process


def ln_normal(x, mu, var):
    """likelihood_helpers.py:232-233."""
    return -0.5 * (np.log(2 * np.pi * var) + (x - mu) ** 2 / var)
