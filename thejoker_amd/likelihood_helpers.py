"""Host-side helpers of the likelihood path (thejoker/likelihood_helpers.py,
thejoker/data_helpers.py, thejoker/utils.py): design-matrix trend columns,
data validation / multi-survey concatenation, batch splitting and the
iterative-rejection growth rule."""
import numpy as np

from . import units as u

__all__ = ["get_constant_term_design_matrix", "get_trend_design_matrix", "validate_prepare_data",
           "batch_tasks", "next_batch_size", "ln_normal"]


def get_constant_term_design_matrix(data, ids=None):
    """The constant-velocity columns of the trend design matrix
    (likelihood_helpers.py:8-25), one per survey: column 0 is 1 on every epoch
    (v0); the survey labels after the first, in sorted order, each get a 0/1
    column marking their epochs (that survey's offset from v0)."""
    n = len(data)
    labels = np.zeros(n, dtype=int) if ids is None else np.asarray(ids)
    surveys = np.unique(labels)
    cols = np.zeros((n, surveys.size))
    cols[:, 0] = 1.0
    cols[:, 1:] = labels[:, None] == surveys[None, 1:]
    return cols


def get_trend_design_matrix(data, ids, poly_trend):
    """Every linear column but the Keplerian one (likelihood_helpers.py:28-37):
    the constant-velocity columns above, then dt, dt^2, ... up to
    dt^(poly_trend - 1) with dt = t - t_ref.  The powers are running products
    (dt, dt*dt, (dt*dt)*dt, ...), the rounding of np.vander's."""
    const = get_constant_term_design_matrix(data, ids)
    dt = np.asarray(data._t_bmjd - data._t_ref_bmjd, dtype=np.float64)
    k = max(int(poly_trend) - 1, 0)
    powers = np.cumprod(np.repeat(dt[:, None], k, axis=1), axis=1)
    return np.concatenate([const, powers], axis=1)


def _as_survey_dict(data):
    """A mapping label -> RVData from a dict-like, or from an iterable
    (labels 0, 1, ... in order)."""
    if hasattr(data, "keys"):
        return data
    try:
        return dict(enumerate(data))
    except Exception as exc:
        raise TypeError("data must be one RVData, an iterable of RVData objects or a mapping "
                        f"whose values are RVData objects; got {type(data)}") from exc


def validate_prepare_data(data, poly_trend, n_offsets):
    """(RVData, survey labels, trend design matrix) of the input
    (data_helpers.py:56-133).  One RVData stands alone (and then takes no v0
    offsets); several surveys -- a list or a dict of RVData -- are merged into
    one RVData in the first survey's RV unit, each epoch labelled by its
    survey's key, and need exactly one v0 offset per survey beyond the first."""
    from .data import RVData

    if isinstance(data, RVData):
        if n_offsets != 0:
            raise ValueError(f"the prior has {n_offsets} v0 offset(s), which need one RVData "
                             "per survey (a list or dict of RVData), not a single RVData")
        return data, np.zeros(len(data), dtype=int), get_trend_design_matrix(data, None, poly_trend)

    surveys = _as_survey_dict(data)
    unit = None
    parts = {"t": [], "rv": [], "err": [], "ids": []}
    for key in surveys.keys():
        d = surveys[key]
        if not isinstance(d, RVData):
            raise TypeError(f"survey '{key}' is a {type(d)}; every survey must be an RVData")
        if d._has_cov:
            raise NotImplementedError("several surveys cannot be combined when one has a full "
                                      "covariance matrix")
        unit = d.rv.unit if unit is None else unit
        parts["t"].append(d._t_bmjd)
        parts["rv"].append(d.rv.to_value(unit))
        parts["err"].append(d.rv_err.to_value(unit))
        parts["ids"].append(np.array([key] * len(d)))
    ids = np.concatenate(parts["ids"])
    n_surveys = np.unique(ids).size
    if n_surveys - 1 != n_offsets:
        raise ValueError(f"{n_surveys} surveys need {n_surveys - 1} v0 offset(s) in the prior; "
                         f"it has {n_offsets}")
    # RVData orders the merged epochs by time while the labels stay in
    # concatenation order, and the offset columns follow the labels as
    # given -- as in data_helpers.py:633-636; it matters only when surveys
    # interleave in time.
    merged = RVData(t=np.concatenate(parts["t"]),
                    rv=u.Quantity(np.concatenate(parts["rv"]), unit),
                    rv_err=u.Quantity(np.concatenate(parts["err"]), unit))
    return merged, ids, get_trend_design_matrix(merged, ids, poly_trend)


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
