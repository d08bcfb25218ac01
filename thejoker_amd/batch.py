"""Many stars against one device-resident prior cache.

The reference evaluates one star at a time (TheJoker._make_joker_helper ->
CJokerHelper, thejoker.py:87-91).  Here a set of stars with the same number
of linear parameters is gathered into a device star table (``jk_batch``) and
evaluated by one persistent work-queue launch over (star, row tile) items
(``jk_lnl_batch``); the rejection step of every star runs as one segmented
launch sequence (``jk_reject_batch``), each star with its own numpy PCG64
stream.  :func:`autotune` picks the launch shape (threads per workgroup,
workgroups, rows per work item) by timing candidates on the actual workload.
"""
import ctypes
import time

import numpy as np

from . import _lib
from .device import (DeviceStar, _dptr, _split, _stream, _torch, pcg64_advance_state, pcg64_state,
                     to_host)

__all__ = ["WORK_ITEM_DTYPE", "SEGMENT_DTYPE", "POLISH_SEGMENT_DTYPE", "SETTLE_SEGMENT_DTYPE",
           "LIVE_SEGMENT_DTYPE", "StarBatch", "work_items", "lnl_batch", "lnl_prune_batch",
           "reject_batch",
           "polish_batch", "settle_batch", "settle_reject_batch", "settle_live_batch",
           "LIVE_MARGIN",
           "LIVE_MARGIN_NONZERO_U",
           "autotune", "LaunchShape", "rejection_sample_many", "SegPlan", "StepGraph"]

WORK_ITEM_DTYPE = np.dtype([("star", "<i4"), ("n_rows", "<i4"), ("row_begin", "<i8"),
                            ("out_offset", "<i8")])
POLISH_SEGMENT_DTYPE = np.dtype([("offset", "<i8"), ("n", "<i8"), ("star", "<i4"),
                                 ("reserved", "<i4"), ("row_begin", "<i8"), ("first", "<i8")])
LIVE_SEGMENT_DTYPE = np.dtype([("offset", "<i8"), ("n", "<i8"), ("first", "<i8"), ("star", "<i4"),
                               ("reserved", "<i4"), ("row_begin", "<i8"), ("floor", "<f8")])
# Settling margins of jk_settle_live_batch: every uniform (exp underflows to
# 0 below -745.14), and every uniform but exactly 0 (exp(-40) < 2^-53).
LIVE_MARGIN = 750.0
LIVE_MARGIN_NONZERO_U = 40.0
# |lnL - lnL(FP64)| every mixed-precision lnL handed back to a caller meets
API_TOL = 1e-3
# Stars with fewer epochs are not probed: the probe (at least 5 epochs) would
# cost about as much as the evaluation it saves.
PRUNE_MIN_EPOCHS = 12
# Nor are stars whose trend-only fit leaves less than this in lnL for an orbit
# to gain (device.trend_signal): their rows all lie near the maximum, and a
# probe would prune next to none of them.
PRUNE_MIN_SIGNAL = 50.0
# Iterative loops in mixed precision: after a round whose live settling
# re-evaluated more than this share of its new rows in FP64, the star leaves
# the live settling (settled rows stay FP64 for every later uniform) for
# settling at each rejection step (the rows that step's uniforms leave open),
# or for FP64 evaluation.  A mixed row costs about 0.65 of an FP64 one and a
# settled row about one more FP64 evaluation, so past about a third settled
# either is cheaper; flat likelihoods (near-constant stars, rows all within
# the margin of the maximum) settle nearly every row.
EXACT_SWITCH = 0.3
SETTLE_SEGMENT_DTYPE = np.dtype([("offset", "<i8"), ("n", "<i8"), ("star", "<i4"),
                                 ("reserved", "<i4"), ("row_begin", "<i8"), ("state_hi", "<u8"),
                                 ("state_lo", "<u8"), ("inc_hi", "<u8"), ("inc_lo", "<u8"),
                                 ("stream_offset", "<u8")])
SEGMENT_DTYPE = np.dtype([("offset", "<i8"), ("n", "<i8"), ("max_keep", "<i8"),
                          ("idx_offset", "<i8"), ("state_hi", "<u8"), ("state_lo", "<u8"),
                          ("inc_hi", "<u8"), ("inc_lo", "<u8"), ("stream_offset", "<u8")])


class LaunchShape:
    """Threads per workgroup, workgroups (0 = 8 per CU), rows per work item."""

    def __init__(self, block=256, grid=0, tile=2048):
        self.block, self.grid, self.tile = int(block), int(grid), int(tile)

    def __repr__(self):
        return f"LaunchShape(block={self.block}, grid={self.grid}, tile={self.tile})"

    def as_dict(self):
        return {"block": self.block, "grid": self.grid, "tile": self.tile}


class StarBatch:
    """Device star table of DeviceStar objects (same L and jitter mode)."""

    def __init__(self, stars, device=0):
        lib = _lib.ensure_device(device)
        self._lib = lib
        self.device = device
        self.stars = list(stars)
        arr = (ctypes.c_void_p * len(self.stars))(*[s.handle.value for s in self.stars])
        h = ctypes.c_void_p()
        _lib.check(lib.jk_batch_create(arr, len(self.stars), ctypes.byref(h)), "jk_batch_create")
        self._h = h
        self.M = np.array([s.n for s in self.stars], dtype=np.int64)
        self.L = self.stars[0].L

    @classmethod
    def slots(cls, capacity, L, jitter_mode="physical", device=0):
        """An empty table of `capacity` slots (jk_batch_create_slots), filled
        with :meth:`set`; a slot keeps its DeviceStar alive until replaced."""
        from .device import JITTER_MODES
        lib = _lib.ensure_device(device)
        self = cls.__new__(cls)
        self._lib = lib
        self.device = device
        self.stars = [None] * int(capacity)
        h = ctypes.c_void_p()
        _lib.check(lib.jk_batch_create_slots(int(capacity), int(L), JITTER_MODES[jitter_mode],
                                             ctypes.byref(h)), "jk_batch_create_slots")
        self._h = h
        self.M = np.zeros(int(capacity), dtype=np.int64)
        self.L = int(L)
        return self

    def set(self, slot, star):
        """Point slot `slot` at DeviceStar `star` (jk_batch_set_star), ordered
        after the work already queued on the device's stream."""
        _lib.check(self._lib.jk_batch_set_star(self._h, int(slot), star.handle,
                                                _stream(self.device)), "jk_batch_set_star")
        self.stars[slot] = star
        self.M[slot] = star.n

    def set_exact(self, slot, exact=True):
        """Evaluate slot `slot` in FP64 inside mixed-precision launches, with
        error bound 0 (jk_batch_set_exact); set_star clears it."""
        _lib.check(self._lib.jk_batch_set_exact(self._h, int(slot), int(bool(exact)),
                                                 _stream(self.device)), "jk_batch_set_exact")

    @property
    def handle(self):
        return self._h

    def __len__(self):
        return len(self.stars)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.jk_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def work_items(n_rows, tile, row_begin=None, order=None, stars=None):
    """Work items covering rows [row_begin[s], row_begin[s] + n_rows[s]) of
    each star s, `tile` rows per item, outputs packed star after star.
    `stars` (optional) gives each star's star-table index (default s);
    `order` (optional) permutes the items (e.g. heaviest first).
    Returns (items, out_offsets, total)."""
    n_rows = np.asarray(n_rows, dtype=np.int64)
    row_begin = np.zeros_like(n_rows) if row_begin is None else np.asarray(row_begin, np.int64)
    out_off = np.concatenate([[0], np.cumsum(n_rows)[:-1]]).astype(np.int64)
    counts = (n_rows + tile - 1) // tile
    pos = np.repeat(np.arange(len(n_rows)), counts)
    first = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
    j = np.arange(int(counts.sum()), dtype=np.int64) - first[pos]
    items = np.zeros(len(pos), dtype=WORK_ITEM_DTYPE)
    items["star"] = pos if stars is None else np.asarray(stars)[pos]
    items["n_rows"] = np.minimum(tile, n_rows[pos] - j * tile)
    items["row_begin"] = row_begin[pos] + j * tile
    items["out_offset"] = out_off[pos] + j * tile
    if order is not None:
        items = items[order]
    return items, out_off, int(n_rows.sum())


def heaviest_first(items, M):
    """Order work items by decreasing cost (rows x epochs), so the queue ends
    on short items."""
    cost = items["n_rows"].astype(np.int64) * np.asarray(M)[items["star"]]
    return np.argsort(-cost, kind="stable")


def upload_items(items, device=0):
    torch = _torch()
    raw = np.frombuffer(items.tobytes(), dtype=np.uint8)
    return torch.as_tensor(raw.copy(), device=torch.device("cuda", device))


class ItemUploader:
    """Work items to the device through one pinned host buffer and one
    device buffer, both kept between calls and grown geometrically: the
    upload is a host memcpy and an asynchronous DMA, instead of a pageable
    copy that stages through the runtime on the host.  The returned device
    view is valid until the next call; a call waits for the previous call's
    copy before it rewrites the host buffer, and its own copy is ordered on
    the stream before any later launch reads the device buffer."""

    def __init__(self, device=0):
        self.device = device
        self._host = self._host_np = self._dev = self._done = None

    def __call__(self, items):
        torch = _torch()
        nb = int(items.nbytes)
        if self._host is None or self._host.numel() < nb:
            cap = max(nb, 2 * (0 if self._host is None else self._host.numel()), 1 << 20)
            if self._done is not None:
                self._done.synchronize()
            self._host = torch.empty(cap, dtype=torch.uint8, pin_memory=True)
            self._host_np = self._host.numpy()
            self._dev = torch.empty(cap, dtype=torch.uint8,
                                    device=torch.device("cuda", self.device))
        elif self._done is not None:
            self._done.synchronize()
        np.frombuffer(self._host_np, dtype=items.dtype, count=len(items))[:] = items
        stream = torch.cuda.current_stream(self.device)
        with torch.cuda.stream(stream):
            self._dev[:nb].copy_(self._host[:nb], non_blocking=True)
        self._done = torch.cuda.Event()
        self._done.record(stream)
        return self._dev[:nb]


def lnl_batch(batch, cache, d_items, n_items, n_out, precision="mixed", shape=None, out=None,
              rows=None, workspace=None, bound=None):
    """lnL of every work item into a device float64 tensor of n_out entries;
    with `bound` (mixed only, a device float32 tensor indexed like `out`) each
    row's bound on |lnL(mixed) - lnL(FP64)| as well."""
    torch = _torch()
    lib = _lib.ensure_device(cache.device)
    shape = shape or LaunchShape()
    if out is None:
        out = torch.empty(n_out, dtype=torch.float64, device=cache.P.device)
    cols = (cache.e, cache.omega, cache.M0, cache.s)
    if precision == "f64":
        prec = 0
    elif precision == "mixed":
        prec = 1
    else:
        raise ValueError("precision must be 'f64' or 'mixed'")
    _lib.check(lib.jk_lnl_batch(batch.handle, prec, int(n_items), _dptr(d_items), _dptr(cache.P),
                                *[_dptr(c) for c in cols], _dptr(rows), _dptr(out), _dptr(bound),
                                shape.block, shape.grid, _dptr(workspace),
                                _stream(cache.device)), "jk_lnl_batch")
    return out


def lnl_prune_batch(batch, probes, prune, cache, d_items, n_items, precision="mixed", shape=None,
                    out=None, bound=None, workspace=None, n_full=None):
    """lnl_batch with pruning (jk_lnl_prune_batch): `probes` is a table whose
    entry k is star k's probe star (probe_star), `prune` a device float64
    tensor of (threshold, add) per table entry.  Rows whose probe bound lies
    below their star's threshold get that bound instead of their lnL, marked
    by bound -1.  `bound` is required in mixed precision; in f64 it is
    optional and receives only the marks (0: an FP64 value, -1: an upper
    bound).  n_full: optional device int64 tensor with an entry per table
    slot, accumulating each star's rows evaluated in full."""
    lib = _lib.ensure_device(cache.device)
    shape = shape or LaunchShape()
    if precision not in ("f64", "mixed"):
        raise ValueError("precision must be 'f64' or 'mixed'")
    if precision == "mixed" and bound is None:
        raise ValueError("mixed precision writes a bound: pass one")
    if out is None or prune.numel() < 2 * len(batch) or prune.dtype != _torch().float64:
        raise ValueError("out is required and prune needs two float64 entries per table slot")
    if n_full is not None and (n_full.numel() < len(batch) or n_full.dtype != _torch().int64):
        raise ValueError("n_full needs an int64 entry per table slot")
    cols = (cache.e, cache.omega, cache.M0, cache.s)
    _lib.check(lib.jk_lnl_prune_batch(batch.handle, probes.handle, _dptr(prune),
                                      1 if precision == "mixed" else 0, int(n_items),
                                      _dptr(d_items), _dptr(cache.P), *[_dptr(c) for c in cols],
                                      _dptr(out), _dptr(bound), _dptr(n_full), shape.block,
                                      shape.grid, _dptr(workspace), _stream(cache.device)),
               "jk_lnl_prune_batch")
    return out


def reject_batch(ll, offsets, sizes, rng_states, max_keep, stream_offsets=None, device=0,
                 zero_check=False):
    """Segmented rejection step.  rng_states: list of (state, inc).  Returns
    (list of index arrays, n_pass array, llmax array), and with zero_check a
    fourth array: 1 for the segments one of whose uniforms is exactly 0."""
    torch = _torch()
    lib = _lib.ensure_device(device)
    n_seg = len(sizes)
    segs = np.zeros(n_seg, dtype=SEGMENT_DTYPE)
    segs["offset"] = offsets
    segs["n"] = sizes
    mk = np.minimum(np.broadcast_to(np.asarray(max_keep, dtype=np.int64), (n_seg,)), sizes)
    segs["max_keep"] = mk
    segs["idx_offset"] = np.concatenate([[0], np.cumsum(mk)[:-1]])
    for i, (st, inc) in enumerate(rng_states):
        segs["state_hi"][i], segs["state_lo"][i] = _split(int(st))
        segs["inc_hi"][i], segs["inc_lo"][i] = _split(int(inc))
    if stream_offsets is not None:
        segs["stream_offset"] = stream_offsets
    # one device block [n_pass (n_seg), max (n_seg), zero (n_seg), idx...],
    # one copy back
    res = torch.empty(3 * n_seg + max(int(mk.sum()), 1), dtype=torch.int64, device=ll.device)
    p = res.data_ptr()
    args = [_dptr(ll), n_seg, segs.ctypes.data_as(ctypes.c_void_p),
            ctypes.c_void_p(p + 24 * n_seg), ctypes.c_void_p(p), ctypes.c_void_p(p + 8 * n_seg)]
    if zero_check:
        _lib.check(lib.jk_reject_batch_zero_check(*args, ctypes.c_void_p(p + 16 * n_seg),
                                                  _stream(device)), "jk_reject_batch_zero_check")
    else:
        _lib.check(lib.jk_reject_batch(*args, _stream(device)), "jk_reject_batch")
    h = to_host(res)
    n_pass = h[:n_seg].copy()
    mx = h[n_seg:2 * n_seg].view(np.float64).copy()
    idx_h = h[3 * n_seg:]
    out = []
    for i in range(n_seg):
        k = int(min(n_pass[i], mk[i]))
        out.append(idx_h[segs["idx_offset"][i]:segs["idx_offset"][i] + k].copy())
    if zero_check:
        return out, n_pass, mx, h[2 * n_seg:3 * n_seg].copy()
    return out, n_pass, mx


def polish_batch(batch, cache, ll, offsets, sizes, stars=None, row_begins=None, rows=None,
                 window=40.0, first=None, device=0):
    """jk_polish_batch: FP64 re-evaluation of every segment's window.  Segment
    i is ll[offsets[i] .. + sizes[i]) of star stars[i] (default i) on cache
    rows row_begins[i] .. (default 0), mixed precision from entry first[i]
    (default 0) on.  Returns a device int64 tensor of per-segment counts."""
    torch = _torch()
    lib = _lib.ensure_device(device)
    n_seg = len(sizes)
    segs = np.zeros(n_seg, dtype=POLISH_SEGMENT_DTYPE)
    segs["offset"] = offsets
    segs["n"] = sizes
    segs["star"] = np.arange(n_seg) if stars is None else stars
    segs["row_begin"] = 0 if row_begins is None else row_begins
    segs["first"] = 0 if first is None else first
    count = torch.zeros(n_seg, dtype=torch.int64, device=ll.device)
    _lib.check(lib.jk_polish_batch(batch.handle, n_seg, segs.ctypes.data_as(ctypes.c_void_p),
                                   _dptr(cache.P), _dptr(cache.e), _dptr(cache.omega),
                                   _dptr(cache.M0), _dptr(cache.s), _dptr(rows), _dptr(ll),
                                   float(window), _dptr(count), _stream(device)),
               "jk_polish_batch")
    return count


def settle_batch(batch, cache, ll, bound, offsets, sizes, rng_states, stars=None,
                 row_begins=None, rows=None, stream_offsets=None, device=0):
    """jk_settle_batch: re-evaluate in FP64 the rows of each mixed-precision
    segment whose rejection decision (with that segment's numpy PCG64 state,
    as passed to :func:`reject_batch` next) the per-row bounds leave open, and
    the rows that could hold the FP64 maximum.  Returns a device int64 tensor
    of per-segment counts."""
    torch = _torch()
    lib = _lib.ensure_device(device)
    n_seg = len(sizes)
    segs = np.zeros(n_seg, dtype=SETTLE_SEGMENT_DTYPE)
    segs["offset"] = offsets
    segs["n"] = sizes
    segs["star"] = np.arange(n_seg) if stars is None else stars
    segs["row_begin"] = 0 if row_begins is None else row_begins
    for i, (st, inc) in enumerate(rng_states):
        segs["state_hi"][i], segs["state_lo"][i] = _split(int(st))
        segs["inc_hi"][i], segs["inc_lo"][i] = _split(int(inc))
    if stream_offsets is not None:
        segs["stream_offset"] = stream_offsets
    count = torch.zeros(n_seg, dtype=torch.int64, device=ll.device)
    _lib.check(lib.jk_settle_batch(batch.handle, n_seg, segs.ctypes.data_as(ctypes.c_void_p),
                                   _dptr(cache.P), _dptr(cache.e), _dptr(cache.omega),
                                   _dptr(cache.M0), _dptr(cache.s), _dptr(rows), _dptr(ll),
                                   _dptr(bound), _dptr(count), _stream(device)),
               "jk_settle_batch")
    return count


def settle_reject_batch(batch, cache, ll, bound, offsets, sizes, rng_states, max_keep,
                        stars=None, row_begins=None, rows=None, stream_offsets=None, device=0):
    """jk_settle_reject_batch: :func:`settle_batch` then :func:`reject_batch`
    in one launch sequence (one uniform per row).  Returns (list of index
    arrays, n_pass array, llmax array, rows settled per segment)."""
    torch = _torch()
    lib = _lib.ensure_device(device)
    n_seg = len(sizes)
    sizes = np.asarray(sizes, dtype=np.int64)
    segs = np.zeros(n_seg, dtype=SETTLE_SEGMENT_DTYPE)
    segs["offset"] = offsets
    segs["n"] = sizes
    segs["star"] = np.arange(n_seg) if stars is None else stars
    segs["row_begin"] = 0 if row_begins is None else row_begins
    for i, (st, inc) in enumerate(rng_states):
        segs["state_hi"][i], segs["state_lo"][i] = _split(int(st))
        segs["inc_hi"][i], segs["inc_lo"][i] = _split(int(inc))
    if stream_offsets is not None:
        segs["stream_offset"] = stream_offsets
    keep = np.minimum(sizes, int(max_keep))
    keep_off = np.concatenate([[0], np.cumsum(keep)[:-1]]).astype(np.int64)
    # one device block [n_pass | llmax | n_settled | kept indices], one copy back
    res = torch.empty(3 * n_seg + max(int(keep.sum()), 1), dtype=torch.int64, device=ll.device)
    p = res.data_ptr()
    _lib.check(lib.jk_settle_reject_batch(
        batch.handle, n_seg, segs.ctypes.data_as(ctypes.c_void_p), int(max_keep), _dptr(cache.P),
        _dptr(cache.e), _dptr(cache.omega), _dptr(cache.M0), _dptr(cache.s), _dptr(rows),
        _dptr(ll), _dptr(bound), ctypes.c_void_p(p + 24 * n_seg), ctypes.c_void_p(p),
        ctypes.c_void_p(p + 8 * n_seg), ctypes.c_void_p(p + 16 * n_seg), _stream(device)),
        "jk_settle_reject_batch")
    h = to_host(res)
    n_pass = h[:n_seg].copy()
    mx = h[n_seg:2 * n_seg].view(np.float64).copy()
    settled = h[2 * n_seg:3 * n_seg].copy()
    idx = h[3 * n_seg:]
    out = [idx[o:o + min(int(k), int(npass))].copy() for o, k, npass in zip(keep_off, keep, n_pass)]
    return out, n_pass, mx, settled


def settle_live_batch(batch, cache, ll, bound, offsets, sizes, firsts, floors, stars=None,
                      row_begins=None, rows=None, device=0, margin=LIVE_MARGIN):
    """jk_settle_live_batch: settle the new entries [firsts[i], sizes[i]) of
    each growing mixed-precision segment for every future uniform (floors[i]:
    the FP64 maximum of the entries before, -inf for none), so that the
    rejection step over the whole segment accepts the FP64 set under any
    stream -- with margin LIVE_MARGIN_NONZERO_U, under any stream none of whose
    uniforms is exactly 0 (``reject_batch(..., zero_check=True)`` tells).
    Returns a device int64 tensor of per-segment counts."""
    torch = _torch()
    lib = _lib.ensure_device(device)
    n_seg = len(sizes)
    segs = np.zeros(n_seg, dtype=LIVE_SEGMENT_DTYPE)
    segs["offset"] = offsets
    segs["n"] = sizes
    segs["first"] = firsts
    segs["star"] = np.arange(n_seg) if stars is None else stars
    segs["row_begin"] = 0 if row_begins is None else row_begins
    segs["floor"] = floors
    count = torch.empty(n_seg, dtype=torch.int64, device=ll.device)
    _lib.check(lib.jk_settle_live_batch(batch.handle, n_seg, segs.ctypes.data_as(ctypes.c_void_p),
                                        _dptr(cache.P), _dptr(cache.e), _dptr(cache.omega),
                                        _dptr(cache.M0), _dptr(cache.s), _dptr(rows), _dptr(ll),
                                        _dptr(bound), float(margin), _dptr(count),
                                        _stream(device)),
               "jk_settle_live_batch")
    return count


def settle_tol_batch(batch, cache, ll, bound, offsets, sizes, firsts=None, stars=None,
                     row_begins=None, rows=None, device=0, tol=API_TOL):
    """jk_settle_tol_batch: re-evaluate in FP64 every entry [firsts[i],
    sizes[i]) of each segment whose error bound exceeds tol (or is an upper
    bound), so that every entry is within tol of FP64.  Returns a device int64
    tensor of per-segment counts."""
    torch = _torch()
    lib = _lib.ensure_device(device)
    n_seg = len(sizes)
    segs = np.zeros(n_seg, dtype=LIVE_SEGMENT_DTYPE)
    segs["offset"] = offsets
    segs["n"] = sizes
    segs["first"] = 0 if firsts is None else firsts
    segs["star"] = np.arange(n_seg) if stars is None else stars
    segs["row_begin"] = 0 if row_begins is None else row_begins
    count = torch.empty(n_seg, dtype=torch.int64, device=ll.device)
    _lib.check(lib.jk_settle_tol_batch(batch.handle, n_seg, segs.ctypes.data_as(ctypes.c_void_p),
                                       _dptr(cache.P), _dptr(cache.e), _dptr(cache.omega),
                                       _dptr(cache.M0), _dptr(cache.s), _dptr(rows), _dptr(ll),
                                       _dptr(bound), float(tol), _dptr(count), _stream(device)),
               "jk_settle_tol_batch")
    return count


def settle_pruned_batch(batch, cache, ll, bound, offsets, sizes, stars=None, row_begins=None,
                        device=0, window=LIVE_MARGIN_NONZERO_U, margin=LIVE_MARGIN_NONZERO_U):
    """jk_settle_pruned_batch: in segments whose entries are values or upper
    bounds (bound -1, from a probe launch), evaluate in FP64 the upper-bound
    rows within `window` of the segment's upper bound on the maximum, then
    those within `margin` of the lower bound that leaves.  Returns a device
    int64 tensor of per-segment counts."""
    torch = _torch()
    lib = _lib.ensure_device(device)
    n_seg = len(sizes)
    segs = np.zeros(n_seg, dtype=LIVE_SEGMENT_DTYPE)
    segs["offset"] = offsets
    segs["n"] = sizes
    segs["first"] = 0
    segs["star"] = np.arange(n_seg) if stars is None else stars
    segs["row_begin"] = 0 if row_begins is None else row_begins
    count = torch.empty(n_seg, dtype=torch.int64, device=ll.device)
    _lib.check(lib.jk_settle_pruned_batch(batch.handle, n_seg,
                                          segs.ctypes.data_as(ctypes.c_void_p), _dptr(cache.P),
                                          _dptr(cache.e), _dptr(cache.omega), _dptr(cache.M0),
                                          _dptr(cache.s), _dptr(None), _dptr(ll), _dptr(bound),
                                          float(window), float(margin), _dptr(count),
                                          _stream(device)),
               "jk_settle_pruned_batch")
    return count


def probe_table(stars, jitter_mode="physical", device=0, min_epochs=PRUNE_MIN_EPOCHS,
                min_signal=PRUNE_MIN_SIGNAL):
    """(probe StarBatch, add per star, probed mask): entry k is star k's
    probe (device.probe_star) when it has at least min_epochs epochs, a
    diagonal noise covariance and a trend_signal of at least min_signal;
    other stars are not probed (their entry is the star itself, never read as
    a probe)."""
    from .device import probe_star, trend_signal
    probes, adds, probed = [], [], []
    for st in stars:
        ok = (st.cov is None and st.n >= min_epochs and
              trend_signal(st.arrays()) >= min_signal)
        if ok:
            pr, add = probe_star(st.arrays(), jitter_mode, device)
        else:
            pr, add = st, 0.0
        probes.append(pr)
        adds.append(add)
        probed.append(ok)
    return StarBatch(probes, device=device), np.asarray(adds), np.asarray(probed)


def prune_reject_batch(batch, probes, adds, probed, cache, d_items, n_items, offsets, sizes,
                       rng_states, max_keep, precision="mixed", shape=None, device=0):
    """Likelihood and rejection step of many stars (rows [0, sizes[k]) of the
    cache for table entry k) with the rows bounded by a probe: a probe launch
    leaves every probed star's rows as upper bounds, the rows that can matter
    are evaluated in FP64 around the maximum (settle_pruned_batch), and the
    settled rejection step decides every row like the FP64 path would, for
    every uniform.  Returns (ll, bound, settle_reject_batch's result, rows
    evaluated per star by the settling)."""
    torch = _torch()
    total = int(offsets[-1] + sizes[-1]) if len(sizes) else 0
    dev = cache.P.device
    prune = np.zeros((len(batch), 2))
    prune[:, 0] = np.where(probed, np.inf, -np.inf)
    prune[:, 1] = np.where(probed, adds, 0.0)
    d_prune = torch.as_tensor(prune.reshape(-1), device=dev)
    ll = torch.empty(total, dtype=torch.float64, device=dev)
    bnd = torch.empty(total, dtype=torch.float32, device=dev)
    lnl_prune_batch(batch, probes, d_prune, cache, d_items, n_items, precision, shape, out=ll,
                    bound=bnd)
    # only the probed stars' segments hold upper bounds
    k = np.flatnonzero(probed)
    if len(k) == len(sizes):
        # every star probed (one star, typically): the counts come back as they are
        cnt = settle_pruned_batch(batch, cache, ll, bnd, np.asarray(offsets), np.asarray(sizes),
                                  stars=k, device=device)
    else:
        cnt = torch.zeros(len(sizes), dtype=torch.int64, device=dev)
        if len(k):
            cnt[torch.as_tensor(k, device=dev)] = settle_pruned_batch(
                batch, cache, ll, bnd, np.asarray(offsets)[k], np.asarray(sizes)[k], stars=k,
                device=device)
    res = settle_reject_batch(batch, cache, ll, bnd, offsets, sizes, rng_states, max_keep,
                              device=device)
    return ll, bnd, res, cnt


def autotune(batch, cache, n_rows, precision="mixed", candidates=None, sample_stars=64,
             reps=3, device=0):
    """Time candidate launch shapes on a subset of the workload (the first
    `sample_stars` stars); return (best LaunchShape, {repr(shape): ms})."""
    torch = _torch()
    n_rows = np.asarray(n_rows, dtype=np.int64)
    if candidates is None:
        props = torch.cuda.get_device_properties(device)
        cus = props.multi_processor_count
        candidates = [LaunchShape(b, g * cus, t) for b in (256, 128) for g in (4, 8, 16)
                      for t in (1024, 4096)]
    sub = min(sample_stars, len(batch))
    results = {}
    best, best_ms = None, np.inf
    ws = torch.zeros(64, dtype=torch.int32, device=cache.P.device)
    for shape in candidates:
        items, _, total = work_items(n_rows[:sub], shape.tile)
        items = items[heaviest_first(items, batch.M)]
        d_items = upload_items(items, device)
        out = torch.empty(total, dtype=torch.float64, device=cache.P.device)
        lnl_batch(batch, cache, d_items, len(items), total, precision, shape, out=out, workspace=ws)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            lnl_batch(batch, cache, d_items, len(items), total, precision, shape, out=out,
                      workspace=ws)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / reps * 1e3
        results[repr(shape)] = ms
        if ms < best_ms:
            best, best_ms = shape, ms
    return best, results


def rejection_sample_many(datas, prior, prior_samples, rngs, max_posterior_samples=256,
                          n_linear_samples=1, precision="mixed", jitter_mode="physical",
                          shape=None, device=0, return_logprobs=False, return_all_logprobs=False,
                          prune=False):
    """TheJoker.rejection_sample(..., in_memory=True) for many stars at once.

    datas: RVData (or per-survey lists) per star; rngs: one numpy Generator
    per star (consumed exactly as rejection_sample_inmem consumes it,
    likelihood_helpers.py:91-127); prior_samples: a PriorCache (or
    JokerSamples) shared by all stars.  Returns one JokerSamples per star
    (and the per-star lnL arrays with return_all_logprobs).  prune: bound
    the rows of stars with at least PRUNE_MIN_EPOCHS epochs by a probe and
    evaluate in full only those that can matter (prune_reject_batch); the
    samples do not change."""
    from .device import PriorCache, linear_internal_units, star_arrays
    from .likelihood_helpers import validate_prepare_data
    from .thejoker import TheJoker

    torch = _torch()
    stars, units, all_data = [], [], []
    rv_unit = None
    for d in datas:
        dd, ids, trend = validate_prepare_data(d, prior.poly_trend, prior.n_offsets)
        if rv_unit is None:
            rv_unit = dd.rv.unit
        elif not dd.rv.unit.is_equivalent(rv_unit) or dd.rv.unit != rv_unit:
            raise ValueError("all stars of a batch must share one RV unit")
        a = star_arrays(dd, prior, trend)
        if a["cov"] is not None:
            raise NotImplementedError("data with a full covariance are evaluated one star at a "
                                      "time (TheJoker, precision='f64')")
        st = DeviceStar(a["t"], a["rv"], a["ivar"], a["trend"], a["mu"], a["lam"], a["t_ref"],
                        a["k_prior_mode"], a["sigma_K0"], a["P0"], a["max_K"],
                        jitter_mode=jitter_mode, device=device)
        st.data = dd
        st.internal_units = linear_internal_units(prior, dd.rv.unit)
        stars.append(st)
        all_data.append(dd)
    if isinstance(prior_samples, PriorCache):
        cache = prior_samples
    else:
        cache = PriorCache.from_samples(prior_samples, rv_unit=rv_unit, device=device)
    n = len(cache)
    batch = StarBatch(stars, device=device)
    shape = shape or LaunchShape()
    items, off, total = work_items([n] * len(stars), shape.tile)
    items = items[heaviest_first(items, batch.M)]
    states = [pcg64_state(g) for g in rngs]
    sizes = np.full(len(stars), n)
    d_items = upload_items(items, device)
    if prune:
        probes, adds, probed = probe_table(stars, jitter_mode, device)
        ll, bound, res, _ = prune_reject_batch(batch, probes, adds, probed, cache, d_items,
                                               len(items), off, sizes, states,
                                               max_posterior_samples, precision, shape, device)
        good = res[0]
        if return_all_logprobs:
            # the pruned rows hold upper bounds: evaluated before they are returned
            settle_tol_batch(batch, cache, ll, bound, off, sizes, device=device)
    elif precision == "mixed":
        bound = torch.empty(total, dtype=torch.float32, device=cache.P.device)
        ll = lnl_batch(batch, cache, d_items, len(items), total, precision, shape, bound=bound)
        # FP64 re-evaluation of exactly the rows whose decision the mixed
        # likelihood's error bound leaves open, and the rejection step
        good = settle_reject_batch(batch, cache, ll, bound, off, sizes, states,
                                   max_posterior_samples, device=device)[0]
        if return_all_logprobs:
            settle_tol_batch(batch, cache, ll, bound, off, sizes, device=device)
    else:
        ll = lnl_batch(batch, cache, d_items, len(items), total, precision, shape)
        good = reject_batch(ll, off, sizes, states, max_posterior_samples, device=device)[0]
    for g in rngs:
        pcg64_advance_state(g, n)
    ll_h = ll.cpu().numpy() if return_all_logprobs else None
    helper = TheJoker(prior, rng=np.random.default_rng(0), jitter_mode=jitter_mode,
                      precision=precision, device=device)
    out = []
    for s, st in enumerate(stars):
        samp = helper._full_samples(st, cache, good[s], rngs[s], n_linear_samples)
        if return_logprobs:
            if cache.ln_prior is not None:
                samp["ln_prior"] = cache.ln_prior[good[s]]
            samp["ln_likelihood"] = TheJoker._fp64_at(st, cache, good[s])
        if return_all_logprobs:
            out.append((samp, ll_h[off[s]:off[s] + n].copy()))
        else:
            out.append(samp)
    return out


class SegPlan:
    """jk_seg_plan: the settle and rejection steps of a fixed set of segments
    with their descriptors and scratch on the device, so that their launches
    can be captured in a HIP graph.  Results of a step: :meth:`results`."""

    def __init__(self, batch, offsets, sizes, rng_states, max_keep, stars=None, row_begins=None,
                 device=0):
        torch = _torch()
        lib = _lib.ensure_device(device)
        self._lib, self.device, self.batch = lib, device, batch
        n_seg = len(sizes)
        self.n_seg = n_seg
        self.sizes = np.asarray(sizes, dtype=np.int64)
        segs = np.zeros(n_seg, dtype=SETTLE_SEGMENT_DTYPE)
        segs["offset"] = offsets
        segs["n"] = sizes
        segs["star"] = np.arange(n_seg) if stars is None else stars
        segs["row_begin"] = 0 if row_begins is None else row_begins
        self._fill(segs, rng_states)
        h = ctypes.c_void_p()
        _lib.check(lib.jk_seg_plan_create(batch.handle, n_seg, segs.ctypes.data_as(ctypes.c_void_p),
                                          int(max_keep), ctypes.byref(h)), "jk_seg_plan_create")
        self._h = h
        self.keep = np.minimum(self.sizes, int(max_keep))
        self.keep_off = np.concatenate([[0], np.cumsum(self.keep)[:-1]]).astype(np.int64)
        # [n_pass | llmax | n_settled | kept indices]
        self.res = torch.zeros(3 * n_seg + max(int(self.keep.sum()), 1), dtype=torch.int64,
                               device=torch.device("cuda", device))

    @staticmethod
    def _fill(segs, rng_states, stream_offsets=None):
        for i, (st, inc) in enumerate(rng_states):
            segs["state_hi"][i], segs["state_lo"][i] = _split(int(st))
            segs["inc_hi"][i], segs["inc_lo"][i] = _split(int(inc))
        if stream_offsets is not None:
            segs["stream_offset"] = stream_offsets

    def set_streams(self, rng_states, stream_offsets=None):
        """New PCG64 streams, ordered on the current stream before later work."""
        q = np.zeros((self.n_seg, 5), dtype=np.uint64)
        for i, (st, inc) in enumerate(rng_states):
            q[i, 0], q[i, 1] = _split(int(st))
            q[i, 2], q[i, 3] = _split(int(inc))
        if stream_offsets is not None:
            q[:, 4] = stream_offsets
        _lib.check(self._lib.jk_seg_plan_set_streams(self._h, q.ctypes.data_as(ctypes.c_void_p),
                                                     _stream(self.device)),
                   "jk_seg_plan_set_streams")

    def settle(self, cache, ll, bound, rows=None):
        p = self.res.data_ptr()
        _lib.check(self._lib.jk_seg_plan_settle(
            self._h, _dptr(cache.P), _dptr(cache.e), _dptr(cache.omega), _dptr(cache.M0),
            _dptr(cache.s), _dptr(rows), _dptr(ll), _dptr(bound),
            ctypes.c_void_p(p + 16 * self.n_seg), _stream(self.device)), "jk_seg_plan_settle")

    def settle_reject(self, cache, ll, bound, rows=None):
        """Settling and rejection in one launch sequence (jk_settle_reject_batch)."""
        p = self.res.data_ptr()
        _lib.check(self._lib.jk_seg_plan_settle_reject(
            self._h, _dptr(cache.P), _dptr(cache.e), _dptr(cache.omega), _dptr(cache.M0),
            _dptr(cache.s), _dptr(rows), _dptr(ll), _dptr(bound),
            ctypes.c_void_p(p + 24 * self.n_seg), ctypes.c_void_p(p),
            ctypes.c_void_p(p + 8 * self.n_seg), ctypes.c_void_p(p + 16 * self.n_seg),
            _stream(self.device)), "jk_seg_plan_settle_reject")

    def reject(self, ll):
        p = self.res.data_ptr()
        _lib.check(self._lib.jk_seg_plan_reject(
            self._h, _dptr(ll), ctypes.c_void_p(p + 24 * self.n_seg), ctypes.c_void_p(p),
            ctypes.c_void_p(p + 8 * self.n_seg), _stream(self.device)), "jk_seg_plan_reject")

    def results(self):
        """(kept index arrays, n_pass, llmax, rows settled in FP64) of the last
        step, after waiting for it."""
        h = to_host(self.res)
        n = self.n_seg
        n_pass = h[:n].copy()
        mx = h[n:2 * n].view(np.float64).copy()
        settled = h[2 * n:3 * n].copy()
        idx = h[3 * n:]
        out = [idx[o:o + min(int(k), int(npass))].copy()
               for o, k, npass in zip(self.keep_off, self.keep, n_pass)]
        return out, n_pass, mx, settled

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.jk_seg_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StepGraph:
    """One rejection-sampling step of many stars -- the persistent lnL launch
    (with error bounds in mixed precision), the FP64 settling and the
    segmented rejection step -- captured once in a HIP graph (through
    torch.cuda.CUDAGraph) and replayed with new numpy PCG64 streams.  Nothing
    in the step allocates, copies from the host or waits, so a replay is a
    single graph launch."""

    def __init__(self, batch, cache, d_items, n_items, total, offsets, sizes, rng_states,
                 max_keep, precision="mixed", shape=None, device=0):
        torch = _torch()
        dev = torch.device("cuda", device)
        self.precision, self.cache = precision, cache
        self.ll = torch.empty(total, dtype=torch.float64, device=dev)
        self.bound = (torch.empty(total, dtype=torch.float32, device=dev)
                      if precision == "mixed" else None)
        self.ws = torch.zeros(64, dtype=torch.int32, device=dev)
        self.plan = SegPlan(batch, offsets, sizes, rng_states, max_keep, device=device)
        shape = shape or LaunchShape()
        if shape.grid <= 0:
            props = torch.cuda.get_device_properties(device)
            shape = LaunchShape(shape.block, 8 * props.multi_processor_count, shape.tile)

        def body():
            lnl_batch(batch, cache, d_items, n_items, total, precision, shape, out=self.ll,
                      workspace=self.ws, bound=self.bound)
            if precision == "mixed":
                self.plan.settle_reject(cache, self.ll, self.bound)
            else:
                self.plan.reject(self.ll)

        body()                      # warm up every launch path once, eagerly
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            body()
        torch.cuda.synchronize()

    def launch(self, rng_states, stream_offsets=None):
        """Queue a replay of the step with these streams and return at once;
        :meth:`results` collects it."""
        self.plan.set_streams(rng_states, stream_offsets)
        self.graph.replay()

    def results(self):
        """SegPlan.results() of the last launched step, after waiting for it."""
        return self.plan.results()

    def run(self, rng_states, stream_offsets=None):
        """Replay the step with these streams; returns SegPlan.results()."""
        self.launch(rng_states, stream_offsets)
        return self.results()
