"""Continuous-batching request loop over a device-resident prior cache.

Each request is one star run through ``iterative_rejection_sample``
(thejoker.py:259-370) against a shared :class:`PriorCache`.  Requests are
admitted while active slots and the HBM budget allow; every step evaluates
the next batch of rows of *every* active request in one persistent
``jk_lnl_batch`` launch (stars of different epoch counts side by side), then
runs each request's FP64 re-evaluation window and rejection step, advances
its iteration state (:class:`thejoker_amd.thejoker.IterativeRejection`, the
same code the single-star API runs) and retires finished requests while new
ones are admitted.  A request's result is bit-for-bit the one
``TheJoker.iterative_rejection_sample`` returns for that star alone: the
batched lnL is bit-identical to the single-star kernel, and the random
stream of each request is its own.
"""
import time

import numpy as np

from .batch import LaunchShape, StarBatch, lnl_batch, upload_items, work_items
from .device import DeviceStar, _torch, linear_internal_units, polish, star_arrays
from .likelihood_helpers import validate_prepare_data
from .thejoker import POLISH_WINDOW, IterativeRejection, TheJoker

__all__ = ["Request", "JokerServer", "run_poisson"]


class Request:
    def __init__(self, rid, data, rng, arrival):
        self.id = rid
        self.data = data
        self.rng = rng
        self.arrival = arrival
        self.t_admit = None
        self.t_done = None
        self.star = None
        self.it = None
        self.ll = None
        self.result = None
        self.error = None
        self.rows_evaluated = 0

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival


class JokerServer:
    """
    Parameters
    ----------
    prior : JokerPrior
    cache : PriorCache
        Device-resident prior rows shared by every request.
    n_requested : int
        Posterior samples per star (iterative_rejection_sample's
        n_requested_samples).
    precision : "mixed" or "f64"
    max_active : int
        Requests evaluated together.
    hbm_budget_bytes : int
        Device memory the per-request lnL buffers may hold together.
    in_memory : bool
        Iteration rule of the in-memory path (safety factor 1) or of the
        default path (safety factor 4).
    """

    def __init__(self, prior, cache, n_requested=256, precision="mixed", jitter_mode="physical",
                 max_active=64, hbm_budget_bytes=16 << 30, in_memory=False, growth_factor=128,
                 n_linear_samples=1, shape=None, device=0, clock=time.perf_counter):
        self.prior = prior
        self.cache = cache
        self.n_requested = int(n_requested)
        self.precision = precision
        self.jitter_mode = jitter_mode
        self.max_active = int(max_active)
        self.budget = int(hbm_budget_bytes)
        self.in_memory = in_memory
        self.growth_factor = growth_factor
        self.n_linear_samples = n_linear_samples
        self.shape = shape or LaunchShape(256, 0, 2048)
        self.device = device
        self.clock = clock
        self.t0 = clock()
        self.pending = []
        self.active = []
        self.finished = []
        self.reserved = 0
        self.steps = 0
        self.sample_epochs = 0
        self._batch = None
        self._batch_key = None
        self._helper = TheJoker(prior, rng=np.random.default_rng(0), jitter_mode=jitter_mode,
                                precision=precision, device=device)
        torch = _torch()
        self._ws = torch.zeros(64, dtype=torch.int32, device=cache.P.device)

    def now(self):
        return self.clock() - self.t0

    def submit(self, data, rng, rid=None, arrival=None):
        r = Request(len(self.finished) + len(self.active) + len(self.pending) if rid is None else rid,
                    data, rng, self.now() if arrival is None else arrival)
        self.pending.append(r)
        return r

    # ------------------------------------------------------------------
    def _need_bytes(self, n_rows):
        return 8 * int(n_rows)

    def _admit(self):
        torch = _torch()
        while self.pending and len(self.active) < self.max_active:
            r = self.pending[0]
            first = self.growth_factor * self.n_requested
            if self.reserved + self._need_bytes(first) > self.budget and self.active:
                break
            self.pending.pop(0)
            try:
                dd, ids, trend = validate_prepare_data(r.data, self.prior.poly_trend,
                                                       self.prior.n_offsets)
                a = star_arrays(dd, self.prior, trend)
                r.star = DeviceStar(a["t"], a["rv"], a["ivar"], a["trend"], a["mu"], a["lam"],
                                    a["t_ref"], a["k_prior_mode"], a["sigma_K0"], a["P0"],
                                    a["max_K"], jitter_mode=self.jitter_mode, device=self.device)
                r.star.data = dd
                r.star.internal_units = linear_internal_units(self.prior, dd.rv.unit)
                r.it = IterativeRejection(len(self.cache), self.n_requested, r.rng,
                                          in_memory=self.in_memory,
                                          growth_factor=self.growth_factor,
                                          device=self.cache.P.device)
            except Exception as exc:  # a bad request fails alone
                r.error = exc
                r.t_done = self.now()
                self.finished.append(r)
                continue
            r.ll = torch.empty(first, dtype=torch.float64, device=self.cache.P.device)
            self.reserved += self._need_bytes(first)
            r.t_admit = self.now()
            self.active.append(r)

    def _grow(self, r, need):
        torch = _torch()
        if need <= r.ll.shape[0]:
            return
        cap = max(need, 2 * r.ll.shape[0])
        cap = min(cap, r.it.max_prior)
        new = torch.empty(cap, dtype=torch.float64, device=r.ll.device)
        new[:r.it.n_evals] = r.ll[:r.it.n_evals]
        self.reserved += self._need_bytes(cap - r.ll.shape[0])
        r.ll = new

    def step(self):
        """Admit, evaluate one round of every active request, retire the
        finished.  Returns the requests finished in this step."""
        self._admit()
        if not self.active:
            return []
        torch = _torch()
        act = self.active
        key = tuple(id(r) for r in act)
        if key != self._batch_key:
            self._batch = StarBatch([r.star for r in act], device=self.device)
            self._batch_key = key
        rows_l, begins, sizes = [], [], []
        for r in act:
            rows, start, n_now = r.it.next_batch()
            self._grow(r, r.it.n_evals + n_now)
            rows_l.append(rows)
            begins.append(start)
            sizes.append(n_now)
        sizes = np.array(sizes, dtype=np.int64)
        if any(x is not None for x in rows_l):
            # Randomised prior order: gather the rows of every request.
            row_arr = torch.cat([x if x is not None else
                                 torch.arange(b, b + n, device=self.cache.P.device)
                                 for x, b, n in zip(rows_l, begins, sizes)])
            items, off, total = work_items(sizes, self.shape.tile,
                                           row_begin=np.concatenate([[0], np.cumsum(sizes)[:-1]]))
        else:
            row_arr = None
            items, off, total = work_items(sizes, self.shape.tile, row_begin=np.array(begins))
        out = lnl_batch(self._batch, self.cache, upload_items(items, self.device), len(items),
                        total, self.precision, self.shape, rows=row_arr, workspace=self._ws)
        self.sample_epochs += int(np.sum(sizes * self._batch.M))
        done = []
        for k, r in enumerate(act):
            n0 = r.it.n_evals
            n_now = int(sizes[k])
            r.ll[n0:n0 + n_now] = out[int(off[k]):int(off[k]) + n_now]
            r.rows_evaluated += n_now
            view = r.ll[:n0 + n_now]
            try:
                if self.precision == "mixed":
                    if r.it.all_idx is None:
                        polish(r.star, self.cache, view, start=0, window=POLISH_WINDOW)
                    else:
                        polish(r.star, self.cache, view, rows=r.it._d_all_idx[:n0 + n_now],
                               window=POLISH_WINDOW)
                r.it.absorb(view, self.device)
            except Exception as exc:
                r.error = exc
                r.it.done = True
            if r.it.done:
                done.append(r)
        for r in done:
            self.active.remove(r)
            self.reserved -= self._need_bytes(r.ll.shape[0])
            if r.error is None:
                full_idx, _ = r.it.result()
                r.result = self._helper._full_samples(r.star, self.cache, full_idx, r.rng,
                                                      self.n_linear_samples,
                                                      spawn_batches=None if self.in_memory else 1)
            r.ll = None
            r.t_done = self.now()
            self.finished.append(r)
        self.steps += 1
        return done

    def busy(self):
        return bool(self.pending or self.active)


def run_poisson(server, arrival_times, datas, rngs, max_wall=None):
    """Drive `server` with requests arriving at `arrival_times` (seconds from
    start): submit each when its time comes, step while anything is pending
    or active.  Returns the finished requests in arrival order."""
    n = len(arrival_times)
    i = 0
    server.t0 = server.clock()
    while i < n or server.busy():
        now = server.now()
        while i < n and arrival_times[i] <= now:
            server.submit(datas[i], rngs[i], rid=i, arrival=float(arrival_times[i]))
            i += 1
        if server.busy():
            server.step()
        elif i < n:
            time.sleep(max(0.0, min(arrival_times[i] - server.now(), 0.01)))
        if max_wall is not None and server.now() > max_wall:
            break
    return sorted(server.finished, key=lambda r: r.id)
