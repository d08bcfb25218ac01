"""Continuous-batching request loop over a device-resident prior cache.

Each request is one star run through ``iterative_rejection_sample``
(thejoker.py:259-370) against a shared :class:`PriorCache`.  Requests are
admitted while active slots and the HBM arena allow; each admitted star takes
a slot of one device star table (``jk_batch_create_slots``) until it
finishes.  Every step evaluates the next batch of rows of *every* active
request in one persistent ``jk_lnl_batch`` launch (stars of different epoch
counts side by side) writing straight into each request's region of one HBM
lnL arena (mixed precision: with each row's error bound in a parallel
arena), then re-evaluates in FP64 the new rows that the bound leaves within
40 of the maximum, which makes every row decide like FP64 in this and all
later rounds under every uniform but exactly 0 (``jk_settle_live_batch``),
and runs every request's rejection step with its own numpy PCG64 stream
(``jk_reject_batch_zero_check``; a zero uniform, probability 2^-53 per row,
settles that request's rows for every uniform and repeats its rejection
step).  The iteration state of each request is
:class:`thejoker_amd.thejoker.IterativeRejection` -- the code the single-star
API runs -- and the batched lnL is bit-identical to the single-star kernel,
so a request's result is exactly what ``TheJoker.iterative_rejection_sample``
returns for that star alone.

When every active request waits for arena memory to grow, the most recently
admitted one is restarted: its region and slot are freed and it is queued
again with its random stream as it was at admission, so it repeats the same
rounds and its result is unchanged.  No request is admitted while another
waits for memory.
"""
import gc
import time

import numpy as np

from .batch import (EXACT_SWITCH, LIVE_MARGIN, LIVE_MARGIN_NONZERO_U, PRUNE_MIN_SIGNAL,
                    ItemUploader, LaunchShape, StarBatch, heaviest_first, lnl_batch,
                    lnl_prune_batch, reject_batch, settle_live_batch, settle_reject_batch,
                    work_items)
from .device import (DeviceStar, _torch, linear_internal_units, pcg64_state,
                     posterior_inputs_table, probe_star, star_arrays, trend_signal)
from .likelihood_helpers import validate_prepare_data
from .thejoker import IterativeRejection, TheJoker

__all__ = ["Request", "RangeAllocator", "JokerServer", "run_poisson", "share_budget"]

# Step-budget cost of an exact row relative to a mixed one.
EXACT_COST = 1.5


class RangeAllocator:
    """First-fit allocator of [0, capacity) element ranges."""

    def __init__(self, capacity):
        self.capacity = int(capacity)
        self.free = [(0, self.capacity)]

    def alloc(self, n):
        n = int(n)
        for k, (a, b) in enumerate(self.free):
            if b - a >= n:
                if b - a == n:
                    self.free.pop(k)
                else:
                    self.free[k] = (a + n, b)
                return a
        return None

    def extend(self, start, n, new_n):
        """Grow the range [start, start + n) in place to new_n elements if the
        free range right after it is large enough; True on success."""
        end, grow = int(start) + int(n), int(new_n) - int(n)
        for k, (a, b) in enumerate(self.free):
            if a == end and b - a >= grow:
                if b - a == grow:
                    self.free.pop(k)
                else:
                    self.free[k] = (a + grow, b)
                return True
        return False

    def take(self, start, n):
        """Reserve exactly [start, start + n), which must be free."""
        a0, b0 = int(start), int(start) + int(n)
        for k, (a, b) in enumerate(self.free):
            if a <= a0 and b0 <= b:
                parts = [(x, y) for x, y in ((a, a0), (b0, b)) if y > x]
                self.free[k:k + 1] = parts
                return
        raise ValueError(f"range [{a0}, {b0}) is not free")

    def release(self, start, n):
        a, b = int(start), int(start) + int(n)
        self.free.append((a, b))
        self.free.sort()
        merged = []
        for x, y in self.free:
            if merged and merged[-1][1] == x:
                merged[-1] = (merged[-1][0], y)
            else:
                merged.append((x, y))
        self.free = merged

    def used(self):
        return self.capacity - sum(b - a for a, b in self.free)


def share_budget(rem, M, budget, min_rows, policy="share"):
    """Rows for each request from a budget of sample-epochs, M[i] being the
    cost of one of request i's rows, in increasing order of remaining work
    rem[i] * M[i].  policy "share": each takes its remainder or an equal
    share of the budget left, whichever is less; "srpt" (shortest remaining
    work first): each takes as much of its remainder as the budget left
    covers.  Either way every request takes at least min_rows (or its
    remainder), so each one advances in every step."""
    rem = np.asarray(rem, dtype=np.int64)
    M = np.maximum(np.asarray(M, dtype=np.float64), 1e-9)
    left = float(budget)
    out = np.zeros(len(rem), dtype=np.int64)
    order = np.argsort(rem * M, kind="stable")
    for k, i in enumerate(order):
        share = left / (len(order) - k) if policy == "share" else left
        rows = min(int(rem[i]), max(int(min_rows), int(max(share, 0.0) // M[i])))
        out[i] = rows
        left -= rows * M[i]
    return out


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
        self.region = None      # (start, capacity) in the arena
        self.slot = None        # star-table slot
        self.post = None        # (packed rows, a, cov) of the accepted rows
        self.result = None
        self.error = None
        self.rows_evaluated = 0
        self.sample_epochs = 0
        self.rng_state0 = None  # the stream at admission, for a restart
        self.preempted = 0      # times restarted to free arena memory
        self.need_hint = 0      # arena values held when last preempted
        self.round_done = 0     # rows of the current round already evaluated
        self.probe = None       # probe star (DeviceStar on a subset of the epochs)
        self.prune_add = None   # its bound's addend; None: the request is not probed
        self.full_seen = 0      # its slot's full-evaluation counter, last read
        self.full_rate = 0.15   # share of its probed rows evaluated in full, last measured
        self.exact = False      # its slot evaluates in FP64 (flat likelihood)
        self.prepared = False   # data laid out and iterative state made, not yet admitted
        self.arrays = self.data_prepared = None
        self.to_probe = False
        self.shell = False      # settles at each rejection step (flat likelihood)
        self.round_settled = 0  # rows of the current round settled in FP64

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
        Posterior samples per star (n_requested_samples).
    precision : "mixed" or "f64"
    max_active : int
        Requests evaluated together.
    hbm_budget_bytes : int
        Size of the lnL arena the active requests share.
    in_memory : bool
        Iteration rule of the in-memory path (safety factor 1) or of the
        default path (safety factor 4).
    profile : bool
        Synchronise the device between the phases of every step and add each
        phase's wall time to ``self.phase_s`` (admit, plan, lnl, polish,
        reject, update, finish).
    max_step_work : int or None
        Sample-epochs one step may evaluate (a probed request's row costs its
        probe's epochs plus its measured share of rows evaluated in full
        times its epochs); None evaluates every active request's whole round
        in each step.  A round larger than a request's share is spread over
        several steps (its rejection step runs when the round is complete),
        so that short requests are not held behind long ones, and new
        arrivals wait at most one short step for admission.
    budget_policy : "srpt" or "share"
        How a step budget is shared: shortest remaining work first (the
        smallest rounds complete at once and the budget left goes to the
        request nearest completion), or equal shares.
    budget_max_pending : int or None
        The budget applies only while fewer requests than this wait for
        admission (default max_active); with a longer queue every round runs
        whole, for throughput.
    min_chunk_rows : int
        Smallest part of a round a step evaluates under a budget.
    prune : bool
        After a request's first round, skip the full evaluation of rows whose
        upper bound from a probe on a subset of the star's epochs
        (probe_star) lies more than the settling margin below the maximum of
        the earlier rounds: such a row cannot pass the rejection step for
        any uniform but 0 (mixed precision, where the zero-uniform check
        covers that case) or at all (f64, margin 750).  It keeps its bound
        as its value.  Results are unchanged.
    prune_min_epochs : int
        Stars with fewer epochs are not probed (the probe would cost about
        as much as the evaluation it saves).
    prune_min_signal : float
        Nor are stars whose trend-only fit leaves less than this for an orbit
        to gain in lnL (device.trend_signal): a probe prunes next to none of
        their rows.
    exact_switch : float or None
        Mixed precision: a request whose round settled more than this share
        of its rows in FP64 (a flat likelihood, nearly every row within the
        margin of the maximum) leaves the live settling for `flat_mode`;
        None never switches.  Results are unchanged.
    flat_mode : "shell" or "exact"
        "shell": the request's later rows stay in mixed precision, unsettled,
        and each rejection step settles what that step's uniforms need
        (settle_reject_batch: the rows that could hold the maximum and those
        whose decision the bound leaves open, a few per million) -- one more
        pass over its rows per round instead of an FP64 evaluation of each.
        "exact": its later rows are evaluated in FP64 (StarBatch.set_exact).
    prepare_extra : int
        While a step's likelihood launch runs, the server prepares (checks
        and lays out the data, sets up the iterative state) the pending
        requests that free slots will admit next, plus this many more.
    freeze_gc : bool
        Call gc.freeze() at construction: the objects alive then (the
        interpreter's modules, torch, the prior cache) move to the garbage
        collector's permanent generation and later collections skip them.
        Otherwise a full collection scans them all (30-80 ms with torch
        loaded) and stalls whichever step it falls in.  Reference counting
        still frees those objects; only a reference cycle among them would
        stay uncollected (gc.unfreeze() undoes it).
    """

    PHASES = ("admit", "plan", "lnl", "polish", "reject", "update", "finish")

    def __init__(self, prior, cache, n_requested=256, precision="mixed", jitter_mode="physical",
                 max_active=64, hbm_budget_bytes=4 << 30, in_memory=False, growth_factor=128,
                 n_linear_samples=1, shape=None, device=0, clock=time.perf_counter,
                 profile=False, max_step_work=None, min_chunk_rows=1 << 16, prune=True,
                 prune_min_epochs=12, prune_min_signal=PRUNE_MIN_SIGNAL, budget_policy="srpt",
                 budget_max_pending=None, exact_switch=EXACT_SWITCH, flat_mode="shell",
                 freeze_gc=True, prepare_extra=32):
        torch = _torch()
        self.prune = bool(prune)
        self.prune_min_epochs = int(prune_min_epochs)
        self.prune_min_signal = float(prune_min_signal)
        self.exact_switch = None if exact_switch is None else float(exact_switch)
        if flat_mode not in ("shell", "exact"):
            raise ValueError("flat_mode must be 'shell' or 'exact'")
        self.flat_mode = flat_mode
        self.prepare_extra = int(prepare_extra)
        if freeze_gc:
            gc.freeze()
        self.max_step_work = None if max_step_work is None else int(max_step_work)
        if budget_policy not in ("srpt", "share"):
            raise ValueError("budget_policy must be 'srpt' or 'share'")
        self.budget_policy = budget_policy
        self.budget_max_pending = (int(max_active) if budget_max_pending is None
                                   else int(budget_max_pending))
        self.min_chunk_rows = int(min_chunk_rows)
        self.prior = prior
        self.cache = cache
        self.n_requested = int(n_requested)
        self.precision = precision
        self.jitter_mode = jitter_mode
        self.max_active = int(max_active)
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
        self.steps = 0
        self.sample_epochs = 0
        # lnL (8 B) and, in mixed precision, its error bound (4 B) per row
        per_row = 12 if precision == "mixed" else 8
        self.arena = torch.empty(int(hbm_budget_bytes) // per_row, dtype=torch.float64,
                                 device=cache.P.device)
        self.arena_b = (torch.empty(self.arena.shape[0], dtype=torch.float32,
                                    device=cache.P.device) if precision == "mixed" else None)
        self.alloc = RangeAllocator(self.arena.shape[0])
        self._table = StarBatch.slots(self.max_active, 1 + prior.poly_trend + prior.n_offsets,
                                      jitter_mode, device=device)
        self._free_slots = list(range(self.max_active - 1, -1, -1))
        self._probes = None
        if self.prune:
            self._probes = StarBatch.slots(self.max_active, self._table.L, jitter_mode,
                                           device=device)
            self._prune_host = np.zeros((self.max_active, 2), dtype=np.float64)
            self._prune_dev = torch.zeros(2 * self.max_active, dtype=torch.float64,
                                          device=cache.P.device)
        self._n_full = torch.zeros(self.max_active, dtype=torch.int64, device=cache.P.device)
        self.rows_probed = 0      # rows of requests whose star was probed
        self.rows_pruned = 0      # of those, the rows that kept their bound
        self._ws = torch.zeros(64, dtype=torch.int32, device=cache.P.device)
        self._uploader = ItemUploader(device)
        self.profile = profile
        self.phase_s = dict.fromkeys(self.PHASES, 0.0)
        self._t_mark = None
        self._cur = {}
        self.step_profile = []    # profile: per step, phase times and sizes
        self.last_step = {}
        self._to_finish = []      # retired requests whose samples are still to be drawn
        self._waiting = 0         # active requests the last step could not grow
        self.preemptions = 0
        self.zero_uniform_events = 0
        self.flat_switches = 0
        self.region_moves = {"extend": 0, "move": 0, "overlap": 0}
        self._helper = TheJoker(prior, rng=np.random.default_rng(0), jitter_mode=jitter_mode,
                                precision=precision, device=device)

    def now(self):
        return self.clock() - self.t0

    def submit(self, data, rng, rid=None, arrival=None):
        if rid is None:
            rid = len(self.finished) + len(self.active) + len(self.pending)
        r = Request(rid, data, rng, self.now() if arrival is None else arrival)
        self.pending.append(r)
        return r

    def busy(self):
        return bool(self.pending or self.active or self._to_finish)

    def _mark(self, phase=None):
        """Profiling: charge the time since the last mark to `phase`."""
        if not self.profile:
            return
        _torch().cuda.synchronize()
        t = time.perf_counter()
        if phase is not None and self._t_mark is not None:
            self.phase_s[phase] += t - self._t_mark
            self._cur[phase] = self._cur.get(phase, 0.0) + t - self._t_mark
        self._t_mark = t

    # ------------------------------------------------------------------
    def _prune_thresholds(self, plan):
        """(threshold, add) per table slot for this step's launch, or None when
        no request in `plan` is probed.  A request is probed once it has a
        maximum from earlier rounds (its floor, which the final maximum can
        only exceed); the threshold is that floor minus the settling margin
        (40 in mixed precision, whose zero-uniform check covers u == 0; 750
        in FP64)."""
        if not self.prune:
            return None
        margin = LIVE_MARGIN_NONZERO_U if self.precision == "mixed" else LIVE_MARGIN
        thr = self._prune_host
        thr[:, 0] = -np.inf
        thr[:, 1] = 0.0
        any_on = False
        for r, _, _ in plan:
            if r.prune_add is not None and np.isfinite(r.it.floor):
                thr[r.slot] = (r.it.floor - margin, r.prune_add)
                any_on = True
        return thr if any_on else None

    def prune_stats(self):
        """(rows probed, rows pruned): the rows of requests whose star was
        probed, and of those the rows that kept their bound instead of a full
        evaluation."""
        return self.rows_probed, self.rows_pruned

    def _probe_feedback(self, probed):
        """After a step: count each probed request's pruned rows from the
        per-slot counters of full evaluations, and keep each request's share
        of rows evaluated in full for the step budget's cost model."""
        full = self._n_full.cpu().numpy()
        for r, slot, size in probed:
            d = int(full[slot]) - r.full_seen
            r.full_seen = int(full[slot])
            if size:
                r.full_rate = d / float(size)
            self.rows_probed += int(size)
            self.rows_pruned += int(size) - d

    def _release(self, r):
        if r.region is not None:
            self.alloc.release(*r.region)
            r.region = None
        if r.slot is not None:
            self._free_slots.append(r.slot)
            r.slot = None

    def _fail(self, r, exc):
        r.error = exc
        r.t_done = self.now()
        self._release(r)
        self.finished.append(r)

    def _prepare(self, r):
        """The host half of admission, which needs neither the device nor a
        slot or arena memory: the star's data checked and laid out, its
        iterative state set up, the decision to probe it.  Raises what a bad
        request raises."""
        if r.rng_state0 is None:
            r.rng_state0 = r.rng.bit_generator.state
        dd, ids, trend = validate_prepare_data(r.data, self.prior.poly_trend,
                                               self.prior.n_offsets)
        a = star_arrays(dd, self.prior, trend)
        if a["cov"] is not None:
            raise NotImplementedError("data with a full covariance are evaluated one "
                                      "star at a time (TheJoker, precision='f64')")
        r.arrays, r.data_prepared = a, dd
        r.it = IterativeRejection(len(self.cache), self.n_requested, r.rng,
                                  in_memory=self.in_memory, growth_factor=self.growth_factor,
                                  device=self.cache.P.device)
        r.to_probe = (self.prune and a["t"].size >= self.prune_min_epochs and
                      trend_signal(a) >= self.prune_min_signal)
        r.prepared = True

    def _create_device_stars(self, r):
        """The device half of admission: the request's device star (and
        probe).  At the start of a step, when the device is idle: an upload
        behind a running persistent launch would wait for it to end."""
        a, dd = r.arrays, r.data_prepared
        r.star = DeviceStar(a["t"], a["rv"], a["ivar"], a["trend"], a["mu"], a["lam"],
                            a["t_ref"], a["k_prior_mode"], a["sigma_K0"], a["P0"],
                            a["max_K"], jitter_mode=self.jitter_mode, device=self.device)
        r.star.data = dd
        r.star.internal_units = linear_internal_units(self.prior, dd.rv.unit)
        r.probe, r.prune_add = None, None
        if r.to_probe:
            r.probe, r.prune_add = probe_star(a, self.jitter_mode, self.device)
        r.arrays = r.data_prepared = None

    def _prepare_ahead(self):
        """The host half of admission for the pending requests that the next
        step will admit, while this step's likelihood launch runs."""
        room = self.max_active - len(self.active)
        for r in list(self.pending[:max(room, 0) + self.prepare_extra]):
            if r.prepared or r.need_hint:
                continue
            try:
                self._prepare(r)
            except Exception as exc:       # a bad request fails alone
                self.pending.remove(r)
                self._fail(r, exc)

    def _admit(self):
        # no admissions while an active request waits for arena memory
        while self.pending and len(self.active) < self.max_active and not self._waiting:
            r = self.pending[0]
            first = self.growth_factor * self.n_requested
            if r.need_hint and self.alloc.capacity - self.alloc.used() < r.need_hint:
                break                      # a restarted request waits for its room
            start = self.alloc.alloc(first)
            if start is None:
                break                      # wait for memory
            self.pending.pop(0)
            r.region = (start, first)
            try:
                if not r.prepared:
                    self._prepare(r)
                self._create_device_stars(r)
                r.slot = self._free_slots.pop()
                self._table.set(r.slot, r.star)
                if r.prune_add is not None:
                    self._probes.set(r.slot, r.probe)
                    self._n_full[r.slot] = 0
                    r.full_seen = 0
            except Exception as exc:       # a bad request fails alone
                self._fail(r, exc)
                continue
            r.t_admit = self.now()
            self.active.append(r)

    def _preempt(self, r):
        """Free r's arena region and slot and queue r again at the front,
        with its random stream as it was at admission: the restart repeats
        the same rounds, so its result does not change."""
        self.active.remove(r)
        r.need_hint = r.region[1]
        self._release(r)
        r.it = None
        r.prepared = False
        r.round_done = 0
        r.round_settled = 0
        r.exact = False
        r.shell = False
        r.rng.bit_generator.state = r.rng_state0
        r.preempted += 1
        self.preemptions += 1
        self.pending.insert(0, r)

    def _ensure_region(self, r, need):
        """Grow r's arena region to hold `need` values; False if the arena is
        full for now."""
        start, cap = r.region
        if need <= cap:
            return True
        # A round that needs a quarter of the cache or more is followed, if at
        # all, by one that needs the rest: take the whole cache's room at
        # once (when the arena has it), so the values are not moved again.
        want = max(need, 2 * cap)
        if need >= r.it.max_prior // 4:
            want = r.it.max_prior
        caps = [min(want, r.it.max_prior)]
        if caps[0] > need:
            caps.append(need)
        # in place, when the range after the region is free
        for c in caps:
            if self.alloc.extend(start, cap, c):
                r.region = (start, c)
                self.region_moves["extend"] += 1
                return True
        # elsewhere, the old region counted as free (the new range may
        # overlap it: the values then move through a temporary copy)
        self.alloc.release(start, cap)
        new = None
        for c in caps:
            new = self.alloc.alloc(c)
            if new is not None:
                break
        if new is None:
            self.alloc.take(start, cap)
            return False
        n = r.it.n_evals
        self.region_moves["move"] += 1
        if n and new != start:
            overlap = new < start + n and start < new + n
            self.region_moves["overlap"] += int(overlap)
            for buf in (self.arena, self.arena_b):
                if buf is not None:
                    src = buf[start:start + n]
                    buf[new:new + n] = src.clone() if overlap else src
        r.region = (new, c)
        return True

    def _build_results(self):
        """Host half of the finished requests' results (linear-parameter draws
        and JokerSamples); returns those requests."""
        done, self._to_finish = self._to_finish, []
        for r in done:
            r.result = self._helper._samples_from_posterior(
                r.star, r.post, r.rng, self.n_linear_samples,
                spawn_batches=None if self.in_memory else 1)
            r.post = None
            r.t_done = self.now()
            self.finished.append(r)
        return done

    def _chunks(self, plan):
        """Rows of each planned round to evaluate in this step.  Without a
        step budget: all that remain.  With one (max_step_work sample-epochs),
        the budget is shared like a processor: requests in increasing order of
        remaining work each take up to an equal share of what is left (at
        least min_chunk_rows), so short requests finish in one short step."""
        rem = np.array([n - r.round_done for r, _, n in plan], dtype=np.int64)
        if self.max_step_work is None or len(self.pending) >= self.budget_max_pending:
            return rem
        M = self._table.M[[r.slot for r, _, _ in plan]].astype(np.float64)
        # a probed round costs the probe's epochs plus the share of rows
        # evaluated in full (measured on the request's earlier rounds)
        for k, (r, _, _) in enumerate(plan):
            if r.prune_add is not None and np.isfinite(r.it.floor):
                M[k] = r.probe.n + r.full_rate * M[k]
            if r.exact:
                M[k] *= EXACT_COST
        return share_budget(rem, M, self.max_step_work, self.min_chunk_rows, self.budget_policy)

    def _reject_ends(self, plan, ends, seg_off, slots, seg_n, states):
        """The rejection step of the requests plan[k], k in `ends`, whose round
        is complete: (accepted index arrays, pass counts, maxima).  Requests in
        shell mode settle and decide in one sequence (settle_reject_batch:
        FP64 for the rows that could hold the maximum or whose decision
        under this round's uniforms the bound leaves open); the others, whose
        rows within the margin of the maximum are FP64 already, take the plain
        step with the zero-uniform check."""
        mixed = self.precision == "mixed"
        ends = np.asarray(ends)
        n = len(ends)
        good, n_pass, llmax = [None] * n, np.zeros(n, dtype=np.int64), np.zeros(n)
        sh = np.array([mixed and plan[k][0].shell for k in ends], dtype=bool)
        lv = np.flatnonzero(~sh)
        if len(lv):
            ks = ends[lv]
            got = reject_batch(self.arena, seg_off[ks], seg_n[lv], [states[j] for j in lv],
                               self.n_requested, device=self.device, zero_check=mixed)
            for t, j in enumerate(lv):
                good[j], n_pass[j], llmax[j] = got[0][t], got[1][t], got[2][t]
            if mixed and got[3].any():
                # A uniform of exactly 0 accepts any row within 745 of the
                # maximum, which the 40 margin did not settle: settle those
                # segments for every uniform and repeat their rejection step.
                z = lv[np.flatnonzero(got[3])]
                ks = ends[z]
                settle_live_batch(self._table, self.cache, self.arena, self.arena_b, seg_off[ks],
                                  seg_n[z], np.zeros(len(z), dtype=np.int64),
                                  np.full(len(z), -np.inf), stars=slots[ks], device=self.device,
                                  margin=LIVE_MARGIN)
                g2, p2, m2 = reject_batch(self.arena, seg_off[ks], seg_n[z],
                                          [states[j] for j in z], self.n_requested,
                                          device=self.device)
                for t, j in enumerate(z):
                    good[j], n_pass[j], llmax[j] = g2[t], p2[t], m2[t]
                self.zero_uniform_events += len(z)
        fl = np.flatnonzero(sh)
        if len(fl):
            ks = ends[fl]
            g3, p3, m3, _ = settle_reject_batch(self._table, self.cache, self.arena, self.arena_b,
                                                seg_off[ks], seg_n[fl], [states[j] for j in fl],
                                                self.n_requested, stars=slots[ks],
                                                device=self.device)
            for t, j in enumerate(fl):
                good[j], n_pass[j], llmax[j] = g3[t], p3[t], m3[t]
        return good, n_pass, llmax

    def step(self):
        """Admit, evaluate the next rows of every active request (a whole
        round, or a share of the step budget), and run the rejection step of
        the requests whose round is complete; retire the finished.  The host
        work of the requests retired in the previous step (their
        linear-parameter draws) runs while this step's likelihood launch is
        on the device.  Returns the requests whose results were completed in
        this call."""
        self._cur = {}
        tick = [time.perf_counter()]
        host = {}

        def lap(name):
            t = time.perf_counter()
            host[name] = host.get(name, 0.0) + t - tick[0]
            tick[0] = t

        self._mark()
        self._admit()
        lap("admit")
        self._mark("admit")
        if not self.active:
            done = self._build_results()
            self._mark("finish")
            return done
        while True:
            plan = []
            for r in self.active:
                rows, start, n_now = r.it.next_batch()
                # a round's region is sized when the round starts
                if r.round_done or self._ensure_region(r, r.it.n_evals + n_now):
                    plan.append((r, start, n_now))
            self._waiting = len(self.active) - len(plan)
            if plan:
                break
            # Every active request waits to grow and the arena is full.  With
            # others beside it, restart the most recently admitted one; alone,
            # the arena cannot hold the request at all.
            if len(self.active) == 1:
                r = self.active.pop()
                self._fail(r, MemoryError("lnL arena too small for this request's next batch"))
                self._waiting = 0
                return [r]
            self._preempt(max(self.active, key=lambda q: q.t_admit))
        lap("region")
        slots = np.array([r.slot for r, _, _ in plan], dtype=np.int64)
        n_round = np.array([n for _, _, n in plan], dtype=np.int64)
        done0 = np.array([r.round_done for r, _, _ in plan], dtype=np.int64)
        sizes = self._chunks(plan)
        begins = np.array([b for _, b, _ in plan], dtype=np.int64) + done0
        before = np.array([r.it.n_evals for r, _, _ in plan], dtype=np.int64)
        seg_off = np.array([r.region[0] for r, _, _ in plan], dtype=np.int64)
        # rows per work item: the launch shape's tile, widened so that a step
        # never queues more than ~64K items (host ordering and upload stay cheap)
        tile = max(self.shape.tile, 256 * -(-int(sizes.sum()) // (256 * 65536)))
        items, off, total = work_items(sizes, tile, row_begin=begins, stars=slots)
        # write each request's rows straight into its arena region
        counts = (sizes + tile - 1) // tile
        items["out_offset"] += np.repeat(seg_off + before + done0 - off, counts)
        items = items[heaviest_first(items, self._table.M)]
        lap("plan")
        d_items = self._uploader(items)
        lap("upload")
        self._mark("plan")
        if self.profile:
            finished = self._build_results()
            self._mark("finish")
        thr = self._prune_thresholds(plan)
        probed = []
        if thr is None:
            lnl_batch(self._table, self.cache, d_items, len(items), 0, self.precision, self.shape,
                      out=self.arena, workspace=self._ws, bound=self.arena_b)
        else:
            self._prune_dev.copy_(_torch().from_numpy(thr.reshape(-1)))
            lnl_prune_batch(self._table, self._probes, self._prune_dev, self.cache, d_items,
                            len(items), self.precision, self.shape, out=self.arena,
                            bound=self.arena_b, workspace=self._ws, n_full=self._n_full)
            probed = [(r, r.slot, int(sizes[k])) for k, (r, _, _) in enumerate(plan)
                      if np.isfinite(thr[r.slot, 0])]
        lap("lnl_launch")
        if not self.profile:
            finished = self._build_results()      # overlaps the launch above
            lap("finish")
            self._prepare_ahead()                 # so does this
            lap("prepare")
        M = self._table.M[slots]
        self.sample_epochs += int(np.sum(sizes * M))
        self._mark("lnl")
        if self.precision == "mixed":
            # FP64 for the new rows within reach of the maximum: with them,
            # every row decides like FP64 in this and all later rounds
            # (requests in shell mode settle at their rejection step instead)
            floors = np.array([r.it.floor for r, _, _ in plan], dtype=np.float64)
            live = np.array([not r.shell for r, _, _ in plan])
            cnt = np.zeros(len(plan), dtype=np.int64)
            if live.any():
                c = settle_live_batch(self._table, self.cache, self.arena, self.arena_b,
                                      seg_off[live], (before + done0 + sizes)[live],
                                      (before + done0)[live], floors[live], stars=slots[live],
                                      device=self.device, margin=LIVE_MARGIN_NONZERO_U)
                cnt[live] = c.cpu().numpy()    # the settling call has synchronised the stream
            lap("settle")
            for k, (r, _, _) in enumerate(plan):
                r.round_settled += int(cnt[k])
            if self.profile:
                self._cur["settled"] = int(cnt.sum())
        self._mark("polish")
        for k, (r, _, _) in enumerate(plan):
            r.round_done += int(sizes[k])
            r.rows_evaluated += int(sizes[k])
            r.sample_epochs += int(sizes[k] * M[k])
        ends = [k for k, (r, _, n) in enumerate(plan) if r.round_done == n]
        retired = []
        if ends:
            seg_n = before[ends] + n_round[ends]
            states = [pcg64_state(plan[k][0].rng) for k in ends]
            mixed = self.precision == "mixed"
            good, n_pass, llmax = self._reject_ends(plan, ends, seg_off, slots, seg_n, states)
            lap("reject")
            self._mark("reject")
            for j, k in enumerate(ends):
                r = plan[k][0]
                r.round_done = 0
                if (mixed and not (r.exact or r.shell) and self.exact_switch is not None and
                        r.round_settled > self.exact_switch * n_round[k]):
                    if self.flat_mode == "exact":
                        self._table.set_exact(r.slot)
                        r.exact = True
                    else:
                        r.shell = True
                    self.flat_switches += 1
                r.round_settled = 0
                try:
                    if self.in_memory and not np.isfinite(llmax[j]):
                        raise RuntimeError("There are NaN or Inf likelihood values in iteration "
                                           f"step {r.it.iteration}!")
                    r.it.floor = float(llmax[j])
                    r.it.update(seg_n[j], good[j], n_pass[j])
                except Exception as exc:
                    r.error = exc
                    r.it.done = True
                if r.it.done:
                    retired.append(r)
        else:
            self._mark("reject")
        ok = [r for r in retired if r.error is None]
        if ok:
            posts = posterior_inputs_table(self._table, [r.slot for r in ok], self.cache,
                                           [r.it.result()[0] for r in ok])
            for r, post in zip(ok, posts):
                r.post = post
        for r in retired:
            self.active.remove(r)
            if r.error is None:
                self._to_finish.append(r)
                self._release(r)
            else:
                self._fail(r, r.error)
        if probed:
            self._probe_feedback(probed)
        self.steps += 1
        lap("update")
        self._mark("update")
        self.last_step = dict(requests=len(plan), rows=int(sizes.sum()),
                              sample_epochs=int(np.sum(sizes * M)), rounds_done=len(ends),
                              finished=len(finished), retired=len(retired),
                              host_s={k: round(v, 4) for k, v in host.items()})
        if self.profile:
            self._cur.update(self.last_step)
            self.step_profile.append(self._cur)
        return finished


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
