"""TheJoker: the drop-in sampler API (thejoker/thejoker.py:23-370).

Same methods, signatures and defaults as the reference --
``marginal_ln_likelihood``, ``rejection_sample``,
``iterative_rejection_sample`` -- with the per-sample likelihood, the
rejection step and the linear-parameter posterior running in libjoker's gfx950
kernels on device-resident prior samples.  The numpy random stream is
consumed exactly as the reference consumes it:

* rejection uniforms: ``rng.uniform(size=n)`` (likelihood_helpers.py:107,
  multiproc_helpers.py:256/373) -- generated on the device from the
  generator's PCG64 state, after which the host generator is advanced by n;
* ``randomize_prior_order``: ``rng.choice`` on the host first
  (multiproc_helpers.py:247, 352);
* linear parameters: ``rng.multivariate_normal`` per accepted row on the host
  (fast_likelihood.pyx:529-530), from ``rng`` itself on the in-memory path
  (likelihood_helpers.py:69-88) or from per-batch children of
  ``rng.bit_generator._seed_seq.spawn`` on the default path
  (multiproc_helpers.py:49-54, 150-182).

``in_memory`` keeps its meaning for the iteration rule: the in-memory path
grows batches with safety_factor 1 (likelihood_helpers.py:145), the default
path with 4 (multiproc_helpers.py:334).  The HDF5 temp-file and process-pool
transport of the default path is replaced by the device-resident cache.
"""
import logging
import os

import numpy as np

from . import units as u
from .device import (DeviceStar, PriorCache, linear_internal_units, lnl, lnl_reject,
                     pcg64_advance_state, pcg64_state, reject, to_host)
from .likelihood_helpers import batch_tasks, next_batch_size, validate_prepare_data
from .prior import JokerPrior
from .samples import JokerSamples

__all__ = ["TheJoker"]

logger = logging.getLogger("thejoker")


class TheJoker:
    """
    Parameters
    ----------
    prior : JokerPrior
    pool : object with ``map`` and ``close`` (optional)
        Accepted for API compatibility; computation runs on the GPU.  Its
        ``size`` sets the default number of batches, which (as in the
        reference) decides how many child generators draw the linear
        parameters on the default path.
    rng : numpy.random.Generator (optional)
    tempfile_path : str (optional)
    jitter_mode : {"physical", "reference"}
        "physical": C = diag(sigma^2 + s^2) (py_likelihood.py:17-29).
        "reference": the jitter is ignored, as in fast_likelihood.pyx:458.
        They agree whenever s = 0 (the default prior).
    precision : {"f64", "mixed"}
        "f64": the plain FP64 kernel.  "mixed": FP32 Kepler / design column
        with FP64 phase and accumulation, plus a per-row bound on the
        difference to FP64; the rows whose rejection decision (or claim to the
        maximum) the bound leaves open are re-evaluated in FP64 before the
        rejection step, so the accepted set equals the FP64 one.  Reported
        ln_likelihood values of accepted samples are FP64, and every value of
        an lnL array returned (marginal_ln_likelihood, return_all_logprobs)
        is within 1e-3 of FP64: rows whose bound exceeds that are
        re-evaluated in FP64 first.
    device : int
        HIP device ordinal.
    prune : bool
        Bound every row by a probe on a subset of the epochs
        (device.probe_star) and skip the full evaluation of the rows no
        uniform can accept.  rejection_sample: the probe bounds every row,
        then the rows within reach of the maximum are evaluated in FP64
        (batch.prune_reject_batch); iterative_rejection_sample: from the
        second round on, rows whose bound lies more than 750 below the
        maximum of the earlier rounds keep the bound as their value.  The
        samples returned do not change.  Stars with fewer than 12 epochs
        (20 in rejection_sample), full covariances and randomize_prior_order
        are not probed.  Off by default, so the default path evaluates every
        row as the reference does.
    """

    PRUNE_MIN_EPOCHS = 12
    # rejection_sample's single pass must find the maximum and settle every
    # row near it: a cost nearly independent of the epoch count (1.1-1.7 ms for
    # 2^24 rows on one MI355X from 16 to 200 epochs) against about 55 us per
    # epoch for the full evaluation of the same rows, so the probe pays from
    # about 20 epochs on (measured: 0.82-0.89x the plain pass at 16 epochs,
    # 1.3x at 24, 2.2x at 50, 5.8-6.0x at 200)
    PRUNE_MIN_EPOCHS_SINGLE = 20

    def __init__(self, prior, pool=None, rng=None, tempfile_path=None, jitter_mode="physical",
                 precision="f64", device=0, random_state=None, prune=False):
        if random_state is not None and rng is None:
            rng = random_state
        if pool is not None and (not hasattr(pool, "map") or not hasattr(pool, "close")):
            raise TypeError("Input pool object must have .map() and .close() methods. We "
                            "recommend using `schwimmbad` pools.")
        self.pool = pool
        if rng is None:
            rng = np.random.default_rng()
        elif not isinstance(rng, np.random.Generator):
            raise TypeError("The input random number generator must be a numpy.random.Generator "
                            "instance.")
        self.rng = rng
        if not isinstance(prior, JokerPrior):
            raise TypeError("The input prior must be a JokerPrior instance.")
        self.prior = prior
        if tempfile_path is None:
            self._tempfile_path = os.path.expanduser("~/.thejoker/")
        else:
            self._tempfile_path = os.path.abspath(os.path.expanduser(tempfile_path))
        if jitter_mode not in ("physical", "reference"):
            raise ValueError("jitter_mode must be 'physical' or 'reference'")
        if precision not in ("f64", "mixed"):
            raise ValueError("precision must be 'f64' or 'mixed'")
        self.jitter_mode = jitter_mode
        self.precision = precision
        self.device = device
        self.prune = bool(prune)

    @property
    def tempfile_path(self):
        os.makedirs(self._tempfile_path, exist_ok=True)
        return self._tempfile_path

    # ------------------------------------------------------------------
    def _make_joker_helper(self, data):
        """thejoker.py:87-91: validate data, build the star on the device."""
        all_data, ids, trend_M = validate_prepare_data(data, self.prior.poly_trend,
                                                       self.prior.n_offsets)
        if all_data._has_cov and self.precision != "f64":
            raise NotImplementedError("data with a full covariance need precision='f64' (the "
                                      "mixed kernels and their error bound are diagonal)")
        star = DeviceStar.from_data_prior(all_data, self.prior, trend_M,
                                          jitter_mode=self.jitter_mode, device=self.device)
        star.data = all_data
        star.internal_units = linear_internal_units(self.prior, all_data.rv.unit)
        return star

    def _cache(self, star, prior_samples, return_logprobs=False):
        """Prior samples (JokerSamples, PriorCache, or an HDF5 or .npz
        filename) as a device-resident PriorCache.  An HDF5 file streams into
        the device columns without a host copy of the whole table."""
        if isinstance(prior_samples, PriorCache):
            return prior_samples
        if isinstance(prior_samples, str):
            if os.path.splitext(prior_samples)[1] in (".hdf5", ".h5"):
                return PriorCache.from_hdf5(prior_samples, rv_unit=star.data.rv.unit,
                                            device=self.device)
            prior_samples = JokerSamples.read(prior_samples)
        if isinstance(prior_samples, JokerSamples):
            return PriorCache.from_samples(prior_samples, rv_unit=star.data.rv.unit,
                                           device=self.device)
        if isinstance(prior_samples, np.ndarray):
            return PriorCache.from_packed(prior_samples, device=self.device)
        raise TypeError("prior_samples must either be a string filename specifying a cache file "
                        "containing prior samples, or must be a JokerSamples instance, not: "
                        f"{type(prior_samples)}")

    def _lnl_rows(self, star, cache, rows=None, start=0, n=None, bound=None):
        """Device lnL at the current precision (mixed: with each row's error
        bound written to `bound`)."""
        return lnl(star, cache, rows=rows, start=start, n=n, precision=self.precision,
                   bound=bound)

    @staticmethod
    def _settle(star, cache, ll, bound, rng_state, rows=None, device=0):
        """Mixed precision: re-evaluate in FP64 the rows of ll (cache rows
        `rows`, or 0 ..) whose rejection decision under rng_state, or whose
        claim to the maximum, the error bound leaves open (jk_settle_batch on
        a one-star table)."""
        from .batch import StarBatch, settle_batch
        table = getattr(star, "_settle_table", None)
        if table is None:
            table = star._settle_table = StarBatch([star], device=device)
        settle_batch(table, cache, ll, bound, [0], [int(ll.shape[0])], [rng_state], rows=rows,
                     device=device)

    @staticmethod
    def _settle_reject(star, cache, ll, bound, rng_state, max_keep, rows=None, device=0):
        """Mixed precision, one rejection step: _settle then the rejection step
        (jk_settle_reject_batch on a one-star table); the accepted indices."""
        from .batch import StarBatch, settle_reject_batch
        table = getattr(star, "_settle_table", None)
        if table is None:
            table = star._settle_table = StarBatch([star], device=device)
        return settle_reject_batch(table, cache, ll, bound, [0], [int(ll.shape[0])], [rng_state],
                                   max_keep, rows=rows, device=device)[0][0]

    @staticmethod
    def _settle_live(star, cache, ll, bound, first, floor, rows=None, device=0):
        """Mixed precision, iterative loop: re-evaluate in FP64 the new rows
        ll[first:] that lie within a 750 margin of the maximum, so that the
        rejection step over ll decides like FP64 for every later uniform
        (jk_settle_live_batch on a one-star table; floor: the maximum of
        ll[:first]).  Returns the number of rows re-evaluated."""
        from .batch import StarBatch, settle_live_batch
        table = getattr(star, "_settle_table", None)
        if table is None:
            table = star._settle_table = StarBatch([star], device=device)
        cnt = settle_live_batch(table, cache, ll, bound, [0], [int(ll.shape[0])], [int(first)],
                                [float(floor)], rows=rows, device=device)
        return int(cnt.cpu()[0])

    @staticmethod
    def _settle_tol(star, cache, ll, bound, rows=None, device=0):
        """Mixed precision, an lnL array handed back to the caller: every row
        whose error bound exceeds 1e-3 re-evaluated in FP64, so that every
        value is within 1e-3 of FP64 (jk_settle_tol_batch on a one-star
        table)."""
        from .batch import StarBatch, settle_tol_batch
        table = getattr(star, "_settle_table", None)
        if table is None:
            table = star._settle_table = StarBatch([star], device=device)
        return settle_tol_batch(table, cache, ll, bound, [0], [int(ll.shape[0])], rows=rows,
                                device=device)

    @staticmethod
    def _fp64_at(star, cache, cache_rows):
        """FP64 lnL of the given cache rows (the values reported for accepted
        samples, whatever the precision of the search)."""
        import torch
        if len(cache_rows) == 0:
            return np.zeros(0)
        r = torch.as_tensor(np.asarray(cache_rows, dtype=np.int64), device=cache.P.device)
        return to_host(lnl(star, cache, rows=r, precision="f64"))

    def _prunable(self, star):
        from .batch import PRUNE_MIN_SIGNAL
        from .device import trend_signal
        return (self.prune and star.cov is None and star.n >= self.PRUNE_MIN_EPOCHS and
                trend_signal(star.arrays()) >= PRUNE_MIN_SIGNAL)

    def _pruned_reject(self, star, cache, n, rng_state, max_keep):
        """rejection_sample's likelihood and rejection step over cache rows
        [0, n) with the rows bounded by a probe (batch.prune_reject_batch on a
        one-star table): (lnL buffer, its bounds, accepted indices)."""
        from .batch import StarBatch, prune_reject_batch, upload_items, work_items
        from .batch import LaunchShape
        from .device import probe_star
        table = getattr(star, "_settle_table", None)
        if table is None:
            table = star._settle_table = StarBatch([star], device=self.device)
        probes = getattr(star, "_probe_table", None)
        if probes is None:
            probe, add = probe_star(star.arrays(), self.jitter_mode, self.device)
            probes = star._probe_table = StarBatch([probe], device=self.device)
            star._probe_add = add
        shape = LaunchShape()
        # the work items of a one-star table depend only on n: planned and
        # uploaded once per sampler and row count
        plan = getattr(self, "_prune_items", None)
        if plan is None or plan[0] != (n, shape.tile):
            items, off, _ = work_items([n], shape.tile)
            plan = self._prune_items = ((n, shape.tile), upload_items(items, self.device),
                                        len(items), off)
        _, d_items, n_items, off = plan
        ll, bnd, res, _ = prune_reject_batch(table, probes, np.array([star._probe_add]),
                                             np.array([True]), cache, d_items, n_items, off,
                                             np.array([n]), [rng_state], max_keep,
                                             self.precision, shape, self.device)
        return ll, bnd, res[0][0]

    def _pruner(self, star, it):
        """The probe launcher of iterative_rejection_sample, or None when the
        star is not probed."""
        if not self._prunable(star) or it._d_all_idx is not None:
            return None
        return _probe_pruner(star, self.precision, self.jitter_mode, self.device)

    def _n_batches(self, n_batches):
        if n_batches is not None:
            return int(n_batches)
        size = getattr(self.pool, "size", 1) if self.pool is not None else 1
        return max(1, int(size or 0))

    def _full_samples(self, star, cache, rows, rng, n_linear_samples, spawn_batches=None):
        """Linear parameters of accepted rows (fast_likelihood.pyx:471-545) and
        JokerSamples.unpack (samples.py:447-478)."""
        post = self._posterior_inputs(star, cache, rows)
        return self._samples_from_posterior(star, post, rng, n_linear_samples, spawn_batches)

    @staticmethod
    def _posterior_inputs(star, cache, rows):
        """The device part of _full_samples: (packed rows, a, cov) of the
        accepted cache rows."""
        rows = np.asarray(rows, dtype=np.int64)
        a, cov, _ = _linear_posterior(star, cache, rows)
        packed = cache.host_rows(rows) if len(rows) else np.zeros((0, 5))
        return packed, a, cov

    def _samples_from_posterior(self, star, post, rng, n_linear_samples, spawn_batches=None):
        """The host part of _full_samples: linear-parameter draws in row order
        from `rng` (or its spawned children) and the JokerSamples."""
        packed, a, cov = post
        n = len(packed)
        n_pars = 5 + star.L
        raw = np.zeros((n, n_linear_samples, n_pars))
        if spawn_batches is None:
            gens = [(np.arange(n), rng)]
        else:
            tasks = batch_tasks(n, n_batches=spawn_batches, arr=np.arange(n))
            from numpy.random import PCG64, Generator
            sg = rng.bit_generator._seed_seq.spawn(len(tasks))
            gens = [(tasks[i][0], Generator(PCG64(sg[i]))) for i in range(len(tasks))]
        for sel, g in gens:
            sel = np.asarray(sel, dtype=np.int64)
            if len(sel):
                raw[sel, :, :5] = packed[sel][:, None, :]
                raw[sel, :, 5:] = mvn_draws(g, a[sel], cov[sel], n_linear_samples)
        raw = raw.reshape(n * n_linear_samples, n_pars)
        return JokerSamples.unpack(raw, star.internal_units, t_ref=star.data.t_ref,
                                   poly_trend=self.prior.poly_trend,
                                   n_offsets=self.prior.n_offsets)

    # ------------------------------------------------------------------
    def marginal_ln_likelihood(self, data, prior_samples, n_batches=None, in_memory=False):
        """thejoker.py:93-138: lnL at every prior sample (np.ndarray)."""
        star = self._make_joker_helper(data)
        cache = self._cache(star, prior_samples)
        if self.precision == "f64":
            return to_host(self._lnl_rows(star, cache)).copy()
        import torch
        bound = torch.empty(len(cache), dtype=torch.float32, device=cache.P.device)
        ll = self._lnl_rows(star, cache, bound=bound)
        self._settle_tol(star, cache, ll, bound, device=self.device)
        return to_host(ll).copy()

    def rejection_sample(self, data, prior_samples, n_prior_samples=None,
                         max_posterior_samples=None, n_linear_samples=1, return_logprobs=False,
                         return_all_logprobs=False, n_batches=None, randomize_prior_order=False,
                         in_memory=False):
        """thejoker.py:140-257."""
        import torch

        star = self._make_joker_helper(data)
        if isinstance(prior_samples, (int, np.integer)) and not isinstance(prior_samples, bool):
            prior_samples = self.prior.sample(size=int(prior_samples),
                                              return_logprobs=return_logprobs)
        cache = self._cache(star, prior_samples)
        n_total = len(cache)
        rng = self.rng

        if in_memory:
            n_use = n_total
            rows = None
            ln_prior = cache.ln_prior if return_logprobs else None
        else:
            if return_logprobs and cache.ln_prior is None:
                raise RuntimeError("return_logprobs=True but ln_prior values not found in prior "
                                   "cache: make sure you generate prior samples with prior.sample "
                                   "(..., return_logprobs=True) before saving the prior samples.")
            if n_prior_samples is None:
                n_use = n_total
            elif n_prior_samples > n_total:
                raise ValueError("Number of prior samples to use is greater than the number of "
                                 "prior samples passed, or cached to a filename specified. "
                                 f"n_prior_samples={n_prior_samples} vs. "
                                 f"n_total_samples={n_total}")
            else:
                n_use = int(n_prior_samples)
            rows = None
            if randomize_prior_order:
                rows = rng.choice(n_total, size=n_use, replace=False)
            ln_prior = cache.ln_prior
        if max_posterior_samples is None:
            max_posterior_samples = n_use

        d_rows = None
        if rows is not None:
            d_rows = torch.as_tensor(np.asarray(rows, dtype=np.int64), device=cache.P.device)
        bound = None
        if self._prunable(star) and star.n >= self.PRUNE_MIN_EPOCHS_SINGLE and rows is None:
            lls, bound, good = self._pruned_reject(star, cache, n_use, pcg64_state(rng),
                                                   max_posterior_samples)
        elif self.precision == "f64":
            # likelihood and rejection step in one call
            lls, good, _, _ = lnl_reject(star, cache, pcg64_state(rng), max_posterior_samples,
                                         rows=d_rows, start=0, n=n_use)
        else:
            bound = torch.empty(n_use, dtype=torch.float32, device=cache.P.device)
            lls = self._lnl_rows(star, cache, rows=d_rows, start=0, n=n_use, bound=bound)
            # the rows the bound leaves open re-evaluated in FP64, and the
            # rejection step, in one launch sequence
            good = self._settle_reject(star, cache, lls, bound, pcg64_state(rng),
                                       max_posterior_samples, rows=d_rows, device=self.device)
        pcg64_advance_state(rng, n_use)
        full_idx = rows[good] if rows is not None else good

        samples = self._full_samples(star, cache, full_idx, rng, n_linear_samples,
                                     spawn_batches=None if in_memory else self._n_batches(n_batches))
        ll_good = None
        if (in_memory and ln_prior is not None and ln_prior is not False) or \
                (not in_memory and return_logprobs):
            ll_good = self._fp64_at(star, cache, full_idx)
        if in_memory:
            if ln_prior is not None and ln_prior is not False:
                samples["ln_prior"] = ln_prior[good]
                samples["ln_likelihood"] = ll_good
        elif return_logprobs:
            samples["ln_likelihood"] = ll_good
            samples["ln_prior"] = cache.ln_prior[full_idx]
        if return_all_logprobs:
            if bound is not None:
                # mixed values to 1e-3, and the probe's upper bounds evaluated
                self._settle_tol(star, cache, lls, bound, rows=d_rows, device=self.device)
            return samples, to_host(lls).copy()
        return samples

    def iterative_rejection_sample(self, data, prior_samples, n_requested_samples,
                                   max_prior_samples=None, n_linear_samples=1,
                                   return_logprobs=False, n_batches=None,
                                   randomize_prior_order=False, init_batch_size=None,
                                   growth_factor=128, in_memory=False):
        """thejoker.py:259-370, with the loop of likelihood_helpers.py:130-229
        (in_memory=True) or multiproc_helpers.py:289-427 (default)."""
        import torch

        star = self._make_joker_helper(data)
        cache = self._cache(star, prior_samples)
        if not in_memory and return_logprobs and cache.ln_prior is None:
            raise RuntimeError("return_logprobs=True but ln_prior values not found in prior "
                               "cache: make sure you generate prior samples with prior.sample "
                               "(..., return_logprobs=True) before saving the prior samples.")
        it = IterativeRejection(len(cache), n_requested_samples, self.rng, in_memory=in_memory,
                                max_prior_samples=max_prior_samples,
                                randomize_prior_order=randomize_prior_order,
                                init_batch_size=init_batch_size, growth_factor=growth_factor,
                                device=cache.P.device)
        dev = cache.P.device
        all_ll = torch.empty(it.max_prior, dtype=torch.float64, device=dev)
        mixed = self.precision == "mixed"
        pruner = self._pruner(star, it)
        # per-row bounds: the mixed error bound, or (f64 with pruning) the
        # marks 0 (an FP64 value) and -1 (a probe's upper bound)
        all_b = (torch.empty(it.max_prior, dtype=torch.float32, device=dev)
                 if mixed or pruner is not None else None)
        # Mixed precision: after a round whose live settling re-evaluated more
        # than EXACT_SWITCH of its rows in FP64 (a flat likelihood), the later
        # rounds leave their rows unsettled and each rejection step settles
        # what its own uniforms need (settle_reject_batch over the buffer).
        from .batch import EXACT_SWITCH, settle_reject_batch
        shell = False
        self.last_shell_from = None     # rows evaluated before the switch, if any
        while not it.done:
            rows, start, n_now = it.next_batch()
            a, b = it.n_evals, it.n_evals + n_now
            part_b = all_b[a:b] if all_b is not None else None
            if pruner is not None and np.isfinite(it.floor):
                pruner(cache, start, n_now, it.floor, all_ll[a:b], part_b)
            else:
                lnl_b = part_b if mixed else None
                if rows is not None:
                    lnl(star, cache, rows=rows, precision=self.precision, out=all_ll[a:b],
                        bound=lnl_b)
                else:
                    lnl(star, cache, start=start, n=n_now, precision=self.precision,
                        out=all_ll[a:b], bound=lnl_b)
                if part_b is not None and not mixed:
                    part_b.zero_()
            if shell:
                it.check_finite(all_ll[:b])
                idx, n_pass, mx, _ = settle_reject_batch(
                    star._settle_table, cache, all_ll[:b], all_b[:b], [0], [b],
                    [pcg64_state(it.rng)], it.n_requested, rows=it._d_all_idx, device=self.device)
                it.floor = float(mx[0])
                it.update(b, idx[0], int(n_pass[0]))
                continue
            if mixed:
                # the new rows, for this and every later round's uniforms
                n_set = self._settle_live(star, cache, all_ll[:b], all_b[:b], a, it.floor,
                                          rows=it._d_all_idx, device=self.device)
                if n_set > EXACT_SWITCH * n_now:
                    shell = True
                    self.last_shell_from = b
            it.absorb(all_ll[:b], self.device)
        # the final lnL buffer (and bounds: 0 for FP64 values, -1 for a probe's
        # upper bound), for callers that inspect the run
        self.last_lnl = (all_ll[:it.n_evals], all_b[:it.n_evals] if all_b is not None else
                         torch.zeros(it.n_evals, dtype=torch.float32, device=dev))
        self.last_star = star
        full_idx, good = it.result()
        samples = self._full_samples(star, cache, full_idx, self.rng, n_linear_samples,
                                     spawn_batches=None if in_memory else self._n_batches(n_batches))
        if return_logprobs:
            ll_good = self._fp64_at(star, cache, full_idx)
            if in_memory:
                if cache.ln_prior is not None:
                    samples["ln_prior"] = cache.ln_prior[full_idx]
                    samples["ln_likelihood"] = ll_good
            else:
                samples["ln_likelihood"] = ll_good
                samples["ln_prior"] = cache.ln_prior[full_idx]
        return samples


def _probe_pruner(star, precision, jitter_mode, device):
    """f(cache, start, n, floor, out, bound) evaluating cache rows
    start .. start + n of `star` into out (and bound) with
    jk_lnl_prune_batch: rows whose probe bound lies more than 750 below
    `floor` keep that bound."""
    import torch

    from .batch import LIVE_MARGIN, LaunchShape, StarBatch, lnl_prune_batch, upload_items, \
        work_items
    from .device import probe_star

    probe, add = probe_star(star.arrays(), jitter_mode, device)
    # the tables hold the stars for as long as the launcher lives
    table, probes = StarBatch([star], device=device), StarBatch([probe], device=device)
    prune = torch.zeros(2, dtype=torch.float64, device=torch.device("cuda", device))
    shape = LaunchShape()

    def run(cache, start, n, floor, out, bound):
        items, _, _ = work_items([n], shape.tile, row_begin=[start], stars=[0])
        prune.copy_(torch.tensor([floor - LIVE_MARGIN, add], dtype=torch.float64))
        lnl_prune_batch(table, probes, prune, cache, upload_items(items, device), len(items),
                        precision, shape, out=out, bound=bound)

    return run


class IterativeRejection:
    """State of one iterative rejection run -- the loop of
    likelihood_helpers.py:130-229 (in_memory, safety_factor 1) and
    multiproc_helpers.py:289-427 (default, safety_factor 4, optional
    randomised prior order, max_prior_samples).  Each round evaluates the next
    batch of rows, then redraws uniforms for every lnL so far, so the rows and
    the random stream consumed depend only on this star's lnL values.

    Use: ``while not it.done: rows, start, n = it.next_batch(); <lnL of those
    rows appended to the star's lnL buffer>; it.absorb(ll_so_far)``."""

    maxiter = 128

    def __init__(self, n_total, n_requested, rng, in_memory=False, max_prior_samples=None,
                 randomize_prior_order=False, init_batch_size=None, growth_factor=128,
                 device=None):
        self.rng = rng
        self.n_requested = int(n_requested)
        self.in_memory = in_memory
        if in_memory:
            self.safety_factor = 1
            self.max_prior = int(n_total)
            self.all_idx = None
        else:
            self.safety_factor = 4
            self.max_prior = int(n_total) if max_prior_samples is None else int(max_prior_samples)
            self.all_idx = None
            if randomize_prior_order:
                self.all_idx = rng.choice(int(n_total), size=self.max_prior, replace=False)
        self.n_process = (growth_factor * self.n_requested if init_batch_size is None
                          else int(init_batch_size))
        if self.n_process > self.max_prior:
            raise ValueError("Prior sample library not big enough! For iterative sampling, you "
                             "have to have at least growth_factor * n_requested_samples = "
                             f"{growth_factor * self.n_requested} samples in the prior samples "
                             f"cache file. You have, or have limited to, {self.max_prior} "
                             "samples.")
        self._d_all_idx = None
        if self.all_idx is not None:
            import torch
            self._d_all_idx = torch.as_tensor(np.asarray(self.all_idx, dtype=np.int64),
                                              device=device)
        self.start_idx = 0
        self.n_evals = 0
        self.iteration = 0
        self.good = None
        self.n_good = 0
        self.done = False
        self.floor = -np.inf     # lnL maximum over the rows evaluated so far

    def next_batch(self):
        """(device rows or None, first cache row, count) of the next round."""
        n_now = min(self.n_process, self.max_prior - self.start_idx)
        if self._d_all_idx is not None:
            return self._d_all_idx[self.start_idx:self.start_idx + n_now], 0, n_now
        return None, self.start_idx, n_now

    def absorb(self, ll_so_far, device=0):
        """Rejection over every lnL evaluated so far (a device tensor of
        n_evals + the new rows), then the growth rule."""
        import torch

        n_evals = int(ll_so_far.shape[0])
        self.check_finite(ll_so_far)
        good, n_good, mx = reject(ll_so_far, pcg64_state(self.rng), self.n_requested,
                                  device=device)
        self.floor = mx
        self.update(n_evals, good, n_good)

    def check_finite(self, ll_so_far):
        """The in-memory loop's check (likelihood_helpers.py:180-183)."""
        import torch

        if self.in_memory and not bool(torch.isfinite(ll_so_far).all()):
            raise RuntimeError(f"There are NaN or Inf likelihood values in iteration step "
                               f"{self.iteration}!")

    def update(self, n_evals, good, n_good):
        """Bookkeeping after the rejection step of a round whose lnL buffer now
        holds n_evals values and whose uniforms were drawn from self.rng
        (advanced here by n_evals)."""
        i = self.iteration
        self.n_evals = int(n_evals)
        pcg64_advance_state(self.rng, self.n_evals)
        logger.log(1, f"iteration {i}, computed {self.n_evals} likelihoods")
        self.good, self.n_good = good, int(n_good)
        self.iteration += 1
        if n_good == 0:
            raise RuntimeError("Failed to find any good samples!")
        logger.log(1, f"{n_good} good samples after rejection sampling")
        if n_good >= self.n_requested:
            self.done = True
            return
        self.start_idx += self.n_process
        self.n_process = next_batch_size(self.n_requested, n_good, self.n_evals, self.start_idx,
                                         self.max_prior, self.safety_factor)
        if self.n_process <= 0:
            self.done = True
            return
        if self.iteration >= self.maxiter:
            raise RuntimeError("Hit maximum number of iterations!")

    def result(self):
        """(cache rows of the kept samples, their positions in the lnL buffer)."""
        good = self.good[:self.n_requested]
        full_idx = self.all_idx[good] if self.all_idx is not None else good
        return full_idx, good


def _linear_posterior(star, cache, rows):
    from .device import linear_posterior
    return linear_posterior(star, cache, rows)


def mvn_draws(rng, means, covs, size):
    """rng.multivariate_normal(means[j], covs[j], size=size) for every j in
    order, as one stacked computation (fast_likelihood.pyx:529-530 draws row
    by row).  Generator.multivariate_normal (method "svd") factors
    cov = u diag(s) vh, draws standard normals of shape (size, L) and returns
    mean + z @ (u sqrt(s))^T; the stacked SVD and product run the same LAPACK
    and BLAS kernels per matrix, and the normals are drawn in the same order,
    so the values and the generator's final state are the same."""
    means = np.asarray(means, dtype=np.float64)
    covs = np.asarray(covs, dtype=np.float64)
    n, L = means.shape
    u, s, _ = np.linalg.svd(covs)
    factor = u * np.sqrt(s)[:, None, :]
    z = rng.standard_normal((n, size, L))
    return means[:, None, :] + np.matmul(z, np.swapaxes(factor, 1, 2))
