"""Benchmark of The Joker's likelihood + rejection path on one MI355X.

One JSON line.  `value` is the single-star plain-FP64 workload: one star with
M = 16 epochs over 1000 d, 2^24 prior rows (P log-uniform on [2, 2048] d,
e ~ Beta(0.867, 3.03), omega, M0 ~ U[0, 2 pi), ln s ~ N(-3, 1) km/s),
K ~ N(0, 30), v0 ~ N(0, 100), physical jitter mode.  One step = the FP64
marginal-likelihood kernel over all 2^24 device-resident rows, the on-device
rejection step (max, numpy-PCG64 uniforms, ordered compaction of at most 256
indices) and the read-back of the accepted indices; value = prior-samples x
epochs / s of the step.  Steps are pipelined: the next step's stream depends
only on the row count, so it is queued before the host collects the previous
step's indices.

Further legs (each an object of the same line; --workload picks one):
  mixed     -- the mixed-precision kernel on the same rows, its error
               against FP64 (max |dlnL| all rows / live set);
  pruned    -- TheJoker(prune=True)'s rejection pass (probe bound, FP64
               rows near the maximum, settled rejection) on the same rows,
               FP64 and mixed, and against the plain pass for stars of
               --prune-epochs epochs on the same rows; accepted rows checked
               against the plain path's;
  batched   -- 1024 stars, M ~ U{3..64}, baselines U[200, 3000] d, half
               binaries, poly_trend = 2, 2^20 shared rows: one autotuned
               persistent launch, mixed vs FP64 rates, errors, acceptance
               flips after FP64 re-evaluation of the live sets;
  serving   -- Poisson arrivals at 200 stars/s, M log-uniform on [3, 200],
               2^26-row device cache, iterative_rejection_sample to 256
               samples per star under continuous batching: p50 / p99
               latency, throughput, flips against FP64 runs of the same
               stars alone;
  stress    -- one star, M = 50 over 3650 d, P log-uniform on [1, 16384] d,
               2^29 rows generated and kept in HBM, evaluated in 2^26-row
               chunks in FP64 and mixed: rates, HBM GB/s, errors, flips.
roofline: the FP64 likelihood kernel's algorithmic FLOP rate (HIP events
around its launches on the launch stream) against the FP64 vector peak;
traffic from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary.
cpu_baseline: the oracle's restatement of the reference algorithm (LU +
Woodbury, fast_likelihood.pyx:255-469) on the host cores, bounded sample.
"""
import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = ("prior-samples×epochs/s and p50/p99 star latency, mixed vs FP64 at "
          "|ΔlnL|≤1e-3")
# Algorithmic FLOPs (weights: add/mul 1, FMA 2, div/sqrt 4, sin/cos/exp/log 8;
# Kepler = starter + 4 Newton steps) per sample-epoch with sampled jitter, and
# per sample for L = 2 / 3.
F_EPOCH_JITTER = 176
F_EPOCH_L3 = 184
F_SAMPLE = {2: 90, 3: 128}
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
BYTES_PER_SAMPLE_F64 = 56   # 5 x f64 read + lnL write + lnL read by the rejection step
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_k_lnl_f64_single_star.json")
VALU_FILE = os.path.join(ROOT, "profiles", "valu_counters.json")
CLOCK_WARM_S = 0.3   # untimed likelihood launches before the single-star warmup steps
PEAK_CLOCK_HZ = 2.4e9      # the clock MI355X's published peaks assume
N_SIMD = 1024              # 256 CUs x 4 SIMDs
WINDOW = 40.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(star, rows, mu, lam, n_sample, gpu_ll=None):
    """Oracle (reference algorithm) on the host cores over the first n_sample
    rows; sample-epochs/s.  gpu_ll: the GPU's FP64 lnL of the same rows (host
    array), compared with the oracle's values (relative error with |lnL|
    floored at 1, rows where the oracle's anomaly solve hit its iteration cap
    excluded)."""
    from oracle import oracle

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    ost = oracle.OracleStar(star.t, star.rv, star.ivar(), star.trend(), mu, lam, star.t_ref,
                            jitter_mode=1)
    sub = rows[:n_sample]
    t0 = time.perf_counter()
    want, capped = oracle.batch_lnl(ost, sub, method="reference", n_threads=cores,
                                    return_capped=True)
    dt = time.perf_counter() - t0
    parity = None
    if gpu_ll is not None:
        import numpy as np
        ok = ~capped
        rel = np.abs(gpu_ll[:n_sample][ok] - want[ok]) / np.maximum(np.abs(want[ok]), 1.0)
        parity = {"max_rel_err_vs_gpu": float(rel.max()), "rows_compared": int(ok.sum()),
                  "rows_capped_excluded": int(capped.sum()),
                  "tolerance": 1e-10, "within_tolerance": bool(rel.max() <= 1e-10)}
    # beside it: one thread of the same, and the O(N L^2) Cholesky restatement
    # on all cores (what the algorithm, not the hardware, is worth)
    n1 = max(1, n_sample // 8)
    t0 = time.perf_counter()
    oracle.batch_lnl(ost, rows[:n1], method="reference", n_threads=1)
    dt1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.batch_lnl(ost, sub, method="chol", n_threads=cores)
    dtc = time.perf_counter() - t0
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")),
                         "")
    except OSError:
        pass
    return {"value": n_sample * star.M / dt, "unit": "sample-epochs/s", "cores": cores,
            "kind": "port",
            "sample": f"first {n_sample} of the same prior rows, reference LU+Woodbury algorithm "
                      f"restated in C (-O3), {cores} threads, {dt:.2f} s wall",
            "single_thread": {"value": n1 * star.M / dt1, "sample": f"first {n1} rows, 1 thread"},
            "cholesky_all_cores": {"value": n_sample * star.M / dtc,
                                   "sample": f"first {n_sample} rows, O(N L^2) form, "
                                             f"{cores} threads"},
            "host": {"cpus_visible": os.cpu_count(), "cpus_used": cores, "model": model},
            **(parity or {})}


def mixed_vs_f64(torch, np, D, dstar, cache, llf, llm, bnd, seed, max_keep=None):
    """Raw mixed-vs-FP64 errors (all rows, and rows within 40 of the maximum),
    the largest error as a fraction of the kernel's bound, then the product
    path -- settle the mixed buffer, reject -- against the FP64 rejection with
    the same uniforms: rows re-evaluated and accepted-set differences."""
    from thejoker_amd.thejoker import TheJoker
    n = int(llf.shape[0])
    d = (llf - llm).abs()
    live = llf >= llf.max() - WINDOW
    res = {"max_abs_dlnl_all": float(d.max()), "max_abs_dlnl_live": float(d[live].max()),
           "live_rows": int(live.sum()),
           "max_dlnl_over_bound": float((d / bnd.double()).max())}
    del d, live
    max_keep = n if max_keep is None else max_keep
    st = D.pcg64_state(np.random.default_rng(seed))
    TheJoker._settle(dstar, cache, llm, bnd, st)
    res["rows_settled_in_fp64"] = int((bnd == 0).sum())
    # what the product path hands on: the settled buffer's live-row error
    live = llf >= llf.max() - WINDOW
    res["max_abs_dlnl_live_settled"] = float((llf - llm).abs()[live].max())
    del live
    a_f, n_f, _ = D.reject(llf, st, max_keep)
    a_m, _, _ = D.reject(llm, st, max_keep)
    res["accepted"] = int(min(n_f, max_keep))
    res["accept_flips_vs_fp64"] = int(len(np.setxor1d(a_f, a_m)))
    return res


def fmt_mixed(r):
    return (f"max|dlnL| all {r['max_abs_dlnl_all']:.2e} live {r['max_abs_dlnl_live']:.2e} "
            f"(settled {r['max_abs_dlnl_live_settled']:.2e}), "
            f"max |dlnL|/bound {r['max_dlnl_over_bound']:.3f}, settled "
            f"{r['rows_settled_in_fp64']} rows, flips {r['accept_flips_vs_fp64']}")


def valu_from_counters(key, kernel_ms):
    """The committed rocprofv3 VALU counters of one likelihood kernel
    (profiles/valu_counters.json, same workload), with the rates taken over
    this run's own kernel time: hardware FP64 / FP32 FLOP/s from the
    FLOPS_* counters, and the issue-cycle model's share of the SIMDs' cycles
    at the published peak clock."""
    if not os.path.exists(VALU_FILE):
        return None
    with open(VALU_FILE) as f:
        k = json.load(f)["kernels"].get(key)
    if k is None:
        return None
    t = kernel_ms * 1e-3
    return {"source": "profiles/valu_counters.json", "counters_kernel_ms": k["dispatch_s"] * 1e3,
            "valu_insts_per_launch": k["valu_insts"],
            "valu_insts_per_wave_epoch": k["valu_insts_per_wave_epoch"],
            "fp64_insts_per_wave_epoch": k["f64_insts_per_wave_epoch"],
            "fp32_insts_per_wave_epoch": k["f32_insts_per_wave_epoch"],
            "hw_fp64_tflops": k["hw_fp64_flops"] / t / 1e12,
            "hw_fp64_frac_of_peak": k["hw_fp64_flops"] / t / 1e12 / PEAK_FP64_TFLOPS,
            "hw_fp32_tflops": k["hw_fp32_flops"] / t / 1e12,
            "hw_fp32_frac_of_peak": k["hw_fp32_flops"] / t / 1e12 / 157.3,
            "issue_frac_at_peak_clock": k["issue_cycles_model"] / (N_SIMD * PEAK_CLOCK_HZ * t),
            "issue_frac_at_profiled_clock": k["issue_frac"]}


def timed(fn, reps, torch):
    """Mean ms of fn() over reps launches, HIP events on the current stream."""
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        fn()
    ev1.record()
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / reps


# ---------------------------------------------------------------------------
def leg_single(args, torch, np, D, syn):
    n = 1 << args.log2_rows
    log(f"[bench] single: M=16 star, {n} prior rows")
    star = syn.single_star()
    cols = syn.prior_columns(n, seed=0, jitter=True)
    mu, lam = np.zeros(2), np.array([900.0, 1e4])
    dstar = D.DeviceStar(star.t, star.rv, star.ivar(), star.trend(), mu, lam, star.t_ref,
                         jitter_mode="physical", device=0)
    cache = D.PriorCache(cols["P"], cols["e"], cols["omega"], cols["M0"], cols["s"], device=0)
    rng = np.random.default_rng(42)
    ll = torch.empty(n, dtype=torch.float64, device="cuda")

    pending = []

    def step():
        # lnL of every row and the rejection step (max reduced in the lnL
        # launch), queued; the stream of the next step depends only on n, so
        # the previous step's accepted indices are collected while this one
        # runs
        pending.append(D.lnl_reject_async(dstar, cache, D.pcg64_state(rng), 256, out=ll))
        D.pcg64_advance_state(rng, n)
        if len(pending) > 1:
            _, idx, n_pass, _ = pending.pop(0).result()
            return idx, n_pass
        return None, None

    def drain():
        _, idx, n_pass, _ = pending.pop(0).result()
        return idx, n_pass

    # Untimed, before the warmup steps: likelihood launches for CLOCK_WARM_S
    # seconds, so that the timed steps run at the device's steady clock even
    # when the warmup is a handful of 1 ms steps.
    t_w = time.perf_counter()
    while time.perf_counter() - t_w < CLOCK_WARM_S:
        D.lnl(dstar, cache, precision="f64", out=ll)
        torch.cuda.synchronize()
    for _ in range(args.warmup):
        step()
    drain()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    idx, n_pass = drain()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_step = dt / args.steps * 1e3
    value = n * star.M / (dt / args.steps)
    log(f"[bench] single: {ms_step:.3f} ms/step, {value:.4e} sample-epochs/s, accepted {n_pass}")

    reps = max(10, min(args.steps, 50))
    k_ms = timed(lambda: D.lnl(dstar, cache, precision="f64", out=ll), reps, torch)
    flops = n * (star.M * F_EPOCH_JITTER + F_SAMPLE[2])
    achieved = flops / (k_ms * 1e-3) / 1e12
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
        if tr.get("rows") == n:
            traffic = tr["hbm_bytes_per_launch"]
    # The kernel is compute-bound, on the FP64 vector ALU (the per-row 2x2 / 3x3
    # systems are not matrix products, so no MFMA).  The schema's compute bound
    # is named "mfma"; MI355X's FP64 vector and FP64 dense-matrix peaks are the
    # same 78.6 TFLOP/s, which is the peak used here.
    roofline = {"bound": "mfma", "compute_unit": "FP64 vector ALU (VALU); no MFMA instructions",
                "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                "kernel": "k_lnl<2,true,false,false,true> (FP64, constant-offset column folded)", "kernel_ms": k_ms,
                "flops_per_launch": flops,
                "hbm_gbs_prior_stream": n * BYTES_PER_SAMPLE_F64 / (k_ms * 1e-3) / 1e9,
                "valu": valu_from_counters("single_f64", k_ms) if n == 1 << 24 else None}
    log(f"[bench] single: lnL kernel {k_ms:.3f} ms -> {achieved:.2f} TFLOP/s algorithmic "
        f"({100 * achieved / PEAK_FP64_TFLOPS:.1f}% of FP64 peak)")

    out = {"metric": METRIC, "value": value, "unit": "sample-epochs/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "clock_warm_s": CLOCK_WARM_S,
           "ms_per_step": ms_step,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": "single star, plain FP64 likelihood + rejection: M=16 epochs, "
                                  f"2^{args.log2_rows} prior rows, ln s ~ N(-3,1), physical "
                                  "jitter, K~N(0,30), v0~N(0,100), max_posterior_samples=256",
                      "epochs": star.M, "prior_rows": n, "linear_params": 2,
                      "parallelism": "single GPU"},
           "roofline": roofline}

    # mixed precision on the same rows
    ll32 = torch.empty(n, dtype=torch.float64, device="cuda")
    bnd = torch.empty(n, dtype=torch.float32, device="cuda")
    D.lnl(dstar, cache, precision="mixed", out=ll32)
    m_ms = timed(lambda: D.lnl(dstar, cache, precision="mixed", out=ll32), reps, torch)
    D.lnl(dstar, cache, precision="mixed", out=ll32, bound=bnd)
    D.lnl(dstar, cache, precision="f64", out=ll)
    out["mixed"] = {"kernel_ms": m_ms, "sample_epochs_per_s": n * star.M / (m_ms * 1e-3),
                    "speedup_vs_f64_kernel": k_ms / m_ms,
                    "valu": valu_from_counters("single_mixed", m_ms) if n == 1 << 24 else None,
                    **mixed_vs_f64(torch, np, D, dstar, cache, ll, ll32, bnd, seed=7)}
    log(f"[bench] mixed: {m_ms:.3f} ms ({k_ms / m_ms:.2f}x), {fmt_mixed(out['mixed'])}")

    # The same rows through the probe bound: TheJoker(prune=True).rejection_
    # sample's likelihood and rejection step (probe launch, FP64 evaluation of
    # the rows near the maximum, settled rejection), against the plain path's
    # accepted rows for one random stream.
    import thejoker_amd as tj
    from thejoker_amd import units as u
    pr = tj.JokerPrior.default(P_min=2 * u.day, P_max=2048 * u.day, sigma_K0=30 * u.km / u.s,
                               sigma_v=100 * u.km / u.s)
    state = D.pcg64_state(np.random.default_rng(99))
    _, ref_idx, ref_pass, _ = D.lnl_reject(dstar, cache, state, 256, out=ll)
    out["pruned"] = {"plain_step_ms": ms_step}
    for prec in ("f64", "mixed"):
        jk = tj.TheJoker(pr, precision=prec, prune=True)
        _, _, idx = jk._pruned_reject(dstar, cache, n, state, 256)
        torch.cuda.synchronize()
        t_p = time.perf_counter()
        for _ in range(reps):
            _, _, idx = jk._pruned_reject(dstar, cache, n, state, 256)
        torch.cuda.synchronize()
        p_ms = (time.perf_counter() - t_p) / reps * 1e3
        out["pruned"][prec] = {"ms": p_ms, "speedup_vs_plain_step": ms_step / p_ms,
                               "accepted_equal_plain": bool(np.array_equal(idx, ref_idx))}
        log(f"[bench] single: pruned {prec} {p_ms:.3f} ms ({ms_step / p_ms:.2f}x the plain step), "
            f"accepted rows equal the plain path's: {np.array_equal(idx, ref_idx)}")

    # Where the probe pays: the same rows against stars of more epochs, the
    # plain pass (lnL + rejection, synchronous) against the pruned one.
    out["pruned"]["by_epochs"] = []
    ll_m = torch.empty_like(ll)     # ll keeps the M=16 star's values for the oracle check
    for M in [int(x) for x in args.prune_epochs.split(",") if x.strip()]:
        s_m = syn.single_star(M=M)
        d_m = D.DeviceStar(s_m.t, s_m.rv, s_m.ivar(), s_m.trend(), mu, lam, s_m.t_ref,
                           jitter_mode="physical", device=0)
        _, ref_m, _, _ = D.lnl_reject(d_m, cache, state, 256, out=ll_m)
        rec = {"epochs": M, "plain_ms": timed(lambda: D.lnl_reject(d_m, cache, state, 256,
                                                                   out=ll_m), reps, torch)}
        for prec in ("f64", "mixed"):
            jk = tj.TheJoker(pr, precision=prec, prune=True)
            _, _, idx = jk._pruned_reject(d_m, cache, n, state, 256)
            p_ms = timed(lambda: jk._pruned_reject(d_m, cache, n, state, 256), reps, torch)
            rec[prec] = {"ms": p_ms, "speedup_vs_plain": rec["plain_ms"] / p_ms,
                         "accepted_equal_plain": bool(np.array_equal(idx, ref_m))}
        out["pruned"]["by_epochs"].append(rec)
        log(f"[bench] single: M={M} plain {rec['plain_ms']:.3f} ms, pruned f64 "
            f"{rec['f64']['ms']:.3f} ms ({rec['f64']['speedup_vs_plain']:.2f}x), mixed "
            f"{rec['mixed']['ms']:.3f} ms ({rec['mixed']['speedup_vs_plain']:.2f}x), accepted "
            f"equal {rec['f64']['accepted_equal_plain'] and rec['mixed']['accepted_equal_plain']}")
    del ll_m

    if not args.no_cpu:
        rows = np.ascontiguousarray(np.stack([cols[k] for k in ("P", "e", "omega", "M0", "s")],
                                             axis=1))
        n_cpu = min(n, 1 << 21)
        gpu_ll = ll[:n_cpu].cpu().numpy()
        out["cpu_baseline"] = cpu_baseline(star, rows, mu, lam, n_cpu, gpu_ll=gpu_ll)
        log(f"[bench] cpu {out['cpu_baseline']['value']:.3e} sample-epochs/s, GPU lnL vs "
            f"oracle max rel err {out['cpu_baseline']['max_rel_err_vs_gpu']:.2e}")
    del cache, ll, ll32
    torch.cuda.empty_cache()
    return out


# ---------------------------------------------------------------------------
def _l3_star(D, np, st):
    return D.DeviceStar(st.t, st.rv, st.ivar(), st.trend(poly_trend=2), np.zeros(3),
                        np.array([900.0, 1e4, 0.25]), st.t_ref, jitter_mode="physical", device=0)


def leg_batched(args, torch, np, D, syn):
    from thejoker_amd import batch as B

    n_stars, n = 1024, 1 << 20
    log(f"[bench] batched: {n_stars} stars x {n} rows")
    stars = syn.many_stars(n_stars, seed=3)
    dstars = [_l3_star(D, np, st) for st in stars]
    cols = syn.prior_columns(n, seed=0, jitter=True)
    cache = D.PriorCache(cols["P"], cols["e"], cols["omega"], cols["M0"], cols["s"], device=0)
    batch = B.StarBatch(dstars)
    M = batch.M
    work = int(M.sum()) * n
    t0 = time.perf_counter()
    shape, table = B.autotune(batch, cache, [n] * n_stars, "mixed", sample_stars=128)
    tune_s = time.perf_counter() - t0
    log(f"[bench] batched: autotuned {shape} in {tune_s:.1f} s")
    items, off, total = B.work_items([n] * n_stars, shape.tile)
    items = items[B.heaviest_first(items, M)]
    d_items = B.upload_items(items)
    ws = torch.zeros(64, dtype=torch.int32, device="cuda")
    llm = torch.empty(total, dtype=torch.float64, device="cuda")
    llf = torch.empty(total, dtype=torch.float64, device="cuda")
    bnd = torch.empty(total, dtype=torch.float32, device="cuda")
    run_m = lambda: B.lnl_batch(batch, cache, d_items, len(items), total, "mixed", shape,
                                out=llm, workspace=ws, bound=bnd)
    run_f = lambda: B.lnl_batch(batch, cache, d_items, len(items), total, "f64", shape,
                                out=llf, workspace=ws)
    run_m()
    run_f()
    m_ms = timed(run_m, 3, torch)
    f_ms = timed(run_f, 2, torch)
    # raw mixed errors (before settling) and how they compare with the bounds
    dmax_all, dmax_live, ratio = 0.0, 0.0, 0.0
    for s in range(n_stars):
        a = llf[off[s]:off[s] + n]
        b = llm[off[s]:off[s] + n]
        dd = (a - b).abs()
        live = a >= a.max() - WINDOW
        dmax_all = max(dmax_all, float(dd.max()))
        dmax_live = max(dmax_live, float(dd[live].max()))
        ratio = max(ratio, float((dd / bnd[off[s]:off[s] + n].double()).max()))
    sizes = np.full(n_stars, n)
    states = lambda: [D.pcg64_state(np.random.default_rng(1000 + s)) for s in range(n_stars)]
    settled = [0]

    def step():
        # lnL with bounds, then settling and rejection in one launch sequence
        run_m()
        good, n_pass, mx, sett = B.settle_reject_batch(batch, cache, llm, bnd, off, sizes,
                                                       states(), 256)
        settled[0] = int(sett.sum())
        return good, n_pass, mx

    def step_f():
        run_f()
        return B.reject_batch(llf, off, sizes, states(), 256)

    step()
    step_f()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(2):
        good_m, np_m, _ = step()
    torch.cuda.synchronize()
    e2e_ms = (time.perf_counter() - t0) / 2 * 1e3
    t0 = time.perf_counter()
    for _ in range(2):
        good_f, np_f, _ = step_f()
    torch.cuda.synchronize()
    e2e_f_ms = (time.perf_counter() - t0) / 2 * 1e3
    flips = sum(len(np.setxor1d(x, y)) for x, y in zip(good_m, good_f))
    # the live-row error of the settled buffer the mixed step handed on
    dmax_live_settled = 0.0
    for s in range(n_stars):
        a = llf[off[s]:off[s] + n]
        live = a >= a.max() - WINDOW
        dmax_live_settled = max(dmax_live_settled,
                                float((a - llm[off[s]:off[s] + n]).abs()[live].max()))
    # the same mixed step captured once in a HIP graph and replayed
    t0 = time.perf_counter()
    graph = B.StepGraph(batch, cache, d_items, len(items), total, off, sizes, states(), 256,
                        "mixed", shape)
    capture_s = time.perf_counter() - t0
    graph.run(states())
    torch.cuda.synchronize()
    # the next step's streams are drawn on the host while the replay runs,
    # as the eager step draws them while its lnL launch runs
    st = states()
    t0 = time.perf_counter()
    for _ in range(2):
        graph.launch(st)
        st = states()
        good_g, _, _, _ = graph.results()
    torch.cuda.synchronize()
    graph_ms = (time.perf_counter() - t0) / 2 * 1e3
    graph_flips = sum(len(np.setxor1d(x, y)) for x, y in zip(good_g, good_f))
    graph.plan.close()
    del graph
    # the lnL arrays the API hands back in mixed precision: every row whose
    # bound exceeds 1e-3 evaluated in FP64 first (jk_settle_tol_batch)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n_tol = int(B.settle_tol_batch(batch, cache, llm, bnd, off, sizes).sum())
    torch.cuda.synchronize()
    tol_ms = (time.perf_counter() - t0) * 1e3
    dmax_all_api = max(float((llf[off[s]:off[s] + n] - llm[off[s]:off[s] + n]).abs().max())
                       for s in range(n_stars))
    api = {"tolerance": 1e-3, "ms": tol_ms, "rows_evaluated_in_fp64": n_tol,
           "rows_fraction": n_tol / total, "max_abs_dlnl_all_after": dmax_all_api}
    log(f"[bench] batched: API arrays gated at 1e-3 in {tol_ms:.1f} ms ({n_tol} rows), max "
        f"|dlnL| all rows {dmax_all_api:.2e}")
    # the mixed step with the rows bounded by a probe (stars with >= 12 epochs)
    del llm, bnd
    torch.cuda.empty_cache()
    probes, adds, probed = B.probe_table(dstars)

    def step_p():
        ll_p, b_p, res_p, cnt_p = B.prune_reject_batch(batch, probes, adds, probed, cache,
                                                       d_items, len(items), off, sizes,
                                                       states(), 256, "mixed", shape)
        return res_p, int(cnt_p.sum()), int((b_p < 0).sum())

    step_p()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(2):
        res_p, n_eval_p, n_left_p = step_p()
    torch.cuda.synchronize()
    e2e_p_ms = (time.perf_counter() - t0) / 2 * 1e3
    flips_p = sum(len(np.setxor1d(x, y)) for x, y in zip(res_p[0], good_f))
    pruned = {"end_to_end_ms": e2e_p_ms,
              "end_to_end_sample_epochs_per_s": work / (e2e_p_ms * 1e-3),
              "stars_probed": int(probed.sum()),
              "rows_left_as_upper_bounds": n_left_p,
              "rows_evaluated_after_probe": n_eval_p,
              "accept_flips_vs_fp64": int(flips_p)}
    log(f"[bench] batched: pruned mixed step {e2e_p_ms:.2f} ms, {n_left_p / total:.3f} of rows "
        f"left as bounds, flips {flips_p}")
    res = {"stars": n_stars, "prior_rows": n, "epochs_total": int(M.sum()), "linear_params": 3,
           "launch_shape": shape.as_dict(), "autotune_s": tune_s,
           "autotune_ms": {k: round(v, 3) for k, v in table.items()},
           "mixed_lnl_ms": m_ms, "f64_lnl_ms": f_ms,
           "valu": {"mixed": valu_from_counters("batched_mixed", m_ms),
                    "f64": valu_from_counters("batched_f64", f_ms)},
           "mixed_sample_epochs_per_s": work / (m_ms * 1e-3),
           "f64_sample_epochs_per_s": work / (f_ms * 1e-3),
           "mixed_speedup_vs_f64": f_ms / m_ms,
           "mixed_end_to_end_ms": e2e_ms,
           "mixed_end_to_end_sample_epochs_per_s": work / (e2e_ms * 1e-3),
           "graph_step_ms": graph_ms, "graph_capture_s": capture_s,
           "graph_accept_flips_vs_fp64": int(graph_flips),
           "f64_end_to_end_ms": e2e_f_ms,
           "f64_end_to_end_sample_epochs_per_s": work / (e2e_f_ms * 1e-3),
           "max_abs_dlnl_all": dmax_all, "max_abs_dlnl_live": dmax_live,
           "max_abs_dlnl_live_settled": dmax_live_settled,
           "max_dlnl_over_bound": ratio,
           "rows_settled_in_fp64": settled[0], "rows_settled_fraction": settled[0] / total,
           "accept_flips_vs_fp64": int(flips),
           "accepted_total": int(np.minimum(np_f, 256).sum()),
           "api_gate": api, "pruned": pruned,
           "speedup_vs_plain_f64": {"mixed": e2e_f_ms / e2e_ms,
                                    "mixed_graph": e2e_f_ms / graph_ms,
                                    "mixed_pruned": e2e_f_ms / e2e_p_ms}}
    log(f"[bench] batched: mixed {m_ms:.2f} ms ({res['mixed_sample_epochs_per_s']:.3e}/s), f64 "
        f"{f_ms:.2f} ms; end to end mixed {e2e_ms:.2f} ms (graph replay {graph_ms:.2f} ms), "
        f"f64 {e2e_f_ms:.2f} ms; |dlnL| all "
        f"{dmax_all:.2e} live {dmax_live:.2e} (settled {dmax_live_settled:.2e}), max "
        f"|dlnL|/bound {ratio:.3f}, settled "
        f"{settled[0] / total:.2e} of rows, flips {flips}")
    del llf, cache, batch, probes
    torch.cuda.empty_cache()
    return res


# ---------------------------------------------------------------------------
def _prune_stats(srv, prune):
    """Rows of the requests whose star was probed and, of those, the rows
    that kept their upper bound instead of a full evaluation."""
    if not prune:
        return None
    probed, pruned = srv.prune_stats()
    return {"rows_probed": int(probed), "rows_pruned": int(pruned),
            "pruned_fraction_of_probed": (pruned / probed) if probed else None}


def leg_serving(args, torch, np, D, syn):
    import thejoker_amd as tj
    from thejoker_amd import serving as S
    from thejoker_amd import units as u

    n_req = args.serving_requests
    log(f"[bench] serving: {n_req} requests at 200/s, 2^26-row cache")
    cache = D.PriorCache.generate(1 << 26, seed=4, P_min=2.0, P_max=2048.0, ln_s=(-3.0, 1.0))
    prior = tj.JokerPrior.default(
        P_min=2 * u.day, P_max=2048 * u.day,
        sigma_v=[100 * u.km / u.s, 0.5 * u.km / u.s / u.day], poly_trend=2,
        pars={"K": tj.distributions.Normal(0.0, 30.0, unit=u.km / u.s),
              "s": tj.distributions.LogNormal(-3.0, 1.0, unit=u.km / u.s)})
    times, stars = syn.arrivals(n_req, rate=200.0, seed=4)
    datas = [st.to_rvdata() for st in stars]
    arena = int(args.serving_arena_gib * (1 << 30))
    step_work = int(args.serving_step_work) if args.serving_step_work > 0 else None
    prune = bool(args.serving_prune)
    # Warm-up, untimed: the first 60 requests of the stream through a server
    # like the measured one, with their own random streams.  Beyond the launch
    # paths, the process pays a one-time host cost of about 80 ms (seen in
    # the step that first plans ~50K work items) that a long-running server
    # pays once; without this warm-up it lands on the first requests of the
    # latency run.
    wsrv = S.JokerServer(prior, cache, precision="mixed", max_active=args.serving_active,
                         hbm_budget_bytes=arena, max_step_work=step_work, prune=prune)
    S.run_poisson(wsrv, times[:60], datas[:60], [np.random.default_rng(s) for s in range(60)])
    del wsrv
    torch.cuda.empty_cache()
    srv = S.JokerServer(prior, cache, precision="mixed", max_active=args.serving_active,
                        hbm_budget_bytes=arena, max_step_work=step_work, prune=prune)
    rngs = [np.random.default_rng(10_000 + i) for i in range(n_req)]
    lat_steps = []
    lat_step0 = srv.step

    def lat_logged_step():
        t = time.perf_counter()
        n_act = len(srv.active) + len(srv.pending)
        out = lat_step0()
        lat_steps.append(dict(srv.last_step, step_s=round(time.perf_counter() - t, 4),
                              active_or_pending=n_act))
        return out

    srv.step = lat_logged_step
    t0 = time.perf_counter()
    done = S.run_poisson(srv, times, datas, rngs)
    wall = time.perf_counter() - t0
    srv.step = lat_step0
    durs = np.array([d["step_s"] for d in lat_steps])
    lat = np.array([r.latency for r in done if r.error is None])
    errs = [r for r in done if r.error is not None]
    res = {"sample_epochs_note": "rows x epochs of every row the iterative rule decides (the "
                                 "reference's work), pruned rows included",
           "requests": n_req, "rate_per_s": 200.0, "completed": int(len(lat)),
           "failed": len(errs), "p50_latency_s": float(np.percentile(lat, 50)),
           "p99_latency_s": float(np.percentile(lat, 99)), "mean_latency_s": float(lat.mean()),
           "wall_s": wall, "stars_per_s": len(lat) / wall,
           "sample_epochs": int(srv.sample_epochs),
           "sample_epochs_per_s": srv.sample_epochs / wall, "steps": srv.steps,
           "max_active": args.serving_active, "prior_rows": len(cache)}
    log(f"[bench] serving: p50 {res['p50_latency_s'] * 1e3:.1f} ms p99 "
        f"{res['p99_latency_s'] * 1e3:.1f} ms, {res['stars_per_s']:.1f} stars/s, "
        f"{res['sample_epochs_per_s']:.3e} sample-epochs/s, {srv.steps} steps, "
        f"{srv.preemptions} restarts")
    # where the slowest requests' time goes: waiting for admission, then
    # their own rounds
    ok = [r for r in done if r.error is None]
    slow = sorted(ok, key=lambda r: -r.latency)[:max(1, len(ok) // 100)]
    res["step_s_quantiles"] = [float(q) for q in np.percentile(durs, [50, 90, 99, 100])]
    res["longest_steps"] = sorted(lat_steps, key=lambda d: -d["step_s"])[:5]
    res["slowest_1pct"] = [{"M": int(len(r.data)), "latency_s": round(r.latency, 4),
                            "queued_s": round(r.t_admit - r.arrival, 4),
                            "rounds": int(r.it.iteration), "rows": int(r.rows_evaluated)}
                           for r in slow]
    res["preemptions"] = srv.preemptions
    res["flat_switches"] = srv.flat_switches
    res["arena_gib"] = args.serving_arena_gib
    res["max_step_work"] = step_work
    res["prune"] = _prune_stats(srv, prune)
    del srv
    torch.cuda.empty_cache()
    # capacity: the whole stream submitted at once
    srv2 = S.JokerServer(prior, cache, precision="mixed", max_active=args.serving_active,
                         hbm_budget_bytes=arena, max_step_work=step_work, prune=prune)
    rngs2 = [np.random.default_rng(10_000 + i) for i in range(n_req)]
    step_log, step_info = [], []
    step0 = srv2.step

    def logged_step():
        t = time.perf_counter()
        n_act = len(srv2.active)
        out = step0()
        step_log.append((time.perf_counter() - t, n_act, srv2._waiting))
        step_info.append(dict(srv2.last_step, step_s=round(step_log[-1][0], 4)))
        return out

    srv2.step = logged_step
    t0 = time.perf_counter()
    done2 = S.run_poisson(srv2, np.zeros(n_req), datas, rngs2)
    wall2 = time.perf_counter() - t0
    sl = np.array(step_log)
    ok2 = [r for r in done2 if r.error is None]
    same = sum(int(np.array_equal(a.result["P"].value, b.result["P"].value))
               for a, b in zip(done, done2) if a.error is None and b.error is None)
    sat = {"preemptions": srv2.preemptions, "sample_epochs_per_s": srv2.sample_epochs / wall2,
           "steps": srv2.steps, "region_moves": dict(srv2.region_moves),
           "flat_switches": srv2.flat_switches,
           "prune": _prune_stats(srv2, prune)}
    del srv2, logged_step, step0
    gc.collect()
    torch.cuda.empty_cache()
    # where the host waits, summed over the steps (no extra synchronisation:
    # the settling call's wait includes the lnL launch before it)
    host_s = {}
    for d in step_info:
        for k, v in d.get("host_s", {}).items():
            host_s[k] = host_s.get(k, 0.0) + v
    res["saturated"] = {"wall_s": wall2, "stars_per_s": len(ok2) / wall2,
                        "completed": len(ok2), "failed": n_req - len(ok2), **sat,
                        "results_equal_latency_run": same,
                        "step_s_quantiles": [float(q) for q in
                                             np.percentile(sl[:, 0], [10, 50, 90, 100])],
                        "step_s_total": float(sl[:, 0].sum()),
                        "active_per_step_mean": float(sl[:, 1].mean()),
                        "waiting_per_step_mean": float(sl[:, 2].mean()),
                        "host_wait_s": {k: round(v, 4) for k, v in host_s.items()},
                        "slowest_steps": sorted(step_info, key=lambda d: -d["step_s"])[:4]}
    log(f"[bench] serving: saturated {res['saturated']['stars_per_s']:.1f} stars/s "
        f"({sat['sample_epochs_per_s']:.3e} sample-epochs/s, {sat['steps']} steps, "
        f"{sat['preemptions']} restarts, {n_req - len(ok2)} failed, {same} results equal the "
        f"latency run's); host time per call {res['saturated']['host_wait_s']}; "
        f"prune {sat['prune']}")
    if prune and not args.serving_quick:
        # the same saturated stream without pruning: its rate and its results
        srv3 = S.JokerServer(prior, cache, precision="mixed", max_active=args.serving_active,
                             hbm_budget_bytes=arena, max_step_work=step_work, prune=False)
        rngs3 = [np.random.default_rng(10_000 + i) for i in range(n_req)]
        t0 = time.perf_counter()
        done3 = S.run_poisson(srv3, np.zeros(n_req), datas, rngs3)
        wall3 = time.perf_counter() - t0
        ok3 = [r for r in done3 if r.error is None]
        same3 = sum(int(np.array_equal(a.result["P"].value, b.result["P"].value) and
                        np.array_equal(a.result["K"].value, b.result["K"].value))
                    for a, b in zip(done2, done3) if a.error is None and b.error is None)
        res["saturated_no_prune"] = {"wall_s": wall3, "stars_per_s": len(ok3) / wall3,
                                     "sample_epochs_per_s": srv3.sample_epochs / wall3,
                                     "results_equal_pruned_run": same3}
        log(f"[bench] serving: saturated without pruning {len(ok3) / wall3:.1f} stars/s; "
            f"{same3} of {len(ok3)} results equal the pruned run's")
        del srv3
        gc.collect()
        torch.cuda.empty_cache()
    # The plain path: the same stars one at a time through
    # TheJoker(precision="f64").iterative_rejection_sample.  Same rows, so
    # the same sample-epochs as in the server; also checks for flips.
    if args.serving_quick:
        del cache
        torch.cuda.empty_cache()
        return res
    flips, checked, se, t_alone = 0, 0, 0, 0.0
    for i in range(0, n_req, max(1, n_req // max(1, args.serving_plain))):
        if done[i].error is not None:
            continue
        joker = tj.TheJoker(prior, rng=np.random.default_rng(10_000 + i), precision="f64")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        alone = joker.iterative_rejection_sample(datas[i], cache, n_requested_samples=256)
        torch.cuda.synchronize()
        t_alone += time.perf_counter() - t0
        se += done[i].sample_epochs
        a, b = alone["P"].value, done[i].result["P"].value
        flips += int(len(np.setxor1d(a, b)))
        checked += 1
    res["plain_f64_one_at_a_time"] = {"stars": checked, "wall_s": t_alone,
                                      "sample_epochs_per_s": se / t_alone,
                                      "stars_per_s": checked / t_alone}
    res["speedup_vs_plain_f64_one_at_a_time"] = res["saturated"]["sample_epochs_per_s"] / (
        se / t_alone)
    res["fp64_alone_checked"] = checked
    res["accept_flips_vs_fp64_alone"] = flips
    # The mixed-precision lnL the served stars' decisions rest on, against
    # FP64 on the same rows: the same iterative path run alone (its samples
    # equal the server's) keeps its final buffer; every row that holds a
    # value (not a probe's upper bound) is compared, and the rows within 40
    # of the maximum separately.
    d_all, d_live, n_cmp, n_same = 0.0, 0.0, 0, 0
    for i in range(0, n_req, max(1, n_req // 64)):
        if done[i].error is not None:
            continue
        joker = tj.TheJoker(prior, rng=np.random.default_rng(10_000 + i), precision="mixed",
                            prune=prune)
        alone = joker.iterative_rejection_sample(datas[i], cache, n_requested_samples=256)
        n_same += int(np.array_equal(alone["P"].value, done[i].result["P"].value))
        ll_m, b_m = joker.last_lnl
        dstar = joker.last_star
        ll_f = D.lnl(dstar, cache, n=int(ll_m.shape[0]), precision="f64")
        val = b_m >= 0
        dd = (ll_f - ll_m).abs()
        d_all = max(d_all, float(dd[val].max()))
        live = val & (ll_f >= ll_f.max() - WINDOW)
        d_live = max(d_live, float(dd[live].max()))
        n_cmp += 1
    res["mixed_vs_fp64"] = {"stars": n_cmp, "results_equal_server": n_same,
                            "max_abs_dlnl_evaluated_rows": d_all,
                            "max_abs_dlnl_live_rows": d_live,
                            "returned_ln_likelihood": "FP64 values (max |dlnL| 0)"}
    log(f"[bench] serving: mixed vs FP64 on {n_cmp} stars' buffers: max |dlnL| {d_all:.2e} "
        f"(live {d_live:.2e}); {n_same} results equal the server's")
    log(f"[bench] serving: plain FP64 one star at a time {se / t_alone:.3e} sample-epochs/s "
        f"({checked / t_alone:.1f} stars/s) -> serving "
        f"{res['speedup_vs_plain_f64_one_at_a_time']:.1f}x; flips vs FP64 alone on {checked} "
        f"stars: {flips}")
    del cache
    torch.cuda.empty_cache()
    return res


# ---------------------------------------------------------------------------
def leg_prior_io(args, torch, np, D, syn):
    """A prior cache file in the reference's HDF5 layout: 2^24 rows written by
    JokerSamples.write, then streamed into the device columns by
    PriorCache.from_hdf5 (the file is in the page cache: this times the
    format decoding, unit conversion and host-to-device copies)."""
    import tempfile

    import thejoker_amd as tj
    from thejoker_amd import units as u

    n = 1 << 24
    cols = syn.prior_columns(n, seed=6, jitter=True)
    js = tj.JokerSamples(poly_trend=1)
    js["P"] = cols["P"] * u.day
    js["e"] = cols["e"]
    js["omega"] = cols["omega"] * u.rad
    js["M0"] = cols["M0"] * u.rad
    js["s"] = cols["s"] * 1000.0 * u.m / u.s
    js["ln_prior"] = np.zeros(n)
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "prior.hdf5")
        t0 = time.perf_counter()
        js.write(path)
        t_write = time.perf_counter() - t0
        size = os.path.getsize(path)
        D.PriorCache.from_hdf5(path, rv_unit=u.km / u.s, block_rows=1 << 20)   # warm the cache
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        c = D.PriorCache.from_hdf5(path, rv_unit=u.km / u.s)
        torch.cuda.synchronize()
        t_read = time.perf_counter() - t0
    same = bool(np.array_equal(c.P.cpu().numpy(), cols["P"]) and
                np.allclose(c.s.cpu().numpy(), cols["s"], rtol=1e-15, atol=0))
    res = {"rows": n, "file_bytes": size, "write_s": t_write, "read_to_device_s": t_read,
           "read_gb_per_s": size / t_read / 1e9, "columns_equal": same}
    log(f"[bench] prior_io: {size / 1e9:.2f} GB HDF5 written in {t_write:.2f} s, streamed into "
        f"HBM in {t_read:.2f} s ({res['read_gb_per_s']:.2f} GB/s), columns equal: {same}")
    del c
    torch.cuda.empty_cache()
    return res


def leg_stress(args, torch, np, D, syn):
    log2 = args.stress_log2_rows
    n = 1 << log2
    chunk = min(n, 1 << 26)
    log(f"[bench] stress: M=50 star, 2^{log2} rows in 2^{chunk.bit_length() - 1}-row chunks")
    star = syn.single_star(seed=5, M=50, baseline=3650.0)
    dstar = D.DeviceStar(star.t, star.rv, star.ivar(), star.trend(), np.zeros(2),
                         np.array([900.0, 1e4]), star.t_ref, jitter_mode="physical", device=0)
    t0 = time.perf_counter()
    cache = D.PriorCache.generate(n, seed=5, P_min=1.0, P_max=16384.0, ln_s=(-3.0, 1.0))
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    llf = torch.empty(n, dtype=torch.float64, device="cuda")
    llm = torch.empty(n, dtype=torch.float64, device="cuda")

    def run(prec, out):
        for s0 in range(0, n, chunk):
            D.lnl(dstar, cache, precision=prec, start=s0, n=min(chunk, n - s0), out=out[s0:s0 + chunk])

    run("f64", llf)
    run("mixed", llm)
    f_ms = timed(lambda: run("f64", llf), 2, torch)
    m_ms = timed(lambda: run("mixed", llm), 2, torch)
    work = n * star.M
    bnd = torch.empty(n, dtype=torch.float32, device="cuda")
    for s0 in range(0, n, chunk):
        D.lnl(dstar, cache, precision="mixed", start=s0, n=min(chunk, n - s0),
              out=llm[s0:s0 + chunk], bound=bnd[s0:s0 + chunk])
    cmp = mixed_vs_f64(torch, np, D, dstar, cache, llf, llm, bnd, seed=5, max_keep=256)
    del bnd
    flops = n * (star.M * F_EPOCH_JITTER + F_SAMPLE[2])
    res = {"prior_rows": n, "epochs": star.M, "generate_s": gen_s,
           "f64_ms": f_ms, "mixed_ms": m_ms,
           "f64_sample_epochs_per_s": work / (f_ms * 1e-3),
           "mixed_sample_epochs_per_s": work / (m_ms * 1e-3),
           "mixed_speedup_vs_f64": f_ms / m_ms,
           "f64_tflops_algorithmic": flops / (f_ms * 1e-3) / 1e12,
           "f64_frac_fp64_peak": flops / (f_ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
           "f64_hbm_gbs": n * 48 / (f_ms * 1e-3) / 1e9,
           "mixed_hbm_gbs": n * 48 / (m_ms * 1e-3) / 1e9,
           **cmp}
    log(f"[bench] stress: f64 {f_ms:.1f} ms mixed {m_ms:.1f} ms ({f_ms / m_ms:.2f}x), "
        f"{fmt_mixed(cmp)}")
    del llf, llm, cache
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # The default warmup covers the device's clock ramp: the single-star step
    # time falls by about 10% over the first ~30 steps of a fresh process.
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--log2-rows", type=int, default=24)
    ap.add_argument("--workload", default="all",
                    choices=["all", "single", "batched", "serving", "stress", "prior_io"])
    ap.add_argument("--serving-requests", type=int, default=1000)
    ap.add_argument("--serving-active", type=int, default=64)
    ap.add_argument("--serving-step-work", type=float, default=4e9,
                    help="sample-epochs one serving step may evaluate, shared shortest "
                         "remaining work first (0: whole rounds)")
    ap.add_argument("--serving-arena-gib", type=float, default=64.0,
                    help="HBM for the server's lnL arena (12 B per evaluated row)")
    ap.add_argument("--serving-prune", type=int, default=1,
                    help="1: the server prunes rows by the probe bound (results unchanged); "
                         "the saturated stream is also run without pruning for comparison")
    ap.add_argument("--serving-plain", type=int, default=1000,
                    help="stars of the serving stream also run one at a time in plain FP64")
    ap.add_argument("--serving-quick", action="store_true",
                    help="serving: latency and saturated runs only (no comparison runs)")
    ap.add_argument("--stress-log2-rows", type=int, default=29)
    ap.add_argument("--prune-epochs", default="24,50,100,200",
                    help="epoch counts of the single-star pruned-vs-plain records (comma "
                         "separated; empty for none, as in the kernel profiles of the M=16 star)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (the "
                         "reference dispatch for scaling FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the likelihood path here runs on a single MI355X "
              f"(--gpus {args.gpus} is not supported; shard stars or prior rows across "
              f"processes instead)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from thejoker_amd import device as D
    from thejoker_amd import synthetic as syn

    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes, "dispatches": 3}), flush=True)
        return

    out = leg_single(args, torch, np, D, syn) if args.workload in ("all", "single") else {
        "metric": METRIC, "value": None, "unit": "sample-epochs/s", "n_gpus": 1}
    legs = {"batched": leg_batched, "serving": leg_serving, "stress": leg_stress,
            "prior_io": leg_prior_io}
    for name, fn in legs.items():
        if args.workload in ("all", name):
            out[name] = fn(args, torch, np, D, syn)
    if "serving" in out and out.get("value"):
        out["serving"]["sample_epochs_per_s_vs_single_f64"] = (
            out["serving"]["saturated"]["sample_epochs_per_s"] / out["value"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
