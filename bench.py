"""Benchmark of The Joker's likelihood + rejection path on one MI355X.

Default workload (N = 1): the single-star plain-FP64 case -- one star with
M = 16 epochs over 1000 d, 2^24 prior rows (P log-uniform on [2, 2048] d,
e ~ Beta(0.867, 3.03), omega, M0 ~ U[0, 2 pi), ln s ~ N(-3, 1) km/s),
K ~ N(0, 30), v0 ~ N(0, 100), physical jitter mode.  One step = the FP64
marginal-likelihood kernel over all 2^24 device-resident rows plus the
on-device rejection step (max, numpy-PCG64 uniforms, ordered compaction of
at most 256 indices) and the read-back of the accepted indices.

value = prior-samples x epochs / s of the whole step.  The JSON line also
carries:
  roofline      -- the likelihood kernel's algorithmic FLOP rate (HIP events
                   around its launches on the launch stream) against the
                   FP64 vector peak;
  cpu_baseline  -- the oracle's restatement of the reference algorithm
                   (LU + Woodbury, fast_likelihood.pyx:255-469) on the host
                   cores over a bounded sample of the same rows;
  mixed         -- the mixed-precision kernel on the same rows: its rate and
                   its error against FP64 (max |dlnL| over all rows and over
                   the live set, acceptance flips after FP64 re-evaluation).
"""
import argparse
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
# Kepler = starter + 4 Newton steps): per sample-epoch with sampled jitter, and
# per sample for L = 2.
F_EPOCH_JITTER = 176
F_EPOCH_NOJITTER = 156
F_SAMPLE_L2 = 90
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
BYTES_PER_SAMPLE_F64 = 56   # 5 x f64 read + lnL write + lnL read by the rejection step


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(star, rows, mu, lam, n_sample):
    """Oracle (reference algorithm) on the host cores over the first n_sample
    rows; sample-epochs/s."""
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
    oracle.batch_lnl(ost, sub, method="reference", n_threads=cores)
    dt = time.perf_counter() - t0
    return {"value": n_sample * star.M / dt, "unit": "sample-epochs/s", "cores": cores,
            "kind": "port",
            "sample": f"first {n_sample} of the same prior rows, reference LU+Woodbury algorithm "
                      f"restated in C (-O3), {cores} threads, {dt:.2f} s wall"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2-rows", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
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

    n = 1 << args.log2_rows
    log(f"[bench] building workload: M=16 star, {n} prior rows")
    star = syn.single_star()
    cols = syn.prior_columns(n, seed=0, jitter=True)
    mu, lam = np.zeros(2), np.array([900.0, 1e4])
    dstar = D.DeviceStar(star.t, star.rv, star.ivar(), star.trend(), mu, lam, star.t_ref,
                         jitter_mode="physical", device=0)
    cache = D.PriorCache(cols["P"], cols["e"], cols["omega"], cols["M0"], cols["s"], device=0,
                         fp32_copy=not args.no_mixed)
    rng = np.random.default_rng(42)
    ll = torch.empty(n, dtype=torch.float64, device="cuda")

    def step():
        D.lnl(dstar, cache, precision="f64", out=ll)
        idx, n_pass, mx = D.reject(ll, D.pcg64_state(rng), 256)
        D.pcg64_advance_state(rng, n)
        return idx, n_pass

    log(f"[bench] warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        idx, n_pass = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_step = dt / args.steps * 1e3
    value = n * star.M / (dt / args.steps)
    log(f"[bench] {ms_step:.3f} ms/step, {value:.4e} sample-epochs/s, accepted {n_pass}")

    # Kernel time of the likelihood launch alone (HIP events on the launch stream).
    reps = max(5, args.steps)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        D.lnl(dstar, cache, precision="f64", out=ll)
    ev1.record()
    torch.cuda.synchronize()
    k_ms = ev0.elapsed_time(ev1) / reps
    flops = n * (star.M * F_EPOCH_JITTER + F_SAMPLE_L2)
    achieved = flops / (k_ms * 1e-3) / 1e12
    roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                "kernel": "k_lnl<2,true,false> (FP64)", "kernel_ms": k_ms,
                "flops_per_launch": flops,
                "hbm_gbs_prior_stream": n * BYTES_PER_SAMPLE_F64 / (k_ms * 1e-3) / 1e9}
    log(f"[bench] lnL kernel {k_ms:.3f} ms -> {achieved:.2f} TFLOP/s algorithmic "
        f"({100 * achieved / PEAK_FP64_TFLOPS:.1f}% of FP64 peak)")

    out = {"metric": METRIC, "value": value, "unit": "sample-epochs/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": "single star, plain FP64 likelihood + rejection: M=16 epochs, "
                                  f"2^{args.log2_rows} prior rows, ln s ~ N(-3,1), physical "
                                  "jitter, K~N(0,30), v0~N(0,100), max_posterior_samples=256",
                      "epochs": star.M, "prior_rows": n, "linear_params": 2,
                      "parallelism": "single GPU"},
           "roofline": roofline}

    if not args.no_mixed:
        ll32 = torch.empty(n, dtype=torch.float64, device="cuda")
        D.lnl(dstar, cache, precision="mixed", out=ll32)
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(reps):
            D.lnl(dstar, cache, precision="mixed", out=ll32)
        ev1.record()
        torch.cuda.synchronize()
        m_ms = ev0.elapsed_time(ev1) / reps
        D.lnl(dstar, cache, precision="f64", out=ll)
        diff = (ll32 - ll).abs()
        live = ll >= ll.max() - 40
        out["mixed"] = {"kernel_ms": m_ms, "sample_epochs_per_s": n * star.M / (m_ms * 1e-3),
                        "speedup_vs_f64_kernel": k_ms / m_ms,
                        "max_abs_dlnl_all": float(diff.max()),
                        "max_abs_dlnl_live": float(diff[live].max()),
                        "live_rows": int(live.sum())}
        log(f"[bench] mixed kernel {m_ms:.3f} ms ({k_ms / m_ms:.2f}x), max|dlnL| all "
            f"{out['mixed']['max_abs_dlnl_all']:.2e} live {out['mixed']['max_abs_dlnl_live']:.2e}")

    if not args.no_cpu:
        rows = np.stack([cols[k] for k in ("P", "e", "omega", "M0", "s")], axis=1)
        n_cpu = min(n, 1 << 21)
        log(f"[bench] cpu baseline on {n_cpu} rows")
        out["cpu_baseline"] = cpu_baseline(star, np.ascontiguousarray(rows), mu, lam, n_cpu)
        log(f"[bench] cpu {out['cpu_baseline']['value']:.3e} sample-epochs/s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
