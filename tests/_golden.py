"""Loaders of the committed golden fixtures (tests/golden/*.npz)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def lnl_stars():
    d = load("lnl_stars.npz")
    out = {}
    for name in d["names"]:
        p = str(name) + "__"
        kp = d[p + "kprior"]
        out[str(name)] = dict(
            t=d[p + "t"], rv=d[p + "rv"], err=d[p + "err"], t_ref=float(d[p + "t_ref"]),
            trend=d[p + "trend"], mu=d[p + "mu"], lam=d[p + "lam"],
            kprior=None if kp.size == 0 else tuple(float(x) for x in kp),
            rows=d[p + "rows"], ll_reference_mode=d[p + "ll_reference_mode"],
            ll_physical_mode=d[p + "ll_physical_mode"], post_a=d[p + "post_a"],
            post_cov=d[p + "post_cov"])
    return out


def oracle_star(st, jitter_mode):
    from oracle import oracle
    kw = {}
    if st["kprior"] is not None:
        kw = dict(k_prior_mode=1, sigma_K0=st["kprior"][0], P0=st["kprior"][1],
                  max_K=st["kprior"][2])
    return oracle.OracleStar(st["t"], st["rv"], 1.0 / st["err"] ** 2, st["trend"], st["mu"],
                             st["lam"], st["t_ref"], jitter_mode=jitter_mode, **kw)


def rel_err(a, b):
    """|a - b| / max(|b|, 1): relative error with the scale floored at 1."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1.0)
