"""Write the golden fixtures of tests/golden/ (numpy + scipy only).

Run from the repository root:  python tests/golden/make_golden.py

Nothing here imports the reference or the repository's oracle: every
expected value comes from an independent computation --
* Kepler's equation by vectorised bisection (brute force to the last ulp);
* the marginal likelihood as the dense Gaussian ln N(y | M mu, C + M Lambda M^T)
  with numpy's LU slogdet / solve (the quantity fast_likelihood.pyx:359-425
  and src/tests/py_likelihood.py:32-107 compute);
* numpy's own PCG64 stream (np.random.default_rng) for the rejection uniforms,
  and np.where(np.exp(ll - ll.max()) > u) for the accepted sets
  (likelihood_helpers.py:107-109).
Star inputs: the 5-epoch star of the reference's doctest (docs/index.rst:52-55),
the first 20 rows of docs/examples/data.ecsv (read as text from the reference
tree when present; the committed fixture holds the numbers), the 3-epoch star
construction of src/tests/test_fast_likelihood.py:24-30, and synthetic stars.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_ECSV = "/root/reference/docs/examples/data.ecsv"


def kepler_bisect(M, e, iters=200):
    M = np.asarray(M, dtype=np.float64)
    e = np.broadcast_to(np.asarray(e, dtype=np.float64), M.shape)
    lo = M - 1.0
    hi = M + 1.0
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        f = mid - e * np.sin(mid) - M
        hi = np.where(f > 0, mid, hi)
        lo = np.where(f > 0, lo, mid)
    return 0.5 * (lo + hi)


def design_column(t, t_ref, row):
    P, e, om, M0 = row[0], row[1], row[2], row[3]
    M = 2 * np.pi * (t - t_ref) / P - M0
    E = kepler_bisect(M, e)
    f = 2 * np.arctan2(np.sqrt(1 + e) * np.sin(E / 2), np.sqrt(1 - e) * np.cos(E / 2))
    return np.cos(om + f) + e * np.cos(om)


def dense_lnl(t, rv, err, t_ref, trend, mu, lam, row, jitter, kprior=None):
    """ln N(y | M mu, C + M Lambda M^T); C = diag(err^2 (+ s^2 if jitter))."""
    M = np.hstack([design_column(t, t_ref, row)[:, None], trend])
    lam = np.array(lam, dtype=np.float64)
    if kprior is not None:
        sK0, P0, maxK = kprior
        v = sK0 ** 2 / (1 - row[1] ** 2) * (row[0] / P0) ** (-2.0 / 3)
        lam[0] = min(maxK ** 2, v)
    var = err ** 2 + (row[4] ** 2 if jitter else 0.0)
    B = np.diag(var) + M @ np.diag(lam) @ M.T
    r = rv - M @ mu
    sign, logdet = np.linalg.slogdet(2 * np.pi * B)
    # B is ill-conditioned (v0 prior variance ~1e4 against sigma^2 ~ 1e-2), so
    # refine the solve twice with residuals computed in long double.
    x = np.linalg.solve(B, r)
    Bl = B.astype(np.longdouble)
    rl = r.astype(np.longdouble)
    for _ in range(2):
        res = rl - Bl @ x.astype(np.longdouble)
        x = x + np.linalg.solve(B, res.astype(np.float64))
    chi2 = float(np.dot(rl, x.astype(np.longdouble)))
    return -0.5 * (chi2 + logdet)


def dense_posterior(t, rv, err, t_ref, trend, mu, lam, row, jitter, kprior=None):
    """a = G^-1 (M^T C^-1 y + Lambda^-1 mu), cov = G^-1 with G = Lambda^-1 +
    M^T C^-1 M; Lambda[0] unclipped (fast_likelihood.pyx:394-423, 519-521)."""
    M = np.hstack([design_column(t, t_ref, row)[:, None], trend])
    lam = np.array(lam, dtype=np.float64)
    if kprior is not None:
        sK0, P0, maxK = kprior
        lam[0] = sK0 ** 2 / (1 - row[1] ** 2) * (row[0] / P0) ** (-2.0 / 3)
    var = err ** 2 + (row[4] ** 2 if jitter else 0.0)
    G = np.diag(1 / lam) + M.T @ np.diag(1 / var) @ M
    rhs = M.T @ (rv / var) + mu / lam
    a = np.linalg.solve(G, rhs)
    a = a + np.linalg.solve(G, (rhs.astype(np.longdouble) - G.astype(np.longdouble) @ a).astype(np.float64))
    return a, np.linalg.inv(G)


def prior_rows(rng, n, P_min, P_max, jitter):
    P = np.exp(rng.uniform(size=n) * (np.log(P_max) - np.log(P_min)) + np.log(P_min))
    e = rng.beta(0.867, 3.03, size=n)
    om = rng.uniform(0, 2 * np.pi, size=n)
    M0 = rng.uniform(0, 2 * np.pi, size=n)
    s = np.exp(rng.normal(-3.0, 1.0, size=n)) if jitter else np.zeros(n)
    return np.stack([P, e, om, M0, s], axis=1)


def special_rows(P_min, P_max):
    rows = [
        [P_min, 0.0, 0.0, 0.0, 0.0],
        [P_max, 0.0, 1.0, 2.0, 0.1],
        [37.3, 0.9, 2.5, 0.1, 0.05],
        [120.0, 0.3, 1.0, 2.0, 0.0],
        [5.1, 0.99, 4.0, 6.2, 0.3],
        [800.0, 0.999, 0.5, 3.14159, 0.02],
        [P_min * 1.0001, 0.5, 6.28, 0.0, 1.5],
    ]
    return np.array(rows, dtype=np.float64)


def read_ecsv_rows(path, n):
    t, rv, err = [], [], []
    with open(path) as f:
        for line in f:
            if line.startswith("#") or line.startswith("bjd"):
                continue
            a, b, c = line.split()[:3]
            t.append(float(a) - 2400000.5)
            rv.append(float(b))
            err.append(float(c))
            if len(t) == n:
                break
    return np.array(t), np.array(rv), np.array(err)


def synth_star(seed, M, baseline):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0.0, baseline, M))
    err = rng.uniform(0.1, 0.5, M)
    t_ref = t.min()
    P, e, om, M0, K, v0 = 120.0, 0.3, 1.0, 2.0, 10.0, -5.0
    m = design_column(t, t_ref, np.array([P, e, om, M0, 0.0]))
    rv = K * m + v0 + rng.normal(0.0, err)
    return t, rv, err, t_ref


def make_stars():
    stars = {}
    # (a) the doctest star, default prior: FixedCompanionMass K (sigma_K0 = 30
    # km/s, P0 = 365.25 d, max_K = 500 km/s), v0 ~ N(0, 100).
    t = np.array([0., 49.452, 95.393, 127.587, 190.408])
    rv = np.array([38.77, 39.70, 37.45, 38.31, 38.31])
    err = np.array([0.184, 0.261, 0.112, 0.155, 0.223])
    stars["doctest"] = dict(t=t, rv=rv, err=err, t_ref=0.0, poly_trend=1,
                            mu=np.zeros(2), lam=np.array([1.0, 100.0 ** 2]),
                            kprior=np.array([30.0, 365.25, 500.0]), P_min=2.0, P_max=256.0)
    # (b) the 16-epoch synthetic star of the plain-method workload.
    t, rv, err, t_ref = synth_star(1, 16, 1000.0)
    stars["cfg1"] = dict(t=t, rv=rv, err=err, t_ref=t_ref, poly_trend=1, mu=np.zeros(2),
                         lam=np.array([30.0 ** 2, 100.0 ** 2]), kprior=None, P_min=2.0,
                         P_max=2048.0)
    # (c) 3-epoch star as in test_fast_likelihood.py:24-30 (rv = cos t).
    rng = np.random.default_rng(seed=42)
    t = np.sort(rng.uniform(0, 250, 3)) + 56831.324
    stars["three"] = dict(t=t, rv=np.cos(t), err=rng.uniform(0.1, 0.2, t.size), t_ref=t.min(),
                          poly_trend=1, mu=np.zeros(2), lam=np.array([10.0 ** 2, 100.0 ** 2]),
                          kprior=None, P_min=8.0, P_max=32768.0)
    # (d) data.ecsv rows, linear trend (L = 3), nonzero prior means.
    if os.path.exists(REF_ECSV):
        t, rv, err = read_ecsv_rows(REF_ECSV, 20)
    else:
        old = np.load(os.path.join(HERE, "lnl_stars.npz"))
        t, rv, err = old["ecsv__t"], old["ecsv__rv"], old["ecsv__err"]
    stars["ecsv"] = dict(t=t, rv=rv, err=err, t_ref=t.min(), poly_trend=2,
                         mu=np.array([0.0, 38.0, 0.01]), lam=np.array([1.0, 10.0 ** 2, 0.5 ** 2]),
                         kprior=np.array([25.0, 365.25, 500.0]), P_min=1.0, P_max=512.0)
    # (e) 50 epochs over 3650 d with quadratic trend (L = 4), K mean nonzero.
    t, rv, err, t_ref = synth_star(5, 50, 3650.0)
    stars["long"] = dict(t=t, rv=rv, err=err, t_ref=t_ref, poly_trend=3,
                         mu=np.array([1.0, 0.0, 0.0, 0.0]),
                         lam=np.array([30.0 ** 2, 100.0 ** 2, 0.5 ** 2, 1e-3 ** 2]),
                         kprior=None, P_min=1.0, P_max=16384.0)
    return stars


def main():
    out = {}
    rng = np.random.default_rng(20240306)
    for name, st in make_stars().items():
        dt = st["t"] - st["t_ref"]
        trend = np.hstack([np.ones((dt.size, 1)),
                           np.vander(dt, N=st["poly_trend"], increasing=True)[:, 1:]])
        rows = np.vstack([special_rows(st["P_min"], st["P_max"]),
                          prior_rows(rng, 249, st["P_min"], st["P_max"], jitter=True)])
        kp = None if st["kprior"] is None else tuple(st["kprior"])
        ll_ref = np.array([dense_lnl(st["t"], st["rv"], st["err"], st["t_ref"], trend, st["mu"],
                                     st["lam"], r, False, kp) for r in rows])
        ll_phys = np.array([dense_lnl(st["t"], st["rv"], st["err"], st["t_ref"], trend, st["mu"],
                                      st["lam"], r, True, kp) for r in rows])
        post = [dense_posterior(st["t"], st["rv"], st["err"], st["t_ref"], trend, st["mu"],
                                st["lam"], r, True, kp) for r in rows[:32]]
        pref = name + "__"
        out[pref + "t"] = st["t"]
        out[pref + "rv"] = st["rv"]
        out[pref + "err"] = st["err"]
        out[pref + "t_ref"] = np.array(st["t_ref"])
        out[pref + "trend"] = trend
        out[pref + "mu"] = st["mu"]
        out[pref + "lam"] = st["lam"]
        out[pref + "kprior"] = np.array([]) if st["kprior"] is None else st["kprior"]
        out[pref + "rows"] = rows
        out[pref + "ll_reference_mode"] = ll_ref
        out[pref + "ll_physical_mode"] = ll_phys
        out[pref + "post_a"] = np.array([p[0] for p in post])
        out[pref + "post_cov"] = np.array([p[1] for p in post])
    out["names"] = np.array(list(make_stars().keys()))
    np.savez_compressed(os.path.join(HERE, "lnl_stars.npz"), **out)

    # Kepler brute force on a grid of (M, e), M spanning several turns.
    Ms = np.concatenate([np.linspace(-40.0, 40.0, 401), [0.0, np.pi, -np.pi, 1e-8, 2 * np.pi]])
    es = np.array([0.0, 1e-6, 0.1, 0.3, 0.5, 0.8, 0.9, 0.95, 0.99, 0.999])
    MM, EE = np.meshgrid(Ms, es, indexing="ij")
    np.savez_compressed(os.path.join(HERE, "kepler.npz"), M=MM.ravel(), e=EE.ravel(),
                        E=kepler_bisect(MM.ravel(), EE.ravel()))

    # numpy PCG64 streams.
    pcg = {}
    for i, seed in enumerate([0, 1, 42, 2 ** 63 + 5, 123456789]):
        g = np.random.default_rng(seed)
        st = g.bit_generator.state["state"]
        pcg[f"s{i}__state"] = np.array([st["state"] >> 64, st["state"] & (2 ** 64 - 1)],
                                       dtype=np.uint64)
        pcg[f"s{i}__inc"] = np.array([st["inc"] >> 64, st["inc"] & (2 ** 64 - 1)], dtype=np.uint64)
        g2 = np.random.default_rng(seed)
        pcg[f"s{i}__raw"] = g2.bit_generator.random_raw(64)
        pcg[f"s{i}__uniform"] = np.random.default_rng(seed).uniform(size=3000)
        g3 = np.random.default_rng(seed)
        g3.bit_generator.advance(10 ** 7 + 3)
        pcg[f"s{i}__uniform_after_1e7p3"] = g3.uniform(size=64)
    pcg["n_seeds"] = np.array(5)
    np.savez_compressed(os.path.join(HERE, "pcg64.npz"), **pcg)

    # Accepted sets: a flat star (many accepted) and the cfg1 star, lnL from
    # the dense formula, uniforms from default_rng(42).
    st = make_stars()["cfg1"]
    rngs = np.random.default_rng(7)
    rows = prior_rows(rngs, 4096, 2.0, 2048.0, jitter=False)
    dt = st["t"] - st["t_ref"]
    trend = np.ones((dt.size, 1))
    acc = {"rows": rows}
    for label, rv in (("informative", st["rv"]), ("flat", 1e-4 * np.cos(st["t"]))):
        ll = np.array([dense_lnl(st["t"], rv, st["err"], st["t_ref"], trend, np.zeros(2),
                                 np.array([900.0, 1e4]), r, False) for r in rows])
        uu = np.random.default_rng(42).uniform(size=ll.size)
        acc[label + "__rv"] = rv
        acc[label + "__ll"] = ll
        acc[label + "__idx"] = np.where(np.exp(ll - ll.max()) > uu)[0]
    acc["t"], acc["err"], acc["t_ref"] = st["t"], st["err"], np.array(st["t_ref"])
    np.savez_compressed(os.path.join(HERE, "accept_sets.npz"), **acc)
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


if __name__ == "__main__":
    sys.exit(main())
