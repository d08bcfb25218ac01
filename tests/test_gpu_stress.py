"""The large-precision-stress workload on the GPU: one star with M = 50
epochs over 3650 d, P log-uniform on [1, 16384] d (short periods need the
FP64 phase reduction), e ~ Beta(0.867, 3.03), ln s ~ N(-3, 1), L = 2, on
2^22 device-generated prior rows.  Checks: the FP64 kernel against the
oracle (the reference algorithm, fast_likelihood.pyx:255-469) to 1e-10 on
2^16 rows; the mixed kernel's error bound on every row; after settling, the
mixed path's accepted set equals the FP64 path's under the same uniforms
(likelihood_helpers.py:107-109), and the API's mixed lnL array is within
1e-3 of FP64 everywhere."""
import numpy as np
import pytest
import torch

from oracle import oracle
from thejoker_amd import device as D
from thejoker_amd import synthetic as syn
from thejoker_amd.thejoker import TheJoker

pytestmark = pytest.mark.gpu

N_ROWS = 1 << 22


@pytest.fixture(scope="module")
def workload(gpu):
    star = syn.single_star(seed=5, M=50, baseline=3650.0)
    mu, lam = np.zeros(2), np.array([900.0, 1e4])
    dstar = D.DeviceStar(star.t, star.rv, star.ivar(), star.trend(), mu, lam, star.t_ref,
                         jitter_mode="physical")
    cache = D.PriorCache.generate(N_ROWS, seed=5, P_min=1.0, P_max=16384.0, ln_s=(-3.0, 1.0))
    return star, dstar, cache, mu, lam


def test_stress_f64_against_oracle(workload):
    star, dstar, cache, mu, lam = workload
    ll = D.lnl(dstar, cache, precision="f64").cpu().numpy()
    assert np.isfinite(ll).all()
    rows = cache.host_rows(np.arange(1 << 16))
    assert rows[:, 0].min() < 2.0 and rows[:, 0].max() > 8000.0
    ost = oracle.OracleStar(star.t, star.rv, star.ivar(), star.trend(), mu, lam, star.t_ref,
                            jitter_mode=1)
    want, capped = oracle.batch_lnl(ost, rows, n_threads=16, return_capped=True)
    ok = ~capped
    rel = np.abs(ll[:1 << 16][ok] - want[ok]) / np.maximum(np.abs(want[ok]), 1.0)
    assert rel.max() < 1e-10, rel.max()


def test_stress_mixed_bound_and_accepted_set(workload):
    star, dstar, cache, _, _ = workload
    llf = D.lnl(dstar, cache, precision="f64")
    bnd = torch.empty(N_ROWS, dtype=torch.float32, device="cuda")
    llm = D.lnl(dstar, cache, precision="mixed", bound=bnd)
    d = (llf - llm).abs()
    # the bound holds on every row
    assert bool((d <= bnd.double()).all()), float((d / bnd.double()).max())
    for seed in (5, 6, 7):
        st = D.pcg64_state(np.random.default_rng(seed))
        m2 = llm.clone()
        b2 = bnd.clone()
        got = TheJoker._settle_reject(dstar, cache, m2, b2, st, 256)
        want, n_want, _ = D.reject(llf, st, 256)
        assert np.array_equal(got, want) and n_want > 0
    # the tolerance gate for arrays handed to a caller
    TheJoker._settle_tol(dstar, cache, llm, bnd)
    assert float((llf - llm).abs().max()) <= 1e-3
