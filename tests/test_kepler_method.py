"""The non-iterative Kepler method of the gfx950 kernels, emulated on the CPU
in numpy (FP32 starter rounded through numpy float32, FP64 table sincos and
one-reciprocal series correction) against a long-double Newton solution: the
error stays within a small multiple of the conditioning ulp(M) / (1 - e cos E)
for e up to 1 - 1e-6 and M down to 1e-12, as a converged Newton iteration's
does.  Pins the method, not the device code (the GPU parity tests do that)."""
import numpy as np
import pytest

f32 = np.float32
LD = np.longdouble
A0 = f32(7.6516382901912925)     # 3 pi^2 / (pi^2 - 6)
A1N = f32(1.29898246041084)      # 1.6 pi / (pi^2 - 6)
PIF = f32(np.pi)
NT = 1024
_m, _ex = np.frexp(np.pi / NT)
H = float(np.ldexp(np.ceil(_m * 2 ** 42), _ex - 42))
TS = np.sin(np.arange(NT + 1) * LD(H)).astype(np.float64)
TC = np.cos(np.arange(NT + 1) * LD(H)).astype(np.float64)


def starter(M, e):
    """Markley's cubic starter in FP32, one reciprocal (as the kernel)."""
    ome = f32(1) - e
    a1 = A1N / (f32(1) + e)
    alpha = a1 * (PIF - M) + A0
    d = alpha * e + f32(3) * ome
    M2 = M * M
    ad = alpha * d
    q = f32(2) * ad * ome - M2
    r = M * (f32(3) * ad * (d - ome) + M2)
    q2 = q * q
    x = np.abs(r) + np.sqrt(np.maximum(q2 * q + r * r, f32(0)))
    w = np.exp2(np.log2(x) * f32(2.0 / 3.0))
    den = w * (w + q) + q2
    return ((f32(2) * r * w + M * den) / (den * d)).astype(f32)


def table_sincos(Ef):
    idx = np.clip((Ef * f32(NT / np.pi)).astype(np.int32), 0, NT)
    d = Ef.astype(np.float64) - idx * H
    d2 = d * d
    sd = d + d * d2 * (d2 / 120 - 1 / 6)
    cd = 1 + d2 * (-0.5 + d2 / 24)
    return TS[idx] * cd + TC[idx] * sd, TC[idx] * cd - TS[idx] * sd


def solve(M, e):
    """E, sin E, cos E for M in [0, pi] (the kernel's sequence of steps)."""
    E1f = starter(M.astype(f32), np.minimum(e.astype(f32), f32(0.99999994)))
    E1 = E1f.astype(np.float64)
    s1, c1 = table_sincos(E1f)
    # fma(-e, s1, E1 - M) and fma(-e, c1, 1): the exact product and sum
    # rounded once (long double holds the product exactly enough)
    f0 = ((E1 - M).astype(LD) - e.astype(LD) * s1.astype(LD)).astype(np.float64)
    r = 1.0 / (1.0 - e.astype(LD) * c1.astype(LD)).astype(np.float64)
    t = -f0 * r
    b2 = (0.5 * e) * s1 * r
    b3 = (r - 1.0) * (1 / 6)
    b22 = b2 * b2
    c3 = 2 * b22 - b3
    c4 = b2 * (5 * (b3 - b22) + 1 / 12)
    d = t * (1 + t * (-b2 + t * (c3 + t * c4)))
    d2 = d * d
    cd = 1 + d2 * (-0.5 + d2 / 24)
    sd = d * (1 - d2 / 6)
    return E1 + d, s1 * cd + c1 * sd, c1 * cd - s1 * sd, d


def newton_ld(M, e, E):
    E = E.astype(LD)
    Ml, el = M.astype(LD), e.astype(LD)
    for _ in range(6):
        E = E - (E - el * np.sin(E) - Ml) / (1 - el * np.cos(E))
    return E


@pytest.mark.parametrize("e0", [0.0, 0.3, 0.7, 0.9, 0.99, 0.999, 0.9999, 0.99999, 1 - 1e-6])
def test_series_correction_reaches_conditioning(e0):
    rng = np.random.default_rng(int(e0 * 1e6))
    k = 40_000
    M = np.concatenate([np.exp(rng.uniform(np.log(1e-12), np.log(np.pi), k)),
                        rng.uniform(0, np.pi, k), [0.0, np.pi]])
    e = np.full(M.shape, e0)
    E, s, c, d = solve(M, e)
    T = newton_ld(M, e, E)
    eps = np.finfo(float).eps
    cond = np.maximum(eps * (np.abs(T) + e * np.abs(np.sin(T)) + M) / (1 - e * np.cos(T)), 1e-300)
    assert np.all(np.abs(d) < 1e-3)                      # no fallback needed
    assert (np.abs((E - T).astype(float)) / cond).max() < 2.5
    assert np.abs((s - np.sin(T)).astype(float)).max() < 4 * eps + 2.5 * cond.max()
    assert np.abs((c - np.cos(T)).astype(float)).max() < 4 * eps + 2.5 * cond.max()


def test_grid_constants_are_the_kernels():
    """The emulation's grid is the one the kernel source defines: N
    intervals, H = pi / N rounded up to 42 significant bits (i H exact in FP64
    for i <= N, N H >= pi), the FP32 step and the FP32 index scale."""
    import os
    import re
    src = open(os.path.join(os.path.dirname(__file__), "..", "thejoker_amd", "csrc",
                            "jk_kernels.hip")).read()
    define = lambda name: re.search(r"#define %s (\S+)" % name, src).group(1)
    assert int(define("JK_TAB_N")) == NT
    assert float.fromhex(define("JK_TAB_H")) == H
    assert float.fromhex(define("JK_TAB_H32").rstrip("f")) == _H32
    assert f32(define("JK_TAB_INV_H_F").rstrip("f")) == f32(NT / np.pi)
    assert NT * H >= np.pi
    from fractions import Fraction
    assert all(Fraction(i) * Fraction(H) == Fraction(i * H) for i in range(NT + 1))


def test_table_sincos_accuracy():
    Ef = np.random.default_rng(0).uniform(0, np.pi, 200_000).astype(f32)
    s, c = table_sincos(Ef)
    x = Ef.astype(LD)
    assert np.abs((s - np.sin(x)).astype(float)).max() < 2.5e-16
    assert np.abs((c - np.cos(x)).astype(float)).max() < 2.5e-16


def test_truncated_series_would_fail():
    """Dropping the t^4 term of the reversion (a plausible slip) breaks the
    bound at e = 0.99, so the test above can see such a mistake."""
    rng = np.random.default_rng(5)
    M = np.concatenate([rng.uniform(0, np.pi, 50_000),
                        np.exp(rng.uniform(np.log(1e-9), 0.0, 50_000))])
    e = np.full(M.shape, 0.99)
    E1f = starter(M.astype(f32), e.astype(f32))
    E1 = E1f.astype(np.float64)
    s1, c1 = table_sincos(E1f)
    es, ec = e * s1, e * c1
    t = -((E1 - M) - es) / (1 - ec)
    b2 = 0.5 * es / (1 - ec)
    b3 = ec / (1 - ec) / 6
    E = E1 + t * (1 + t * (-b2 + t * (2 * b2 * b2 - b3)))
    T = newton_ld(M, e, E)
    cond = np.finfo(float).eps * (np.abs(T) + e * np.abs(np.sin(T)) + M) / (1 - e * np.cos(T))
    assert (np.abs((E - T).astype(float)) / cond).max() > 10


_H32 = np.float64(np.float32(np.pi / NT))
_G32 = np.arange(NT + 1, dtype=LD) * LD(_H32)
_TS32 = np.sin(_G32).astype(f32)
_TC32 = np.cos(_G32).astype(f32)


def _sincos_tab_f32(x):
    """The mixed kernel's FP32 sin/cos on [0, pi]: the FP32 grid i * H32
    (H32 = pi / 1024 rounded to FP32, values rounded once from long double),
    d = x - i H32 rounded once (fma), sin d to d^3, cos d to d^2."""
    E = x.astype(f32)
    idx = np.clip((E * f32(NT / np.pi)).astype(np.int32), 0, NT)
    d = (E.astype(np.float64) - idx * _H32).astype(f32)
    d2 = d * d
    sd = (d.astype(np.float64) * d2 * (-1 / 6) + d).astype(f32)
    cd = (d2.astype(np.float64) * -0.5 + 1).astype(f32)
    ts, tc = _TS32[idx], _TC32[idx]
    s = (ts.astype(np.float64) * cd + (tc * sd)).astype(f32)
    c = (tc.astype(np.float64) * cd - (ts * sd)).astype(f32)
    return s, c


def test_fp32_table_sincos_accuracy():
    E = np.random.default_rng(3).uniform(0, np.pi, 300_000).astype(f32)
    s, c = _sincos_tab_f32(E)
    El = E.astype(LD)
    assert np.abs((s - np.sin(El)).astype(float)).max() < 1.2e-7
    assert np.abs((c - np.cos(El)).astype(float)).max() < 1.2e-7


def _fmaf(a, b, c):
    """FP32 fused multiply-add: the exact product and sum rounded once."""
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(f32)


def _mixed_design(M, e, om, newton=True):
    """The mixed row's design value m: FP32 starter, FP32 table sin/cos, one
    first-order Newton step in FP32 (fused multiply-adds rounded once),
    1 / (1 - e cos E) in FP32, and the combination
    m = fma(fma(cos w, cos E - e, -(sin w sqrt(1-e^2)) sin E), den, e cos w)
    in FP32 from FP32 copies of the per-sample constants."""
    ef = np.minimum(e.astype(f32), f32(0.99999994))
    E1 = starter(M.astype(f32), ef)
    s1, c1 = _sincos_tab_f32(E1)
    if newton:
        f0 = (E1.astype(np.float64) - M.astype(f32)).astype(f32)
        f0 = (f0 - ef.astype(np.float64) * s1).astype(f32)
        f1 = (1.0 - ef.astype(np.float64) * c1).astype(f32)
        t = (-f0 * (f32(1) / f1)).astype(f32)
        sE = (c1.astype(np.float64) * t + s1).astype(f32)
        cE = (-s1.astype(np.float64) * t + c1).astype(f32)
    else:
        sE, cE = s1, c1
    den = (f32(1) / _fmaf(-ef, cE, f32(1))).astype(f32)
    cwf = np.cos(om).astype(f32)
    swsqf = (np.sin(om) * np.sqrt((1 - e) * (1 + e))).astype(f32)
    ecwf = (e * np.cos(om)).astype(f32)
    inner = _fmaf(cwf, (cE - ef).astype(f32), (-swsqf * sE).astype(f32))
    m = _fmaf(inner, den, ecwf)
    return m.astype(np.float64), den.astype(np.float64)


@pytest.mark.parametrize("dist", ["beta", 0.9, 0.999, 0.99999, 1 - 1e-6])
def test_mixed_design_error_within_bound_model(dist):
    """The per-epoch error model of the mixed error bound,
    eps = 2^-22 (1 + d)^2 (1 + d / 1000) with d = 1 / (1 - e cos E), covers
    the FP32 chain's design-column error with a factor 2 to spare; without
    the Newton step the starter's own error breaks it for e up to 0.999 (the
    check has teeth; nearer 1 the model's (1 + d)^2 grows faster than the
    starter's error)."""
    rng = np.random.default_rng(11)
    n = 300_000
    e = rng.beta(0.867, 3.03, n) if dist == "beta" else np.full(n, dist)
    M = np.concatenate([rng.uniform(0, np.pi, n // 2), np.exp(rng.uniform(np.log(1e-9), 0, n - n // 2))])
    om = rng.uniform(0, 2 * np.pi, n)
    _, s64, c64, _ = solve(M, e)
    den64 = 1 / (1 - e * c64)
    m64 = (np.cos(om) * (c64 - e) - np.sin(om) * np.sqrt(1 - e * e) * s64) * den64 + e * np.cos(om)
    m, den = _mixed_design(M, e, om)
    model = 2.0 ** -22 * (1 + den) ** 2 * (1 + den / 1000)
    assert np.max(np.abs(m - m64) / model) < 0.5
    if dist in ("beta", 0.9, 0.999):
        m0, den0 = _mixed_design(M, e, om, newton=False)
        model0 = 2.0 ** -22 * (1 + den0) ** 2 * (1 + den0 / 1000)
        assert np.max(np.abs(m0 - m64) / model0) > 1.0
