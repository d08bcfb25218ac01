"""The natural logarithm of the likelihood finish (jk_log_tab in
jk_device.h), emulated in numpy: x = m 2^k with m in [1, 2), a 128-entry
table of R_i = 1 / (1 + (i + 1/2) / 128) and -ln R_i indexed by m's top seven
fraction bits, r = m R_i - 1 by one fused multiply-add, ln(1 + r) to r^6, and
k ln 2.  Checked against a long-double log: absolute error below 1e-15 plus
the rounding of k ln 2.  Pins the method and its constants, not the device
code (the GPU parity tests against the oracle do that)."""
import os
import re

import numpy as np

LD = np.longdouble
SRC = open(os.path.join(os.path.dirname(__file__), "..", "thejoker_amd", "csrc",
                        "jk_device.h")).read()
N = int(re.search(r"#define JK_LOG_TAB_N (\d+)", SRC).group(1))
LN2 = float(re.search(r"#define JK_LN2 (\S+)", SRC).group(1))
R = (1.0 / (1.0 + (np.arange(N, dtype=LD) + 0.5) / N)).astype(np.float64)
NL = (-np.log(R.astype(LD))).astype(np.float64)


def fma(a, b, c):
    """a * b + c rounded once (the long-double product of two doubles is
    exact to 64 bits, enough for these operands)."""
    return (np.asarray(a, dtype=LD) * np.asarray(b, dtype=LD) + np.asarray(c, dtype=LD)).astype(
        np.float64)


def jk_log_tab(x, order=6, bits=7):
    m, k = np.frexp(x)
    m = 2.0 * m
    i = (m.view(np.uint64) >> np.uint64(45)).astype(np.int64) & ((1 << bits) - 1)
    i = i * (N >> bits) + (N >> bits) // 2 if bits < 7 else i
    r = fma(m, R[i], -1.0)
    coef = {6: [-1.0 / 6.0, 0.2, -0.25, 1.0 / 3.0, -0.5], 4: [-0.25, 1.0 / 3.0, -0.5]}[order]
    q = np.full_like(r, coef[0])
    for c in coef[1:]:
        q = fma(r, q, np.full_like(r, c))
    return fma((k - 1).astype(np.float64), np.full_like(r, LN2), NL[i] + fma(r * r, q, r))


def samples():
    rng = np.random.default_rng(2)
    return np.concatenate([
        np.exp(rng.uniform(-700, 700, 200_000)),
        rng.uniform(0.5, 2.0, 200_000),
        1.0 + rng.uniform(-1e-6, 1e-6, 10_000),
        np.ldexp(1.0, np.arange(-1074, 1024)),
        np.nextafter(np.ldexp(1.0, np.arange(-1000, 1000)), 0),
        [1.0, 2.0, 0.5, 5e-324, 2.2250738585072014e-308],
    ])


def err(got, x):
    ref = np.log(x.astype(LD))
    k = np.frexp(x)[1] - 1
    floor = 1e-15 + np.abs(k) * 1.2e-16       # the table and series; k ln 2 rounded once
    return (np.abs(got.astype(LD) - ref).astype(float) / floor).max()


def test_ln2_and_table():
    assert abs(LD(LN2) - np.log(LD(2))) < 1e-16
    assert N == 128
    m = 1.0 + (np.arange(N) + 0.5) / N
    assert np.abs(m * R - 1.0).max() < 1e-15


def test_log_tab_absolute_error():
    x = samples()
    assert err(jk_log_tab(x), x) < 1.0
    assert abs(jk_log_tab(np.array([1.0]))[0]) < 1e-17     # ln 1 to the truncation


def test_log_tab_method_has_teeth():
    """Plausible slips fail the bound: the series stopped at r^4, or a
    table of 16 intervals indexed by m's top four bits."""
    x = samples()
    assert err(jk_log_tab(x, order=4), x) > 1.0
    assert err(jk_log_tab(x, bits=4), x) > 1.0
