"""The helicity chain's FP64 log and exp kernels (log_core_f64 and
exp_core_f64 in phys.h), pinned on the host from their source text:

* the constants by what the mathematics fixes: the log's polynomial P
  against R(z)/z = sum 2 z^(j-1) / (2j + 1) and the exp's Q against
  (e^r - 1 - r) / r^2, in 50-digit arithmetic over the reduced ranges (the
  fits' errors below 5e-18 in ln m and 4e-19 in e^r), LN2_HI + LN2_LO = ln 2
  with k LN2_HI exact for every binary exponent, 1/ln 2 and the split point
  sqrt(1/2) correctly rounded;
* the steps by running them in IEEE double on the host (Python floats, fma
  exact through rationals, 1/b correctly rounded where the device starts
  from its reciprocal instruction) against ln x and e^x to 50 digits: within
  an ulp, with the reductions' edges included.

The device's own results are checked against the host's long double log in
the GPU tests (test_gpu_math.py).
"""
import math
import os
import re
from fractions import Fraction

import numpy as np
import pytest

mp = pytest.importorskip('mpmath')

PHYS = os.path.join(os.path.dirname(__file__), '..', 'xcape_amd', 'csrc', 'phys.h')


def _constants():
    src = open(PHYS).read()
    body = src[src.index('double log_core_f64('):]
    body = body[:body.index('\n}\n')]
    vals = dict((k, float.fromhex(v)) for k, v in
                re.findall(r'\b(LN2_HI|LN2_LO|C[1-7])\s*=\s*(0x[0-9a-fA-Fp.+-]+)', body))
    split = float.fromhex(re.search(r'if \(m < (0x[0-9a-fA-Fp.+-]+)\)', body).group(1))
    assert len(vals) == 9, sorted(vals)
    return vals, split


def _fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _log_core(x, K, split):
    """log_core_f64's steps for a positive finite double x."""
    m, k = math.frexp(x)
    if m < split:
        m = m + m
        k = k - 1
    f = m - 1.0
    b = 2.0 + f
    y = 1.0 / b
    y = _fma(_fma(-b, y, 1.0), y, y)
    y = _fma(_fma(-b, y, 1.0), y, y)
    s = f * y
    z = s * s
    p = _fma(z, K['C7'], K['C6'])
    for c in ('C5', 'C4', 'C3', 'C2', 'C1'):
        p = _fma(z, p, K[c])
    r = z * p
    hfsq = 0.5 * f * f
    dk = float(k)
    return dk * K['LN2_HI'] - ((hfsq - (s * (hfsq + r) + dk * K['LN2_LO'])) - f)


def test_log_kernel_constants():
    K, split = _constants()
    mp.mp.dps = 50
    assert split == float(mp.sqrt(mp.mpf(1) / 2))
    ln2 = mp.log(2)
    assert abs(mp.mpf(K['LN2_HI']) + mp.mpf(K['LN2_LO']) - ln2) < mp.mpf(2) ** -80
    # k LN2_HI exact for |k| <= 1075: LN2_HI has at most 42 significant bits
    mant = Fraction(K['LN2_HI']) * 2 ** 42
    assert mant.denominator == 1
    smax = (mp.sqrt(2) - 1) / (mp.sqrt(2) + 1)
    worst = mp.mpf(0)
    for i in range(1, 801):
        s = smax * i / 800
        z = s * s
        g = (mp.log((1 + s) / (1 - s)) - 2 * s) / (s * z)
        p = mp.mpf(0)
        for c in ('C7', 'C6', 'C5', 'C4', 'C3', 'C2', 'C1'):
            p = p * z + mp.mpf(K[c])
        worst = max(worst, abs(s * z * (p - g)) / mp.log((1 + s) / (1 - s)))
    assert worst < 5e-18, worst


def test_log_kernel_steps_within_an_ulp():
    K, split = _constants()
    mp.mp.dps = 50
    rng = np.random.default_rng(23)
    xs = list(10.0 ** rng.uniform(-300, 300, 1500))
    xs += list(rng.uniform(0.85, 1.2, 1500)) + list(rng.uniform(150, 350, 500))
    for e in rng.integers(-1000, 1000, 300):
        for edge in (math.sqrt(0.5), 1.0, math.sqrt(2.0)):
            for j in (-3, -1, 0, 1, 3):
                xs.append(math.ldexp(edge * (1.0 + j * 2.0 ** -52), int(e)))
    xs += [1.0, 2.0, 0.5, 5e-324, 2.2250738585072014e-308, 1.7976931348623157e308]
    worst = 0.0
    for x in xs:
        got = _log_core(x, K, split)
        exact = mp.log(mp.mpf(x))
        ulp = math.ulp(float(exact)) if exact != 0 else 5e-324
        err = float(abs(mp.mpf(got) - exact)) / ulp
        worst = max(worst, err)
    print(f'log kernel steps: max {worst:.3f} ulp over {len(xs)} arguments')
    assert worst < 1.0


def test_a_wrong_constant_is_caught():
    """The steps' check fails when a coefficient loses its last digits."""
    K, split = _constants()
    mp.mp.dps = 50
    bad = dict(K, C2=float(np.float32(K['C2'])))
    x = 1.3
    err = abs(mp.mpf(_log_core(x, bad, split)) - mp.log(mp.mpf(x))) / math.ulp(math.log(x))
    assert err > 1.0


def _exp_constants():
    src = open(PHYS).read()
    body = src[src.index('double exp_core_f64('):]
    body = body[:body.index('\n}\n')]
    vals = dict((k, float.fromhex(v)) for k, v in
                re.findall(r'\b(INV_LN2|LN2_HI|LN2_LO|E10|E[0-9])\s*=\s*(0x[0-9a-fA-Fp.+-]+)', body))
    assert len(vals) == 14, sorted(vals)
    return vals


def _exp_core(x, K):
    """exp_core_f64's steps for -745.2 < x < 709.8."""
    k = float(round(x * K['INV_LN2']))         # Python rounds half to even, as rint
    hi = _fma(-k, K['LN2_HI'], x)
    assert Fraction(hi) == Fraction(x) - Fraction(k) * Fraction(K['LN2_HI'])   # exact
    lo = k * K['LN2_LO']
    r = hi - lo
    c = (hi - r) - lo
    q = _fma(r, K['E10'], K['E9'])
    for e in ('E8', 'E7', 'E6', 'E5', 'E4', 'E3', 'E2', 'E1', 'E0'):
        q = _fma(r, q, K[e])
    p = 1.0 + (r + _fma(r * r, q, c))
    return math.ldexp(p, int(k))


def test_exp_kernel_constants():
    K = _exp_constants()
    mp.mp.dps = 50
    ln2 = mp.log(2)
    assert K['INV_LN2'] == float(1 / ln2)
    assert K['LN2_HI'] == _constants()[0]['LN2_HI'] and K['LN2_LO'] == _constants()[0]['LN2_LO']
    half = ln2 / 2
    worst = mp.mpf(0)
    for i in range(-800, 801):
        r = half * i / 800
        q = mp.mpf(0)
        for c in ('E10', 'E9', 'E8', 'E7', 'E6', 'E5', 'E4', 'E3', 'E2', 'E1', 'E0'):
            q = q * r + mp.mpf(K[c])
        worst = max(worst, abs(1 + r + r * r * q - mp.exp(r)) / mp.exp(r))
    assert worst < 4e-19, worst


def test_exp_kernel_steps_within_an_ulp():
    K = _exp_constants()
    mp.mp.dps = 50
    rng = np.random.default_rng(29)
    xs = list(rng.uniform(-708.0, 709.0, 4000)) + list(rng.uniform(-30.0, 0.0, 4000))
    xs += list(rng.uniform(-1.0, 1.0, 1000))
    for j in rng.integers(-1000, 1000, 300):
        for d in (-3, -1, 0, 1, 3):
            xs.append((float(j) + 0.5) * math.log(2.0) * (1.0 + d * 2.0 ** -52))
    xs += [0.0, 1.0, -1.0, 709.7, -708.3]
    worst = 0.0
    for x in xs:
        got = _exp_core(x, K)
        exact = mp.exp(mp.mpf(x))
        err = float(abs(mp.mpf(got) - exact)) / math.ulp(float(exact))
        worst = max(worst, err)
    print(f'exp kernel steps: max {worst:.3f} ulp over {len(xs)} arguments')
    assert worst < 1.0

