"""Device exp / log / constant-divisor division against the CPU, bit for bit.

The engine's FP64 form is bit-exact with the oracle only if its exp and log
equal the oracle's and its constant-divisor division equals IEEE division.
The divisors checked are every literal divisor the kernels use (parsed from
the sources), on random operands over the whole exponent range plus zeros,
subnormals, infinities and NaN (which take the IEEE fallback).
"""
import os
import re

import numpy as np
import pytest

from oracle_lib import ROOT, oracle

pytestmark = pytest.mark.gpu


def _divisors():
    src = ""
    for f in ("hvc_neuron.hip.h", "hvc_kernels.hip"):
        src += open(os.path.join(ROOT, "unsupervised_growth_amd", "csrc", f)).read()
    vals = set(float(x) for x in re.findall(r"\bCD\([^;]*?,\s*(?:HR\()?([0-9.]+)\)", src))
    vals |= set(float(x) for x in re.findall(r"HVC_CDIV\([^;]*?,\s*([0-9.]+)\)", src))
    vals |= set(float(x) for x in re.findall(r"\bCK\(-?\d+,\s*[\w.]+,\s*(\d+)\)", src))
    return sorted(vals)


def test_divisor_list_parsed():
    d = _divisors()
    assert 5000.0 in d and 614563906.0 in d and 10000.0 in d and len(d) > 50


def _engine():
    from unsupervised_growth_amd.engine import Engine
    return Engine()


def test_device_exp_log_equal_oracle():
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-60, 60, 400000), rng.uniform(-750, 715, 100000),
                         [0.0, -0.0, 1.0, -1.0, 709.78, 709.79, -745.1, -745.2, np.inf, -np.inf, 800.0, -900.0]])
    ys = np.concatenate([rng.uniform(0, 1, 200000), 10 ** rng.uniform(-308, 308, 50000),
                         [5e-324, 1e-310, 1.0, 0.5, 2.0, np.inf, 0.0]])
    lib = oracle()
    with _engine() as eng:
        de = eng.eval_math(0, xs)
        dl = eng.eval_math(1, ys)
    oe = np.array([lib.ora_exp(float(x)) for x in xs])
    ol = np.array([lib.ora_log(float(y)) for y in ys])
    np.testing.assert_array_equal(de, oe)
    np.testing.assert_array_equal(dl, ol)
    np.testing.assert_array_equal(np.signbit(de), np.signbit(oe))


def test_device_constant_division_is_ieee_division():
    rng = np.random.default_rng(4)
    divs = _divisors()
    n = 20000
    a_parts, b_parts = [], []
    for B in divs:
        mant = rng.uniform(1, 2, n) * rng.choice([-1.0, 1.0], n)
        a = np.ldexp(mant, rng.integers(-1020, 1020, n))
        a[:200] = np.round(rng.uniform(-1e6, 1e6, 200)) * B          # exact quotients
        a[200:210] = [0.0, -0.0, 5e-324, -1e-310, np.inf, -np.inf, np.nan, 1e308, -1e308, 2.2250738585072014e-308]
        a_parts.append(a)
        b_parts.append(np.full(n, B))
    a = np.concatenate(a_parts)
    b = np.concatenate(b_parts)
    with _engine() as eng:
        q = eng.eval_math(2, a, b)
    with np.errstate(all="ignore"):
        ref = a / b
    np.testing.assert_array_equal(q, ref)
    # the sign of a zero quotient too (assert_array_equal takes -0.0 == 0.0)
    np.testing.assert_array_equal(np.signbit(q), np.signbit(ref))
