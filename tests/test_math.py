"""The shared transcendental library (mfc_math.h) against the platform libm and the IEEE 754
special cases. The device copy is checked bit for bit against this host copy in
test_gpu_parity.py::test_device_math_bitwise."""
import math

import numpy as np
import pytest

import oracle_binding as ob

EXP, LOG, POW, ATAN2, SIN, COS, ATAN2F, COSF, SINF, SQRT, SQRTF, DIV, TANH = range(13)


def ulps(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    ia = np.atleast_1d(a).view(np.int64)
    ib = np.atleast_1d(b).view(np.int64)
    d = np.abs(ia - ib)  # exact in int64 for same-sign operands
    d = np.where(np.signbit(np.atleast_1d(a)) != np.signbit(np.atleast_1d(b)), np.int64(1) << 62, d)
    d = np.where((np.atleast_1d(a) == np.atleast_1d(b)) | (np.isnan(np.atleast_1d(a)) & np.isnan(np.atleast_1d(b))), 0, d)
    return d


@pytest.mark.parametrize("fn,ref,gen,bound", [
    (EXP, np.exp, lambda r: r.uniform(-740, 709, 300000), 1),
    (LOG, np.log, lambda r: np.exp(r.uniform(-700, 700, 300000)), 1),
    (SIN, np.sin, lambda r: r.uniform(-50, 50, 300000), 1),
    (COS, np.cos, lambda r: r.uniform(-50, 50, 300000), 1),
    (TANH, np.tanh, lambda r: r.uniform(-25, 25, 300000), 3),
])
def test_unary_ulp(fn, ref, gen, bound):
    x = gen(np.random.default_rng(fn))
    assert ulps(ob.host_math(fn, x), ref(x)).max() <= bound


def test_pow_ulp_on_johnson_cook_ranges():
    r = np.random.default_rng(3)
    base = np.concatenate([np.exp(r.uniform(-20, 1, 200000)), 1 + np.exp(r.uniform(-30, 5, 100000))])
    expo = r.uniform(0.0, 3.0, base.size)
    assert ulps(ob.host_math(POW, base, expo), np.power(base, expo)).max() <= 2


def test_atan2_ulp():
    r = np.random.default_rng(4)
    y = r.uniform(-1, 1, 300000) * np.exp(r.uniform(-20, 20, 300000))
    x = r.uniform(-1, 1, 300000) * np.exp(r.uniform(-20, 20, 300000))
    assert ulps(ob.host_math(ATAN2, y, x), np.arctan2(y, x)).max() <= 3


def test_float_variants_are_rounded_double_results():
    r = np.random.default_rng(5)
    y = r.uniform(-3, 3, 100000)
    x = r.uniform(-3, 3, 100000)
    f = ob.host_math(ATAN2F, y, x)
    ref = np.arctan2(y.astype(np.float32).astype(np.float64), x.astype(np.float32).astype(np.float64))
    assert np.all(f == f.astype(np.float32).astype(np.float64))  # representable in float
    assert np.max(np.abs(f - ref) / np.abs(ref)) < 1.2e-7
    t = r.uniform(-np.pi / 2, np.pi / 2, 100000)
    c, s = ob.host_math(COSF, t), ob.host_math(SINF, t)
    tf = t.astype(np.float32).astype(np.float64)
    assert np.max(np.abs(c - np.cos(tf))) < 1.2e-7 and np.max(np.abs(s - np.sin(tf))) < 1.2e-7


def test_pow_special_cases():
    cases = [(0.0, 1.1, 0.0), (0.0, 0.34, 0.0), (1.0, np.nan, 1.0), (np.nan, 0.0, 1.0), (2.0, 0.0, 1.0),
             (0.0, -1.0, np.inf), (-8.0, 1.0 / 3.0, np.nan), (-2.0, 3.0, -8.0), (-2.0, 2.0, 4.0),
             (np.inf, 0.5, np.inf), (0.5, np.inf, 0.0), (2.0, -np.inf, 0.0), (-1.0, np.inf, 1.0)]
    x = np.array([c[0] for c in cases])
    y = np.array([c[1] for c in cases])
    out = ob.host_math(POW, x, y)
    want = np.array([c[2] for c in cases])
    assert np.array_equal(out, want, equal_nan=True), (out, want)


def test_atan2_special_cases():
    pz, nz = 0.0, -0.0
    y = np.array([pz, nz, pz, nz, 1.0, -1.0, np.inf, np.inf, -np.inf, 1.0, 1.0])
    x = np.array([pz, pz, nz, nz, 0.0, 0.0, np.inf, -np.inf, 1.0, np.inf, -np.inf])
    out = ob.host_math(ATAN2, y, x)
    want = np.arctan2(y, x)
    assert np.array_equal(out, want) and np.array_equal(np.signbit(out), np.signbit(want))


def test_log_exp_edges():
    assert ob.host_math(LOG, np.array([0.0]))[0] == -np.inf
    assert np.isnan(ob.host_math(LOG, np.array([-1.0]))[0])
    assert ob.host_math(LOG, np.array([1.0]))[0] == 0.0
    assert ob.host_math(EXP, np.array([0.0]))[0] == 1.0
    assert ob.host_math(EXP, np.array([1000.0]))[0] == np.inf
    assert ob.host_math(EXP, np.array([-1000.0]))[0] == 0.0
    sub = ob.host_math(EXP, np.array([-740.0]))[0]
    assert sub > 0 and ulps(sub, math.exp(-740.0)) <= 1


def test_sqrtf_is_correctly_rounded_float():
    x = np.random.default_rng(6).uniform(0, 1e-4, 100000)
    want = np.sqrt(x.astype(np.float32)).astype(np.float64)
    assert np.array_equal(ob.host_math(SQRTF, x), want)
