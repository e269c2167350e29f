"""The shared scalar transcendental functions (tinytrace_amd/csrc/tt_math.h),
pinned against correctly rounded double-precision libm results."""
import numpy as np
import pytest

from oracle import binding


def ulp_error(got, exact):
    got = got.astype(np.float64)
    spacing = np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
    return np.abs(got - exact) / spacing


@pytest.mark.parametrize("name,fn,lo,hi,max_ulp", [
    ("sin", np.sin, -8.0, 8.0, 4.0),
    ("cos", np.cos, -8.0, 8.0, 4.0),
    ("exp", np.exp, -80.0, 80.0, 3.0),
    ("acos", np.arccos, -1.0, 1.0, 4.0),
    ("asin", np.arcsin, -1.0, 1.0, 4.0),
])
def test_against_libm(name, fn, lo, hi, max_ulp):
    x = np.random.default_rng(0).uniform(lo, hi, 200000).astype(np.float32)
    got = binding.math_f32(name, x)
    exact = fn(x.astype(np.float64))
    if name in ("sin", "cos"):
        # absolute error near zeros of sin/cos, relative elsewhere
        err = np.abs(got - exact)
        assert np.all(err <= np.maximum(max_ulp * np.spacing(np.abs(exact).astype(np.float32)), 2e-7))
    else:
        assert ulp_error(got, exact).max() <= max_ulp


def test_log_against_libm():
    rng = np.random.default_rng(1)
    x = np.exp(rng.uniform(-80, 80, 200000)).astype(np.float32)
    got = binding.math_f32("log", x)
    exact = np.log(x.astype(np.float64))
    err = np.abs(got - exact)
    assert np.all(err <= np.maximum(3 * np.spacing(np.abs(exact).astype(np.float32)), 1e-7))


def test_pow_small_integral_exponents():
    x = np.random.default_rng(2).uniform(0, 1, 10000).astype(np.float32)
    got = binding.math_f32("pow", x, np.full_like(x, 3.0))
    np.testing.assert_allclose(got, x.astype(np.float64) ** 3, rtol=2e-6, atol=1e-30)


def test_special_values():
    nan, inf = np.float32(np.nan), np.float32(np.inf)
    assert np.isnan(binding.math_f32("sin", np.array([nan, inf], np.float32))).all()
    assert binding.math_f32("exp", np.array([-200.0], np.float32))[0] == 0.0
    assert np.isinf(binding.math_f32("exp", np.array([100.0], np.float32))[0])
    assert np.isneginf(binding.math_f32("log", np.array([0.0], np.float32))[0])
    assert np.isnan(binding.math_f32("log", np.array([-1.0], np.float32))[0])
    assert binding.math_f32("cos", np.array([0.0], np.float32))[0] == 1.0
    assert binding.math_f32("acos", np.array([1.0], np.float32))[0] == 0.0
    # subnormal input
    assert abs(binding.math_f32("log", np.array([1e-40], np.float32))[0] - np.log(1e-40)) < 1e-5


@pytest.mark.parametrize("name,fn", [("sinpi", np.sin), ("cospi", np.cos)])
def test_sincospi_exact_reduction(name, fn):
    """sin(pi v) / cos(pi v) with the argument reduced exactly: within
    2.5e-7 of the double-precision value up to v = 4096 (the position
    encoding's 2^11 pi u), where sin(pi * v) of the rounded product pi v
    would be off by up to 2^-11 pi; exact at the integers and half-integers
    (zeros and +-1)."""
    rng = np.random.default_rng(5)
    v = np.concatenate([rng.uniform(-2, 2, 100000), rng.uniform(0, 4096, 100000),
                        np.ldexp(rng.uniform(0, 1, 1000), 11)]).astype(np.float32)
    got = binding.math_f32(name, v).astype(np.float64)
    exact = fn(np.pi * v.astype(np.float64))
    assert np.abs(got - exact).max() < 2.5e-7
    k = np.arange(-64, 65, dtype=np.float32) / 2
    got = binding.math_f32(name, k)
    want = np.round(fn(np.pi * k.astype(np.float64)))
    np.testing.assert_array_equal(got, want.astype(np.float32))
    # A planted slip: reducing pi v by the rounded product misses by far more.
    coarse = fn(np.float32(np.pi) * v.astype(np.float32))
    assert np.abs(coarse - fn(np.pi * v.astype(np.float64))).max() > 1e-4
