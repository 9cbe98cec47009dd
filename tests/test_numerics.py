"""Floating-point identities the kernels rely on, checked on the CPU.

The sampler stages phi_k = fl(x / den_k) (x = fl(fl(n_wk) + beta)) as the
FP32 rounding of the FP64 product x * RN64(1 / den_k).  For FP32 x, d the
quotient x / d is never an FP32 rounding midpoint and lies at least 2^-49
(relative) from every midpoint, while the FP64 product is within 2^-52 of
it, so both round to the same FP32 number.  Checked here against IEEE FP32
division (numpy) on random operands over the ranges counts and
denominators take, on exact quotients, and on quotients placed next to
midpoints.
"""
import numpy as np


def recip_quotient(x, d):
    r = 1.0 / d.astype(np.float64)
    return (x.astype(np.float64) * r).astype(np.float32)


def test_reciprocal_product_equals_fp32_division_random():
    rng = np.random.default_rng(7)
    n = 10_000_000
    # x = fl(count + beta), count 0 .. 2^24; d = fl(n_k + beta V)
    cnt = rng.integers(0, 1 << 24, n).astype(np.float32)
    beta = np.float32(0.01)
    x = (cnt + beta).astype(np.float32)
    d = (rng.integers(0, 1 << 30, n).astype(np.float32) + np.float32(1000.0)).astype(np.float32)
    q = (x / d).astype(np.float32)
    assert np.array_equal(recip_quotient(x, d), q)
    # arbitrary FP32 significands and exponents
    x = np.exp2(rng.uniform(-20, 30, n)).astype(np.float32)
    d = np.exp2(rng.uniform(-10, 40, n)).astype(np.float32)
    assert np.array_equal(recip_quotient(x, d), (x / d).astype(np.float32))


def test_reciprocal_product_near_midpoints_and_exact_quotients():
    rng = np.random.default_rng(11)
    n = 2_000_000
    d = (rng.integers(1 << 23, 1 << 24, n)).astype(np.float32)
    # x = d * m for m at a midpoint (25-bit odd significand), rounded to FP32:
    # quotients as close to a midpoint as FP32 operands allow
    m = (rng.integers(1 << 24, 1 << 25, n) | 1).astype(np.float64) / float(1 << 24)
    x = (d.astype(np.float64) * m).astype(np.float32)
    assert np.array_equal(recip_quotient(x, d), (x / d).astype(np.float32))
    # exact quotients
    q = rng.integers(1, 1 << 12, n).astype(np.float32)
    dd = rng.integers(1, 1 << 12, n).astype(np.float32)
    x = (q * dd).astype(np.float32)
    assert np.array_equal(recip_quotient(x, dd), q)
