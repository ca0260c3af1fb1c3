"""Exhaustive check of the kernels' division routines against IEEE division.

The density kernel divides by a fixed h*h (Smooth, kernels.inl:32-34) with a
Markstein-corrected product by 1/(h*h); P2G and the resampler divide a fixed h*h
by max(r2, 1e-5) (Sharpen, kernels.inl:36-38) with a reciprocal-based
Newton-Raphson sequence (P2G at the reference radius 1.4 with the hardware
reciprocal and one correction). All must return the IEEE quotient bit for bit, since
the stage outputs are compared bit-exactly with the oracle. Every float in each
range is tried on the device; the divisors are those of the dam-break and pool
workloads (density 0.5 at the listed grid sizes) plus random ones.
"""
import ctypes

import numpy as np
import pytest

from ariel_amd import _abi

pytestmark = pytest.mark.gpu

SIZES = [16, 32, 48, 64, 80, 96, 128, 256, 512]


def _check(which, a, lo, hi):
    lib = _abi.load()
    bad = ctypes.c_int64(-1)
    rc = lib.ariel_selftest_division(which, a, lo, hi, ctypes.byref(bad))
    assert rc == 0, lib.ariel_last_error(None).decode()
    return bad.value


def _f32(x):
    return float(np.float32(x))


def _fixed_divisors():
    out = []
    for n in SIZES:
        h = np.float32(np.float32(4.0) * np.float32(0.5) / np.float32(n))
        out.append(_f32(h * h))                      # ComputeDensity
        re = np.float32(np.float32(0.5) / np.float32(n))
        out.append(_f32(re * re))                    # Resample spring
    rng = np.random.default_rng(7)
    out += [_f32(x) for x in 10.0 ** rng.uniform(-7, 1, size=6)]
    return out


@pytest.mark.parametrize("d", _fixed_divisors())
def test_fixed_divisor_matches_ieee(d):
    # every float x in [0, 4 d]: beyond d the weight is 0 either way
    assert _check(0, d, 0.0, _f32(4.0 * d)) == 0


def _sharpen_numerators():
    out = [_f32(np.float32(1.4) * np.float32(1.4)), 1.0, _f32(np.float32(1.2) ** 2)]
    for n in (16, 32, 64, 128):
        re = np.float32(np.float32(0.5) / np.float32(n))
        out.append(_f32(re * re))
    return out


@pytest.mark.parametrize("hh", _sharpen_numerators())
def test_sharpen_divide_matches_ieee(hh):
    # every float divisor max(r2, 1e-5) in [1e-5, 4]
    assert _check(1, hh, _f32(1e-5), 4.0) == 0


def test_p2g_sharpen_divide_matches_ieee():
    """P2G's Sharpen at the reference radius (RE = 1.4, particlegridoperations.inl:149)
    divides 1.96f by max(r2, 1e-5) with the hardware reciprocal and one residual
    correction: equal to IEEE division for every float divisor in [1e-5, 4]
    (beyond 4 the quotient is below 1/2, where the weight is 0 either way)."""
    assert _check(2, 1.0, _f32(1e-5), 4.0) == 0


def test_division_check_detects_a_wrong_divisor_range():
    # sanity of the checker: outside the routine's stated domain (divisors far
    # below 2^-60 with a large numerator overflow the quotient) mismatches appear
    assert _check(1, 1.0e30, 1.0e-30, 1.0e-28) > 0


def test_bf16_rounding_matches_round_to_nearest_even():
    """The network kernels round FP32 to BF16 with the hardware's packed
    conversion; it must equal round-to-nearest-even on the bit pattern (the
    rounding the PyTorch restatement and the trainer use) for every finite float."""
    lib = _abi.load()
    bad = ctypes.c_int64(-1)
    rc = lib.ariel_selftest_bf16(ctypes.byref(bad))
    assert rc == 0, lib.ariel_last_error(None).decode()
    assert bad.value == 0
