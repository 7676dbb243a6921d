"""The rounding-tie checker the GPU parity tests use (tests/pictures.py assert_ties), on planted cases: a difference
whose two float values straddle the half-step passes; one whose values are far apart, or a difference of two, fails."""
import numpy as np
import pytest

from pictures import assert_ties, float_soft


def _case(xa, xb):
    """One sub-channel (640 slots) of QPSK-like values whose slot 3 holds 8 k y = xa on one path and xb on the other
    (the other path given by its coefficients, whose k float_soft forms; slot 3 is moved until its own k puts it at
    xb). Returns (a, b, soft_a, yt_b)."""
    rng = np.random.default_rng(1)
    yt = (np.where(rng.random(640) < 0.5, -1.0, 1.0) + rng.standard_normal(640) * 0.8).astype(np.float32)
    for _ in range(50):
        x = float_soft(yt)
        yt[3] = np.float32(yt[3] * xb / x[3])
    xb_real = float_soft(yt)
    soft_a = (xb_real / 8.0).astype(np.float32)
    soft_a[3] = np.float32(xa / 8.0)
    a = np.rint(8.0 * soft_a.astype(np.float64)).astype(np.int8)
    b = np.rint(xb_real).astype(np.int8)
    return a, b, soft_a, yt


def test_straddling_values_are_a_tie():
    a, b, s, yb = _case(2.5005, 2.4995)
    assert a[3] != b[3]
    frac, rel = assert_ties(a, b, s, yb)
    assert frac == pytest.approx(1 / 640) and rel < 1e-3


def test_equal_roundings_pass():
    a, b, s, yb = _case(2.3, 2.31)
    assert assert_ties(a, b, s, yb) == (0.0, 0.0)


def test_values_far_apart_fail():
    a, b, s, yb = _case(2.6, 2.4)  # straddle 2.5, but 0.2 apart: a real discrepancy, not rounding
    assert a[3] != b[3]
    with pytest.raises(AssertionError):
        assert_ties(a, b, s, yb)


def test_a_value_not_its_floats_rounding_fails():
    a, b, s, yb = _case(2.51, 2.51)
    a = a.copy()
    a[3] -= 1  # this path wrote 2 where its own float rounds to 3
    with pytest.raises(AssertionError):
        assert_ties(a, b, s, yb)


def test_differences_of_two_fail():
    a, b, s, yb = _case(2.5, 2.5)
    a = a.copy()
    a[3] += 2
    with pytest.raises(AssertionError):
        assert_ties(a, b, s, yb)
