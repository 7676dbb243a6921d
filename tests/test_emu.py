"""The LDPC kernel's per-code lane code (lizard_amd/csrc/ldpc_tmpl.hpp), walked lane by lane on the CPU
(tests/emu/ldpc_emu.cpp), against the oracle: final posteriors, messages, decisions, iterations."""
import ctypes as C
import os

import numpy as np
import pytest

from _libs import RATES, ptr
from ldpc_cases import code, word

EMU = os.path.join(os.path.dirname(os.path.abspath(__file__)), "emu", "libldpc_emu.so")


@pytest.fixture(scope="module")
def emu():
    if not os.path.exists(EMU):
        pytest.skip("tests/emu not built")
    lib = C.CDLL(EMU)
    lib.emu_ldpc_decode.restype = C.c_int
    lib.emu_ldpc_decode.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_void_p,
                                    C.c_void_p]
    return lib


@pytest.mark.parametrize("rate", RATES)
def test_lanes_match_oracle(emu, rate):
    c = code(rate)
    rng = np.random.default_rng(70 + rate)
    for w in range(12):
        _, llr = word(rate, (1.0, 2.0, 3.0, 4.0)[w % 4], rng)
        for mi, st in ((30, 9), (20, 0), (2, 0)):
            its, bits, L, R = c.decode(llr, mi, st, 0.95, trace=True)
            out = np.zeros(c.n, np.uint8)
            post = np.zeros(c.n, np.int16)
            msg = np.zeros(c.slots_total * 128, np.int8)
            got = emu.emu_ldpc_decode(rate, ptr(llr), ptr(out), mi, st, 0.95, ptr(post), ptr(msg))
            assert got == its
            assert np.array_equal(out, bits) and np.array_equal(post, L) and np.array_equal(msg, R)


@pytest.mark.parametrize("rate", RATES)
def test_lanes_match_oracle_on_ties(emu, rate):
    """LLRs of two magnitudes only: most checks see equal minima, so the first-minimum rule decides every message."""
    c = code(rate)
    rng = np.random.default_rng(90 + rate)
    for w in range(6):
        llr = (rng.choice([-16, -8, 8, 16], c.n, p=[0.1, 0.15, 0.45, 0.3])).astype(np.int8)
        its, bits, L, R = c.decode(llr, 12, 0, 0.95, trace=True)
        out = np.zeros(c.n, np.uint8)
        post = np.zeros(c.n, np.int16)
        msg = np.zeros(c.slots_total * 128, np.int8)
        got = emu.emu_ldpc_decode(rate, ptr(llr), ptr(out), 12, 0, 0.95, ptr(post), ptr(msg))
        assert got == its
        assert np.array_equal(out, bits) and np.array_equal(post, L) and np.array_equal(msg, R)


@pytest.mark.parametrize("rate", RATES)
def test_lanes_match_oracle_saturated(emu, rate):
    """Channel values at +-80 with 0-3% of them wrong: most checks hold several magnitudes far above the message cap
    (equal or not), and posteriors reach the largest sums of messages the column degrees allow."""
    c = code(rate)
    rng = np.random.default_rng(110 + rate)
    for w in range(4):
        if w < 2:
            llr = rng.choice([-80, 80], c.n).astype(np.int8)  # no codeword near: every iteration runs
        else:
            _, llr = word(rate, 4.0, rng)
            llr = (np.sign(llr.astype(np.int16)) * 80).astype(np.int8)
            flip = rng.random(c.n) < 0.03 * (w - 1)
            llr[flip] = -llr[flip]
        its, bits, L, R = c.decode(llr, 20, 0, 0.95, trace=True)
        out = np.zeros(c.n, np.uint8)
        post = np.zeros(c.n, np.int16)
        msg = np.zeros(c.slots_total * 128, np.int8)
        got = emu.emu_ldpc_decode(rate, ptr(llr), ptr(out), 20, 0, 0.95, ptr(post), ptr(msg))
        assert got == its
        assert np.array_equal(out, bits) and np.array_equal(post, L) and np.array_equal(msg, R)
