"""The product library's host side, no GPU needed: it loads, exports every symbol its header declares, and its format
tables, encoder and CRC agree with the oracle."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import lizard_amd
from _libs import RATES, OrCode, oracle, ptr


def header_symbols():
    txt = open(lizard_amd.HEADER).read()
    return sorted(set(re.findall(r"LIZ_API\s+[\w\s\*]+?\b(liz\w+)\s*\(", txt)))


def test_exports_match_header():
    L = lizard_amd.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", lizard_amd.LIB_PATH], capture_output=True, text=True).stdout
    exported = sorted(set(re.findall(r" T (\w+)", out)))
    assert exported == syms, "exports beyond or short of the header"


def test_errors_are_per_call():
    L = lizard_amd.lib()
    info = np.zeros(6, np.int32)
    assert L.lizx_format_info(3, lizard_amd._p(info)) == lizard_amd.LIZ_E_FORMAT
    assert L.liz_last_error_code() == lizard_amd.LIZ_E_FORMAT and L.liz_last_error()
    assert L.lizx_format_info(4, lizard_amd._p(info)) == 0
    assert L.liz_last_error_code() == 0
    assert L.lizx_ldpc_dims(5, lizard_amd._p(np.zeros(12, np.int32))) == lizard_amd.LIZ_E_ARG


@pytest.mark.parametrize("version", list(range(4, 65)))
def test_format_info(version):
    o = oracle()
    subch = 16 * version
    n = o.or_n_for(subch)
    f = o.or_format_new(n, subch)
    io = np.zeros(6, np.int32)
    o.or_format_info(f, ptr(io))
    got = lizard_amd.format_info(version)
    assert [got[k] for k in ("n", "subch", "blocks", "tiers", "vblocks", "used")] == list(io)
    o.or_format_free(f)


@pytest.mark.parametrize("rate", RATES)
def test_dims_and_encoder(rate):
    c = OrCode(rate)
    d = lizard_amd.ldpc_dims(rate)
    assert (d["n"], d["k"], d["m"], d["mb"], d["kb"], d["np"], d["nt"], d["edges"], d["slots_total"]) == (
        c.n, c.k, c.m, c.mb, c.kb, c.np, c.nt, c.edges, c.slots_total)
    rng = np.random.default_rng(rate)
    data = rng.integers(0, 2, (6, c.k)).astype(np.uint8)
    cw = lizard_amd.ldpc_encode(rate, data)
    for i in range(6):
        assert np.array_equal(cw[i], c.encode(data[i]))


def test_crc32c():
    o = oracle()
    rng = np.random.default_rng(9)
    assert lizard_amd.crc32c(np.frombuffer(b"123456789", np.uint8)) == 0xE3069283
    for n in (1, 17, 476, 1000):
        b = rng.integers(0, 256, n).astype(np.uint8)
        assert lizard_amd.crc32c(b) == o.or_crc32c(ptr(b), n)


def test_no_gpu_is_an_error_not_a_fallback():
    """Without a device the decoder constructors fail loudly (no CPU path behind them)."""
    if lizard_amd.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(lizard_amd.LizError):
        lizard_amd.LdpcDecoder()
    with pytest.raises(lizard_amd.LizError):
        lizard_amd.Decoder()


def test_finder_entries_check_their_arguments():
    """The finder's entries refuse a missing decoder or output before any device work (no GPU needed)."""
    L = lizard_amd.lib()
    img = np.zeros((64, 64), np.uint8)
    H, nodes = np.zeros(9, np.float32), np.zeros(3200, np.float32)
    assert L.lizx_find(None, lizard_amd._p(img), 64, 64, 64, None, lizard_amd._p(H), lizard_amd._p(nodes), None,
                       None) == lizard_amd.LIZ_E_ARG
    ptrs = (C.c_void_p * 1)(img.ctypes.data)
    spans, found = np.zeros(1, np.int32), np.zeros(1, np.int32)
    assert L.lizx_find_batch(None, 1, ptrs, 64, 64, 64, lizard_amd._p(spans), lizard_amd._p(H), lizard_amd._p(nodes),
                             lizard_amd._p(found), None) == lizard_amd.LIZ_E_ARG
    assert L.liz_last_error_code() == lizard_amd.LIZ_E_ARG
