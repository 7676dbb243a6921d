"""lizx_ldpc_decode_device_noalloc (the LDPC device decode with its task list and queue counter in the library's
fixed per-device work area) against lizx_ldpc_decode_device, through the C ABI: the same iterations, hard decisions,
CRC verdicts and bytes on two calls in a row, the second on fewer codewords (the area keeps the first call's tasks
past the second's count, which must not matter); counts beyond the area refused."""
import numpy as np
import pytest

import lizard_amd
from _libs import RATES, awgn_llr, oracle, ptr
from ldpc_cases import code

pytestmark = pytest.mark.gpu


def _words(count, seed):
    rng = np.random.default_rng(seed)
    o = oracle()
    rates = np.array([RATES[(i * 3) % 4] for i in range(count)], np.uint8)
    llrs, offs, at = [], [], 0
    for i, r in enumerate(rates):
        c = code(int(r))
        by = rng.integers(0, 256, 476).astype(np.uint8)
        crc = o.or_crc32c(ptr(by), 476)
        data = np.zeros(c.k, np.uint8)
        data[:3808] = np.unpackbits(by)
        data[3808:3840] = [(crc >> (31 - b)) & 1 for b in range(32)]
        llr = awgn_llr(c.n, c.k, c.encode(data), (1.5, 2.5, 3.5, 4.5)[i % 4], 3840.0 / c.nt, rng,
                       slice(3840, c.k) if c.k > 3840 else None)
        llrs.append(llr), offs.append(at)
        at += c.n
    return rates, np.concatenate(llrs), np.array(offs, np.int64)


def _run(dec, fn, rates, llr, offs):
    D = lizard_amd.DeviceArray
    n = len(rates)
    d_rt, d_llr, d_off = D.of(rates), D.of(llr), D.of(offs)
    d_bits, d_its, d_by, d_ok = D(llr.size), D(n), D(n * 476), D(n)
    getattr(dec, fn)(n, d_rt, d_llr, d_off, d_bits, d_its, max_iter=20, stall_it=0, crc=True, d_bytes=d_by,
                     d_crc_ok=d_ok)
    dec.sync()
    return (d_its.get(np.int8, n), d_bits.get(np.uint8, llr.size), d_ok.get(np.uint8, n),
            d_by.get(np.uint8, n * 476))


def test_noalloc_equals_device_entry_twice():
    dec = lizard_amd.LdpcDecoder()
    try:
        for count, seed in ((1500, 3), (700, 4)):  # the second call smaller than the first
            w = _words(count, seed)
            want = _run(dec, "decode_device", *w)
            got = _run(dec, "decode_device_noalloc", *w)
            for g, x in zip(got, want):
                assert np.array_equal(g, x)
            assert 0 < want[2].sum() < count
        # and the small grid, where workgroups take several codewords each from the fixed queue
        dec.shape(1)
        got = _run(dec, "decode_device_noalloc", *w)
        for g, x in zip(got, want):
            assert np.array_equal(g, x)
    finally:
        dec.close()


def test_noalloc_refuses_counts_beyond_the_area():
    dec = lizard_amd.LdpcDecoder()
    try:
        D = lizard_amd.DeviceArray
        d = D(16)
        with pytest.raises(lizard_amd.LizError):
            dec.decode_device_noalloc(16385, d, d, d, None, d)
        dec.decode_device_noalloc(0, d, d, d, None, d)  # nothing to do
    finally:
        dec.close()
