"""The GPU LDPC decoder (lizx_ldpc_decode) against the oracle: bit-exact hard decisions, iteration counts and verdicts,
final posteriors and messages, CRC verdicts and bytes, for every code, persistent and plain launches."""
import numpy as np
import pytest

import lizard_amd
from _libs import RATES, oracle, ptr
from ldpc_cases import batch, code

pytestmark = pytest.mark.gpu

DBS = (1.0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0, 4.5)


def _oracle_all(rates, llr, offs, max_iter, stall_it, ratio=0.95):
    res = []
    for r, o in zip(rates, offs):
        c = code(int(r))
        res.append(c.decode(llr[o:o + c.n], max_iter, stall_it, ratio, trace=True))
    return res


def _check(out, ref, rates, offs, trace=True):
    for i, (r, o) in enumerate(zip(rates, offs)):
        c = code(int(r))
        its, bits, L, R = ref[i]
        assert out["its"][i] == its, "codeword %d (rate %d): iterations %d, oracle %d" % (i, r, out["its"][i], its)
        assert np.array_equal(out["bits"][o:o + c.n], bits), "codeword %d (rate %d): hard decisions" % (i, r)
        if trace:
            assert np.array_equal(out["post"][o:o + c.n], L), "codeword %d (rate %d): posteriors" % (i, r)
            assert np.array_equal(out["msg"][i][:c.slots_total * 128], R), "codeword %d (rate %d): messages" % (i, r)


@pytest.mark.parametrize("max_iter,stall_it", [(30, 9), (20, 0), (2, 0), (1, 0), (5, 1)])
def test_mixed_batch_trace(ldpc_gpu, max_iter, stall_it):
    rates = [r for r in RATES for _ in range(len(DBS) * 3)]
    dbs = [d for _ in RATES for d in DBS * 3]
    rt, llr, offs, _ = batch(rates, dbs, seed=100 + max_iter + stall_it)
    out = ldpc_gpu.decode(rt, llr, offs, max_iter=max_iter, stall_it=stall_it, trace=True)
    ref = _oracle_all(rt, llr, offs, max_iter, stall_it)
    _check(out, ref, rt, offs)


def test_plain_equals_persistent(ldpc_gpu):
    rates = [r for r in RATES for _ in range(6)]
    dbs = [d for _ in RATES for d in (1.5, 2.5, 3.0, 3.5, 4.0, 6.0)]
    rt, llr, offs, _ = batch(rates, dbs, seed=7)
    a = ldpc_gpu.decode(rt, llr, offs, trace=True)
    b = ldpc_gpu.decode(rt, llr, offs, trace=True, plain=True)
    for k in ("its", "bits", "post", "msg"):
        assert np.array_equal(a[k], b[k]), k
    _check(a, _oracle_all(rt, llr, offs, 30, 9), rt, offs)


def test_order_independent(ldpc_gpu):
    rates = [RATES[i % 4] for i in range(64)]
    dbs = [DBS[i % len(DBS)] for i in range(64)]
    rt, llr, offs, _ = batch(rates, dbs, seed=11)
    a = ldpc_gpu.decode(rt, llr, offs, max_iter=20, stall_it=0)
    perm = np.random.default_rng(3).permutation(64)
    b = ldpc_gpu.decode(rt[perm], llr, offs[perm], max_iter=20, stall_it=0)
    assert np.array_equal(a["its"][perm], b["its"])
    assert np.array_equal(a["bits"], b["bits"])


def _crc_words(rate, db, rng, flip=False):
    """A codeword whose info bits are 476 random bytes and their CRC-32C (SPEC 7.1), through BI-AWGN."""
    o = oracle()
    c = code(rate)
    by = rng.integers(0, 256, 476).astype(np.uint8)
    crc = o.or_crc32c(ptr(by), 476)
    bits = np.unpackbits(by)
    cb = np.array([(crc >> (31 - i)) & 1 for i in range(32)], np.uint8)
    if flip:
        cb[5] ^= 1  # a word whose CRC does not hold: must decode (it is a codeword) and fail the CRC
    data = np.zeros(c.k, np.uint8)
    data[:3808] = bits
    data[3808:3840] = cb
    cw = c.encode(data)
    from _libs import awgn_llr
    llr = awgn_llr(c.n, c.k, cw, db, 3840.0 / c.nt, rng, slice(3840, c.k) if c.k > 3840 else None)
    return by, llr


def test_crc(ldpc_gpu):
    rng = np.random.default_rng(21)
    rates, llrs, offs, bys, flips, at = [], [], [], [], [], 0
    for i in range(48):
        r = RATES[i % 4]
        flip = i % 7 == 3
        by, llr = _crc_words(r, (2.0, 3.0, 4.0, 8.0)[i // 12], rng, flip)
        rates.append(r), llrs.append(llr), offs.append(at), bys.append(by), flips.append(flip)
        at += len(llr)
    rt, llr, off = np.array(rates, np.uint8), np.concatenate(llrs), np.array(offs, np.int64)
    out = ldpc_gpu.decode(rt, llr, off, max_iter=30, stall_it=9, crc=True)
    o = oracle()
    for i in range(48):
        c = code(rates[i])
        its, bits = c.decode(llr[offs[i]:offs[i] + c.n], 30, 9, 0.95)
        assert out["its"][i] == its
        packed = np.packbits(bits[:3808])
        rx = int("".join(map(str, bits[3808:3840])), 2)
        want = its > 0 and rx == o.or_crc32c(ptr(packed), 476)
        assert bool(out["crc_ok"][i]) == want, "codeword %d" % i
        if its > 0:
            assert np.array_equal(out["bytes"][i], packed)
        if flips[i]:
            assert not out["crc_ok"][i]


def test_edges(ldpc_gpu):
    """Clean words stop after one iteration; all-zero and saturated inputs; the int8 extremes."""
    rates, llrs, offs, at = [], [], [], 0
    rng = np.random.default_rng(5)
    for r in RATES:
        c = code(r)
        data = np.zeros(c.k, np.uint8)
        data[:3840] = rng.integers(0, 2, 3840)
        cw = c.encode(data)
        cases = [np.where(cw, -127, 127).astype(np.int8),          # clean, saturated
                 np.zeros(c.n, np.int8),                          # no information
                 np.where(cw, -1, 1).astype(np.int8),             # clean, weakest
                 rng.integers(-128, 128, c.n).astype(np.int8),    # noise over the whole int8 range
                 np.full(c.n, -128, np.int8)]                     # every bit certain 1 at the extreme
        for llr in cases:
            rates.append(r), llrs.append(llr), offs.append(at)
            at += len(llr)
    rt, llr, off = np.array(rates, np.uint8), np.concatenate(llrs), np.array(offs, np.int64)
    for mi, st in ((30, 9), (1, 0), (20, 0)):
        out = ldpc_gpu.decode(rt, llr, off, max_iter=mi, stall_it=st, trace=True)
        _check(out, _oracle_all(rt, llr, off, mi, st), rt, off)
    out = ldpc_gpu.decode(rt, llr, off)
    assert all(out["its"][i] == 1 for i in range(0, len(rates), 5)), "clean words take one iteration"


def test_empty_and_bad_args(ldpc_gpu):
    out = ldpc_gpu.decode(np.zeros(0, np.uint8), np.zeros(8, np.int8), np.zeros(0, np.int64))
    assert out["its"].size == 0
    with pytest.raises(lizard_amd.LizError):
        ldpc_gpu.decode(np.array([5], np.uint8), np.zeros(7680, np.int8), np.zeros(1, np.int64))
    with pytest.raises(lizard_amd.LizError):
        ldpc_gpu.decode(np.array([2], np.uint8), np.zeros(7680, np.int8), np.zeros(1, np.int64), max_iter=0)


def test_device_entry(ldpc_gpu):
    """The device-memory entry on HBM-resident inputs gives what the host entry gives."""
    rates = [RATES[i % 4] for i in range(96)]
    dbs = [DBS[i % len(DBS)] for i in range(96)]
    rt, llr, offs, _ = batch(rates, dbs, seed=31)
    a = ldpc_gpu.decode(rt, llr, offs, max_iter=20, stall_it=0)
    D = lizard_amd.DeviceArray
    d_rt, d_llr, d_off = D.of(rt), D.of(llr), D.of(offs)
    d_bits, d_its = D(llr.size), D(len(rt))
    ldpc_gpu.decode_device(len(rt), d_rt, d_llr, d_off, d_bits, d_its, max_iter=20, stall_it=0)
    ms = ldpc_gpu.sync()
    assert ms > 0
    assert np.array_equal(d_its.get(np.int8, len(rt)), a["its"])
    bits = d_bits.get(np.uint8, llr.size)
    for r, o in zip(rt, offs):
        n = code(int(r)).n
        assert np.array_equal(bits[o:o + n], a["bits"][o:o + n])


def test_tuned_grid_same_bits():
    """The autotuner's grid size (workgroups a CU) changes only the launch shape: the decode after tuning gives the
    bits and iteration counts the untuned decoder gives."""
    rates = [RATES[i % 4] for i in range(256)]
    dbs = [DBS[i % len(DBS)] for i in range(256)]
    rt, llr, offs, _ = batch(rates, dbs, seed=37)
    dec = lizard_amd.LdpcDecoder()
    try:
        a = dec.decode(rt, llr, offs, max_iter=20, stall_it=0)
        D = lizard_amd.DeviceArray
        d_rt, d_llr, d_off = D.of(rt), D.of(llr), D.of(offs)
        d_bits, d_its = D(llr.size), D(len(rt))
        pc = dec.tune(len(rt), d_rt, d_llr, d_off, d_its, max_iter=20, stall_it=0, reps=1)
        assert pc in (0, 2, 3, 4, 6, 8, 10, 12, 16)
        with pytest.raises(lizard_amd.LizError):
            dec.tune(len(rt), d_rt, d_llr, d_off, d_its, reps=0)
        b = dec.decode(rt, llr, offs, max_iter=20, stall_it=0)
        assert np.array_equal(a["bits"], b["bits"]) and np.array_equal(a["its"], b["its"])
        dec.decode_device(len(rt), d_rt, d_llr, d_off, d_bits, d_its, max_iter=20, stall_it=0)
        dec.sync()
        assert np.array_equal(d_its.get(np.int8, len(rt)), a["its"])
        bits = d_bits.get(np.uint8, llr.size)
        for r, o in zip(rt, offs):
            n = code(int(r)).n
            assert np.array_equal(bits[o:o + n], a["bits"][o:o + n])
    finally:
        dec.close()
