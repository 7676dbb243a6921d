"""The persistent LDPC queue with several codewords a workgroup: the grid forced small (lizx_ldpc_shape) so that every
workgroup takes codeword after codeword off the queue, reusing its posteriors' LDS, its reduction slots and its task
slot between them. Held to the oracle codeword by codeword (iterations, hard decisions, final posteriors and messages,
CRC verdicts and bytes), at the sizes of the LDPC-only workload (4096 codewords, 7/8:3/4:2/3:1/2 = 1597:1065:614:820,
Eb/N0 1.5-4.0 dB, 20 iterations; liblizard/src/ldpc.c:1080-1172)."""
import numpy as np
import pytest

import lizard_amd
from _libs import RATES, awgn_llr, oracle, ptr
from ldpc_cases import batch, code

pytestmark = pytest.mark.gpu

EBN0 = (1.5, 2.0, 2.5, 3.0, 3.5, 4.0)
MIX = ((6, 1597), (4, 1065), (3, 614), (2, 820))


def _oracle(rt, llr, offs, max_iter, stall_it, trace):
    return [code(int(r)).decode(llr[o:o + code(int(r)).n], max_iter, stall_it, 0.95, trace=trace)
            for r, o in zip(rt, offs)]


def _same(out, ref, rt, offs, trace):
    for i, (r, o) in enumerate(zip(rt, offs)):
        c = code(int(r))
        assert out["its"][i] == ref[i][0], "codeword %d (rate %d): iterations %d, oracle %d" % (
            i, r, out["its"][i], ref[i][0])
        assert np.array_equal(out["bits"][o:o + c.n], ref[i][1]), "codeword %d (rate %d): hard decisions" % (i, r)
        if trace:
            assert np.array_equal(out["post"][o:o + c.n], ref[i][2]), "codeword %d: posteriors" % i
            assert np.array_equal(out["msg"][i][:c.slots_total * 128], ref[i][3]), "codeword %d: messages" % i


@pytest.mark.parametrize("per_cu", [1, 2])
def test_queue_several_codewords_a_workgroup(per_cu):
    """1024 mixed-rate codewords at 1.5-4.0 dB on a grid of CUs x per_cu workgroups (256 or 512 on an MI355X): each
    workgroup decodes four or two codewords in turn, of different codes and iteration counts."""
    rates = [RATES[(i * 7) % 4] for i in range(1024)]
    dbs = [EBN0[(i * 5) % 6] for i in range(1024)]
    rt, llr, offs, _ = batch(rates, dbs, seed=900 + per_cu)
    dec = lizard_amd.LdpcDecoder()
    try:
        assert dec.shape(per_cu) == per_cu
        out = dec.decode(rt, llr, offs, max_iter=20, stall_it=0, trace=True)
        _same(out, _oracle(rt, llr, offs, 20, 0, True), rt, offs, True)
        # the PHY's setting (30 iterations, stall rule) through the same small grid
        out = dec.decode(rt, llr, offs, max_iter=30, stall_it=9, trace=True)
        _same(out, _oracle(rt, llr, offs, 30, 9, True), rt, offs, True)
        assert dec.shape(0) >= 1
    finally:
        dec.close()


def _workload(seed):
    """The LDPC-only workload's 4096 codewords: 476 random bytes and their CRC-32C a codeword, rates in a seeded
    order, Eb/N0 cycling over the six points codeword by codeword."""
    rng = np.random.default_rng(seed)
    o = oracle()
    rates = np.concatenate([np.full(c, r, np.uint8) for r, c in MIX])
    rates = rates[rng.permutation(rates.size)]
    llrs, offs, bys, at = [], [], [], 0
    for i, r in enumerate(rates):
        c = code(int(r))
        by = rng.integers(0, 256, 476).astype(np.uint8)
        crc = o.or_crc32c(ptr(by), 476)
        data = np.zeros(c.k, np.uint8)
        data[:3808] = np.unpackbits(by)
        data[3808:3840] = [(crc >> (31 - b)) & 1 for b in range(32)]
        llr = awgn_llr(c.n, c.k, c.encode(data), EBN0[i % 6], 3840.0 / c.nt, rng,
                       slice(3840, c.k) if c.k > 3840 else None)
        llrs.append(llr), offs.append(at), bys.append(by)
        at += c.n
    return rates, np.concatenate(llrs), np.array(offs, np.int64), np.array(bys)


@pytest.fixture(scope="module")
def workload():
    return _workload(4096)


def test_full_workload_bit_exact(workload):
    """All 4096 codewords of the workload in one persistent launch (the grid at its largest, then forced to one
    workgroup a CU): iterations, hard decisions, CRC verdicts and bytes equal the oracle's and the CRC's, codeword
    by codeword; the plain per-codeword launches give the same."""
    rt, llr, offs, sent = workload
    ref = _oracle(rt, llr, offs, 20, 0, False)
    o = oracle()
    want_ok = np.zeros(len(rt), bool)
    for i, (its, bits) in enumerate(ref):
        packed = np.packbits(bits[:3808])
        rx = int("".join(map(str, bits[3808:3840])), 2)
        want_ok[i] = its > 0 and rx == o.or_crc32c(ptr(packed), 476)
    dec = lizard_amd.LdpcDecoder()
    try:
        outs = []
        for per_cu in (0, 1):
            dec.shape(per_cu)
            outs.append(dec.decode(rt, llr, offs, max_iter=20, stall_it=0, crc=True))
        outs.append(dec.decode(rt, llr, offs, max_iter=20, stall_it=0, crc=True, plain=True))
        for out in outs:
            _same(out, ref, rt, offs, False)
            assert np.array_equal(out["crc_ok"].astype(bool), want_ok)
            v = want_ok
            assert np.array_equal(out["bytes"][v], np.array([np.packbits(ref[i][1][:3808]) for i in np.nonzero(v)[0]]))
        # a verified block is the block sent
        assert np.array_equal(outs[0]["bytes"][want_ok], sent[want_ok])
        assert 0.3 < want_ok.mean() < 1.0  # the sweep holds both failing and clean codewords
    finally:
        dec.close()


def test_device_entry_small_grid(workload):
    """The device-memory entry on the resident workload, grid at one workgroup a CU, equals the host entry."""
    rt, llr, offs, _ = workload
    dec = lizard_amd.LdpcDecoder()
    try:
        want = dec.decode(rt, llr, offs, max_iter=20, stall_it=0, crc=True)
        dec.shape(1)
        D = lizard_amd.DeviceArray
        d_rt, d_llr, d_off = D.of(rt), D.of(llr), D.of(offs)
        d_bits, d_its, d_by, d_ok = D(llr.size), D(len(rt)), D(len(rt) * 476), D(len(rt))
        for _ in range(2):  # twice on the same buffers: the queue restarts from zero each call
            dec.decode_device(len(rt), d_rt, d_llr, d_off, d_bits, d_its, max_iter=20, stall_it=0, crc=True,
                              d_bytes=d_by, d_crc_ok=d_ok)
            dec.sync()
            assert np.array_equal(d_its.get(np.int8, len(rt)), want["its"])
            assert np.array_equal(d_ok.get(np.uint8, len(rt)), want["crc_ok"])
            assert np.array_equal(d_bits.get(np.uint8, llr.size), want["bits"])
            assert np.array_equal(d_by.get(np.uint8, len(rt) * 476).reshape(-1, 476), want["bytes"])
    finally:
        dec.close()


def test_shape_arguments():
    dec = lizard_amd.LdpcDecoder()
    try:
        with pytest.raises(lizard_amd.LizError):
            dec.shape(-1)
        full = dec.shape(0)
        assert full >= 1 and dec.shape(10 ** 6) == full
    finally:
        dec.close()
