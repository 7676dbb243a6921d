"""The oracle held to the reference decoder itself, compiled from its sources (oracle/_ref, built by
oracle/ref/build_ref.sh): LDPC trajectories (hard decisions, iterations, final posteriors and messages) of the scalar
form, decisions of the x86-64 SIMD form, the encoder, format tables and the soft-value finish path. Skipped where
oracle/_ref was not built."""
import hashlib

import numpy as np
import pytest

from _libs import RATES, OrCode, oracle, ptr, ref, ref_decode
from ldpc_cases import word

pytestmark = pytest.mark.skipif(ref("scalar") is None, reason="oracle/_ref not built")


@pytest.mark.parametrize("rate", RATES)
def test_ldpc_trajectories(rate):
    r, rx = ref("scalar"), ref("x86")
    c = OrCode(rate)
    rng = np.random.default_rng(50 + rate)
    for w in range(24):
        _, llr = word(rate, (1.0, 1.75, 2.5, 3.25, 4.0, 5.0)[w % 6], rng)
        for mi, st in ((30, 9), (20, 0), (1, 0), (3, 2)):
            a = c.decode(llr, mi, st, 0.95, trace=True)
            b = ref_decode(r, rate, llr, mi, st, 0.95, trace=True)
            assert a[0] == b[0]
            for x, y in zip(a[1:], b[1:]):
                assert np.array_equal(x, y)
            if rx is not None:
                bx = ref_decode(rx, rate, llr, mi, st, 0.95)
                assert bx[0] == a[0] and np.array_equal(bx[1], a[1])


@pytest.mark.parametrize("rate", RATES)
def test_encoder(rate):
    r = ref("scalar")
    c = OrCode(rate)
    rng = np.random.default_rng(rate)
    for _ in range(8):
        data = rng.integers(0, 2, c.k).astype(np.uint8)
        cw = np.zeros(c.n, np.uint8)
        r.ref_ldpc_encode(rate, ptr(data), ptr(cw))
        assert np.array_equal(cw, c.encode(data))


def test_crc():
    r, o = ref("scalar"), oracle()
    rng = np.random.default_rng(2)
    for n in (0, 1, 9, 63, 476, 601):
        b = rng.integers(0, 256, max(n, 1)).astype(np.uint8)
        assert r.ref_crc32c(ptr(b), n) == o.or_crc32c(ptr(b), n)


@pytest.mark.parametrize("subch", [64, 80, 144, 320, 656, 720, 1024])
def test_format_and_slots(subch):
    r, o = ref("scalar"), oracle()
    n = o.or_n_for(subch)
    fr, fo = r.ref_format_new(n, subch, 256), o.or_format_new(n, subch)
    ir, io = np.zeros(10, np.int32), np.zeros(6, np.int32)
    r.ref_format_info(fr, ptr(ir))
    o.or_format_info(fo, ptr(io))
    assert list(ir[:6]) == list(io)
    S = subch * 320
    assert np.array_equal(np.ctypeslib.as_array(r.ref_format_pos(fr), (S,)), np.ctypeslib.as_array(o.or_format_pos(fo), (S,)))
    B = int(io[2])
    a = [np.zeros(B, np.int32) for _ in range(3)]
    b = [np.zeros(B, np.int32) for _ in range(3)]
    r.ref_format_blocks(fr, *map(ptr, a))
    o.or_format_blocks(fo, *map(ptr, b))
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    rng = np.random.default_rng(subch)
    blocks = rng.integers(0, 256, B * 476).astype(np.uint8)
    for cnt in range(4):
        so, sr = np.zeros(2 * S, np.uint8), np.zeros(2 * S, np.uint8)
        drive = np.zeros(int(ir[8]) ** 2, np.float32)
        o.or_frame_slots(fo, ptr(blocks), cnt, ptr(so))
        assert r.ref_encode(fr, ptr(blocks), cnt, ptr(drive), ptr(sr)) == 0
        assert np.array_equal(so, sr)
    r.ref_format_free(fr)
    o.or_format_free(fo)


def test_spec74_against_reference_encoder():
    """SPEC 7.4's frame through the reference's own encoder: the slots the oracle makes for every pilot count."""
    import test_oracle_vectors as V
    r, o = ref("scalar"), oracle()
    fr, fo = r.ref_format_new(384, 64, 256), o.or_format_new(384, 64)
    blocks = V.spec74_blocks()
    for cnt in range(4):
        so, sr = np.zeros(40960, np.uint8), np.zeros(40960, np.uint8)
        drive = np.zeros(1626 * 1626, np.float32)
        o.or_frame_slots(fo, ptr(blocks), cnt, ptr(so))
        r.ref_encode(fr, ptr(blocks), cnt, ptr(drive), ptr(sr))
        assert np.array_equal(so, sr)
    sha0 = hashlib.sha256(np.packbits(so if cnt == 0 else sr).tobytes()).hexdigest()
    assert sha0
    r.ref_format_free(fr)
    o.or_format_free(fo)


@pytest.mark.parametrize("version,sigma", [(4, 0.5), (5, 0.8), (20, 0.6), (45, 0.55), (64, 0.7)])
def test_finish_llr(version, sigma):
    r, o = ref("scalar"), oracle()
    subch = 16 * version
    n = o.or_n_for(subch)
    fr, fo = r.ref_format_new(n, subch, 256), o.or_format_new(n, subch)
    io = np.zeros(6, np.int32)
    o.or_format_info(fo, ptr(io))
    B = int(io[2])
    rng = np.random.default_rng(version)
    blocks = rng.integers(0, 256, B * 476).astype(np.uint8)
    slots = np.zeros(subch * 640, np.uint8)
    o.or_frame_slots(fo, ptr(blocks), 2, ptr(slots))
    y = (1.0 - 2.0 * slots) + sigma * rng.standard_normal(slots.size)
    llr = np.clip(np.rint(8 * 2 * y / sigma ** 2), -80, 80).astype(np.int8)
    est = rng.uniform(0.02, 1.0, subch).astype(np.float32)
    outs = []
    for lib, f, fn in ((o, fo, o.or_finish_llr), (r, fr, r.ref_finish_llr)):
        bl, ok, its, be = np.zeros(B * 476, np.uint8), np.zeros(B, np.uint8), np.zeros(B, np.int8), np.zeros(B, np.float32)
        g = fn(f, ptr(llr), ptr(est), ptr(bl), ptr(ok), ptr(its), ptr(be))
        outs.append((g, bl.reshape(B, 476), ok, its, be))
    a, b = outs
    assert a[0] == b[0]
    for x, y in zip(a[2:], b[2:]):
        assert np.array_equal(x, y)
    for blk in range(B):
        if a[2][blk]:
            assert np.array_equal(a[1][blk], b[1][blk])
    r.ref_format_free(fr)
    o.or_format_free(fo)


@pytest.mark.parametrize("version,span", [(4, 128), (5, 128), (20, 256), (64, 512)])
def test_modulator_matches_reference_picture(version, span):
    """The FP64 modulator (oracle/picture.py) against the picture the reference's own encoder makes before it
    resamples it into the drive (lizard_encode's rs_pic), both scalar and x86 builds: within float rounding."""
    import lizard_amd as L
    from pictures import picture as opic
    for form in ("scalar", "x86"):
        r = ref(form)
        if r is None:
            continue
        subch = 16 * version
        n = L.format_info(version)["n"]
        f = r.ref_format_new(n, subch, span)
        info = np.zeros(10, np.int32)
        r.ref_format_info(f, ptr(info))
        rng = np.random.default_rng(version)
        blocks = rng.integers(0, 256, int(info[2]) * 476).astype(np.uint8)
        u, v = L.format_coefs(version)
        for cnt in (0, 3):
            drive = np.zeros(int(info[8]) ** 2, np.float32)
            assert r.ref_encode(f, ptr(blocks), cnt, ptr(drive), None) == 0
            pic = np.zeros(n * n, np.float32)
            assert r.ref_encoded_picture(f, ptr(pic)) == 0
            want = opic.modulate(n, subch, L.frame_slots(version, blocks, cnt).astype(np.float64), u, v)
            assert np.abs(pic.reshape(n, n) - want).max() < 1e-6
        r.ref_format_free(f)
