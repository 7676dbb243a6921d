"""The transfer's BLAKE3 tree (lizx_b3_*): the host hash against BLAKE3's official test vectors (tests/golden), and the
host's and the GPU's chunk chaining values, roots and hashes against the reference's own (liblizard/src/xfer.c over
the vendored BLAKE3, compiled in oracle/_ref), for chunk sizes 2^10 .. 2^20 and files of every awkward length."""
import json
import os

import numpy as np
import pytest

import lizard_amd as L
from _libs import ptr, ref

HERE = os.path.dirname(os.path.abspath(__file__))
VECTORS = json.load(open(os.path.join(HERE, "golden", "blake3_vectors.json")))["cases"]
xfer_needed = pytest.mark.skipif(ref("xfer") is None, reason="oracle/_ref not built")


def pattern(n):
    return (np.arange(n) % 251).astype(np.uint8)


def test_host_hash_matches_official_vectors():
    for c in VECTORS:
        assert L.b3_hash(pattern(c["input_len"])).hex() == c["hash32"], c["input_len"]


def test_host_root_is_the_hash_of_the_whole():
    """The root merged from 2^k-byte chunks' values equals the whole file's hash (the property the transfer relies
    on), with the chunks' values from the reference."""
    r = ref("xfer")
    if r is None:
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(3)
    for length, k in ((2048, 10), (5000, 10), (70000, 12), (1 << 16, 14), (300001, 16)):
        data = rng.integers(0, 256, length).astype(np.uint8)
        chunks = (length + (1 << k) - 1) >> k
        cvs = np.zeros((chunks, 32), np.uint8)
        for i in range(chunks):
            part = np.ascontiguousarray(data[i << k:(i + 1) << k])
            assert r.xfer_chunk_cv(ptr(part), part.size, i, k, ptr(cvs[i])) == 0
        assert L.b3_root(cvs) == L.b3_hash(data)
        want = np.zeros(32, np.uint8)
        assert r.xfer_root(ptr(cvs), chunks, ptr(want)) == 0
        assert L.b3_root(cvs) == want.tobytes()


@xfer_needed
def test_host_hash_matches_reference_on_random_lengths():
    r = ref("xfer")
    rng = np.random.default_rng(5)
    for length in (0, 1, 63, 64, 65, 1023, 1024, 1025, 4095, 4096, 4097, 65535, 65537, 262144 + 77):
        data = rng.integers(0, 256, max(length, 1)).astype(np.uint8)[:length]
        want = np.zeros(32, np.uint8)
        r.xfer_b3_hash(ptr(data) if length else None, length, ptr(want))
        assert L.b3_hash(data) == want.tobytes(), length


@pytest.mark.gpu
def test_gpu_hash_matches_vectors_and_reference(decoder_gpu):
    for c in VECTORS:
        n = c["input_len"]
        d = L.DeviceArray.of(pattern(max(n, 1)))
        assert decoder_gpu.b3_hash_device(d, n).hex() == c["hash32"], n
        d.free()
    r = ref("xfer")
    rng = np.random.default_rng(9)
    for length in (1 << 20, (1 << 24) + 12345, 3 * (1 << 22) - 1):
        data = rng.integers(0, 256, length).astype(np.uint8)
        d = L.DeviceArray.of(data)
        got = decoder_gpu.b3_hash_device(d, length)
        d.free()
        if r is not None:
            want = np.zeros(32, np.uint8)
            r.xfer_b3_hash(ptr(data), length, ptr(want))
            assert got == want.tobytes(), length
        else:
            assert got == L.b3_hash(data)


@pytest.mark.gpu
@pytest.mark.parametrize("length,k", [(1, 10), (1000, 10), (1024, 10), (5000, 10), (70000, 12), (1 << 20, 16),
                                      ((1 << 22) + 3, 20), (9 * (1 << 20) - 17, 20), ((1 << 16) * 5 + 1024, 16)])
def test_gpu_chunk_values_match_reference(decoder_gpu, length, k):
    rng = np.random.default_rng(length + k)
    data = rng.integers(0, 256, length).astype(np.uint8)
    d = L.DeviceArray.of(data)
    cvs = decoder_gpu.b3_chunk_cvs(d, length, k)
    d.free()
    chunks = (length + (1 << k) - 1) >> k
    assert cvs.shape == (chunks, 32)
    r = ref("xfer")
    for i in range(chunks):
        part = np.ascontiguousarray(data[i << k:(i + 1) << k])
        if r is not None:
            want = np.zeros(32, np.uint8)
            assert r.xfer_chunk_cv(ptr(part), part.size, i, k, ptr(want)) == 0
            assert np.array_equal(cvs[i], want), i
    if chunks >= 2:
        assert L.b3_root(cvs) == L.b3_hash(data)


@pytest.mark.gpu
def test_gpu_argument_errors(decoder_gpu):
    d = L.DeviceArray(4096)
    with pytest.raises(L.LizError):
        decoder_gpu.b3_chunk_cvs(d, 4096, 9)
    with pytest.raises(L.LizError):
        decoder_gpu.b3_chunk_cvs(d, 0, 10)
    d.free()


def test_stream_fill_is_shake256_of_the_id():
    """The test stream's payload (liz_stream_fill) against Python's own SHAKE256 and the reference's stream_fill
    (liblizard/src/shake.c, compiled in oracle/_ref), over awkward ids."""
    import hashlib
    r = ref("xfer")
    ids = [0, 1, 255, 256, 0xFFFF, 0x10000, 0x7FFFFFFF, 0xFFFFFFFF, 0xFFFC0000, 12345678] + list(range(1000, 1100))
    for i in ids:
        got = L.stream_fill(i)
        assert got.tobytes() == hashlib.shake_256(int(i).to_bytes(4, "little")).digest(472), i
        if r is not None:
            want = np.zeros(472, np.uint8)
            r.stream_fill(i, ptr(want), 472)
            assert np.array_equal(got, want), i
    assert L.lib().liz_simd() == 1
