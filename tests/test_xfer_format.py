"""The transfer's control blocks (lizx_xfer_*: the header and manifest payloads of the wire format, version 3) held
byte for byte and return code for return code to the reference's own xfer.c compiled from its sources
(oracle/_ref/libref_xfer.so): valid headers and manifests written and read back, random corruptions of them read,
headers that break a rule refused on writing, chaining-value lists checked against roots. Host only."""
import ctypes as C

import numpy as np
import pytest

import lizard_amd
from _libs import ref

xfer_needed = pytest.mark.skipif(ref("xfer") is None, reason="oracle/_ref not built")


class Header(C.Structure):
    # lizx_xfer_header and xfer_header_t: the same fields in the same order
    _fields_ = [("version", C.c_uint8), ("hash", C.c_uint8), ("chunk_log2", C.c_uint8), ("codec", C.c_uint8),
                ("seed", C.c_uint8), ("name_len", C.c_uint8), ("type_len", C.c_uint8), ("length", C.c_uint64),
                ("chunks", C.c_uint32), ("sent", C.c_uint32), ("root", C.c_uint8 * 32), ("name", C.c_uint8 * 255),
                ("type", C.c_uint8 * 158)]


P = C.c_void_p


def libs():
    mine, theirs = lizard_amd.lib(), ref("xfer")
    for lib, pre in ((mine, "lizx_xfer_"), (theirs, "xfer_")):
        getattr(lib, pre + "chunk_count").restype = C.c_uint32
        getattr(lib, pre + "chunk_count").argtypes = [C.c_uint64, C.c_int]
        getattr(lib, pre + "blocks").restype = C.c_uint32
        getattr(lib, pre + "blocks").argtypes = [C.c_uint32]
        getattr(lib, pre + "manifest_blocks").argtypes = [C.c_uint32]
        getattr(lib, pre + "sent_ok").argtypes = [P, C.c_uint32, C.c_uint32, C.c_int]
        getattr(lib, pre + "header_init").argtypes = [P, C.c_uint64, C.c_int]
        getattr(lib, pre + "header_write").argtypes = [P, P]
        getattr(lib, pre + "header_parse").argtypes = [P, P]
        getattr(lib, pre + "manifest_write").argtypes = [P, P, P, C.c_uint32, P, C.c_uint32, P]
        getattr(lib, pre + "manifest_parse").argtypes = [P, P, C.c_uint32, P, P, P]
        getattr(lib, pre + "manifest_check").argtypes = [P, C.c_uint32, P]
    mine.lizx_xfer_chunk_len.restype = C.c_uint64
    mine.lizx_xfer_chunk_len.argtypes = [P, C.c_uint32]
    theirs.xfer_chunk_len.restype = C.c_size_t
    theirs.xfer_chunk_len.argtypes = [P, C.c_uint32]
    mine.lizx_xfer_id_kind.argtypes = [C.c_uint32]
    return mine, theirs


def p(a):
    return a.ctypes.data if isinstance(a, np.ndarray) else C.addressof(a)


def same_header(a, b):
    return bytes(a) == bytes(b)


def random_header(rng, lib):
    """A header a sender could write: the file's length and chunk size, a codec, a name and a type, and for a file of
    one chunk its sent bytes and seed attempt."""
    h = Header()
    log2 = int(rng.integers(10, 25))
    kind = rng.integers(0, 4)
    length = (0 if kind == 0 else int(rng.integers(1, (1 << log2) + 1)) if kind == 1
              else int(rng.integers(1 << log2, 16383 << log2)))
    assert lib.lizx_xfer_header_init(C.byref(h), length, log2) == 0
    h.codec = int(rng.integers(0, 2))
    h.root[:] = rng.integers(0, 256, 32).tolist()
    h.name_len = int(rng.integers(0, 256))
    h.name[:h.name_len] = rng.integers(1, 256, h.name_len).tolist()
    h.type_len = int(rng.integers(0, 159))
    h.type[:h.type_len] = rng.integers(1, 256, h.type_len).tolist()
    if h.chunks == 1:
        h.sent = length if h.codec == 0 else int(rng.integers(1, length + 1))
        h.seed = int(rng.integers(0, 256)) if (h.sent + 471) // 472 >= 2 else 0
    return h


@xfer_needed
def test_counts_and_kinds():
    mine, theirs = libs()
    rng = np.random.default_rng(1)
    for _ in range(3000):
        length, log2 = int(rng.integers(0, 1 << 40)) >> int(rng.integers(0, 40)), int(rng.integers(8, 27))
        assert mine.lizx_xfer_chunk_count(length, log2) == theirs.xfer_chunk_count(length, log2)
        sent = int(rng.integers(0, 1 << 25))
        assert mine.lizx_xfer_blocks(sent) == theirs.xfer_blocks(sent)
        chunks = int(rng.integers(0, 20000))
        assert mine.lizx_xfer_manifest_blocks(chunks) == theirs.xfer_manifest_blocks(chunks)
    for length in (0, 1, (1 << 64) - 1, 16383 << 24, (16383 << 24) + 1):
        for log2 in (9, 10, 24, 25):
            assert mine.lizx_xfer_chunk_count(length, log2) == theirs.xfer_chunk_count(length, log2)
    # ids: chunk in bits 31..18, symbol in 17..0; chunk 16383 is control: 0 the header, 1..1366 manifest blocks
    ctl = 16383 << 18
    assert mine.lizx_xfer_id_kind(ctl) == 1 and mine.lizx_xfer_id_kind(ctl + 1) == 2
    assert mine.lizx_xfer_id_kind(ctl + 1366) == 2 and mine.lizx_xfer_id_kind(ctl + 1367) == -1
    assert mine.lizx_xfer_id_kind(0) == 0 and mine.lizx_xfer_id_kind(ctl - 1) == 0


@xfer_needed
def test_headers_written_and_read_as_the_reference():
    mine, theirs = libs()
    rng = np.random.default_rng(2)
    a, b = np.zeros(472, np.uint8), np.zeros(472, np.uint8)
    for _ in range(400):
        h = random_header(rng, mine)
        assert mine.lizx_xfer_header_write(C.byref(h), p(a)) == 0
        assert theirs.xfer_header_write(C.byref(h), p(b)) == 0
        assert np.array_equal(a, b)
        hm, ht = Header(), Header()
        assert mine.lizx_xfer_header_parse(p(a), C.byref(hm)) == 0 == theirs.xfer_header_parse(p(a), C.byref(ht))
        assert same_header(hm, ht) and same_header(hm, h)
        for i in range(int(h.chunks) + 1):
            assert mine.lizx_xfer_chunk_len(C.byref(h), i) == theirs.xfer_chunk_len(C.byref(h), i)
        # corruptions: one to four bytes changed anywhere, or a field set out of range
        for _ in range(8):
            c = a.copy()
            for _ in range(int(rng.integers(1, 5))):
                c[int(rng.integers(0, 472))] = int(rng.integers(0, 256))
            rm, rt = mine.lizx_xfer_header_parse(p(c), C.byref(hm)), theirs.xfer_header_parse(p(c), C.byref(ht))
            assert rm == rt, (rm, rt)
            assert same_header(hm, ht)
        # headers that break a rule are refused on writing by both
        g = Header.from_buffer_copy(bytes(h))
        field = rng.integers(0, 6)
        if field == 0:
            g.chunks += 1
        elif field == 1:
            g.sent = int(rng.integers(0, 1 << 32))
        elif field == 2:
            g.seed = int(rng.integers(1, 256))
        elif field == 3:
            g.codec = int(rng.integers(2, 256))
        elif field == 4:
            g.name_len = max(0, int(g.name_len) - 1)
            g.name[g.name_len] = 7
        else:
            g.chunk_log2 = int(rng.choice([9, 25, 31]))
        assert mine.lizx_xfer_header_write(C.byref(g), p(a)) == theirs.xfer_header_write(C.byref(g), p(b))


@xfer_needed
def test_manifests_written_read_and_checked_as_the_reference():
    mine, theirs = libs()
    rng = np.random.default_rng(3)
    a, b = np.zeros(472, np.uint8), np.zeros(472, np.uint8)
    for _ in range(60):
        h = Header()
        log2 = int(rng.integers(10, 16))
        chunks = int(rng.integers(2, 60))
        length = (chunks - 1) * (1 << log2) + int(rng.integers(1, (1 << log2) + 1))
        assert mine.lizx_xfer_header_init(C.byref(h), length, log2) == 0 and h.chunks == chunks
        h.codec = int(rng.integers(0, 2))
        cvs = rng.integers(0, 256, (chunks, 32)).astype(np.uint8)
        root = np.zeros(32, np.uint8)
        assert theirs.xfer_root(p(cvs), chunks, p(root)) == 0
        h.root[:] = root.tolist()
        lens = np.array([theirs.xfer_chunk_len(C.byref(h), i) for i in range(chunks)], np.int64)
        sent = np.where(h.codec == 1, rng.integers(1, lens + 1), lens).astype(np.uint32)
        seeds = np.where((sent + 471) // 472 >= 2, rng.integers(0, 256, chunks), 0).astype(np.uint8)
        assert mine.lizx_xfer_manifest_check(p(cvs), chunks, p(root)) == 0 == theirs.xfer_manifest_check(p(cvs), chunks,
                                                                                                      p(root))
        bad = root.copy()
        bad[int(rng.integers(0, 32))] ^= 1 << int(rng.integers(0, 8))
        assert mine.lizx_xfer_manifest_check(p(cvs), chunks, p(bad)) == theirs.xfer_manifest_check(p(cvs), chunks,
                                                                                                 p(bad)) == -1
        nb = mine.lizx_xfer_manifest_blocks(chunks)
        for m in range(nb + 1):
            rm = mine.lizx_xfer_manifest_write(p(cvs), p(sent), p(seeds), chunks, p(root), m, p(a))
            rt = theirs.xfer_manifest_write(p(cvs), p(sent), p(seeds), chunks, p(root), m, p(b))
            assert rm == rt and (rm != 0 or np.array_equal(a, b))
            if rm:
                continue
            for trial in range(6):
                c = a.copy()
                if trial:
                    for _ in range(int(rng.integers(1, 4))):
                        c[int(rng.integers(0, 472))] = int(rng.integers(0, 256))
                outs = [(np.zeros((chunks, 32), np.uint8), np.zeros(chunks, np.uint32), np.zeros(chunks, np.uint8))
                        for _ in range(2)]
                r1 = mine.lizx_xfer_manifest_parse(p(c), C.byref(h), m, *[p(x) for x in outs[0]])
                r2 = theirs.xfer_manifest_parse(p(c), C.byref(h), m, *[p(x) for x in outs[1]])
                assert r1 == r2, (trial, r1, r2)
                for x, y in zip(*outs):
                    assert np.array_equal(x, y)
                if trial == 0:
                    assert r1 == 0
                    first, n = 12 * m, min(12, chunks - 12 * m)
                    assert np.array_equal(outs[0][0][first:first + n], cvs[first:first + n])
