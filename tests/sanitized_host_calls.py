"""The library's host entries driven hard, for a run under the host sanitizers (tests/test_host_sanitized.py runs this
script with the sanitized build, lizard_amd/_san/, and the sanitizers' runtime preloaded): the replay reader on
truncated, bit-flipped and ill-formed archives and meta.json files; the format word's decoder on random and
adversarial soft words; the registration model's fills on random borders and anchors; the format arithmetic, the
encoder and CRC, the test stream and the host BLAKE3. Every call must return or raise LizError; nothing may read or
write out of bounds or hit undefined behaviour. Exits 0 when every call came back."""
import json
import os
import sys
import tempfile

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import lizard_amd as L  # noqa: E402


def replays(rng, tmp):
    frames = [rng.integers(0, 256, (int(rng.integers(1, 40)), int(rng.integers(1, 60)))).astype(np.uint8)
              for _ in range(4)]
    good = os.path.join(tmp, "good.lizard")
    w = L.ReplayWriter(good)
    for f in frames:
        w.add(f)
    w.finish(json.dumps({"frames": len(frames), "sizes": [[f.shape[1], f.shape[0]] for f in frames]}), '{"a": 1}\n')
    raw = bytearray(open(good, "rb").read())
    metas = [{"frames": 1e300}, {"frames": -0.0, "w": 3, "h": 3}, {"sizes": [[1e30, 5]]}, {"sizes": [[5]]},
             {"sizes": "x"}, {"frames": 2, "w": "a", "h": 2}, {"frames": 3, "sizes": [[1, 1], None, [2, 2]]},
             {"frames": 65536, "w": 1, "h": 1}, {"w": float("nan"), "h": 1, "frames": 1}]
    cases = [raw[:k] for k in (0, 1, 21, 22, 60, len(raw) // 2, len(raw) - 1)]
    for _ in range(300):
        b = bytearray(raw)
        for _ in range(int(rng.integers(1, 6))):
            b[int(rng.integers(0, len(b)))] ^= int(rng.integers(1, 256))
        cases.append(b)
    for k, m in enumerate(metas):
        p = os.path.join(tmp, "m%d.lizard" % k)
        w = L.ReplayWriter(p)
        w.add(frames[0])
        w.finish(json.dumps(m).replace("NaN", "1e999"))
        cases.append(bytearray(open(p, "rb").read()))
    seen = 0
    for k, data in enumerate(cases):
        p = os.path.join(tmp, "c%d.lizard" % k)
        open(p, "wb").write(bytes(data))
        try:
            r = L.Replay(p)
            for i in range(len(r)):
                r.info(i)
                r.frame(i)
            r.meta(0), r.meta(1), r.meta(2)
            r.close()
            seen += 1
        except L.LizError:
            pass
    return seen


def words(rng):
    n = 0
    for span in (128, 256, 512):
        cells, nbytes = L.word_geometry(span)
        for k in range(200):
            q = rng.uniform(-0.5, 0.5, 4 * cells).astype(np.float32)
            if k % 4 == 1:
                q = np.sign(q).astype(np.float32) * 0.5
            elif k % 4 == 2:
                q[:] = 0
            elif k % 4 == 3:
                L.word_encode(span, int(rng.integers(4, 65)), int(rng.integers(0, 256)))
                q = rng.normal(0, 0.3, 4 * cells).astype(np.float32)
            if L.word_decode(span, q) is not None:
                n += 1
    return n


def fills(rng):
    n = 0
    for span in (128, 256, 512):
        pos = L.node_lattice(span)
        nx = len(pos)
        for k in range(20):
            H = np.array([rng.uniform(1, 4), rng.uniform(-0.3, 0.3), rng.uniform(0, 200), rng.uniform(-0.3, 0.3),
                          rng.uniform(1, 4), rng.uniform(0, 200), rng.uniform(-1e-4, 1e-4), rng.uniform(-1e-4, 1e-4),
                          1.0], np.float32)
            nodes = rng.normal(0, [0.01, 1, 30][k % 3], (nx, nx, 2)).astype(np.float32)
            use = (rng.random((nx, nx)) < [1.0, 0.5, 0.0][k % 3]).astype(np.uint8)
            an = np.zeros((49, 5), np.float32)
            an[:, 0] = rng.uniform(0, span + 30, 49)
            an[:, 1] = rng.uniform(0, span + 30, 49)
            an[:, 2:4] = rng.normal(0, 2, (49, 2))
            an[:, 4] = rng.random(49) < [0.9, 0.3, 0.0][k % 3]
            try:
                L.border_fill(span, H, 1920, 1080, nodes, use=use)
                L.anchor_fill(span, H, 1920, 1080, nodes, an, float(rng.uniform(0.5, 8)), use=use)
                n += 1
            except L.LizError:
                pass
    return n


def misc(rng):
    for r in (2, 3, 4, 6):
        k = L.ldpc_dims(r)["k"]
        data = rng.integers(0, 2, (3, k)).astype(np.uint8)
        L.ldpc_encode(r, data)
    L.crc32c(rng.integers(0, 256, 1000).astype(np.uint8))
    for b in (1, 7, 64, 200):
        for ring in (-1, 0, 1, 2):
            try:
                L.geometry_of(b, ring)
                L.room_for(b, ring)
                L.pick(1920.0, 1080.0, 1, ring)
            except L.LizError:
                pass
    for i in (0, 1, 2 ** 31, 2 ** 32 - 1):
        L.stream_fill(i)
    for n in (0, 1, 1023, 1024, 1025, 65537):
        L.b3_hash(rng.integers(0, 256, n).astype(np.uint8))
    for w, h in ((1, 1), (1920, 1080), (100, 4000)):
        for lay in (0, 1, 2):
            try:
                L.layout_rects(w, h, lay)
            except L.LizError:
                pass


def xfer_blocks(rng):
    """Random and half-valid 472-byte payloads through the transfer's header and manifest readers (untrusted input)."""
    import ctypes as C
    lib = L.lib()
    h = (C.c_uint8 * 512)()  # room for lizx_xfer_header (472 bytes)
    # outputs for the most chunks a header can name (16,383), whatever the parsed header says
    cvs, sent, seeds = np.zeros((16383, 32), np.uint8), np.zeros(16383, np.uint32), np.zeros(16383, np.uint8)
    n = 0
    for _ in range(300):
        pay = rng.integers(0, 256, 472).astype(np.uint8)
        if rng.random() < 0.5:
            pay[0], pay[1], pay[7] = 3, 1, 0
            pay[469:] = 0
        lib.lizx_xfer_header_parse(pay.ctypes.data_as(C.c_void_p), h)
        lib.lizx_xfer_manifest_parse(pay.ctypes.data_as(C.c_void_p), h, C.c_uint32(int(rng.integers(0, 2000))),
                                     cvs.ctypes.data_as(C.c_void_p), sent.ctypes.data_as(C.c_void_p),
                                     seeds.ctypes.data_as(C.c_void_p))
        n += 1
    return n


def main():
    rng = np.random.default_rng(int(sys.argv[1]) if len(sys.argv) > 1 else 1)
    xfer_blocks(rng)
    with tempfile.TemporaryDirectory() as tmp:
        r = replays(rng, tmp)
    w = words(rng)
    f = fills(rng)
    misc(rng)
    print("host calls done: %d replays read, %d words decoded, %d fills" % (r, w, f))


main()
