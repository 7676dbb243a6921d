"""Replays (lizx_replay_*; research/SPEC.md 15, liblizard/core/rx/lizfile.h): the archive as Python's own zipfile
reads it (stored entries, their CRC-32s, the comment, the dates, the entries' names and order), each frame's entry
against the median predictor's residuals computed here from the SPEC's rule (decompressed through the same system
zstd), round trips bit for bit, and the refusals the SPEC asks for. The reference's own reader (lizfile.cpp) needs
zstd's headers, which this image lacks, so these checks pin the format instead; and, on the GPU, a replay of painted
captures read back and decoded with nothing held."""
import ctypes as C
import json
import zipfile

import numpy as np
import pytest

import lizard_amd as L


def residuals(img):
    """The SPEC's r = (p - pred) mod 256, pred the median edge predictor over a (left), b (up), c (up-left), 0 outside."""
    p = img.astype(np.int32)
    a = np.zeros_like(p)
    b = np.zeros_like(p)
    c = np.zeros_like(p)
    a[:, 1:] = p[:, :-1]
    b[1:, :] = p[:-1, :]
    c[1:, 1:] = p[:-1, :-1]
    mx, mn = np.maximum(a, b), np.minimum(a, b)
    pred = np.where(c >= mx, mn, np.where(c <= mn, mx, a + b - c))
    return ((p - pred) % 256).astype(np.uint8)


def zstd_decompress(z, n):
    lib = C.CDLL("libzstd.so.1")
    lib.ZSTD_decompress.restype = C.c_size_t
    lib.ZSTD_decompress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]
    out = C.create_string_buffer(n)
    got = lib.ZSTD_decompress(out, n, z, len(z))
    assert got == n
    return np.frombuffer(out.raw, np.uint8)


def frames_for_test(rng):
    yield rng.integers(0, 256, (37, 53)).astype(np.uint8)  # noise
    yield np.tile(np.arange(200, dtype=np.uint8), (120, 1))  # a ramp
    yield np.full((1, 1), 7, np.uint8)  # one pixel
    yield (np.add.outer(np.arange(90), np.arange(160)) % 256).astype(np.uint8)
    yield np.zeros((64, 3), np.uint8)


def write_replay(path, frames, stats=""):
    w = L.ReplayWriter(path)
    for f in frames:
        w.add(f)
    meta = {"frames": len(frames), "w": int(frames[0].shape[1]), "h": int(frames[0].shape[0]),
            "sizes": [[int(f.shape[1]), int(f.shape[0])] for f in frames], "ms": [10.0 * i for i in range(len(frames))],
            "layout": "1:1", "run": "test"}
    n = w.finish(json.dumps(meta), stats)
    return n, meta


def test_round_trip_and_zip_structure(tmp_path):
    rng = np.random.default_rng(1)
    frames = list(frames_for_test(rng))
    path = str(tmp_path / "run.lizard")
    n, meta = write_replay(path, frames, stats='{"t": 1}\n')
    assert n == (tmp_path / "run.lizard").stat().st_size and not (tmp_path / "run.lizard.part").exists()
    with zipfile.ZipFile(path) as z:
        assert z.comment == b"LIZARD replay 1"
        assert z.testzip() is None  # every CRC-32 holds
        names = z.namelist()
        assert names == ["%04d.zst" % i for i in range(len(frames))] + ["stats.jsonl", "meta.json"]
        for zi in z.infolist():
            assert zi.compress_type == zipfile.ZIP_STORED and zi.date_time == (1980, 1, 1, 0, 0, 0)
        for i, f in enumerate(frames):
            got = zstd_decompress(z.read("%04d.zst" % i), f.size).reshape(f.shape)
            assert np.array_equal(got, residuals(f)), i
        assert json.loads(z.read("meta.json")) == meta
    r = L.Replay(path)
    assert len(r) == len(frames) and r.meta(2) == "1:1" and r.meta(1) == '{"t": 1}\n'
    for i, f in enumerate(frames):
        assert r.info(i) == (f.shape[1], f.shape[0], 10.0 * i)
        assert np.array_equal(r.frame(i), f)
    r.close()


def test_padded_rows_and_sizes_from_w_h(tmp_path):
    base = np.random.default_rng(2).integers(0, 256, (40, 70)).astype(np.uint8)
    padded = np.zeros((40, 96), np.uint8)
    padded[:, :70] = base
    path = str(tmp_path / "p.lizard")
    w = L.ReplayWriter(path)
    w.add(padded[:, :70])  # rows 96 bytes apart
    w.add(base)
    w.finish(json.dumps({"frames": 2, "w": 70, "h": 40}))
    r = L.Replay(path)
    assert np.array_equal(r.frame(0), base) and np.array_equal(r.frame(1), base)
    assert np.isnan(r.info(0)[2]) and r.meta(2) == "1:1"
    r.close()


def test_refusals(tmp_path):
    rng = np.random.default_rng(3)
    frames = [rng.integers(0, 256, (20, 30)).astype(np.uint8) for _ in range(3)]
    good = str(tmp_path / "g.lizard")
    write_replay(good, frames)
    raw = bytearray(open(good, "rb").read())

    def refused(data, name):
        p = str(tmp_path / name)
        open(p, "wb").write(bytes(data))
        with pytest.raises(L.LizError):
            r = L.Replay(p)
            for i in range(len(r)):
                r.frame(i)

    refused(raw.replace(b"LIZARD replay 1", b"LIZARD replay 2"), "v2.lizard")  # a version not read
    refused(raw.replace(b"LIZARD replay 1", b"NOTANY replay 1"), "other.lizard")  # not a replay
    bad = bytearray(raw)
    bad[40] ^= 0xFF  # inside frame 0's entry: its CRC-32 no longer holds
    refused(bad, "crc.lizard")
    refused(raw[:len(raw) // 2], "short.lizard")
    # meta.json naming a frame the archive lacks
    p = str(tmp_path / "m.lizard")
    w = L.ReplayWriter(p)
    w.add(frames[0])
    w.finish(json.dumps({"frames": 2, "w": 30, "h": 20}))
    with pytest.raises(L.LizError):
        L.Replay(p)
    # no meta.json at all: a plain ZIP with the comment
    p = str(tmp_path / "nometa.lizard")
    with zipfile.ZipFile(p, "w", zipfile.ZIP_STORED) as z:
        z.writestr("0000.zst", b"x")
        z.comment = b"LIZARD replay 1"
    with pytest.raises(L.LizError):
        L.Replay(p)
    with pytest.raises(L.LizError):
        L.Replay(str(tmp_path / "missing.lizard"))


def test_frame_count_from_sizes_and_sizes_range_checked(tmp_path):
    """As the reference's reader takes meta.json (liblizard/core/rx/lizfile.cpp:200-212 framesOf): with no "frames" the
    count is the number of "sizes"; a size that is negative, fractional, zero or past 16384 is refused (it is checked
    before any conversion)."""
    rng = np.random.default_rng(4)
    frames = [rng.integers(0, 256, (20, 30)).astype(np.uint8), rng.integers(0, 256, (11, 17)).astype(np.uint8)]

    def made(name, meta):
        p = str(tmp_path / name)
        w = L.ReplayWriter(p)
        for f in frames:
            w.add(f)
        w.finish(json.dumps(meta))
        return p

    r = L.Replay(made("nocount.lizard", {"sizes": [[30, 20], [17, 11]]}))
    assert len(r) == 2 and r.info(1)[:2] == (17, 11)
    assert np.array_equal(r.frame(1), frames[1])
    r.close()
    for k, meta in enumerate([{"sizes": [[30, 20], [-17, 11]]}, {"sizes": [[30, 20], [17.5, 11]]},
                              {"frames": 2, "w": 30, "h": 0}, {"frames": 2, "w": 4294967326.0, "h": 20},
                              {"frames": -1, "w": 30, "h": 20}, {"frames": 1.5, "w": 30, "h": 20},
                              {"w": 30, "h": 20}]):
        with pytest.raises(L.LizError):
            L.Replay(made("bad%d.lizard" % k, meta))


@pytest.mark.gpu
def test_replay_of_painted_captures_decodes_blind(tmp_path, decoder_gpu):
    """Captures of symbols painted on the GPU (lizx_capture_symbol) packed into a replay, read back frame by frame and
    decoded with nothing held: each frame's blocks, in order."""
    from lizard_amd import sim
    rng = np.random.default_rng(12)
    path = str(tmp_path / "caps.lizard")
    w = L.ReplayWriter(path)
    sent = []
    for k in range(3):
        blocks = sim.random_blocks(20, rng)
        dpic = decoder_gpu.modulate(20, blocks, cnt=k % 4)
        cam = sim.Camera(1920, 1080, rot_deg=float(rng.uniform(-8, 8)), scale=0.85, persp=0.02, k1=0.02, blur=0.8,
                         noise=3.0)
        dimg = decoder_gpu.capture(20, 256, dpic, 1920, 1080, cam.H(256), k1=cam.k1, blur=cam.blur, noise=cam.noise,
                                   seed=k, fps=30)
        w.add(dimg.get(np.uint8, 1920 * 1080).reshape(1080, 1920))
        dimg.free()
        dpic.free()
        sent.append(blocks)
    w.finish(json.dumps({"frames": 3, "w": 1920, "h": 1080, "layout": "1:1"}))
    r = L.Replay(path)
    dec = L.Decoder(0)
    for k in range(3):
        n, got, out, held = dec.decode(r.frame(k))
        assert out.found == 1 and out.word == 1 and out.blocks == 20
        assert n == L.format_info(20)["blocks"] and np.array_equal(got.ravel(), sent[k].ravel())
    dec.close()
    r.close()
