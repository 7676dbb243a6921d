"""The sender's picture on the GPU (lizx_modulate) against the FP64 modulator (oracle/picture.py; itself held to the
reference encoder's own picture in tests/test_oracle_ref.py), and the round trip: pictures made on the GPU decode
through the batch path to the sent blocks."""
import numpy as np
import pytest

import lizard_amd as L
from pictures import picture as opic

pytestmark = pytest.mark.gpu

SPAN = {4: 128, 5: 128, 13: 256, 20: 256, 45: 512, 64: 512}


def blocks_for(version, frames, rng):
    B = L.format_info(version)["blocks"]
    return rng.integers(0, 256, (frames, B * 476)).astype(np.uint8)


@pytest.mark.parametrize("version", [4, 5, 13, 20, 45, 64])
def test_pictures_match_fp64(decoder_gpu, version):
    rng = np.random.default_rng(100 + version)
    info = L.format_info(version)
    n, subch = info["n"], info["subch"]
    u, v = L.format_coefs(version)
    frames = 3
    blocks = blocks_for(version, frames, rng)
    for cnt in (0, 2):
        d = decoder_gpu.modulate(version, blocks, cnt)
        pics = d.get(np.float32, frames * n * n).reshape(frames, n, n)
        d.free()
        for f in range(frames):
            want = opic.modulate(n, subch, L.frame_slots(version, blocks[f], cnt).astype(np.float64), u, v)
            err = np.abs(pics[f] - want)
            assert err.max() < 2e-6, (version, cnt, f, err.max())
            assert pics[f].min() >= 0.0 and pics[f].max() <= 1.0


def test_gpu_pictures_decode_to_the_sent_blocks(decoder_gpu):
    """Four sizes of picture made on the GPU, with channel noise, decoded as grids in one batch: every block verified,
    every verified block the one sent."""
    rng = np.random.default_rng(7)
    versions = [5, 20, 45, 64]
    spans = [SPAN[v] for v in versions]
    sent, keep, items = [], [], []
    for ver in versions:
        info = L.format_info(ver)
        bl = blocks_for(ver, 1, rng)
        d = decoder_gpu.modulate(ver, bl)
        pic = d.get(np.float32, info["n"] ** 2)
        d.free()
        g = L.DeviceArray.of((pic + rng.standard_normal(pic.size) * 0.01).astype(np.float32))
        keep.append(g)
        items.append(dict(grid=g))
        sent.append(bl.reshape(-1, 476))
    b = L.Batch().plan(versions, spans, source=L.SRC_GRID, stages=L.STAGE_BLOCKS)
    b.run(L.make_frames(items), L.RUN_GRAPH)
    b.sync()
    out = b.blocks_out()
    b.close()
    want = np.concatenate(sent)
    assert out["ok"].all()
    assert np.array_equal(out["blocks"], want)
    for g in keep:
        g.free()


def test_modulate_argument_errors(decoder_gpu):
    B = L.format_info(5)["blocks"]
    d = L.DeviceArray(384 * 384 * 4)
    with pytest.raises(L.LizError):
        decoder_gpu.modulate(3, np.zeros(B * 476, np.uint8), out=d)
    with pytest.raises(L.LizError):
        decoder_gpu.modulate(5, np.zeros(B * 476, np.uint8), cnt=4, out=d)
    d.free()


@pytest.mark.parametrize("version", [4, 5, 13, 20, 45, 64])
def test_device_frame_slots_match_host(decoder_gpu, version):
    """The bit side on the GPU (CRC-32C, QC-IRA encoding by a prefix XOR across lanes, bit map, pilots, whitening)
    equals the host encoder's frame slots, themselves the oracle's and the reference encoder's (test_sim,
    test_oracle_ref)."""
    rng = np.random.default_rng(200 + version)
    info = L.format_info(version)
    frames, S2 = 3, info["subch"] * 640
    blocks = blocks_for(version, frames, rng)
    db = L.DeviceArray.of(blocks)
    ds = L.DeviceArray(frames * S2)
    for cnt in range(4):
        decoder_gpu.frame_slots_device(version, db, frames, cnt, ds)
        got = ds.get(np.uint8, frames * S2).reshape(frames, S2)
        for f in range(frames):
            assert np.array_equal(got[f], L.frame_slots(version, blocks[f], cnt)), (version, cnt, f)
    db.free()
    ds.free()
