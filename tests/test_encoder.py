"""liblizard's format arithmetic and per-frame encoder (liz_ring_cells, liz_geometry_of, liz_room_for, liz_blocks_for,
liz_pick; liz_encoder_*): the geometry against the reference codec's own lizard_init for every size and ring, the room
against the value the reference states (LIZARD-64 in the default ring needs 349.2 px, liblizard/sim/lizard_pick.mjs:257)
and its square-root law; on the GPU, the page against the reference encoder's page (lizard_encode +
lizard_paint_grey) from the same blocks: the border, word, marks, anchors and quiet zone to the grey level, the
resampled picture but for rounding ties, Lanczos-3 included."""
import numpy as np
import pytest

import lizard_amd as L
from _libs import ptr, ref


def test_ring_cells_and_blocks_for():
    assert [L.ring_cells(r) for r in (0, 1, 2, -1)] == [64, 128, 256, 128]
    with pytest.raises(L.LizError):
        L.ring_cells(3)
    for b in range(4, 65):
        assert L.blocks_for(b) == L.format_info(b)["blocks"]
    assert L.blocks_for(3) == 0 and L.blocks_for(65) == 0


@pytest.mark.skipif(ref("scalar") is None, reason="oracle/_ref not built")
def test_geometry_matches_reference_codec():
    r = ref("scalar")
    info = np.zeros(10, np.int32)
    for ring in (0, 1, 2):
        for b in range(4, 65):
            g = L.geometry_of(b, ring, fps=30, codes=1)
            f = r.ref_format_new(g["n"], 16 * b, g["span"])
            assert f
            r.ref_format_info(f, ptr(info))
            r.ref_format_free(f)
            assert (g["n"], g["pxm"], g["frame_blocks"]) == (info[0], info[7], info[2])
            assert g["side"] == info[8] + 4 * info[7] == g["width"] == g["height"]
            g2 = L.geometry_of(b, ring, fps=30, codes=2)
            assert g2["gap"] == 12 * g["pxm"] and g2["width"] == 2 * g["side"] + g2["gap"]
            assert g2["frame_blocks"] == 2 * g["frame_blocks"]
    for bad in (dict(blocks=3), dict(blocks=65), dict(blocks=8, fps=0), dict(blocks=8, codes=3), dict(blocks=8, ring=3)):
        kw = dict(blocks=8, ring=-1, fps=30, codes=1)
        kw.update(bad)
        with pytest.raises(L.LizError):
            L.geometry_of(**kw)


def test_room_and_pick():
    assert abs(L.room_for(4) - 349.2) < 0.05
    for ring in (0, 1, 2):
        for b in range(4, 17):
            assert abs(L.room_for(4 * b, ring) / L.room_for(b, ring) - 2.0) < 1e-12  # the radius goes as root size
    # the rings differ only by their border's share: (span + 34) / span
    assert abs(L.room_for(20, 0) / L.room_for(20, 2) - (162 / 128) / (546 / 512)) < 1e-12
    for codes in (1, 2):
        for ring in (0, 1, 2):
            for top in (4, 26, 64):
                last = 0
                for wh in range(200, 4200, 173):
                    p = L.pick(wh, wh, codes, ring, top)
                    assert 4 <= p <= top and p >= last  # more room never picks less
                    last = p
                    gap = (codes - 1) * 12 / (2 * L.ring_cells(ring) + 34)
                    room = max(64.0, min(wh / (codes + gap), wh))
                    assert L.room_for(p, ring) <= room or p == 4
                    if p < top:
                        assert L.room_for(p + 1, ring) > room
    assert L.pick(1e6, 1e6, 1, -1, 64) == 64 and L.pick(50, 50) == 4


PAGES = [(5, 0), (20, 1), (64, 2), (45, 2), (5, 1), (64, 0), (20, 0), (45, 1)]


def reference_page(blocks_n, ring, fps, blocks, picture):
    r = ref("scalar")
    g = L.geometry_of(blocks_n, ring, fps)
    f = r.ref_format_new(g["n"], 16 * blocks_n, g["span"])
    assert r.ref_set_fps(f, fps) == 0
    info = np.zeros(10, np.int32)
    r.ref_format_info(f, ptr(info))
    px, pxm = int(info[8]), int(info[7])
    drive = np.zeros(px * px, np.float32)
    assert r.ref_encode(f, ptr(blocks), picture & 3, ptr(drive), None) == 0
    side = px + 4 * pxm
    page = np.zeros(side * side, np.uint8)
    r.ref_paint_grey(f, ptr(drive), ptr(page))
    r.ref_format_free(f)
    return page.reshape(side, side)


@pytest.mark.gpu
@pytest.mark.skipif(ref("scalar") is None, reason="oracle/_ref not built")
@pytest.mark.parametrize("blocks_n,ring", PAGES)
def test_encoder_page_matches_reference(blocks_n, ring):
    fps, picture = 30, 5 + blocks_n
    g = L.geometry_of(blocks_n, ring, fps)
    blocks = np.random.default_rng(blocks_n * 3 + ring).integers(0, 256, g["frame_blocks"] * 476).astype(np.uint8)
    want = reference_page(blocks_n, ring, fps, blocks, picture)
    enc = L.Encoder(blocks_n, ring, fps)
    got = enc.paint(blocks, picture)
    enc.close()
    assert got.shape == want.shape == (g["side"], g["side"])
    pxm, L_ = g["pxm"], g["span"] + 30
    kinds = L.border_kinds(g["span"], blocks_n, fps)
    module = np.full((g["side"], g["side"]), 255, np.int32)
    module[2 * pxm:g["side"] - 2 * pxm, 2 * pxm:g["side"] - 2 * pxm] = np.repeat(np.repeat(kinds, pxm, 0), pxm, 1)
    fixed = module != 0
    assert np.array_equal(got[fixed], want[fixed])
    d = np.abs(got.astype(np.int32) - want)
    assert d.max() <= 1 and (d > 0).mean() < 2e-3, (d.max(), (d > 0).mean())
    # the page reads back: its word names the format and the rate
    H = np.array([[pxm, 0, 2 * pxm], [0, pxm, 2 * pxm], [0, 0, 1]], np.float32).ravel()
    nx = len(L.node_lattice(g["span"]))
    assert L.read_word(g["span"], got, H, np.zeros((nx, nx, 2), np.float32)) == (blocks_n, fps)


@pytest.mark.gpu
def test_encoder_two_codes_and_rgba():
    g = L.geometry_of(8, 0, 24, codes=2)
    blocks = np.random.default_rng(2).integers(0, 256, g["frame_blocks"] * 476).astype(np.uint8)
    two = L.Encoder(8, 0, 24, codes=2)
    page = two.paint(blocks, 1)
    rgba = two.paint(blocks, 1, fmt=3)
    two.close()
    one = L.Encoder(8, 0, 24)
    half = g["frame_blocks"] // 2 * 476
    a = one.paint(blocks[:half], 1)
    b = one.paint(blocks[half:], 1)
    one.close()
    side, gap = g["side"], g["gap"]
    assert page.shape == (side, 2 * side + gap)
    assert np.array_equal(page[:, :side], a) and np.array_equal(page[:, side + gap:], b)
    assert (page[:, side:side + gap] == page[:, side - 1:side]).all()  # the gap takes the row's last pixel
    assert np.array_equal(rgba[:, :, 0], page) and np.array_equal(rgba[:, :, 2], page) and (rgba[:, :, 3] == 255).all()
    with pytest.raises(L.LizError):
        L.Encoder(8, 0, 24, codes=3)


def test_layout_rects():
    assert L.layout_rects(1920, 1080) == [(420, 0, 1080, 1080)]
    assert L.layout_rects(1080, 1920) == [(0, 420, 1080, 1080)]
    assert L.layout_rects(1920, 1080, 2) == [(0, 60, 960, 960), (960, 60, 960, 960)]
    assert L.layout_rects(2000, 800, 2) == [(200, 0, 800, 800), (1000, 0, 800, 800)]
    assert L.layout_rects(801, 2001, 2) == [(0, 199, 801, 801), (0, 1000, 801, 801)]  # (2001 - 1602) // 2
    assert L.layout_rects(7, 7) == [(0, 0, 7, 7)]
    for bad in ((0, 10, 1), (10, 10, 3)):
        with pytest.raises(L.LizError):
            L.layout_rects(*bad)


@pytest.mark.gpu
def test_two_codes_through_crops_and_blind_decode():
    """A sender's two codes (liz_encoder_paint, codes 2) shown in a wide frame, cut by liz_layout_rects and each crop
    decoded with nothing held (liz_decode finds, reads the word, decodes): every block of both codes, in order."""
    g = L.geometry_of(8, 0, 30, codes=2)
    blocks = np.random.default_rng(6).integers(0, 256, g["frame_blocks"] * 476).astype(np.uint8)
    enc = L.Encoder(8, 0, 30, codes=2)
    page = enc.paint(blocks, 0)
    enc.close()
    H, W = 2 * g["side"] // 2 + 200, 2 * g["side"] + 400
    frame = np.full((H, W), 255, np.uint8)
    y0, x0 = (H - g["height"]) // 2, (W - g["width"]) // 2
    frame[y0:y0 + g["height"], x0:x0 + g["width"]] = page
    rects = L.layout_rects(W, H, 2)
    assert len(rects) == 2
    dec = L.Decoder(1536)
    got = []
    for x, y, w, h in rects:
        k, b, out, held = dec.decode(frame[y:y + h, x:x + w])
        assert out.found == 1 and out.word == 1 and out.blocks == 8 and out.fps == 30
        got.append(b)
    dec.close()
    half = g["frame_blocks"] // 2
    assert np.array_equal(np.concatenate(got).ravel(), blocks)
    assert all(len(b) == half for b in got)
