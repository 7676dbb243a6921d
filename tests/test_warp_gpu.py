"""The warp on the GPU (lizx_warp) against the reference's picture sampler compiled from its sources: the scalar
grid walk to the bit, the x86-64 integer read within 1/200 of the grey range (liblizard/test/sample_grid.c:137); and
the whole decode of a capture through its registration (lizx_decode_reg, liz_decode with a held registration)
against the reference's lizard_finish on the same capture and registration: soft values equal but at rounding ties,
and the reference's integer path on the GPU's own soft values gives the GPU's verdicts, iterations and bytes."""
import numpy as np
import pytest

import lizard_amd
from _libs import ptr, ref
from pictures import Capture, assert_finish_equal, ref_finish_soft, ref_grid_finish, ties_against_reference

pytestmark = pytest.mark.gpu


def wavy(scale):
    def f(node_x):
        X, Y = np.meshgrid(node_x, node_x, indexing="xy")
        return np.stack([scale * np.sin(X / 23.0) * np.cos(Y / 31.0), scale * np.cos(X / 17.0 + Y / 29.0)], -1)
    return f


WARPS = [(5, 128, 0.0), (5, 128, 1.5), (20, 256, 0.7), (64, 512, 1.0), (45, 512, 0.0), (8, 256, 2.0)]


@pytest.mark.parametrize("version,span,wave", WARPS)
def test_warp_matches_reference_sampler(decoder_gpu, version, span, wave):
    C = Capture(version, span, seed=version, W=1920 if span == 512 else 1280, Hh=1080 if span == 512 else 720,
                node_fn=wavy(wave) if wave else None)
    grid = decoder_gpu.warp(version, span, C.img, C.H, C.nodes)
    r = ref("scalar")
    if r is None:
        pytest.skip("oracle/_ref not built")
    f = r.ref_format_new(C.n, C.subch, span)
    want = np.zeros((C.n, C.n), np.float32)
    r.ref_sample_grid(f, ptr(C.img), C.img.shape[1], C.img.shape[0], ptr(C.H), ptr(C.nodes), 0, ptr(want))
    r.ref_format_free(f)
    assert np.array_equal(grid, want), np.abs(grid - want).max()
    rx = ref("x86")
    if rx is not None:
        f = rx.ref_format_new(C.n, C.subch, span)
        lin = np.zeros((C.n, C.n), np.float32)
        rx.ref_sample_grid(f, ptr(C.img), C.img.shape[1], C.img.shape[0], ptr(C.H), ptr(C.nodes), 1, ptr(lin))
        rx.ref_format_free(f)
        assert np.abs(grid - lin).max() <= 1.0 / 200


def test_warp_clamps_outside_the_frame(decoder_gpu):
    """Points mapped far outside the capture read its edge pixels (the clamp), never out of bounds."""
    C = Capture(5, 128, seed=2)
    H = C.H.copy()
    H[2] += 5000.0  # the whole picture lands right of the frame
    grid = decoder_gpu.warp(5, 128, C.img, H, C.nodes)
    assert np.isfinite(grid).all() and grid.min() >= 0 and grid.max() <= 1


DECODES = [(5, 128, 1, 0.8, 4.0, 0.0), (4, 128, 2, 0.5, 2.0, 0.5), (20, 256, 3, 1.0, 6.0, 0.7), (45, 512, 4, 0.8, 3.0, 0.0)]


@pytest.mark.parametrize("version,span,seed,blur,noise,wave", DECODES)
def test_decode_reg_matches_reference(decoder_gpu, version, span, seed, blur, noise, wave):
    big = span == 512
    C = Capture(version, span, seed=seed, W=1920 if big else 1280, Hh=1080 if big else 720, blur=blur, noise=noise,
                node_fn=wavy(wave) if wave else None)
    out = decoder_gpu.decode_reg(version, span, C.img, C.H, C.nodes)
    B = C.blocks_n
    for b in range(B):
        if out["ok"][b]:
            assert np.array_equal(out["blocks"][b], C.blocks.reshape(B, 476)[b])
    r = ref("scalar")
    if r is None:
        return
    f = r.ref_format_new(C.n, C.subch, span)
    bl, ok, its, be = np.zeros(B * 476, np.uint8), np.zeros(B, np.uint8), np.zeros(B, np.int8), np.zeros(B, np.float32)
    llr, al = np.zeros(C.subch * 640, np.int8), np.zeros(2, np.float32)
    good = r.ref_decode_reg(f, ptr(C.img), C.img.shape[1], C.img.shape[0], ptr(C.H), ptr(C.nodes), ptr(bl), ptr(ok),
                            ptr(its), ptr(be), ptr(llr), ptr(al))
    r.ref_format_free(f)
    assert abs(out["align"][0] - al[0]) < 1e-3 and abs(out["align"][1] - al[1]) < 1e-3
    # the soft values differ from the reference's only at rounding ties (the reference's finish of its own sampler's
    # grid, which is the decode's, gives its coefficients)
    f = r.ref_format_new(C.n, C.subch, span)
    grid = np.zeros((C.n, C.n), np.float32)
    r.ref_sample_grid(f, ptr(C.img), C.img.shape[1], C.img.shape[0], ptr(C.H), ptr(C.nodes), 1, ptr(grid))
    r.ref_format_free(f)
    rg = ref_grid_finish(r, version, span, grid)
    assert np.array_equal(rg["llr"], llr)
    tie = ties_against_reference(version, out, rg)
    assert tie < 1e-3, tie
    if tie == 0:
        assert out["good"] == good and np.array_equal(out["its"], its) and np.array_equal(out["ok"], ok)
    # the reference's integer path on the GPU's own soft values: the same verdicts, iterations and bytes
    assert_finish_equal(out, ref_finish_soft(r, version, span, out["llr"], out["sub_est"]))


@pytest.mark.parametrize("fmt", [1, 4])
def test_liz_decode_held(fmt):
    """liblizard's liz_decode, through a held registration, on GREY8 and BGRA8 captures: every block verified, packed
    in block order (api/codec.cpp:143-147)."""
    C = Capture(5, 128, seed=9, blur=0.6, noise=2.0)
    dec = lizard_amd.Decoder(384)
    dec.hold(5, 128, C.H, C.nodes)
    img = C.img if fmt == 1 else np.repeat(C.img[:, :, None], 4, axis=2).copy()
    k, blocks, out, held = dec.decode(img, fmt=fmt)
    dec.close()
    assert k == C.blocks_n and out.verified == k and out.total == C.blocks_n and held == 5
    assert np.array_equal(blocks.ravel(), C.blocks)
    # the pilots (painted count 0): both parities read near +1 over the 7/8 and 3/4 blocks (lizard_pilot)
    rates = lizard_amd.format_blocks(5)[0]
    assert out.pilot_blocks == int(np.sum((rates == 6) | (rates == 4)))
    assert out.pilot_r[0] > 0.5 and out.pilot_r[1] > 0.5
    assert 0 <= out.pilot_sd[0] < 0.2 and 0 <= out.pilot_sd[1] < 0.2


def test_liz_decode_without_registration_finds_nothing_in_a_blank_frame():
    """No registration held: liz_decode looks for the symbol itself (lizx_find); a blank frame has none."""
    dec = lizard_amd.Decoder(384)
    k, blocks, out, held = dec.decode(np.zeros((10, 10), np.uint8))
    assert k == 0 and out.found == 0 and out.ring == -1 and held == 0
    k, blocks, out, held = dec.decode(np.full((200, 300), 255, np.uint8), held=7)
    assert k == 0 and out.found == 0 and held == 7
    dec.close()


@pytest.mark.parametrize("version,span,W,Hh,stale", [(5, 128, 1280, 720, 0), (20, 256, 1920, 1080, 64)])
def test_liz_decode_word_names_the_format(version, span, W, Hh, stale):
    """A hold that states no format: the frame's own painted word names it (liblizard/src/any.c:48-51), over a stale
    *held too; *held follows the word (api/codec.cpp:143), out->fps is the word's, and every block verifies."""
    C = Capture(version, span, W=W, Hh=Hh, scale=0.85, seed=17, blur=0.7, noise=3.0)
    want = lizard_amd.read_word(span, C.img, C.H, C.nodes)
    assert want is not None and want[0] == version
    dec = lizard_amd.Decoder(1536)
    dec.hold(0, span, C.H, C.nodes)
    k, blocks, out, held = dec.decode(C.img, held=stale)
    dec.close()
    assert out.found == 1 and out.word == 1 and out.held_used == 0
    assert out.blocks == version and held == version and out.fps == want[1]
    assert out.ring == {128: 0, 256: 1, 512: 2}[span] and out.n == lizard_amd.format_info(version)["n"]
    assert k == C.blocks_n == out.total == out.verified
    assert np.array_equal(blocks.ravel(), C.blocks)
    # the corner nodes in the image, top left, top right, bottom right, bottom left (nodes zero here: H alone)
    H = C.H.reshape(3, 3).astype(np.float64)
    x0, x1 = C.node_x[0], C.node_x[-1]
    for c, (mx, my) in enumerate([(x0, x0), (x1, x0), (x1, x1), (x0, x1)]):
        p = H @ np.array([mx, my, 1.0])
        assert abs(out.quad[2 * c] - p[0] / p[2]) < 1e-3 and abs(out.quad[2 * c + 1] - p[1] / p[2]) < 1e-3


def test_liz_decode_without_a_word():
    """Frames with no word band (a grey ring): decoded at *held when it names a format (held_used), else at the version
    the hold states, else not decoded; *held is never changed without a word (api/codec.cpp:143)."""
    from lizard_amd import sim
    rng = np.random.default_rng(4)
    blocks = sim.random_blocks(5, rng)
    img, H, nodes = sim.capture(sim.picture(5, blocks), 128, sim.Camera(1280, 720, blur=0.7, noise=3.0), seed=4)
    img = img.numpy()
    assert lizard_amd.read_word(128, img, H, nodes) is None
    dec = lizard_amd.Decoder(384)
    dec.hold(0, 128, H, nodes)
    k, got, out, held = dec.decode(img, held=5)
    assert out.word == 0 and out.held_used == 1 and out.blocks == 5 and held == 5 and out.fps == 0
    assert k == out.total and np.array_equal(got.ravel(), np.asarray(blocks).ravel())
    k, got, out, held = dec.decode(img, held=0)
    assert k == 0 and out.found == 1 and out.n == 0 and out.blocks == 0 and out.total == 0 and held == 0
    k, got, out, held = dec.decode(img, held=99)  # not a version: as none held
    assert k == 0 and out.n == 0 and held == 99
    dec.hold(5, 128, H, nodes)
    k, got, out, held = dec.decode(img, held=0)
    assert out.word == 0 and out.held_used == 0 and out.blocks == 5 and held == 0 and k == out.total
    # a word-less frame at a picture beyond this decoder's nmax is not decoded
    dec.hold(0, 128, H, nodes)
    k, got, out, held = dec.decode(img, held=64)
    assert k == 0 and out.n == 0 and out.held_used == 1
    dec.close()


def test_single_frame_graph_replays_equal_eager(decoder_gpu):
    """The single-frame decode's kernels replayed as a captured graph (the default) give the launch-by-launch
    decode's bits on every replay, with other eager work of the library between replays, across two formats (each
    change of format or work area captures anew) and back."""
    caps = [Capture(5, 128, seed=31, blur=0.7, noise=3.0), Capture(5, 128, seed=32, blur=0.9, noise=6.0),
            Capture(20, 256, seed=33, W=1280, Hh=720, scale=0.9, blur=0.7, noise=4.0)]
    dec = lizard_amd.Decoder(1536)
    try:
        dec.graphs(False)
        want = [dec.decode_reg(c.version, c.span, c.img, c.H, c.nodes, soft=False) for c in caps]
        dec.graphs(True)
        for rnd in range(3):
            for i in (0, 1, 2, 0):
                c = caps[i]
                got = dec.decode_reg(c.version, c.span, c.img, c.H, c.nodes, soft=False)
                for k in ("its", "ok", "est"):
                    assert np.array_equal(got[k], want[i][k]), (rnd, i, k)
                v = want[i]["ok"].astype(bool)
                assert np.array_equal(got["blocks"][v], want[i]["blocks"][v])
                # eager work of the library between replays: a grid decode with its own launches
                decoder_gpu.decode_reg(caps[1].version, caps[1].span, caps[1].img, caps[1].H, caps[1].nodes)
        # liz_decode through a hold takes the same graph path
        dec.hold(5, 128, caps[0].H, caps[0].nodes)
        k, blocks, out, held = dec.decode(caps[0].img)
        assert k == int(want[0]["ok"].sum())
        assert np.array_equal(blocks, want[0]["blocks"][want[0]["ok"].astype(bool)])
    finally:
        dec.close()


def test_single_frame_shapes_equal(decoder_gpu):
    """Every single-frame launch shape (waves a workgroup of the row FFT, column FFT and demap; lizx_decoder_shape)
    gives the same soft values, verdicts, iterations, estimates and bytes, graphs on and off; a shape outside 1, 2, 4
    is refused."""
    c = Capture(20, 256, seed=35, W=1280, Hh=720, scale=0.9, blur=0.8, noise=5.0)
    dec = lizard_amd.Decoder(1536)
    try:
        want = None
        for graphs in (True, False):
            dec.graphs(graphs)
            for sh in ((4, 4, 4), (1, 2, 1), (1, 1, 1), (2, 4, 2), (4, 1, 2)):
                dec.shape(*sh)
                got = dec.decode_reg(c.version, c.span, c.img, c.H, c.nodes, soft=True)
                if want is None:
                    want = got
                    assert want["ok"].sum() > 0
                    continue
                for k in ("its", "ok", "est", "llr"):
                    assert np.array_equal(got[k], want[k]), (graphs, sh, k)
                v = want["ok"].astype(bool)
                assert np.array_equal(got["blocks"][v], want["blocks"][v]), (graphs, sh)
        with pytest.raises(lizard_amd.LizError):
            dec.shape(3, 1, 1)
    finally:
        dec.close()


def test_sampler_math_equals_plain_forms():
    """The warp sampler's cheap float forms (lizard_amd/csrc/sample_math.hpp) equal the plain operators bit for bit on
    the device: 1/d by v_rcp and one Newton step against the IEEE quotient for all 2^32 float patterns (the forms
    split at |d| in [2^-126, 2^126)), x - (int) x as v_fract for every float in [0, 2^31), g / 255 for all 256
    levels."""
    import ctypes as C
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "devcheck", "libsample_math.so")
    assert os.path.exists(path), "build the device check first (python __graft_entry__.py build)"
    lib = C.CDLL(path)
    counts, first = (C.c_ulonglong * 3)(), (C.c_uint32 * 12)()
    assert lib.lz_check_sample_math(counts, first) == 0
    assert list(counts) == [0, 0, 0], (list(counts), ["%08x" % v for v in first])


def test_decode_padded_rows_equal_tight():
    """liz_decode and lizx_decode_reg on a capture whose rows are padded (stride > w bytes apart, the cropped upload
    gathering the box's rows) give the tight capture's verdicts and bytes; a capture of 2^31 pixels or more is refused
    before anything is read."""
    import ctypes as C
    c = Capture(5, 128, seed=36, blur=0.8, noise=4.0)
    dec = lizard_amd.Decoder(384)
    try:
        want = dec.decode_reg(c.version, c.span, c.img, c.H, c.nodes, soft=False)
        assert want["ok"].sum() > 0
        h, w = c.img.shape
        pad = np.full((h, w + 77), 201, np.uint8)
        pad[:, :w] = c.img
        L = lizard_amd.lib()
        reg = lizard_amd.registration(c.H, c.nodes)
        B = want["ok"].size
        blocks, ok = np.zeros(B * 476, np.uint8), np.zeros(B, np.uint8)
        its, be = np.zeros(B, np.int8), np.zeros(B, np.float32)
        good = L.lizx_decode_reg(dec.h, c.version, c.span, ptr(pad), w, h, w + 77, C.byref(reg), ptr(blocks), ptr(ok),
                                 ptr(its), ptr(be), None)
        assert good == int(want["ok"].sum())
        assert np.array_equal(ok, want["ok"]) and np.array_equal(its, want["its"]) and np.array_equal(be, want["est"])
        v = ok.astype(bool)
        assert np.array_equal(blocks.reshape(B, 476)[v], want["blocks"][v])
        # liz_decode through a hold on the padded rows
        dec.hold(c.version, c.span, c.H, c.nodes)
        ver = np.zeros(dec.max_blocks() * 476, np.uint8)
        held, out = C.c_int(0), lizard_amd._Decoded()
        k = L.liz_decode(dec.h, ptr(pad), w, h, w + 77, 1, C.byref(held), ptr(ver), C.byref(out))
        assert k == good and np.array_equal(ver[:k * 476].reshape(k, 476), want["blocks"][v])
        # 65536 x 32768 pixels claimed over a small buffer: refused by its size alone
        rc = L.lizx_decode_reg(dec.h, c.version, c.span, ptr(pad), 65536, 32768, 65536, C.byref(reg), ptr(blocks),
                               ptr(ok), ptr(its), ptr(be), None)
        assert rc == -1 and b"2^31" in L.liz_last_error()
    finally:
        dec.close()
