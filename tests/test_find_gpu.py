"""The finder (lizx_find; liz_decode with no registration held) against the truth of the channel that made each
capture and against the reference's own finder (lizard_acquire, compiled from its sources) on the same frames: the
ring and orientation, the registration at the node lattice (as close to the truth as the reference's or closer),
the lens, and the blind decode (as many blocks as the reference's finder and decoder verify); frames with no symbol
are not found; a padded capture gives what the tight one does."""
import ctypes as C

import numpy as np
import pytest

import lizard_amd as L
from lizard_amd import sim
from _libs import ptr, ref
from pictures import Capture

pytestmark = pytest.mark.gpu


def node_err(H, nodes, Ht, nodes_t, node_x):
    X, Y = np.meshgrid(node_x, node_x)
    P = np.stack([X.ravel(), Y.ravel(), np.ones(X.size)])

    def at(Hm, nd):
        q = np.asarray(Hm, np.float64).reshape(3, 3) @ P
        return q[:2] / q[2] + np.asarray(nd, np.float64).reshape(-1, 2).T

    e = np.hypot(*(at(H, nodes) - at(Ht, nodes_t)))
    return float(e.mean()), float(e.max())


def reference_path(version, span, img):
    """The reference's finder and decoder on a capture: (registration error inputs, verified blocks)."""
    r = ref("x86")
    info = L.format_info(version)
    f = r.ref_format_new(info["n"], info["subch"], span)
    nx = len(L.node_lattice(span))
    H = np.zeros(9, np.float32)
    nodes = np.zeros(2 * nx * nx, np.float32)
    quad = np.zeros(8, np.float32)
    o = np.zeros(1, np.int32)
    h, w = img.shape
    got = r.ref_acquire(f, ptr(img), w, h, ptr(H), ptr(nodes), ptr(quad), ptr(o))
    verified = 0
    if got:
        B = info["blocks"]
        blocks = np.zeros(B * 476, np.uint8)
        ok = np.zeros(B, np.uint8)
        its = np.zeros(B, np.int8)
        est = np.zeros(B, np.float32)
        r.ref_decode_reg(f, ptr(img), w, h, ptr(H), ptr(nodes), ptr(blocks), ptr(ok), ptr(its), ptr(est), None, None)
        verified = int(ok.sum())
    r.ref_format_free(f)
    return (H, nodes) if got else None, verified


PAINTED = [(5, 128, 1280, 720, 5.0), (20, 256, 1920, 1080, -8.0), (64, 512, 1920, 1080, 12.0),
           (20, 256, 1920, 1080, 100.0), (8, 256, 1280, 720, 190.0), (45, 512, 1920, 1080, 275.0)]


@pytest.mark.parametrize("version,span,W,Hh,rot", PAINTED)
def test_finder_on_reference_painted_symbols(decoder_gpu, version, span, W, Hh, rot):
    C = Capture(version, span, W=W, Hh=Hh, rot_deg=rot, scale=0.85, blur=0.8, noise=4.0, seed=version)
    img = np.ascontiguousarray(C.img)
    f = decoder_gpu.find(img)
    assert f is not None and f["span"] == span
    mean, mx = node_err(f["H"], f["nodes"], C.H, C.nodes, C.node_x)
    assert mean < 0.1 and mx < 0.3, (mean, mx)
    assert abs(f["k1"]) < 0.003 and f["track"] > 0.8
    reg, ref_verified = reference_path(version, span, img)
    assert reg is not None
    ref_mean, _ = node_err(reg[0], reg[1], C.H, C.nodes, C.node_x)
    assert mean <= max(ref_mean, 0.04) + 0.02, (mean, ref_mean)
    # blind: the finder, the word, the decode
    dec = L.Decoder(1536)
    k, blocks, out, held = dec.decode(img)
    dec.close()
    assert out.found == 1 and out.word == 1 and out.blocks == version and held == version
    assert out.ring == {128: 0, 256: 1, 512: 2}[span]
    assert k >= ref_verified - 1, (k, ref_verified)
    truth = decoder_gpu.decode_reg(version, span, img, C.H, C.nodes, soft=False)
    assert k >= truth["good"] - 1
    assert sent_in_order(blocks, k, C.blocks)


def sent_in_order(got, k, sent):
    """The k verified blocks liz_decode packed (in block order, api/codec.cpp:143-147) are blocks that were sent, in
    the order they were sent: every verified byte is a sent byte."""
    sent = np.asarray(sent, np.uint8).reshape(-1, 476)
    got = np.asarray(got, np.uint8).reshape(-1, 476)[:k]
    j = 0
    for blk in got:
        while j < len(sent) and not np.array_equal(sent[j], blk):
            j += 1
        if j == len(sent):
            return False
        j += 1
    return True


LENS = [(64, 512, 4.0, 0.04, 0.8, 4.0, 0.85), (20, 256, 95.0, 0.03, 1.0, 5.0, 0.8),
        (45, 512, 181.0, -0.02, 0.6, 3.0, 0.9), (64, 512, -7.0, -0.05, 1.0, 4.0, 0.8)]


@pytest.mark.parametrize("version,span,rot,k1,blur,noise,scale", LENS)
def test_finder_recovers_the_lens(decoder_gpu, version, span, rot, k1, blur, noise, scale):
    """Symbols painted on the GPU behind a radial lens: k1 and the registration from the border alone."""
    rng = np.random.default_rng(version)
    blocks = sim.random_blocks(version, rng)
    dpic = decoder_gpu.modulate(version, blocks)
    cam = sim.Camera(1920, 1080, rot_deg=rot, scale=scale, persp=0.03, k1=k1, blur=blur, noise=noise)
    Hm = cam.H(span)
    dimg = decoder_gpu.capture(version, span, dpic, 1920, 1080, Hm, k1=k1, blur=blur, noise=noise, seed=7, fps=30)
    img = dimg.get(np.uint8, 1920 * 1080).reshape(1080, 1920)
    dimg.free()
    dpic.free()
    f = decoder_gpu.find(img)
    assert f is not None and f["span"] == span
    assert abs(f["k1"] - k1) < 0.006, f["k1"]
    Ht = (Hm / Hm[2, 2]).astype(np.float32).ravel()
    mean, mx = node_err(f["H"], f["nodes"], Ht, sim.nodes_for(cam, span, Hm), L.node_lattice(span))
    assert mean < 0.12 and mx < 0.4, (mean, mx)
    dec = L.Decoder(1536)
    k, got, out, held = dec.decode(img)
    dec.close()
    assert out.word == 1 and out.blocks == version and out.fps == 30
    truth = decoder_gpu.decode_reg(version, span, img, Ht, sim.nodes_for(cam, span, Hm), soft=False)
    assert k >= truth["good"] - 2, (k, truth["good"])
    assert sent_in_order(got, k, blocks)


def test_finder_padded_rows_match_tight(decoder_gpu):
    C = Capture(20, 256, W=1280, Hh=720, rot_deg=-3.0, scale=0.9, blur=0.7, noise=3.0, seed=2)
    tight = np.ascontiguousarray(C.img)
    padded = np.zeros((720, 1280 + 96), np.uint8)
    padded[:, :1280] = tight
    a = decoder_gpu.find(tight)
    b = decoder_gpu.find(padded[:, :1280])  # rows 1376 bytes apart
    assert a is not None and b is not None
    assert np.array_equal(a["H"], b["H"]) and a["k1"] == b["k1"] and np.array_equal(a["nodes"], b["nodes"])


def test_frames_without_a_symbol_are_not_found(decoder_gpu):
    rng = np.random.default_rng(0)
    noise = rng.integers(0, 256, (720, 1280)).astype(np.uint8)
    assert decoder_gpu.find(noise) is None
    flat = np.full((720, 1280), 200, np.uint8)
    assert decoder_gpu.find(flat) is None
    # a picture with no border round it (grey ring): nothing to find
    blocks = sim.random_blocks(5, rng)
    img, H, nodes = sim.capture(sim.picture(5, blocks), 128, sim.Camera(1280, 720, blur=0.7, noise=3.0), seed=1)
    assert decoder_gpu.find(img.numpy()) is None
    dec = L.Decoder(384)
    k, got, out, held = dec.decode(noise)
    assert k == 0 and out.found == 0 and out.ring == -1 and held == 0
    k, got, out, held = dec.decode(np.zeros((8, 8), np.uint8))
    assert k == 0 and out.found == 0
    dec.close()


def test_hold_then_release(decoder_gpu):
    """A held registration is used while held (even a wrong one); released, each frame's own symbol is found."""
    C = Capture(5, 128, seed=3, blur=0.6, noise=2.0)
    img = np.ascontiguousarray(C.img)
    dec = L.Decoder(384)
    wrong = C.H.copy()
    wrong[2] += 40.0  # shifted 40 pixels
    dec.hold(0, 128, wrong, C.nodes)
    k, got, out, held = dec.decode(img)
    assert out.found == 1 and k < C.blocks_n
    dec.release()
    k, got, out, held = dec.decode(img)
    dec.close()
    assert out.found == 1 and out.word == 1 and k == C.blocks_n and np.array_equal(got.ravel(), C.blocks)


def test_finder_against_reference_finder_on_random_hard_captures(decoder_gpu):
    """24 GPU-painted captures over the whole range (any rotation, scale 0.6-0.9 of the height, perspective up to
    0.05, k1 +-0.05, blur 0.5-1.5 px, AWGN 2-8 grey; three formats and rings): this finder finds at least as many as
    the reference's own finder and registers them at least as closely on the frames both find."""
    rng = np.random.default_rng(2024)
    r = ref("x86")
    mine = theirs = 0
    err_mine, err_ref = [], []
    for i in range(24):
        version, span = [(5, 128), (20, 256), (64, 512)][i % 3]
        blocks = sim.random_blocks(version, rng)
        cam = sim.Camera(1920, 1080, rot_deg=float(rng.uniform(0, 360)), scale=float(rng.uniform(0.6, 0.9)),
                         persp=float(rng.uniform(0, 0.05)), k1=float(rng.uniform(-0.05, 0.05)),
                         blur=float(rng.uniform(0.5, 1.5)), noise=float(rng.uniform(2, 8)))
        Hm = cam.H(span)
        dpic = decoder_gpu.modulate(version, blocks)
        dimg = decoder_gpu.capture(version, span, dpic, 1920, 1080, Hm, k1=cam.k1, blur=cam.blur, noise=cam.noise,
                                   seed=i, fps=30)
        img = dimg.get(np.uint8, 1920 * 1080).reshape(1080, 1920)
        dimg.free()
        dpic.free()
        Ht = (Hm / Hm[2, 2]).astype(np.float32).ravel()
        nt = sim.nodes_for(cam, span, Hm)
        node_x = L.node_lattice(span)
        f = decoder_gpu.find(img)
        ok_mine = f is not None and f["span"] == span
        reg, _ = reference_path(version, span, img)
        ok_ref = reg is not None
        mine += ok_mine
        theirs += ok_ref
        if ok_mine and ok_ref:
            err_mine.append(node_err(f["H"], f["nodes"], Ht, nt, node_x)[0])
            err_ref.append(node_err(reg[0], reg[1], Ht, nt, node_x)[0])
    assert mine >= theirs and mine >= 22, (mine, theirs)
    assert np.median(err_mine) <= np.median(err_ref), (np.median(err_mine), np.median(err_ref))


def test_find_batch_equals_single(decoder_gpu):
    """lizx_find_batch over mixed frames (three rings, rotations, lenses, a frame with no symbol) gives each frame what
    lizx_find gives it alone, bit for bit: the stages are the same code, their launches only shared."""
    rng = np.random.default_rng(77)
    imgs = []
    for i, (version, span) in enumerate([(5, 128), (20, 256), (64, 512), (45, 512), (8, 256), (20, 256)]):
        blocks = sim.random_blocks(version, rng)
        cam = sim.Camera(1920, 1080, rot_deg=float(rng.uniform(0, 360)), scale=float(rng.uniform(0.65, 0.9)),
                         persp=float(rng.uniform(0, 0.04)), k1=float(rng.uniform(-0.04, 0.04)),
                         blur=float(rng.uniform(0.6, 1.2)), noise=float(rng.uniform(2, 6)))
        dpic = decoder_gpu.modulate(version, blocks)
        dimg = decoder_gpu.capture(version, span, dpic, 1920, 1080, cam.H(span), k1=cam.k1, blur=cam.blur,
                                   noise=cam.noise, seed=i, fps=30)
        imgs.append(dimg.get(np.uint8, 1920 * 1080).reshape(1080, 1920))
        dimg.free()
        dpic.free()
    imgs.insert(3, rng.integers(0, 256, (1080, 1920)).astype(np.uint8))  # nothing to find
    imgs.append(np.full((1080, 1920), 255, np.uint8))
    batch = decoder_gpu.find_batch(imgs, words=True)
    assert len(batch) == len(imgs)
    for i, img in enumerate(imgs):
        one = decoder_gpu.find(img)
        if one is None:
            assert batch[i] is None, i
            continue
        assert batch[i] is not None and batch[i]["span"] == one["span"], i
        assert np.array_equal(batch[i]["H"], one["H"]) and np.array_equal(batch[i]["nodes"], one["nodes"]), i
        # the word read on the finder's thread is the one lizx_read_word reads through the same registration
        word = L.read_word(one["span"], img, one["H"], one["nodes"])
        assert batch[i]["version"] == (None if word is None else word[0]), i
    assert sum(b is not None for b in batch) == 6


def test_find_batch_padded_rows_and_reuse(decoder_gpu):
    """lizx_find_batch through the C ABI on frames whose rows are padded (stride > w) gives what the tight frames give,
    and a smaller batch after a larger one on the same handle gives each frame the same again."""
    rng = np.random.default_rng(91)
    imgs = []
    for i, (version, span) in enumerate([(20, 256), (5, 128), (64, 512)]):
        blocks = sim.random_blocks(version, rng)
        cam = sim.Camera(1280, 720, rot_deg=float(rng.uniform(-20, 20)), scale=float(rng.uniform(0.7, 0.9)),
                         persp=0.02, k1=0.01, blur=0.8, noise=3.0)
        dpic = decoder_gpu.modulate(version, blocks)
        dimg = decoder_gpu.capture(version, span, dpic, 1280, 720, cam.H(span), k1=cam.k1, blur=cam.blur,
                                   noise=cam.noise, seed=40 + i, fps=30)
        imgs.append(dimg.get(np.uint8, 1280 * 720).reshape(720, 1280))
        dimg.free()
        dpic.free()
    tight = decoder_gpu.find_batch(imgs)
    assert all(t is not None for t in tight)
    pad = 1280 + 96
    bufs = [np.full((720, pad), 17, np.uint8) for _ in imgs]
    for b, im in zip(bufs, imgs):
        b[:, :1280] = im
    F = len(imgs)
    ptrs = (C.c_void_p * F)(*[b.ctypes.data for b in bufs])
    spans, found = np.zeros(F, np.int32), np.zeros(F, np.int32)
    H, nodes = np.zeros((F, 9), np.float32), np.zeros((F, 3200), np.float32)
    got = L.lib().lizx_find_batch(decoder_gpu.h, F, ptrs, 1280, 720, pad, L._p(spans),
                                           L._p(H), L._p(nodes), L._p(found), None)
    assert got == F
    for f in range(F):
        nx = tight[f]["nodes"].shape[0]
        assert spans[f] == tight[f]["span"] and np.array_equal(H[f], tight[f]["H"])
        assert np.array_equal(nodes[f, :nx * nx * 2].reshape(nx, nx, 2), tight[f]["nodes"])
    again = decoder_gpu.find_batch(imgs[1:])
    for f, a in enumerate(again):
        assert a is not None and np.array_equal(a["H"], tight[f + 1]["H"]), f


def test_find_batch_at_its_largest(decoder_gpu):
    """256 frames (the entry's most) in one call give each frame what lizx_find gives it; 257 are refused before any
    device work."""
    rng = np.random.default_rng(123)
    base = []
    for i, (version, span) in enumerate([(5, 128), (20, 256), (8, 256), (5, 128)]):
        blocks = sim.random_blocks(version, rng)
        cam = sim.Camera(960, 720, rot_deg=float(rng.uniform(0, 360)), scale=float(rng.uniform(0.75, 0.9)),
                         persp=0.02, k1=0.0, blur=0.8, noise=3.0)
        dpic = decoder_gpu.modulate(version, blocks)
        dimg = decoder_gpu.capture(version, span, dpic, 960, 720, cam.H(span), k1=cam.k1, blur=cam.blur,
                                   noise=cam.noise, seed=60 + i, fps=30)
        base.append(dimg.get(np.uint8, 960 * 720).reshape(720, 960))
        dimg.free()
        dpic.free()
    singles = [decoder_gpu.find(b) for b in base]
    assert all(s is not None for s in singles)
    imgs = [base[f % 4] for f in range(256)]
    batch = decoder_gpu.find_batch(imgs)
    for f, got in enumerate(batch):
        one = singles[f % 4]
        assert got is not None and got["span"] == one["span"], f
        assert np.array_equal(got["H"], one["H"]) and np.array_equal(got["nodes"], one["nodes"]), f
    with pytest.raises(L.LizError):
        decoder_gpu.find_batch(imgs + imgs[:1])


def test_find_batch_from_two_threads_at_once():
    """Two handles' batches searched at the same time from two caller threads (each handle with its own threads and
    streams) give what each gives alone."""
    import concurrent.futures as cf
    rng = np.random.default_rng(131)
    d0, d1 = L.Decoder(0), L.Decoder(0)
    try:
        imgs = []
        for i, (version, span) in enumerate([(5, 128), (20, 256), (8, 256), (45, 512)]):
            blocks = sim.random_blocks(version, rng)
            cam = sim.Camera(1280, 720, rot_deg=float(rng.uniform(0, 360)), scale=float(rng.uniform(0.75, 0.9)),
                             persp=0.02, k1=0.01, blur=0.8, noise=3.0)
            dpic = d0.modulate(version, blocks)
            dimg = d0.capture(version, span, dpic, 1280, 720, cam.H(span), k1=cam.k1, blur=cam.blur, noise=cam.noise,
                              seed=80 + i, fps=30)
            imgs.append(dimg.get(np.uint8, 1280 * 720).reshape(720, 1280))
            dimg.free()
            dpic.free()
        alone = d0.find_batch(imgs * 4)
        with cf.ThreadPoolExecutor(2) as ex:
            a, b = ex.submit(d0.find_batch, imgs * 4), ex.submit(d1.find_batch, imgs * 4)
            got = [a.result(), b.result()]
        for g in got:
            for f, (x, y) in enumerate(zip(g, alone)):
                assert (x is None) == (y is None), f
                if x is not None:
                    assert x["span"] == y["span"] and np.array_equal(x["H"], y["H"]), f
                    assert np.array_equal(x["nodes"], y["nodes"]), f
        assert sum(x is not None for x in alone) == 16
    finally:
        d0.close()
        d1.close()


def test_blind_decodes_on_many_threads_at_once():
    """liz_decode on bare frames from 12 caller threads at once, a decoder each, every decoder's first frame of each
    format capturing its frame graph while the others capture, replay and wait: each frame's verified blocks are the
    ones a single decoder gives it alone, in the same order."""
    import concurrent.futures as cf
    rng = np.random.default_rng(141)
    d0 = L.Decoder(0)
    frames = []
    try:
        for i, (version, span) in enumerate([(5, 128), (20, 256), (45, 512), (8, 256)]):
            blocks = sim.random_blocks(version, rng)
            cam = sim.Camera(1280, 720, rot_deg=float(rng.uniform(-10, 10)), scale=float(rng.uniform(0.8, 0.9)),
                             persp=0.02, k1=0.01, blur=0.8, noise=3.0)
            dpic = d0.modulate(version, blocks)
            dimg = d0.capture(version, span, dpic, 1280, 720, cam.H(span), k1=cam.k1, blur=cam.blur, noise=cam.noise,
                              seed=90 + i, fps=30)
            frames.append(dimg.get(np.uint8, 1280 * 720).reshape(720, 1280))
            dimg.free()
            dpic.free()
        alone = [d0.decode(f) for f in frames]
    finally:
        d0.close()
    decs = [L.Decoder(0) for _ in range(12)]
    try:
        def job(t):
            return [decs[t].decode(frames[(t + r) % len(frames)]) for r in range(2 * len(frames))]

        with cf.ThreadPoolExecutor(12) as ex:
            outs = list(ex.map(job, range(12)))
        for t, out in enumerate(outs):
            for r, (k, got, o, _) in enumerate(out):
                ka, ga, oa, _ = alone[(t + r) % len(frames)]
                assert k == ka and o.blocks == oa.blocks, (t, r, k, ka)
                assert np.array_equal(np.asarray(got).ravel()[:k * 476], np.asarray(ga).ravel()[:ka * 476]), (t, r)
        assert all(a[0] > 0 for a in alone)
    finally:
        for d in decs:
            d.close()
