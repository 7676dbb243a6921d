"""The camera channel on the GPU (lizx_capture) against the host model it restates (lizard_amd.sim.capture, float64
geometry in torch): noise-free renders agree to within a grey level at rounding boundaries; the noise is Gaussian of
the asked sigma, reproducible by seed; and a capture made entirely on the GPU (sender picture, then channel) decodes
through the registration the channel reports."""
import numpy as np
import pytest

import lizard_amd as L
from lizard_amd import sim

pytestmark = pytest.mark.gpu


def render_pair(dec, version, span, cam, seed=0):
    rng = np.random.default_rng(seed)
    pic = sim.picture(version, sim.random_blocks(version, rng))
    want, H, nodes = sim.capture(pic, span, cam, seed=seed)
    dpic = L.DeviceArray.of(pic.numpy())
    dimg, H2, nodes2 = sim.capture_gpu(dec, version, span, dpic, cam, seed=seed)
    got = dimg.get(np.uint8, cam.W * cam.Hh).reshape(cam.Hh, cam.W)
    dimg.free()
    dpic.free()
    assert np.array_equal(H, H2) and np.array_equal(nodes, nodes2)
    return got, want.numpy()


@pytest.mark.parametrize("blur,k1", [(0.0, 0.0), (1.0, 0.03), (1.5, -0.04)])
def test_noise_free_render_matches_host(decoder_gpu, blur, k1):
    cam = sim.Camera(1280, 720, rot_deg=7.0, scale=0.8, persp=0.03, k1=k1, blur=blur, noise=0.0)
    got, want = render_pair(decoder_gpu, 20, 256, cam, seed=3)
    d = np.abs(got.astype(np.int16) - want.astype(np.int16))
    assert d.max() <= 1
    assert (d > 0).mean() < 0.005


def test_noise_statistics_and_seed(decoder_gpu):
    rng = np.random.default_rng(9)
    pic = L.DeviceArray.of(sim.picture(5, sim.random_blocks(5, rng)).numpy())
    cam = sim.Camera(1280, 720, blur=0.8, noise=0.0)
    Hm = cam.H(128)
    clean = decoder_gpu.capture(5, 128, pic, cam.W, cam.Hh, Hm, blur=0.8).get(np.uint8, cam.W * cam.Hh)
    imgs = [decoder_gpu.capture(5, 128, pic, cam.W, cam.Hh, Hm, blur=0.8, noise=4.0, seed=s).get(np.uint8,
                                                                                                  cam.W * cam.Hh)
            for s in (1, 1, 2)]
    pic.free()
    assert np.array_equal(imgs[0], imgs[1]) and not np.array_equal(imgs[0], imgs[2])
    inside = (clean > 20) & (clean < 235)  # away from the clamp at 0 and 255
    diff = imgs[0].astype(np.float64)[inside] - clean.astype(np.float64)[inside]
    assert abs(diff.mean()) < 0.05
    assert 3.9 < diff.std() < 4.15


def test_gpu_sender_and_channel_decode(decoder_gpu):
    """Blocks -> GPU picture (lizx_modulate) -> GPU capture (lizx_capture) -> liz decode through the reported
    registration: every block verified and equal to the one sent."""
    version, span = 20, 256
    rng = np.random.default_rng(21)
    blocks = sim.random_blocks(version, rng)
    dpic = decoder_gpu.modulate(version, blocks)
    cam = sim.Camera(1920, 1080, rot_deg=-6.0, scale=0.85, persp=0.03, k1=0.03, blur=1.0, noise=4.0)
    dimg, H, nodes = sim.capture_gpu(decoder_gpu, version, span, dpic, cam, seed=5)
    img = dimg.get(np.uint8, cam.W * cam.Hh).reshape(cam.Hh, cam.W)
    dimg.free()
    dpic.free()
    out = decoder_gpu.decode_reg(version, span, img, H, nodes, soft=False)
    info = L.format_info(version)
    assert out["good"] == info["blocks"]
    assert np.array_equal(out["blocks"], blocks.reshape(-1, 476))


def test_capture_argument_errors(decoder_gpu):
    pic = L.DeviceArray(384 * 384 * 4)
    with pytest.raises(L.LizError):
        decoder_gpu.capture(5, 100, pic, 64, 64, np.eye(3))
    with pytest.raises(L.LizError):
        decoder_gpu.capture(5, 128, pic, 64, 64, np.zeros((3, 3)))
    with pytest.raises(L.LizError):
        decoder_gpu.capture(5, 128, pic, 64, 64, np.eye(3), blur=9.0)
    pic.free()


SYMBOLS = [(5, 128), (20, 256), (64, 512), (45, 512)]


@pytest.mark.parametrize("version,span", SYMBOLS)
def test_symbol_paint_matches_reference_page(decoder_gpu, version, span):
    """lizx_capture_symbol through a camera that puts each paint pixel on one capture pixel (3 a module, the 2-module
    quiet zone) against the reference encoder's own page (lizard_encode + lizard_paint_grey), from the same picture:
    the border, word, corner marks, anchors and quiet zone to the grey level; where the format's picture is 3 samples
    a module (the reference then paints the samples themselves) the picture and its wrapped guard too, but for
    rounding ties."""
    from _libs import ptr, ref
    r = ref("scalar")
    info = L.format_info(version)
    n = info["n"]
    f = r.ref_format_new(n, info["subch"], span)
    fi = np.zeros(10, np.int32)
    r.ref_format_info(f, ptr(fi))
    pxm, px = int(fi[7]), int(fi[8])
    assert pxm == 3 and px == 3 * (span + 30)
    blocks = np.random.default_rng(version).integers(0, 256, info["blocks"] * 476).astype(np.uint8)
    drive = np.zeros(px * px, np.float32)
    assert r.ref_encode(f, ptr(blocks), 0, ptr(drive), None) == 0
    Wg = px + 4 * pxm
    page = np.zeros(Wg * Wg, np.uint8)
    r.ref_paint_grey(f, ptr(drive), ptr(page))
    pic = np.zeros(n * n, np.float32)
    assert r.ref_encoded_picture(f, ptr(pic)) == 0
    r.ref_format_free(f)
    page = page.reshape(Wg, Wg)
    Hm = np.array([[3.0, 0, 6.0], [0, 3.0, 6.0], [0, 0, 1.0]])
    nx = len(L.node_lattice(span))
    word = L.read_word(span, page, Hm.astype(np.float32).ravel(), np.zeros((nx, nx, 2), np.float32))
    assert word is not None and word[0] == version
    dpic = L.DeviceArray.of(pic)
    dimg = decoder_gpu.capture(version, span, dpic, Wg, Wg, Hm, fps=word[1])
    img = dimg.get(np.uint8, Wg * Wg).reshape(Wg, Wg)
    dimg.free()
    dpic.free()
    kinds = L.border_kinds(span, version, word[1])
    module = np.full((Wg, Wg), 255, np.int32)  # kind per capture pixel; 255 the quiet zone
    module[6:Wg - 6, 6:Wg - 6] = np.repeat(np.repeat(kinds, 3, 0), 3, 1)
    fixed = module != 0
    assert np.array_equal(img[fixed], page[fixed])
    d = np.abs(img.astype(np.int32) - page)
    if n == 3 * span:
        assert d.max() <= 1 and (d > 0).mean() < 1e-4
    # the anchors too (inside the picture's modules, exactly black or white in both)
    a = r.ref_anchor_side(span)
    m0 = [r.ref_anchor_m0(span, i) for i in range(7)]
    for j in m0:
        for i in m0:
            ys, xs = 6 + 3 * j, 6 + 3 * i
            assert np.array_equal(img[ys:ys + 3 * a, xs:xs + 3 * a], page[ys:ys + 3 * a, xs:xs + 3 * a])


def test_symbol_capture_decodes_and_names_itself(decoder_gpu):
    """A symbol painted on the GPU and seen through a lens: its word reads (version and fps), and it decodes through
    the channel's registration to the sent blocks."""
    version, span, fps = 20, 256, 30
    rng = np.random.default_rng(8)
    blocks = sim.random_blocks(version, rng)
    dpic = decoder_gpu.modulate(version, blocks)
    cam = sim.Camera(1920, 1080, rot_deg=8.0, scale=0.85, persp=0.03, k1=-0.03, blur=0.9, noise=4.0)
    Hm = cam.H(span)
    dimg = decoder_gpu.capture(version, span, dpic, cam.W, cam.Hh, Hm, k1=cam.k1, blur=cam.blur, noise=cam.noise,
                               seed=3, fps=fps)
    img = dimg.get(np.uint8, cam.W * cam.Hh).reshape(cam.Hh, cam.W)
    dimg.free()
    dpic.free()
    H = (Hm / Hm[2, 2]).astype(np.float32).ravel()
    nodes = sim.nodes_for(cam, span, Hm)
    assert L.read_word(span, img, H, nodes) == (version, fps)
    out = decoder_gpu.decode_reg(version, span, img, H, nodes, soft=False)
    assert out["good"] == L.format_info(version)["blocks"]
    assert np.array_equal(out["blocks"], blocks.reshape(-1, 476))
