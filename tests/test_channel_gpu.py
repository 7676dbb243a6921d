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
