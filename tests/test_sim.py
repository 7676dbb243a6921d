"""The synthetic sender and channel (lizard_amd.sim) and the encoder-side exports it stands on: frame slots and the
coefficient order equal the oracle's; the picture equals the FP64 modulator of research/SPEC.md 6.5-6.7; a capture
rendered through a homography and a radial lens decodes with the compiled reference's own decoder through the
registration the channel reports (the channel and its registration agree)."""
import numpy as np
import pytest

import lizard_amd as L
from lizard_amd import sim
from _libs import oracle, ptr, ref

from pictures import picture as opic


@pytest.mark.parametrize("version", [4, 5, 13, 20, 45, 64])
def test_frame_slots_and_coefficients_match_oracle(version):
    o = oracle()
    subch = 16 * version
    n = o.or_n_for(subch)
    f = o.or_format_new(n, subch)
    info = np.zeros(6, np.int32)
    o.or_format_info(f, ptr(info))
    rng = np.random.default_rng(version)
    blocks = rng.integers(0, 256, int(info[2]) * 476).astype(np.uint8)
    for cnt in range(4):
        want = np.zeros(subch * 640, np.uint8)
        o.or_frame_slots(f, ptr(blocks), cnt, ptr(want))
        assert np.array_equal(L.frame_slots(version, blocks, cnt), want)
    pos = np.ctypeslib.as_array(o.or_format_pos(f), (subch * 320,)).copy()
    o.or_format_free(f)
    u, v = L.format_coefs(version)
    uu, vv = opic.coef_uv(pos, n)
    assert np.array_equal(u, uu) and np.array_equal(v, vv)


def test_frame_slots_errors():
    with pytest.raises(L.LizError):
        L.frame_slots(3, np.zeros(476, np.uint8))
    with pytest.raises(L.LizError):
        L.frame_slots(5, np.zeros(9 * 476, np.uint8), cnt=4)


@pytest.mark.parametrize("version", [5, 20])
def test_picture_matches_fp64_modulator(version):
    info = L.format_info(version)
    blocks = sim.random_blocks(version, np.random.default_rng(3))
    slots = L.frame_slots(version, blocks, 2)
    u, v = L.format_coefs(version)
    want = opic.modulate(info["n"], info["subch"], slots, u, v)
    got = sim.picture(version, blocks, cnt=2).numpy()
    assert np.abs(got - want).max() < 1e-6


def test_random_blocks_carry_ids():
    b = sim.random_blocks(5, np.random.default_rng(0)).reshape(-1, 476)
    assert b.shape == (9, 476)
    assert [int.from_bytes(r[:4].tobytes(), "big") for r in b] == list(range(9))


def test_camera_geometry():
    cam = sim.Camera(1280, 720, rot_deg=0.0, scale=0.5, persp=0.0, k1=0.0)
    H = cam.H(128)
    side = 158
    c = np.array([[0, 0, 1], [side, side, 1]], float) @ H.T
    c = c[:, :2] / c[:, 2:]
    assert np.allclose(c[0], [640 - 180, 360 - 180]) and np.allclose(c[1], [640 + 180, 360 + 180])
    assert np.abs(sim.nodes_for(cam, 128, H)).max() == 0
    cam.k1 = 0.05
    d = sim.nodes_for(cam, 128, H)
    # the radial field points outward for k1 > 0 and grows with the distance from the axis
    assert d[0, 0, 0] < 0 and d[0, 0, 1] < 0 and d[-1, -1, 0] > 0 and d[-1, -1, 1] > 0
    mid = d.shape[0] // 2
    assert np.abs(d[mid, mid]).max() < np.abs(d[0, 0]).max()


@pytest.mark.parametrize("version,span,k1", [(5, 128, 0.0), (5, 128, 0.04), (20, 256, -0.03)])
def test_capture_decodes_with_reference(version, span, k1):
    r = ref("scalar")
    if r is None:
        pytest.skip("oracle/_ref not built")
    info = L.format_info(version)
    B = info["blocks"]
    blocks = sim.random_blocks(version, np.random.default_rng(span))
    cam = sim.Camera(1280 if span == 128 else 1920, 720 if span == 128 else 1080, rot_deg=-6.0, scale=0.85,
                     persp=0.03, k1=k1, blur=0.8, noise=4.0)
    img, H, nodes = sim.capture(sim.picture(version, blocks), span, cam, seed=1)
    img = img.numpy()
    f = r.ref_format_new(info["n"], info["subch"], span)
    out, ok = np.zeros(B * 476, np.uint8), np.zeros(B, np.uint8)
    its, be = np.zeros(B, np.int8), np.zeros(B, np.float32)
    good = r.ref_decode_reg(f, ptr(img), img.shape[1], img.shape[0], ptr(H), ptr(np.ascontiguousarray(nodes)),
                            ptr(out), ptr(ok), ptr(its), ptr(be), None, None)
    r.ref_format_free(f)
    assert good == B
    assert np.array_equal(out, blocks)
