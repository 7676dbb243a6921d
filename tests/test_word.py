"""The format word (lizx_word_*, lizx_read_word; host code) held to the reference's own: the band's geometry and the
codeword against fmt.c/rs.c's encoder, the decode of soft words (clean, noisy, faded sides, garbage) against
ob_fmt_decode, and the read from captures the reference's encoder painted against ob_thin_read_fmt + ob_fmt_decode.
Skipped where oracle/_ref was not built; the word's own invariants run regardless."""
import numpy as np
import pytest

import lizard_amd as L
from _libs import ptr, ref

SPANS = (128, 256, 512)


def soft_of(cw, cells, rng, amp=0.4, noise=0.0):
    """Soft values of a codeword laid out as the band holds it (byte k on side k % 4, cells 8 (k / 4) ..), dark
    positive."""
    q = np.zeros(4 * cells, np.float32)
    for side in range(4):
        for i in range(cells):
            k = 4 * (i >> 3) + side
            bit = (int(cw[k]) >> (7 - (i & 7))) & 1 if k < len(cw) else 0
            q[side * cells + i] = amp if bit else -amp
    if noise:
        q += rng.normal(0, noise, q.shape).astype(np.float32)
    return np.clip(q, -0.5, 0.5).astype(np.float32)


@pytest.mark.parametrize("span", SPANS)
def test_word_round_trip(span):
    cells, nbytes = L.word_geometry(span)
    assert (cells, nbytes) == {128: (32, 16), 256: (64, 32), 512: (128, 64)}[span]
    rng = np.random.default_rng(span)
    for version, fps in ((4, 0), (20, 30), (64, 255), (45, 60)):
        cw = L.word_encode(span, version, fps)
        assert len(cw) == nbytes and cw[0] == 0x4C and cw[1] == version and cw[2] == fps
        assert L.word_decode(span, soft_of(cw, cells, rng)) == (version, fps)
        # (nbytes - 3) // 2 byte errors: still the word
        q = soft_of(cw, cells, rng)
        bad = rng.choice(nbytes, (nbytes - 3) // 2, replace=False)
        for k in bad:
            side, c0 = k % 4, 8 * (k // 4)
            q[side * cells + c0:side * cells + c0 + 8] *= -1
        assert L.word_decode(span, q) == (version, fps)
    with pytest.raises(L.LizError):
        L.word_encode(span, 3, 0)


ref_needed = pytest.mark.skipif(ref("scalar") is None, reason="oracle/_ref not built")


@ref_needed
@pytest.mark.parametrize("span", SPANS)
def test_word_geometry_and_codeword_match_reference(span):
    r = ref("scalar")
    c, b = np.zeros(1, np.int32), np.zeros(1, np.int32)
    r.ref_fmt_geom(span + 30, ptr(c), ptr(b))
    assert L.word_geometry(span) == (int(c[0]), int(b[0]))
    for version in (4, 5, 13, 20, 45, 64):
        for fps in (0, 1, 24, 60, 240, 255):
            want = np.zeros(int(b[0]), np.uint8)
            assert r.ref_fmt_encode(version, fps, ptr(want), int(b[0])) == 0
            assert np.array_equal(L.word_encode(span, version, fps), want)


@ref_needed
@pytest.mark.parametrize("span", SPANS)
def test_word_decode_matches_reference(span):
    """Clean, noisy, partly flipped, faded-side and garbage soft words: the same verdict and word as the reference's
    decoder (hard decisions, then erasures by confidence)."""
    r = ref("scalar")
    cells, nbytes = L.word_geometry(span)
    rng = np.random.default_rng(7 + span)
    agree = decoded = 0
    for trial in range(300):
        version, fps = int(rng.integers(4, 65)), int(rng.integers(0, 256))
        q = soft_of(L.word_encode(span, version, fps), cells, rng, amp=float(rng.uniform(0.1, 0.5)),
                    noise=float(rng.choice([0.0, 0.1, 0.2, 0.3])))
        kind = trial % 5
        if kind == 1:  # whole bytes flipped
            for k in rng.choice(nbytes, int(rng.integers(0, nbytes // 2)), replace=False):
                s0 = (k % 4) * cells + 8 * (k // 4)
                q[s0:s0 + 8] *= -1
        elif kind == 2:  # one or two sides faded to near nothing
            for side in rng.choice(4, int(rng.integers(1, 3)), replace=False):
                q[side * cells:(side + 1) * cells] = rng.normal(0, 0.02, cells)
        elif kind == 3:  # garbage
            q = rng.uniform(-0.5, 0.5, q.shape).astype(np.float32)
        elif kind == 4:  # a faded side plus flipped bytes elsewhere
            q[cells:2 * cells] = rng.normal(0, 0.02, cells)
            for k in rng.choice(nbytes, int(rng.integers(0, nbytes // 4)), replace=False):
                if k % 4 != 1:
                    s0 = (k % 4) * cells + 8 * (k // 4)
                    q[s0:s0 + 8] *= -1
        q = np.ascontiguousarray(q, np.float32)
        v, f = np.zeros(1, np.int32), np.zeros(1, np.int32)
        rr = r.ref_fmt_decode(ptr(q), cells, nbytes, ptr(v), ptr(f))
        want = (int(v[0]), int(f[0])) if rr == 0 else None
        got = L.word_decode(span, q)
        agree += got == want
        decoded += want is not None
        assert got == want, (trial, kind, got, want)
    assert decoded > 100


@ref_needed
@pytest.mark.parametrize("version,span,blur,noise", [(5, 128, 0.8, 4.0), (20, 256, 1.0, 6.0), (64, 512, 0.6, 3.0)])
def test_read_word_matches_reference_on_painted_captures(version, span, blur, noise):
    """Captures of the reference encoder's painted symbol (border, word and all): the same soft values and word."""
    from pictures import Capture
    r = ref("x86")
    cap = Capture(version, span, W=1920, Hh=1080, scale=0.85, blur=blur, noise=noise, seed=version)
    info = L.format_info(version)
    f = r.ref_format_new(info["n"], info["subch"], span)
    cells, nbytes = L.word_geometry(span)
    q_ref = np.zeros(4 * cells, np.float32)
    dims, v, fp = np.zeros(2, np.int32), np.zeros(1, np.int32), np.zeros(1, np.int32)
    img = np.ascontiguousarray(cap.img)
    rr = r.ref_read_word(f, ptr(img), img.shape[1], img.shape[0], ptr(cap.H), ptr(cap.nodes), ptr(q_ref), ptr(dims),
                         ptr(v), ptr(fp))
    r.ref_format_free(f)
    assert rr == 1 and int(v[0]) == version
    got, q = L.read_word(span, img, cap.H, cap.nodes, soft=True)
    assert (int(dims[0]), int(dims[1])) == (cells, nbytes)
    assert np.abs(q - q_ref).max() < 1e-5
    assert got == (int(v[0]), int(fp[0]))


@ref_needed
@pytest.mark.parametrize("span", SPANS)
def test_border_kinds_match_reference_layout(span):
    """The border's module map, word cells and all, equals the reference's layout (ob_layout_init +
    ob_layout_set_fmt) module for module; with no word, its cells light as the reference's before the word is set."""
    r = ref("scalar")
    side = span + 30
    want = np.zeros(side * side, np.uint8)
    assert r.ref_layout_kinds(side, None, ptr(want)) == 0
    got = kinds = L.border_kinds(span)
    assert np.array_equal(got.ravel(), want)
    for version, fps in ((4, 0), (37, 60), (64, 255)):
        cw = L.word_encode(span, version, fps)
        assert r.ref_layout_kinds(side, ptr(cw), ptr(want)) == 0
        got = L.border_kinds(span, version, fps)
        assert np.array_equal(got.ravel(), want), (version, fps)
    # what the map must show whatever the reference: corner marks, the solid line, the guard, word cells flagged
    k, S = kinds, side
    assert k[0, :].max() == 1 and (k[1:3, 12:S - 12] == 2).all() and (k[3:5, 12:S - 12] == 1).all()
    assert (k[12:S - 12, 12:S - 12] == 0).all() and (k[5:11, 5:11] == 2).all() and k[11, 11] == 1
    assert ((got & 8) != 0).sum() == 4 * 2 * 2 * L.word_geometry(span)[0] and not (kinds & 8).any()
