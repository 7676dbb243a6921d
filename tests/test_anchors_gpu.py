"""The anchors and the border's nodes on the GPU against the reference's own anchor search, registration and node
refinement, compiled from its sources in their scalar form (anchor_find, anchor_register, refine_node in
acquire_core's order): every record (centre, displacement, peak, place, found), the registration's homography, corners
and lens model, and every border node (displacement, score) to the bit, on symbols painted through a two-term radial
lens; the records and nodes against the channel's truth; and the finder's registration from the anchors outward
against the finder's radial lens alone on the same captures (node error to the truth, blocks verified)."""
import ctypes as C

import numpy as np
import pytest

import lizard_amd as L
from lizard_amd import sim
from _libs import ptr, ref

pytestmark = pytest.mark.gpu

P = C.c_void_p
VERSION = {128: 5, 256: 20, 512: 64}


def ref_lib():
    r = ref("scalar")
    if r is None:
        pytest.skip("oracle/_ref not built")
    r.ref_anchor_find.restype = C.c_int
    r.ref_anchor_find.argtypes = [P, P, C.c_int, C.c_int, P, P, P]
    r.ref_anchor_register.restype = C.c_int
    r.ref_anchor_register.argtypes = [P, P, C.c_int, C.c_int, P, C.c_float, P, P, P]
    r.ref_refine_border.restype = C.c_int
    r.ref_refine_border.argtypes = [P, P, C.c_int, C.c_int, P, C.c_float, P, P]
    return r


def painted(dec, span, cam, seed):
    """A symbol of this ring painted through cam on the GPU: (host image, H as 9 float32 (h8 = 1), truth nodes)."""
    version = VERSION[span]
    blocks = sim.random_blocks(version, np.random.default_rng(seed))
    dpic = dec.modulate(version, blocks)
    Hm = cam.H(span)
    dimg = dec.capture(version, span, dpic, cam.W, cam.Hh, Hm, k1=cam.k1, k2=cam.k2, blur=cam.blur, noise=cam.noise,
                       seed=seed, fps=30)
    img = dimg.get(np.uint8, cam.W * cam.Hh).reshape(cam.Hh, cam.W)
    dimg.free()
    dpic.free()
    return img, (Hm / Hm[2, 2]).astype(np.float32).ravel(), sim.nodes_for(cam, span, Hm), blocks


def ref_format(r, span):
    info = L.format_info(VERSION[span])
    return r.ref_format_new(info["n"], info["subch"], span)


def truth_place(cam, Hm, X, Y):
    """Where the channel puts module point (X, Y): H, then the lens."""
    q = np.asarray(Hm, np.float64).reshape(3, 3) @ np.array([X, Y, 1.0])
    return cam.distort(q[:2] / q[2])


CASES = [(128, 1280, 720, 6.0, 0.85, 0.02, 0.0, 0.0, 0.8, 3.0),
         (256, 1920, 1080, -9.0, 0.85, 0.03, 0.04, -0.03, 1.0, 4.0),
         (512, 1920, 1080, 3.0, 0.9, 0.02, -0.03, 0.02, 0.7, 3.0),
         (256, 1280, 720, 91.0, 0.8, 0.04, 0.02, 0.05, 1.2, 6.0)]


@pytest.mark.parametrize("span,W,Hh,rot,scale,persp,k1,k2,blur,noise", CASES)
def test_anchor_find_equals_reference(decoder_gpu, span, W, Hh, rot, scale, persp, k1, k2, blur, noise):
    """Every anchor's record through H alone and through H and a lens model equals the reference's anchor_find; the
    anchors found through the true lens sit where the channel put them."""
    r = ref_lib()
    cam = sim.Camera(W, Hh, rot_deg=rot, scale=scale, persp=persp, k1=k1, blur=blur, noise=noise, k2=k2)
    img, H, _, _ = painted(decoder_gpu, span, cam, seed=span + int(rot))
    f = ref_format(r, span)
    try:
        a, R = cam.axis()
        # the channel's own lens as a model about its axis (two radial terms, nothing affine)
        model = np.zeros(15)
        model[0], model[1], model[2], model[3] = 2, a[0], a[1], R
        model[4 + 8], model[4 + 9] = k1 * R, k2 * R
        for m in (None, model):
            found, got = decoder_gpu.anchor_find(span, img, H, m)
            want = np.zeros((49, 8), np.float32)
            rf = r.ref_anchor_find(f, ptr(img), W, Hh, ptr(H), None if m is None else ptr(m), ptr(want))
            assert found == rf
            assert np.array_equal(got, want), (m is None, np.abs(got - want).max(), np.argwhere(got != want)[:4])
        ok = got[:, 7] == 1
        assert ok.sum() >= 40, ok.sum()
        Hm = cam.H(span)
        err = [np.hypot(*(got[k, 5:7] - truth_place(cam, Hm, got[k, 0], got[k, 1]))) for k in np.flatnonzero(ok)]
        mpx = scale * Hh / (span + 30)
        # within a tenth of a module in the median, a third at worst (2 px modules, blur 1.2 px, noise 6 levels
        # the hardest case here)
        assert np.median(err) < 0.1 * mpx and np.max(err) < 0.33 * mpx, (np.median(err) / mpx, np.max(err) / mpx)
    finally:
        r.ref_format_free(f)


def lattice_quad(span, H, nodes):
    pos = L.node_lattice(span)
    nx = len(pos)
    quad, src = np.zeros(8, np.float32), np.zeros(8, np.float32)
    for c in range(4):
        i = nx - 1 if c in (1, 2) else 0
        j = nx - 1 if c >= 2 else 0
        x, y = np.float32(pos[i]), np.float32(pos[j])
        w = H[6] * x + H[7] * y + np.float32(1)
        quad[2 * c] = (H[0] * x + H[1] * y + H[2]) / w + nodes[j, i, 0]
        quad[2 * c + 1] = (H[3] * x + H[4] * y + H[5]) / w + nodes[j, i, 1]
        src[2 * c], src[2 * c + 1] = x, y
    mpx = np.float32(0)
    for k in range(4):
        k1 = (k + 1) & 3
        mpx += np.float32(np.hypot(quad[2 * k1] - quad[2 * k], quad[2 * k1 + 1] - quad[2 * k + 1]) /
                          np.hypot(src[2 * k1] - src[2 * k], src[2 * k1 + 1] - src[2 * k + 1]) / 4)
    return quad, float(mpx)


@pytest.mark.parametrize("span,W,Hh,rot,scale,persp,k1,k2,blur,noise", CASES)
def test_anchor_register_equals_reference(decoder_gpu, span, W, Hh, rot, scale, persp, k1, k2, blur, noise):
    """The registration from the anchors outward (rounds of search through the anchors' homography and lens model,
    the corners judged by that map, the lens model of corners and anchors) equals the reference's anchor_register
    from the same H and corners (no corner is moved here: each sits where the anchors' map puts it)."""
    r = ref_lib()
    cam = sim.Camera(W, Hh, rot_deg=rot, scale=scale, persp=persp, k1=k1, blur=blur, noise=noise, k2=k2)
    img, H, nodes, _ = painted(decoder_gpu, span, cam, seed=2 * span + int(rot))
    quad, mpx = lattice_quad(span, H, nodes)
    f = ref_format(r, span)
    try:
        got = decoder_gpu.anchor_register(span, img, H, mpx, quad)
        Hr, qr = H.copy(), quad.copy()
        rec = np.zeros((49, 8), np.float32)
        model = np.zeros(15)
        have = r.ref_anchor_register(f, ptr(img), W, Hh, ptr(Hr), mpx, ptr(qr), ptr(rec), ptr(model))
        assert have == got["have"] == 1
        assert got["moved"] == 0
        assert np.array_equal(got["H"], Hr) and np.array_equal(got["quad"], qr)
        assert np.array_equal(got["records"], rec), np.abs(got["records"] - rec).max()
        assert np.array_equal(got["model"], model), (got["model"], model)
        assert (rec[:, 7] == 1).sum() >= 40
    finally:
        r.ref_format_free(f)


@pytest.mark.parametrize("span,W,Hh,rot,scale,persp,k1,k2,blur,noise", CASES)
def test_border_refine_equals_reference(decoder_gpu, span, W, Hh, rot, scale, persp, k1, k2, blur, noise):
    """Every marked border node refined level by level on the GPU (through H alone, and through H and the anchors'
    lens model) equals the reference's refine_node run node after node in its order, on the border band blurred as
    the reference blurs it; the nodes it trusts sit within a tenth of a module of the channel's truth."""
    r = ref_lib()
    cam = sim.Camera(W, Hh, rot_deg=rot, scale=scale, persp=persp, k1=k1, blur=blur, noise=noise, k2=k2)
    img, H, nodes_t, _ = painted(decoder_gpu, span, cam, seed=5 * span + int(rot))
    quad, mpx = lattice_quad(span, H, nodes_t)
    reg = decoder_gpu.anchor_register(span, img, H, mpx, quad)
    assert reg["have"] == 1
    f = ref_format(r, span)
    try:
        for m in (None, reg["model"]):
            got = decoder_gpu.border_refine(span, img, reg["H"], mpx, m)
            nx = got.shape[0]
            want = np.zeros((nx, nx, 4), np.float32)
            assert r.ref_refine_border(f, ptr(img), W, Hh, ptr(reg["H"]), mpx, None if m is None else ptr(m),
                                       ptr(want)) == nx
            assert np.array_equal(got, want), (m is None, np.abs(got - want).max(), np.argwhere(got != want)[:4])
        border = np.zeros((nx, nx), bool)
        border[0, :] = border[-1, :] = border[:, 0] = border[:, -1] = True
        assert np.all(got[border, 3] == 1) and not np.any(got[~border, 3])
        marked = border.copy()
        marked[0, 0] = marked[0, -1] = marked[-1, 0] = marked[-1, -1] = False  # the corners: held at 0, no template
        assert np.all(got[~marked, :3] == 0)
        trusted = marked & (got[..., 2] >= 0.35)
        assert trusted.sum() >= 0.9 * marked.sum(), (trusted.sum(), marked.sum())
        Ht = H.astype(np.float64)
        pos = L.node_lattice(span)
        X, Y = np.meshgrid(pos, pos)
        w = Ht[6] * X + Ht[7] * Y + 1
        w2 = reg["H"][6] * X + reg["H"][7] * Y + 1
        tu = (Ht[0] * X + Ht[1] * Y + Ht[2]) / w + nodes_t[..., 0]
        tv = (Ht[3] * X + Ht[4] * Y + Ht[5]) / w + nodes_t[..., 1]
        gu = (reg["H"][0] * X + reg["H"][1] * Y + reg["H"][2]) / w2 + got[..., 0]
        gv = (reg["H"][3] * X + reg["H"][4] * Y + reg["H"][5]) / w2 + got[..., 1]
        e = np.hypot(gu - tu, gv - tv)[trusted] / mpx
        assert np.median(e) < 0.1 and e.max() < 0.35, (np.median(e), e.max())
    finally:
        r.ref_format_free(f)


FINDS = [(256, 1920, 1080, 7.0, 0.85, 0.03, 0.03, 0.06, 0.9, 4.0),
         (128, 1280, 720, -4.0, 0.9, 0.02, -0.02, -0.06, 0.8, 3.0),
         (512, 1920, 1080, 2.0, 0.85, 0.02, 0.02, 0.04, 0.7, 3.0)]


@pytest.mark.parametrize("span,W,Hh,rot,scale,persp,k1,k2,blur,noise", FINDS)
def test_finder_registers_from_the_anchors(decoder_gpu, span, W, Hh, rot, scale, persp, k1, k2, blur, noise):
    """Behind a lens of two radial terms, which the finder's single k1 does not hold, the registration from the anchors
    outward (over the finder's border nodes, and over the border refined by its templates) is closer to the channel's
    truth than the border's lens alone, and decodes at least as many blocks."""
    cam = sim.Camera(W, Hh, rot_deg=rot, scale=scale, persp=persp, k1=k1, blur=blur, noise=noise, k2=k2)
    img, Ht, nodes_t, blocks = painted(decoder_gpu, span, cam, seed=3 * span)
    pos = L.node_lattice(span)
    X, Y = np.meshgrid(pos, pos)
    Pm = np.stack([X.ravel(), Y.ravel(), np.ones(X.size)])

    def place(Hm, nd):
        q = np.asarray(Hm, np.float64).reshape(3, 3) @ Pm
        return q[:2] / q[2] + np.asarray(nd, np.float64).reshape(-1, 2).T

    mpx = scale * Hh / (span + 30)
    res = {}
    for mode in (0, 1, 2):
        decoder_gpu.anchors(mode)
        try:
            fr = decoder_gpu.find(img)
        finally:
            decoder_gpu.anchors(1)
        assert fr is not None and fr["span"] == span
        e = np.hypot(*(place(fr["H"], fr["nodes"]) - place(Ht, nodes_t))) / mpx
        dec = decoder_gpu.decode_reg(VERSION[span], span, img, fr["H"], fr["nodes"], soft=False)
        res[mode] = (float(np.sqrt(np.mean(e * e))), float(e.max()), dec["good"])
    # the default (the anchors over the finder's border) within a fortieth of a module; the reference's chain, its
    # border from the nodes' templates, less close than the edges' fit but closer than the border's lens alone
    assert res[1][0] < 0.6 * res[0][0] and res[1][0] < 0.03 and res[1][1] < 0.2, res
    assert res[2][0] < 0.75 * res[0][0] and res[2][0] < 0.07 and res[2][1] < 0.25, res
    assert res[1][2] >= res[0][2] and res[2][2] >= res[0][2], res
