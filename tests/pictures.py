"""Test pictures: seeded random blocks through the oracle's encoder and FP64 modulator, then a channel on the picture
grid: a brightness ramp (what the detrend removes), junk on the anchors' squares (what the erase removes), a sub-sample
grid shift (what the pilots recover) and AWGN."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import picture  # noqa: E402

from _libs import oracle, ptr  # noqa: E402


class Frame:
    def __init__(self, version, span=256, seed=0, snr_db=20.0, shift=(0.0, 0.0), ramp=0.05, junk=True, cnt=0,
                 blur=0.0):
        o = oracle()
        self.version, self.span, self.subch = version, span, 16 * version
        self.n = n = o.or_n_for(self.subch)
        f = o.or_format_new(n, self.subch)
        info = np.zeros(6, np.int32)
        o.or_format_info(f, ptr(info))
        self.blocks_n = B = int(info[2])
        S = self.subch * 320
        self.pos = np.ctypeslib.as_array(o.or_format_pos(f), (S,)).copy()
        self.u, self.v = picture.coef_uv(self.pos, n)
        rate, s0, ns = np.zeros(B, np.int32), np.zeros(B, np.int32), np.zeros(B, np.int32)
        o.or_format_blocks(f, ptr(rate), ptr(s0), ptr(ns))
        self.rate, self.slot0, self.nslots = rate, s0, ns
        rng = np.random.default_rng(seed)
        self.blocks = rng.integers(0, 256, B * 476).astype(np.uint8)
        self.slots = np.zeros(2 * S, np.uint8)
        o.or_frame_slots(f, ptr(self.blocks), cnt, ptr(self.slots))
        self.white = np.zeros(2 * S, np.uint8)
        o.or_whiten(ptr(self.white), 2 * S)
        o.or_format_free(f)
        pic = picture.modulate(n, self.subch, self.slots, self.u, self.v, shift)
        self.blur = blur
        if blur > 0:
            pic = picture.blur(pic, blur)
        yy, xx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        pic = pic + ramp * ((xx - n / 2) / n + 0.5 * ((yy - n / 2) / n) ** 2)
        if junk:
            lo, hi = picture.erase_ranges(n, span)
            for a, b in zip(lo, hi):
                for c_, d in zip(lo, hi):
                    pic[a:b + 1, c_:d + 1] = rng.uniform(0, 1, (b - a + 1, d - c_ + 1))
        sig = np.std(pic)
        pic = pic + rng.standard_normal(pic.shape) * sig * 10 ** (-snr_db / 20)
        self.grid = pic.astype(np.float32)
        # the pilots: tail coefficients of the 7/8 and 3/4 blocks (SPEC 7.3)
        nt = {6: 4480, 4: 5120, 3: 5760, 2: 7680}
        pil = []
        for b in range(B):
            first = s0[b] // 2
            for s in range(nt[int(rate[b])] // 2, ns[b] // 2):
                pil.append(first + s)
        self.pilots = np.array(pil, np.int64)

    def fp64(self, align=None, otf_sigma=0.0):
        return picture.finish64(self.grid, self.n, self.span, self.u, self.v, self.white, self.pilots, align, otf_sigma)


def lattice(span):
    side = span + 30
    pos = []
    c = np.float32(3.5)
    while c < np.float32(side) - np.float32(3.5) - np.float32(8) and len(pos) < 39:
        pos.append(float(c))
        c = np.float32(c + 16)
    pos.append(float(np.float32(side) - np.float32(3.5)))
    return np.array(pos, np.float32)


def homography(src, dst):
    """The 3x3 projective map taking four src points to four dst points (h8 = 1), in FP64."""
    A, b = [], []
    for (x, y), (u, v) in zip(src, dst):
        A.append([x, y, 1, 0, 0, 0, -u * x, -u * y]); b.append(u)
        A.append([0, 0, 0, x, y, 1, -v * x, -v * y]); b.append(v)
    h = np.linalg.solve(np.array(A, float), np.array(b, float))
    return np.append(h, 1.0).reshape(3, 3)


class Capture:
    """A symbol painted by the reference encoder (oracle/_ref: lizard_encode + lizard_paint_grey), projected into a
    camera frame by a homography G (grey-paint pixel edge coordinates -> capture pixel edge coordinates), blurred,
    with AWGN, as uint8; with the registration the decoder needs (H = G composed with the module -> paint pixel map,
    nodes zero, or a given node displacement field). Needs the compiled reference."""

    def __init__(self, version, span, W=1280, Hh=720, rot_deg=5.0, persp=0.02, scale=0.8, blur=0.8, noise=4.0,
                 seed=0, cnt=0, node_fn=None):
        import scipy.ndimage as ndi
        from _libs import ref
        r = ref("scalar")
        o = oracle()
        self.version, self.span, self.subch = version, span, 16 * version
        self.n = n = o.or_n_for(self.subch)
        f = r.ref_format_new(n, self.subch, span)
        info = np.zeros(10, np.int32)
        r.ref_format_info(f, ptr(info))
        self.blocks_n, pxm, px = int(info[2]), int(info[7]), int(info[8])
        rng = np.random.default_rng(seed)
        self.blocks = rng.integers(0, 256, self.blocks_n * 476).astype(np.uint8)
        drive = np.zeros(px * px, np.float32)
        r.ref_encode(f, ptr(self.blocks), cnt, ptr(drive), None)
        Wg = px + 4 * pxm
        grey = np.zeros(Wg * Wg, np.uint8)
        r.ref_paint_grey(f, ptr(drive), ptr(grey))
        r.ref_format_free(f)
        grey = grey.reshape(Wg, Wg).astype(np.float64)
        # G: the paint's square onto the frame, rotated, scaled to `scale` of the frame height, a little perspective
        s = scale * Hh / Wg
        th = np.deg2rad(rot_deg)
        cxy = np.array([W / 2, Hh / 2])
        corners = np.array([[0, 0], [Wg, 0], [Wg, Wg], [0, Wg]], float) - Wg / 2
        R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
        dst = (corners @ R.T) * s + cxy
        dst[1] += np.array([persp, -persp]) * Hh * 0.5
        dst[2] -= np.array([persp, persp]) * Hh * 0.25
        G = homography(corners + Wg / 2, dst)
        Gi = np.linalg.inv(G)
        Y, X = np.meshgrid(np.arange(Hh) + 0.5, np.arange(W) + 0.5, indexing="ij")
        den = Gi[2, 0] * X + Gi[2, 1] * Y + Gi[2, 2]
        qx = (Gi[0, 0] * X + Gi[0, 1] * Y + Gi[0, 2]) / den - 0.5
        qy = (Gi[1, 0] * X + Gi[1, 1] * Y + Gi[1, 2]) / den - 0.5
        img = ndi.map_coordinates(grey, [qy, qx], order=1, mode="constant", cval=255.0)
        if blur > 0:
            img = ndi.gaussian_filter(img, blur)
        img = img + rng.standard_normal(img.shape) * noise
        self.img = np.clip(np.rint(img), 0, 255).astype(np.uint8)
        M = np.array([[pxm, 0, 2 * pxm], [0, pxm, 2 * pxm], [0, 0, 1]], float)
        Hm = G @ M
        self.H = (Hm / Hm[2, 2]).astype(np.float32).ravel()
        self.node_x = lattice(span)
        nx = len(self.node_x)
        self.nodes = np.zeros((nx, nx, 2), np.float32)
        if node_fn is not None:
            self.nodes = node_fn(self.node_x).astype(np.float32)


def format_tables(version):
    """The oracle's tables of a format: n, subch, blocks, (u, v) of each coefficient, the whitening of every slot and
    the pilot coefficients (the tails of the 7/8 and 3/4 blocks, SPEC 7.3)."""
    o = oracle()
    subch = 16 * version
    n = o.or_n_for(subch)
    f = o.or_format_new(n, subch)
    info = np.zeros(6, np.int32)
    o.or_format_info(f, ptr(info))
    B = int(info[2])
    S = subch * 320
    pos = np.ctypeslib.as_array(o.or_format_pos(f), (S,)).copy()
    rate, s0, ns = np.zeros(B, np.int32), np.zeros(B, np.int32), np.zeros(B, np.int32)
    o.or_format_blocks(f, ptr(rate), ptr(s0), ptr(ns))
    o.or_format_free(f)
    white = np.zeros(2 * S, np.uint8)
    o.or_whiten(ptr(white), 2 * S)
    nt = {6: 4480, 4: 5120, 3: 5760, 2: 7680}
    pil = [s0[b] // 2 + s for b in range(B) for s in range(nt[int(rate[b])] // 2, ns[b] // 2)]
    u, v = picture.coef_uv(pos, n)
    return dict(n=n, subch=subch, blocks=B, u=u, v=v, white=white, pilots=np.array(pil, np.int64))


def soft64(grid, version, span):
    """The FP64 restatement's finish of a row-major n x n grid (oracle/picture.py finish64)."""
    t = format_tables(version)
    return picture.finish64(grid, t["n"], span, t["u"], t["v"], t["white"], t["pilots"])


def turned(version, coef, align):
    """Coefficients (re, im interleaved, coefficient order) turned by a pilot shift through the table index the
    float path takes for it (oracle/picture.py turn_index), the angle exact."""
    t = format_tables(version)
    y = np.asarray(coef, np.float64)[0::2] + 1j * np.asarray(coef, np.float64)[1::2]
    if align[0] or align[1]:
        y = y * np.exp(-2j * np.pi * picture.turn_index(t["u"], t["v"], t["n"], *align) / picture.ALIGN_TURNS)
    out = np.empty(2 * y.size)
    out[0::2], out[1::2] = y.real, y.imag
    return out


def float_soft(yt):
    """A float path's soft values 8 clamp(k y, +-10) from its turned coefficients (re, im interleaved, float32 values):
    each sub-channel's k from the moments of its 320 coefficients in double, as the reference's scalar path forms it
    (liblizard/src/lizard_phy.c:1322-1332, :1760-1768), rounded to float, then k y and 8 (.) in float."""
    f32 = np.float32
    y = np.asarray(yt, np.float64).reshape(-1, 640)
    e = y[:, 0::2] ** 2 + y[:, 1::2] ** 2
    m2, m4 = e.sum(1) / 320, (e * e).sum(1) / 320
    a2 = np.sqrt(np.maximum(2 * m2 * m2 - m4, 0))
    nv = np.maximum(m2 - a2, 0.02 * m2)
    k = (0.7 * 2 * np.sqrt(2 * a2) / nv).astype(f32)
    ky = (k[:, None] * np.asarray(yt, f32).reshape(-1, 640)).ravel()
    return (f32(8.0) * np.clip(ky, f32(-10.0), f32(10.0))).astype(np.float64)


def assert_ties(a, b, soft_a, yt_b, spread=1e-2):
    """Two int8 soft-value arrays a (this build's; soft_a its floats k y before the clamp) and b (the reference's)
    may differ by one, and only at rounding ties: where they differ, this build's 8 k y and the reference's own
    (float_soft of its turned coefficients yt_b, which must round to b) lie on either side of the half-step between a
    and b, within `spread` of each other (relative, at least absolute). Returns (fraction of slots that differ, the
    largest relative distance between the two floats there)."""
    a = np.asarray(a, np.int64)
    b = np.asarray(b, np.int64)
    d = a != b
    if not d.any():
        return 0.0, 0.0
    assert np.abs(a - b).max() <= 1
    f32 = np.float32
    idx = np.nonzero(d)[0]
    xa = (f32(8.0) * np.clip(np.asarray(soft_a, f32), f32(-10.0), f32(10.0))).astype(np.float64)[idx]
    xb = float_soft(yt_b)[idx]
    assert np.array_equal(np.rint(xa).astype(np.int64), a[idx]), "this build's floats do not round to its values"
    assert np.array_equal(np.rint(xb).astype(np.int64), b[idx]), "the reference's floats are not reproduced"
    h = (a[idx] + b[idx]) / 2.0
    rel = np.abs(xa - xb) / np.maximum(1.0, np.abs(xa))
    bad = ~(((xa - h) * (xb - h) <= 0) & (rel <= spread))
    assert not bad.any(), ("%d of %d differing soft values are not rounding ties: slots %s, 8 k y here %s, there %s"
                           % (int(bad.sum()), int(d.sum()), idx[bad][:8].tolist(), xa[bad][:8].tolist(),
                              xb[bad][:8].tolist()))
    return float(d.mean()), float(rel.max())


def ref_grid_finish(r, version, span, grid):
    """The compiled reference's finish of a row-major grid (ref_finish_grid, lizard_finish with the grid): its soft
    values, estimates, pilot shift, coefficients (before the turn), verdicts."""
    t = format_tables(version)
    B, S, n = t["blocks"], t["subch"] * 320, t["n"]
    f = r.ref_format_new(n, t["subch"], span)
    strip = picture.to_strips(grid, n)
    bl, ok, its, be, bp = (np.zeros(B * 476, np.uint8), np.zeros(B, np.uint8), np.zeros(B, np.int8),
                           np.zeros(B, np.float32), np.zeros(B, np.float32))
    llr, est, rms, al, coef = (np.zeros(2 * S, np.int8), np.zeros(t["subch"]), np.zeros(t["subch"], np.float32),
                               np.zeros(2, np.float32), np.zeros(2 * S, np.float32))
    good = r.ref_finish_grid(f, ptr(strip), ptr(bl), ptr(ok), ptr(its), ptr(be), ptr(bp), ptr(llr), ptr(est),
                             ptr(rms), ptr(al), ptr(coef))
    r.ref_format_free(f)
    return dict(good=good, blocks=bl.reshape(B, 476), ok=ok, its=its, blk_est=be, blk_pilot=bp, llr=llr, sub_est=est,
                sub_rms=rms, align=(float(al[0]), float(al[1])), coef=coef)


def ties_against_reference(version, out, ref_out):
    """assert_ties of this build's soft values against the reference's (its coefficients turned by its own pilot
    shift); also the two paths' float soft values within 1e-4 relative L2 over the frame. Returns the fraction of
    slots whose int8 values differ."""
    xb = float_soft(turned(version, ref_out["coef"], ref_out["align"]))
    assert np.array_equal(np.rint(xb).astype(np.int64), np.asarray(ref_out["llr"], np.int64)), "reference floats"
    f32 = np.float32
    xa = (f32(8.0) * np.clip(np.asarray(out["soft"], f32), f32(-10.0), f32(10.0))).astype(np.float64)
    rel = np.linalg.norm(xa - xb) / np.linalg.norm(xb)
    assert rel < 1e-4, rel
    frac, _ = assert_ties(out["llr"], ref_out["llr"], out["soft"], turned(version, ref_out["coef"], ref_out["align"]))
    return frac


def ref_finish_soft(r, version, span, llr, est):
    """The compiled reference's decline gate, LDPC and CRC (lizard_finish_ext with ext_llr, ext_est,
    liblizard/src/lizard_phy.c:1566) on given soft values and sub-channel estimates: good, ok, its, blocks, blk_est."""
    t = format_tables(version)
    B = t["blocks"]
    f = r.ref_format_new(t["n"], t["subch"], span)
    bl, ok, its, be = np.zeros(B * 476, np.uint8), np.zeros(B, np.uint8), np.zeros(B, np.int8), np.zeros(B, np.float32)
    good = r.ref_finish_llr(f, ptr(np.ascontiguousarray(llr, np.int8)), ptr(np.ascontiguousarray(est, np.float32)),
                            ptr(bl), ptr(ok), ptr(its), ptr(be))
    r.ref_format_free(f)
    return dict(good=good, ok=ok, its=its, blocks=bl.reshape(B, 476), blk_est=be)


def assert_finish_equal(out, want):
    """The GPU's verdicts, iterations and verified bytes equal a reference finish of the same soft values."""
    assert np.array_equal(out["its"], want["its"]), np.nonzero(out["its"] != want["its"])
    assert np.array_equal(out["ok"].astype(bool), want["ok"].astype(bool))
    v = want["ok"].astype(bool)
    assert np.array_equal(np.asarray(out["blocks"]).reshape(-1, 476)[v], want["blocks"][v])
