"""The registration model and its refinement loss on the CPU: the FP64 restatement (oracle/register.py) pinned by the
reference's own interpolation compiled from its sources, by the spline's exactness properties and by finite
differences of its own loss; the product's node field (lizx_reg_nodes) against it; the synthetic channel's lens nodes
against the model's."""
import ctypes as C

import numpy as np
import pytest

import lizard_amd
import lizard_amd as L
from lizard_amd import sim
from _libs import ptr, ref

import register as R  # oracle/register.py


def ref_lib():
    r = ref("scalar")
    if r is None:
        pytest.skip("oracle/_ref not built")
    r.ref_anch_herm4.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_void_p]
    r.ref_anchor_m0.restype = C.c_int
    r.ref_anchor_m0.argtypes = [C.c_int, C.c_int]
    return r


def test_hermite_weights_match_reference():
    r = ref_lib()
    for span in (128, 256, 512):
        t = R.anchor_knots(span)
        for i in range(7):
            assert R.anchor_m0(span, i) == r.ref_anchor_m0(span, i)
        for x in np.concatenate([np.linspace(0, 1, 97), t, t[1:-1] - 1e-9]):
            b = np.zeros(1, np.int32)
            w = np.zeros(4)
            r.ref_anch_herm4(ptr(t), len(t), float(x), ptr(b), ptr(w))
            bb, ww = R.herm4(t, x)
            assert bb == b[0]
            assert np.allclose(ww, w, rtol=0, atol=1e-15)


def test_hermite_interpolates_knots_and_lines():
    t = R.anchor_knots(256)
    v = np.random.default_rng(1).standard_normal(len(t))
    lin = 0.3 + 2.5 * t
    for k, x in enumerate(t):
        b, w = R.herm4(t, x)
        val = sum(w[a] * v[b + a] for a in range(4) if 0 <= b + a < len(t))
        assert abs(val - v[k]) < 1e-12
    for x in np.linspace(0, 1, 50):
        b, w = R.herm4(t, x)
        val = sum(w[a] * lin[b + a] for a in range(4) if 0 <= b + a < len(t))
        assert abs(val - (0.3 + 2.5 * x)) < 1e-12  # slopes from neighbouring knots reproduce a line exactly


def test_anchor_mesh_is_zero_on_the_lattice_edge_and_linear():
    phi = R.anchor_basis(256)
    assert np.all(phi[0] == 0) and np.all(phi[-1] == 0) and np.all(phi[:, 0] == 0) and np.all(phi[:, -1] == 0)
    a = np.random.default_rng(2).standard_normal((7, 7, 2))
    H0 = np.diag([3.0, 3.0, 1.0])
    p1 = np.zeros(R.NPARAM); p1[:8] = H0.ravel()[:8]; p1[9:] = a.ravel()
    p2 = p1.copy(); p2[9:] *= 2.5
    n1, n2 = R.nodes(256, H0, 1920, 1080, p1), R.nodes(256, H0, 1920, 1080, p2)
    assert np.allclose(n2, 2.5 * n1)


@pytest.mark.parametrize("span", [128, 256, 512])
def test_product_nodes_match_oracle(span):
    rng = np.random.default_rng(span)
    H = np.array([[2.9, 0.2, 300.0], [-0.15, 3.1, 120.0], [2e-5, -1e-5, 1.0]])
    p = L.reg_params(H, rng.uniform(-0.05, 0.05), rng.normal(0, 0.4, 98))
    got = L.reg_nodes(span, H.astype(np.float32), 1920, 1080, p)
    want = R.nodes(span, H, 1920, 1080, p.astype(np.float64))
    assert np.abs(got - want).max() < 1e-5 * max(1.0, np.abs(want).max())


def test_channel_lens_nodes_are_the_models():
    cam = sim.Camera(1920, 1080, rot_deg=7.0, scale=0.8, persp=0.03, k1=0.04)
    Hm = cam.H(256)
    want = sim.nodes_for(cam, 256, Hm)
    p = np.zeros(R.NPARAM); p[:8] = (Hm / Hm[2, 2]).ravel()[:8]; p[8] = cam.k1
    assert np.abs(R.nodes(256, Hm / Hm[2, 2], 1920, 1080, p) - want).max() < 1e-4


@pytest.fixture(scope="module")
def small_case():
    version, span = 5, 128
    rng = np.random.default_rng(8)
    cam = sim.Camera(1280, 720, rot_deg=-4.0, scale=0.8, persp=0.02, k1=0.03, blur=0.8, noise=3.0)
    blocks = sim.random_blocks(version, rng)
    img, Hn, nodes = sim.capture(sim.picture(version, blocks), span, cam, seed=2)
    n = L.format_info(version)["n"]
    ref_pic = sim.remodulate(version, blocks, np.ones(9, bool), 0, 1.0).astype(np.float64)
    H0 = Hn.reshape(3, 3).astype(np.float64)
    p = np.zeros(R.NPARAM)
    p[:8] = H0.ravel()[:8] + rng.normal(0, 1e-3, 8) * np.array([1, 1, 100, 1, 1, 100, 1e-4, 1e-4])
    p[8] = cam.k1
    p[9:] = rng.normal(0, 0.2, 98)
    return dict(img=img.numpy(), n=n, span=span, version=version, H0=H0, p=p, ref=ref_pic)


def test_oracle_gradient_matches_finite_differences(small_case):
    c = small_case
    f = lambda p: R.loss_grad(c["img"], c["n"], c["span"], c["H0"], p, c["ref"], 0.9, 0.05)["loss"]  # noqa: E731
    g = R.loss_grad(c["img"], c["n"], c["span"], c["H0"], c["p"], c["ref"], 0.9, 0.05)["grad"]
    # step per parameter in its own units: pixels for translation and the lens / anchors, 1e-4 for the linear part
    steps = np.full(R.NPARAM, 1e-3)
    steps[[0, 1, 3, 4]] = 1e-6
    steps[[6, 7]] = 1e-9
    steps[8] = 1e-5
    for k in [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 57, 58, 105, 106]:
        e = np.zeros(R.NPARAM); e[k] = steps[k]
        fd = (f(c["p"] + e) - f(c["p"] - e)) / (2 * steps[k])
        assert abs(fd - g[k]) <= 2e-3 * max(abs(g[k]), 1e-3 * np.abs(g).max()) + 1e-6, (k, fd, g[k])


def test_oracle_spectral_gradient_matches_finite_differences(small_case):
    """The loss on masked coefficients (the DFT by numpy's FFT) and its gradient through the inverse transform and the
    warp: central differences of the same FP64 loss; and over a mask of every coefficient of a picture that carries
    only listed coefficients, the loss is the picture loss (Parseval over each listed coefficient and its partner)."""
    c = small_case
    u, v = L.format_coefs(c["version"])
    mask = (np.random.default_rng(4).random(u.size) < 0.5).astype(np.uint8)
    sp = (u, v, mask)
    f = lambda p: R.loss_grad(c["img"], c["n"], c["span"], c["H0"], p, c["ref"], 0.9, 0.05,  # noqa: E731
                              spectral=sp)["loss"]
    g = R.loss_grad(c["img"], c["n"], c["span"], c["H0"], c["p"], c["ref"], 0.9, 0.05, spectral=sp)["grad"]
    steps = np.full(R.NPARAM, 1e-3)
    steps[[0, 1, 3, 4]] = 1e-6
    steps[[6, 7]] = 1e-9
    steps[8] = 1e-5
    for k in [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 57, 58, 105, 106]:
        e = np.zeros(R.NPARAM); e[k] = steps[k]
        fd = (f(c["p"] + e) - f(c["p"] - e)) / (2 * steps[k])
        assert abs(fd - g[k]) <= 2e-3 * max(abs(g[k]), 1e-3 * np.abs(g).max()) + 1e-6, (k, fd, g[k])
    # Parseval: a residual made of listed coefficients only (none at DC or on its own partner) is wholly seen
    n = c["n"]
    rng = np.random.default_rng(5)
    Z = np.zeros((n, n), complex)
    h = rng.normal(size=u.size) + 1j * rng.normal(size=u.size)
    Z[v % n, u % n] = h
    Z[(-v) % n, (-u) % n] = np.conj(h)
    pic = np.real(np.fft.ifft2(Z)) * n  # a real picture whose spectrum is exactly the listed coefficients and partners
    Fe = np.fft.fft2(pic)
    spec = 2.0 / (n * n) * np.sum(np.abs(Fe[v % n, u % n]) ** 2)
    assert abs(spec - np.sum(pic * pic)) <= 1e-9 * np.sum(pic * pic)


def _border_case(span, W, Hh, rot, scale, persp, cx, cy, seed, K_true, noise):
    """A homography and a border field: the lens model's own terms (K_true radial ones) about the frame's centre plus
    a smooth residual and noise, on every node (the fill reads the border only); every fifth marked node untrusted."""
    from lizard_amd import sim
    cam = sim.Camera(W, Hh, rot_deg=rot, scale=scale, persp=persp)
    cam.cx, cam.cy = cx, cy
    Hm = cam.H(span)
    H = (Hm / Hm[2, 2]).astype(np.float32).ravel()
    pos = lizard_amd.node_lattice(span)
    nx = len(pos)
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(pos, pos, indexing="xy")
    w = H[6] * X + H[7] * Y + 1
    u = (H[0] * X + H[1] * Y + H[2]) / w
    v = (H[3] * X + H[4] * Y + H[5]) / w
    R = np.hypot(W / 2, Hh / 2)
    qx, qy = (u - W / 2) / R, (v - Hh / 2) / R
    s = qx * qx + qy * qy
    t = rng.normal(0, 0.5, 8)
    dx = t[0] * qx + t[1] * qy + t[4] - t[6] * qx * qx - t[7] * qx * qy
    dy = t[2] * qx + t[3] * qy + t[5] - t[6] * qx * qy - t[7] * qy * qy
    sk = s.copy()
    for k in range(K_true):
        c = rng.normal(0, 3.0)
        dx, dy = dx + c * sk * qx, dy + c * sk * qy
        sk = sk * s
    dx = dx + 0.3 * np.sin(X / 40.0) + rng.normal(0, noise, X.shape)
    dy = dy + 0.3 * np.cos(Y / 35.0) + rng.normal(0, noise, X.shape)
    nodes = np.stack([dx, dy], -1).astype(np.float32)
    use = np.ones((nx, nx), np.uint8)
    use.ravel()[::5] = 0
    return H, nodes, use


BORDERS = [(128, 1280, 720, 5.0, 0.8, 0.02, 0, 0, 1, 1, 0.02), (256, 1920, 1080, -8.0, 0.85, 0.03, 0, 0, 2, 2, 0.05),
           (512, 1920, 1080, 3.0, 0.9, 0.01, 0, 0, 3, 3, 0.02), (256, 1920, 1080, 12.0, 0.6, 0.04, 500, 250, 4, 3, 0.01),
           (512, 1920, 1080, -2.0, 0.7, 0.0, -600, 0, 5, 2, 0.1), (128, 1280, 720, 0.0, 0.5, 0.0, 300, 180, 6, 0, 0.0)]


@pytest.mark.parametrize("span,W,Hh,rot,scale,persp,cx,cy,seed,K,noise", BORDERS)
def test_border_fill_equals_reference(span, W, Hh, rot, scale, persp, cx, cy, seed, K, noise):
    """lizx_reg_border_fill (the lens model of the border, K by AICc, and the Coons patch) gives the reference's
    border_fill_axis nodes to the bit, about the frame's centre and about given axes inside and outside the quad;
    the lens field of the model it returns is the reference's lensm_at at every node."""
    r = ref("scalar")
    if r is None:
        pytest.skip("oracle/_ref not built")
    H, nodes, use = _border_case(span, W, Hh, rot, scale, persp, cx, cy, seed, K, noise)
    nx = nodes.shape[0]
    version = {128: 5, 256: 20, 512: 64}[span]
    info = lizard_amd.format_info(version)
    f = r.ref_format_new(info["n"], info["subch"], span)
    for axis in (None, (W * 0.3, Hh * 0.6), (W * 1.4, -Hh * 0.2)):
        want = np.ascontiguousarray(nodes.copy())
        ax, ay = (float("nan"), float("nan")) if axis is None else axis
        r.ref_border_fill_axis(f, ptr(H), W, Hh, ax, ay, ptr(want), ptr(use))
        got, Kgot, model = lizard_amd.border_fill(span, H, W, Hh, nodes, use=use, axis=axis)
        assert np.array_equal(got, want), (axis, np.abs(got - want).max())
        assert -1 <= Kgot <= 3
        # the border itself is left as it was
        border = np.ones((nx, nx), bool)
        border[1:-1, 1:-1] = False
        assert np.array_equal(got[border], nodes[border])
        field = lizard_amd.lens_field(span, H, model)
        pos = lizard_amd.node_lattice(span)
        t = np.ascontiguousarray(model[4:], np.float64)
        for j in (0, nx // 2, nx - 1):
            for i in (0, nx // 3, nx - 1):
                w = H[6] * pos[i] + H[7] * pos[j] + np.float32(1)
                u = np.float32(H[0] * pos[i] + H[1] * pos[j] + H[2]) / w
                v = np.float32(H[3] * pos[i] + H[4] * pos[j] + H[5]) / w
                dx, dy = C.c_float(), C.c_float()
                r.ref_lensm_at(int(model[0]), ptr(t), model[1], model[2], model[3], float(u), float(v), C.byref(dx),
                               C.byref(dy))
                if Kgot >= 0:
                    assert abs(field[j, i, 0] - dx.value) <= 1e-5 * max(1.0, abs(dx.value))
                    assert abs(field[j, i, 1] - dy.value) <= 1e-5 * max(1.0, abs(dy.value))
    r.ref_format_free(f)


def test_border_fill_reproduces_a_lens_field():
    """Border nodes that are exactly a lens model's field (no residual): the interior is that field, to the float
    rounding of the border's values carried through the least squares (the border spans only the outer radii, so the
    fit extrapolates inward: a hundredth of a pixel), since the Coons patch carries nothing inward."""
    H, nodes, use = _border_case(256, 1920, 1080, 4.0, 0.8, 0.02, 0, 0, 9, 2, 0.0)
    model = np.zeros(15)
    model[0], model[1], model[2], model[3] = 2, 960.0, 540.0, np.hypot(960.0, 540.0)
    model[4:14] = [0.3, -0.2, 0.1, 0.4, 0.5, -0.7, 0.05, -0.03, 2.0, -1.0]
    field = lizard_amd.lens_field(256, H, model)
    got, K, fitted = lizard_amd.border_fill(256, H, 1920, 1080, field)
    assert K >= 1
    assert np.abs(got - field).max() < 0.02


def _anchors(span, H, rng, field_fn, found_frac=0.85, noise=0.1):
    """The 7 x 7 anchors' readings: centres in modules, displacement from H = field at the anchor's place plus noise,
    found with probability found_frac."""
    ha = 0.5 * (span // 64)
    out = np.zeros((49, 5), np.float32)
    for k in range(49):
        i, j = k % 7, k // 7
        X = R.anchor_m0(span, i) + ha
        Y = R.anchor_m0(span, j) + ha
        fx, fy = field_fn(X, Y)
        out[k] = (X, Y, fx + rng.normal(0, noise), fy + rng.normal(0, noise), float(rng.random() < found_frac))
    return out


@pytest.mark.parametrize("span,W,Hh,rot,scale,persp,cx,cy,seed,K,noise", BORDERS[:5])
def test_anchor_fill_equals_reference(span, W, Hh, rot, scale, persp, cx, cy, seed, K, noise):
    """lizx_reg_anchor_fill (the robust lens refit over border and anchors, the Coons patch, the shrunk anchor
    residuals' Hermite mesh) gives the reference's border_fill_anchors nodes to the bit; an outlying border node is
    dropped there and here alike; with no anchor found the nodes are the border fill's but for the model's refit."""
    r = ref("scalar")
    if r is None:
        pytest.skip("oracle/_ref not built")
    H, nodes, use = _border_case(span, W, Hh, rot, scale, persp, cx, cy, seed, K, noise)
    nx = nodes.shape[0]
    nodes[0, nx // 2] += 6.0  # an outlier on the top side
    rng = np.random.default_rng(seed + 100)
    field = lambda X, Y: (0.8 * np.sin(X / 60.0), -0.6 * np.cos(Y / 50.0))  # noqa: E731
    version = {128: 5, 256: 20, 512: 64}[span]
    info = L.format_info(version)
    f = r.ref_format_new(info["n"], info["subch"], span)
    mpx = float(np.hypot(H[0], H[3]))
    for found, axis in ((0.85, None), (1.0, (W * 0.3, Hh * 0.6)), (0.3, (W * 1.4, -Hh * 0.2)), (0.0, None)):
        an = _anchors(span, H, rng, field, found)
        want = np.ascontiguousarray(nodes.copy())
        ax, ay = (float("nan"), float("nan")) if axis is None else axis
        r.ref_border_fill_anchors(f, ptr(H), W, Hh, ax, ay, ptr(want), ptr(use), ptr(an), mpx)
        got, Kgot, model = L.anchor_fill(span, H, W, Hh, nodes, an, mpx, use=use, axis=axis)
        assert np.array_equal(got, want), (found, axis, np.abs(got - want).max())
        assert 0 <= Kgot <= 3
    r.ref_format_free(f)
