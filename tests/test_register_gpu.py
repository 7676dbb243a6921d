"""The hand-written adjoint of the warp and the refinement loss on the GPU (lizx_reg_loss_grad) against the FP64
restatement (oracle/register.py), wrapped in a torch.autograd.Function and checked by torch's gradcheck against finite
differences of the GPU loss; Gauss-Newton refinement (lizx_reg_refine) from a homography whose corners are 1.5 px off,
and the blocks it recovers; the same over H and the lens's k1 together (lizx_reg_refine_k1) from a registration with
no lens."""
import numpy as np
import pytest

import lizard_amd as L
from lizard_amd import sim

import register as R  # oracle/register.py

pytestmark = pytest.mark.gpu

SIDE = {128: 158, 256: 286, 512: 542}


def corners(span):
    s = SIDE[span]
    return np.array([[0, 0], [s, 0], [s, s], [0, s]], float)


def perturbed(Ht, span, rng, px=1.5):
    c = corners(span)
    cp = np.stack(R.project(Ht, c[:, 0], c[:, 1]), -1)
    return sim.homography(c, cp + rng.uniform(-px, px, (4, 2)))


def corner_error(Ht, params, span):
    H = np.append(np.asarray(params[:8], np.float64), 1.0).reshape(3, 3)
    c = corners(span)
    a = np.stack(R.project(Ht, c[:, 0], c[:, 1]), -1)
    b = np.stack(R.project(H, c[:, 0], c[:, 1]), -1)
    return float(np.abs(a - b).max())


def make_case(version, span, seed, W=1920, Hh=1080, k1=0.03, known=None, k2=0.0):
    rng = np.random.default_rng(seed)
    cam = sim.Camera(W, Hh, rot_deg=rng.uniform(-8, 8), scale=0.85, persp=0.03, k1=k1, blur=1.0, noise=4.0, k2=k2)
    blocks = sim.random_blocks(version, rng)
    img, Hn, nodes = sim.capture(sim.picture(version, blocks), span, cam, seed=seed)
    Ht = cam.H(span)
    Ht = Ht / Ht[2, 2]
    n = L.format_info(version)["n"]
    s = np.sqrt(abs(np.linalg.det(Ht[:2, :2]))) * span / n  # capture pixels a picture sample
    blur = np.sqrt(cam.blur ** 2 + 1.0 / 3.0) / s
    B = L.format_info(version)["blocks"]
    ref_pic = sim.remodulate(version, blocks, np.ones(B, bool) if known is None else known, 0, blur)
    H0 = perturbed(Ht, span, rng)
    return dict(version=version, span=span, img=img.numpy(), Ht=Ht, H0=H0, k1=k1, blocks=blocks, ref=ref_pic, n=n,
                nodes_true=nodes)


@pytest.fixture(scope="module")
def case384():
    return make_case(5, 128, 3, W=1280, Hh=720)


@pytest.mark.parametrize("which", ["384", "768"])
def test_adjoint_matches_fp64(decoder_gpu, case384, which):
    c = case384 if which == "384" else make_case(20, 256, 4)
    rng = np.random.default_rng(11)
    p = L.reg_params(c["H0"], c["k1"], rng.normal(0, 0.3, 98))
    loss, grad = decoder_gpu.reg_loss_grad(c["version"], c["span"], c["img"], c["H0"], p, c["ref"], 0.9, 0.04)
    want = R.loss_grad(c["img"], c["n"], c["span"], c["H0"].astype(np.float32).astype(np.float64),
                       p.astype(np.float64), c["ref"].astype(np.float64), 0.9, 0.04)
    assert abs(loss - want["loss"]) <= 1e-4 * want["loss"]
    g, gw = grad, want["grad"]
    # each group of parameters against its own scale (their units differ by orders of magnitude)
    for sl in (slice(0, 2), slice(2, 3), slice(3, 5), slice(5, 6), slice(6, 8), slice(8, 9), slice(9, None)):
        err = np.linalg.norm(g[sl] - gw[sl]) / max(np.linalg.norm(gw[sl]), 1e-30)
        assert err < 2e-3, (sl, err)


def test_autograd_gradcheck(decoder_gpu, case384):
    """torch.autograd.gradcheck of the GPU loss (through the C ABI adjoint) against its own finite differences, over
    a scaled perturbation of 12 parameters (one unit ~ one pixel of movement for each)."""
    torch = pytest.importorskip("torch")
    c = case384
    p0 = L.reg_params(c["H0"], c["k1"], np.random.default_rng(5).normal(0, 0.3, 98)).astype(np.float64)
    idx = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 58, 106])
    scale = np.array([1 / 300, 1 / 300, 1.0, 1 / 300, 1 / 300, 1.0, 3e-6, 3e-6, 1 / 300, 1.0, 1.0, 1.0])

    class RegLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, delta):
            p = p0.copy()
            p[idx] += delta.detach().cpu().numpy() * scale
            loss, grad = decoder_gpu.reg_loss_grad(c["version"], c["span"], c["img"], c["H0"], p.astype(np.float32),
                                                   c["ref"], 0.9, 0.04)
            ctx.save_for_backward(torch.as_tensor(grad[idx] * scale, dtype=torch.float64))
            return torch.tensor(loss, dtype=torch.float64)

        @staticmethod
        def backward(ctx, g):
            (gr,) = ctx.saved_tensors
            return g * gr

    delta = torch.zeros(len(idx), dtype=torch.float64, requires_grad=True)
    out = RegLoss.apply(delta)
    out.backward()
    assert torch.isfinite(delta.grad).all() and delta.grad.abs().max() > 0
    # central differences at 0.02 units agree with the adjoint to 4e-4 of the largest component (measured on this
    # case: at most 0.16 against components up to 393, each within 0.2 % of itself)
    assert torch.autograd.gradcheck(RegLoss.apply, (delta,), eps=0.02, atol=0.05, rtol=0.005, nondet_tol=1e-3)


def _half_mask(version, seed):
    u, _ = L.format_coefs(version)
    return (np.random.default_rng(seed).random(u.size) < 0.5).astype(np.uint8)


@pytest.mark.parametrize("which", ["384", "768"])
def test_spectral_adjoint_matches_fp64(decoder_gpu, case384, which):
    """The loss on masked coefficients and its adjoint through the FFT and the warp (lizx_reg_loss_grad_spectral)
    against the FP64 restatement (numpy's FFT, the same chain rule): the loss within 1e-4, each parameter group's
    gradient within 2e-3 relative."""
    c = case384 if which == "384" else make_case(20, 256, 4)
    rng = np.random.default_rng(12)
    p = L.reg_params(c["H0"], c["k1"], rng.normal(0, 0.3, 98))
    mask = _half_mask(c["version"], 7)
    u, v = L.format_coefs(c["version"])
    loss, grad = decoder_gpu.reg_loss_grad_spectral(c["version"], c["span"], c["img"], c["H0"], p, c["ref"], mask,
                                                    0.9, 0.04)
    want = R.loss_grad(c["img"], c["n"], c["span"], c["H0"].astype(np.float32).astype(np.float64),
                       p.astype(np.float64), c["ref"].astype(np.float64), 0.9, 0.04, spectral=(u, v, mask))
    assert abs(loss - want["loss"]) <= 1e-4 * want["loss"], (loss, want["loss"])
    g, gw = grad, want["grad"]
    for sl in (slice(0, 2), slice(2, 3), slice(3, 5), slice(5, 6), slice(6, 8), slice(8, 9), slice(9, None)):
        err = np.linalg.norm(g[sl] - gw[sl]) / max(np.linalg.norm(gw[sl]), 1e-30)
        assert err < 2e-3, (sl, err)
    # no coefficient: no loss, no gradient
    z, gz = decoder_gpu.reg_loss_grad_spectral(c["version"], c["span"], c["img"], c["H0"], p, c["ref"],
                                               np.zeros_like(mask), 0.9, 0.04)
    assert z == 0 and not np.any(gz)


def test_spectral_autograd_gradcheck(decoder_gpu, case384):
    """torch.autograd.gradcheck of the spectral loss through the C ABI adjoint against its own finite differences."""
    torch = pytest.importorskip("torch")
    c = case384
    p0 = L.reg_params(c["H0"], c["k1"], np.random.default_rng(6).normal(0, 0.3, 98)).astype(np.float64)
    mask = _half_mask(c["version"], 8)
    idx = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 58, 106])
    scale = np.array([1 / 300, 1 / 300, 1.0, 1 / 300, 1 / 300, 1.0, 3e-6, 3e-6, 1 / 300, 1.0, 1.0, 1.0])

    class SpecLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, delta):
            p = p0.copy()
            p[idx] += delta.detach().cpu().numpy() * scale
            loss, grad = decoder_gpu.reg_loss_grad_spectral(c["version"], c["span"], c["img"], c["H0"],
                                                            p.astype(np.float32), c["ref"], mask, 0.9, 0.04)
            ctx.save_for_backward(torch.as_tensor(grad[idx] * scale, dtype=torch.float64))
            return torch.tensor(loss, dtype=torch.float64)

        @staticmethod
        def backward(ctx, g):
            (gr,) = ctx.saved_tensors
            return g * gr

    delta = torch.zeros(len(idx), dtype=torch.float64, requires_grad=True)
    out = SpecLoss.apply(delta)
    out.backward()
    assert torch.isfinite(delta.grad).all() and delta.grad.abs().max() > 0
    # measured on this case: differences at most 0.14 against components up to 198, each within 0.25 % of itself
    assert torch.autograd.gradcheck(SpecLoss.apply, (delta,), eps=0.02, atol=0.05, rtol=0.005, nondet_tol=1e-3)


def test_spectral_refine_recovers_the_registration(decoder_gpu):
    """Refinement on the spectral loss (lizx_reg_refine_spectral) from corners 1.5 px off and no lens, over the
    coefficients of a third of the blocks (the decoded ones in a decision-directed round; the rest unknown): the loss
    falls step by step, the corners come within 0.3 px of the truth and k1 toward the camera's."""
    B = L.format_info(20)["blocks"]
    known = np.zeros(B, bool)
    known[::3] = True
    c = make_case(20, 256, 9, k1=0.04, known=known)
    mask = sim.known_mask(20, known)
    p = L.reg_params(c["H0"], 0.0)
    e0 = corner_error(c["Ht"], p, c["span"])
    p1, tr = decoder_gpu.reg_refine_spectral(20, 256, c["img"], c["H0"], p, c["ref"], mask, steps=12)
    assert np.all(np.diff(tr) <= 0) and tr[-1] < 0.7 * tr[0], tr
    e1 = corner_error(c["Ht"], p1, c["span"])
    assert e1 < 0.3 < e0, (e0, e1)
    assert abs(p1[8] - 0.04) < abs(0.0 - 0.04), p1[8]
    assert np.array_equal(p1[9:], p[9:])  # the anchors held


def test_spectral_refine_moves_the_anchors_for_a_second_lens_term(decoder_gpu):
    """Behind a lens of two radial terms (k2, which the H and k1 model cannot hold) the spectral refinement that also
    moves the anchors (lizx_reg_refine_spectral with anchors) ends at a lower loss than the one over H and k1 alone,
    and its node lattice is closer to the channel's (rms over the interior nodes)."""
    B = L.format_info(20)["blocks"]
    known = np.ones(B, bool)
    c = make_case(20, 256, 13, k1=0.02, known=known, k2=0.05)
    mask = sim.known_mask(20, known)
    p = L.reg_params(c["Ht"], 0.02)
    W, Hh = c["img"].shape[1], c["img"].shape[0]
    res = {}
    for anchors in (False, True):
        p1, tr = decoder_gpu.reg_refine_spectral(20, 256, c["img"], c["Ht"], p, c["ref"], mask, steps=8,
                                                 anchors=anchors)
        nodes = L.reg_nodes(256, c["Ht"].astype(np.float32), W, Hh, p1).reshape(c["nodes_true"].shape)
        Hp = np.append(p1[:8], 1.0).reshape(3, 3)
        # each interior node's place in the capture, against the channel's
        node = np.asarray(R.lattice(256))
        X, Y = np.meshgrid(node, node, indexing="xy")
        px, py = R.project(Hp, X, Y)
        tx, ty = R.project(c["Ht"], X, Y)
        got = np.stack([px + nodes[..., 0], py + nodes[..., 1]], -1)
        want = np.stack([tx + c["nodes_true"][..., 0], ty + c["nodes_true"][..., 1]], -1)
        err = np.sqrt(np.mean(np.sum((got - want)[1:-1, 1:-1] ** 2, -1)))
        res[anchors] = (float(tr[-1]), err)
        assert np.all(np.diff(tr) <= 0)
        if not anchors:
            assert np.array_equal(p1[9:], p[9:])
    assert res[True][0] < res[False][0], res
    assert res[True][1] < res[False][1], res


def test_refine_recovers_the_homography(decoder_gpu):
    c = make_case(20, 256, 6)
    p = L.reg_params(c["H0"], c["k1"])
    e0 = corner_error(c["Ht"], p, c["span"])
    assert e0 > 0.5
    p1, trace = decoder_gpu.reg_refine(c["version"], c["span"], c["img"], c["H0"], p, c["ref"], steps=10)
    assert corner_error(c["Ht"], p1, c["span"]) < 0.1
    assert trace[-1] < 0.5 * trace[0] and np.all(np.isfinite(trace))
    # the FP64 Gauss-Newton of the oracle lands on the same homography
    q = p.astype(np.float64)
    r = R.loss_grad(c["img"], c["n"], c["span"], c["H0"].astype(np.float32).astype(np.float64), q,
                    c["ref"].astype(np.float64))
    sw, sr, swr, srr, cnt = r["stats"]
    al = (cnt * swr - sw * sr) / (cnt * srr - sr * sr)
    be = (sw - al * sr) / cnt
    for _ in range(4):
        q, rr = R.gn_step(c["img"], c["n"], c["span"], c["H0"].astype(np.float32).astype(np.float64), q,
                          c["ref"].astype(np.float64), al, be)
        sw, sr, swr, srr, cnt = rr["stats"]
        al = (cnt * swr - sw * sr) / (cnt * srr - sr * sr)
        be = (sw - al * sr) / cnt
    H1 = np.append(p1[:8], 1.0).reshape(3, 3)
    Hq = np.append(q[:8], 1.0).reshape(3, 3)
    cc = corners(c["span"])
    d = np.stack(R.project(H1, cc[:, 0], cc[:, 1]), -1) - np.stack(R.project(Hq, cc[:, 0], cc[:, 1]), -1)
    assert np.abs(d).max() < 0.05


def test_refine_recovers_the_lens(decoder_gpu):
    """Behind a radial lens (k1 = 0.04) from a registration that knows no lens (k1 = 0) and whose homography corners
    are 1.5 px off: 12 Gauss-Newton steps over H and k1 together (lizx_reg_refine_k1) recover k1 within 0.002 and the
    corners within a tenth of a pixel, land where the FP64 restatement's 12 steps over the same nine parameters land,
    and leave a smaller loss than the steps over H alone."""
    c = make_case(20, 256, 12, k1=0.04)
    span = c["span"]
    p = L.reg_params(c["H0"], 0.0)
    pk, tk = decoder_gpu.reg_refine(c["version"], span, c["img"], c["H0"], p, c["ref"], steps=12, k1=True)
    ph, th = decoder_gpu.reg_refine(c["version"], span, c["img"], c["H0"], p, c["ref"], steps=12)
    assert abs(pk[8] - 0.04) < 0.002, pk[8]
    assert ph[8] == 0.0
    assert corner_error(c["Ht"], pk, span) < 0.1
    assert tk[-1] < 0.7 * th[-1] and np.all(np.isfinite(tk))
    q = p.astype(np.float64)
    H0 = c["H0"].astype(np.float32).astype(np.float64)
    r = R.loss_grad(c["img"], c["n"], span, H0, q, c["ref"].astype(np.float64))
    for _ in range(12):  # the same steps, each with the intensity fit of the pass before it, as on the GPU
        sw, sr, swr, srr, cnt = r["stats"]
        al = (cnt * swr - sw * sr) / (cnt * srr - sr * sr)
        be = (sw - al * sr) / cnt
        q, r = R.gn_step(c["img"], c["n"], span, H0, q, c["ref"].astype(np.float64), al, be, k1=True)
    assert abs(q[8] - pk[8]) < 5e-4, (q[8], pk[8])
    cc = corners(span)
    Hk, Hq = np.append(pk[:8], 1.0).reshape(3, 3), np.append(q[:8], 1.0).reshape(3, 3)
    d = np.stack(R.project(Hk, cc[:, 0], cc[:, 1]), -1) - np.stack(R.project(Hq, cc[:, 0], cc[:, 1]), -1)
    assert np.abs(d).max() < 0.05


def test_refine_recovers_blocks(decoder_gpu):
    """Decision-directed: decode through the perturbed registration, re-modulate the blocks whose CRC passed, refine,
    decode again: no block lost, and the blocks the perturbation cost come back."""
    c = make_case(20, 256, 9)
    span, version = c["span"], c["version"]
    p = L.reg_params(c["H0"], c["k1"])
    nodes0 = L.reg_nodes(span, c["H0"].astype(np.float32), c["img"].shape[1], c["img"].shape[0], p)
    before = decoder_gpu.decode_reg(version, span, c["img"], c["H0"], nodes0, soft=False)
    info = L.format_info(version)
    n = info["n"]
    Ht = c["Ht"]
    s = np.sqrt(abs(np.linalg.det(Ht[:2, :2]))) * span / n
    ref_pic = sim.remodulate(version, c["blocks"], before["ok"].astype(bool), 0, np.sqrt(1.0 + 1.0 / 3.0) / s)
    p1, trace = decoder_gpu.reg_refine(version, span, c["img"], c["H0"], p, ref_pic, steps=10)
    H1 = np.append(p1[:8], 1.0).reshape(3, 3)
    after = decoder_gpu.decode_reg(version, span, c["img"], H1, nodes0, soft=False)
    assert after["good"] >= before["good"]
    assert after["good"] == info["blocks"] or after["good"] > before["good"]
    v = after["ok"].astype(bool)
    assert np.array_equal(after["blocks"][v], c["blocks"].reshape(-1, 476)[v])


def test_refine_argument_errors(decoder_gpu, case384):
    c = case384
    p = L.reg_params(c["H0"], c["k1"])
    with pytest.raises(L.LizError):
        decoder_gpu.reg_refine(c["version"], c["span"], c["img"], c["H0"], p, c["ref"], steps=101)
    with pytest.raises(L.LizError):
        decoder_gpu.reg_loss_grad(3, c["span"], c["img"], c["H0"], p, c["ref"])
