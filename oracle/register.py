"""ORACLE — test infrastructure only (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it).

The registration model, the refinement loss and its gradient restated in FP64 numpy, the checker for the GPU's
hand-written adjoint:
  nodes       the node lattice's displacements (liblizard/src/layout.c:110-118 lattice) from a radial lens term k1 about
              the image centre, k1 |q|^2 q D with q the offset over the farthest corner's distance D
              (liblizard/src/acquire.c:1470-1512, lens_rows), evaluated at each node's place under a reference
              homography H0, plus the 7 x 7 anchors' mesh: on each axis a cubic Hermite through 9 knots (the
              lattice's ends, value 0, and the anchors' centres), slopes from the neighbouring knots, one-sided at the
              ends (acquire.c:2196-2210 anch_herm4, :2319-2352), interior nodes only.
  warp        picture sample (x, y) at module 15 + (t + 1/2) span / n (research/SPEC.md 6.9) through H plus the
              nodes' bilinear blend (acquire.c:5095-5103), read bilinearly from the uint8 capture with the reference's
              half-pixel shift and clamp (liblizard/src/internal.h:11-23).
  loss        sum over samples of (w - alpha r - beta)^2, w the warped capture, r a reference picture (the blurred
              re-modulation of what is known), alpha and beta held.
  gradient    with respect to the 8 homography parameters, k1 and the 49 anchors' (dx, dy), by the chain rule in FP64.
  spectral    the loss on the format's masked coefficients, (2 / n^2) sum |E(u, v)|^2 with E the unscaled DFT of
              w - alpha r - beta (numpy's FFT), and its gradient: dL/dw = (4 / n^2) Re sum E exp(+2 pi i (u x + v y) / n)
              carried through the same chain rule.
Pinned by: finite differences of this same FP64 loss (tests/test_register.py), a capture rendered through a known
registration that the model reproduces (the nodes of the synthetic channel equal k1 L exactly), and the mesh's
interpolation properties (knot values reproduced at the knots, zero at the lattice's ends).
"""
import math

import numpy as np

ANCH_K = 7
BORDER = 15
NPARAM = 8 + 1 + 2 * ANCH_K * ANCH_K


def lattice(span):
    side = span + 2 * BORDER
    pos = []
    c = np.float32(3.5)
    while c < np.float32(side) - np.float32(3.5) - np.float32(8) and len(pos) < 39:
        pos.append(float(c))
        c = np.float32(c + 16)
    pos.append(float(np.float32(side) - np.float32(3.5)))
    return np.array(pos, np.float64)


def sample_coords(n, span):
    return BORDER + (np.arange(n) + 0.5) * span / n


def cells(m, node):
    c = np.searchsorted(node, m, side="right") - 1
    return np.clip(c, 0, len(node) - 2)


def anchor_m0(span, i):
    a = span // 64
    return math.floor(BORDER + span * (i + 0.5) / ANCH_K - a / 2.0 + 0.5)


def anchor_knots(span):
    node = lattice(span)
    X0, SX = node[0], node[-1] - node[0]
    ha = 0.5 * (span // 64)
    t = [0.0] + [(anchor_m0(span, i) + ha - X0) / SX for i in range(ANCH_K)] + [1.0]
    return np.array(t)


def herm4(t, x):
    n = len(t)
    i = 0
    while i < n - 2 and x >= t[i + 1]:
        i += 1
    hc = t[i + 1] - t[i]
    s = (x - t[i]) / hc
    s2, s3 = s * s, s * s * s
    h00, h01 = 2 * s3 - 3 * s2 + 1, -2 * s3 + 3 * s2
    A, Bc = (s3 - 2 * s2 + s) * hc, (s3 - s2) * hc
    w = [0.0, h00, h01, 0.0]
    if i > 0:
        d = t[i + 1] - t[i - 1]
        w[0] -= A / d
        w[2] += A / d
    else:
        w[1] -= A / hc
        w[2] += A / hc
    if i + 2 < n:
        d = t[i + 2] - t[i]
        w[1] -= Bc / d
        w[3] += Bc / d
    else:
        w[1] -= Bc / hc
        w[2] += Bc / hc
    return i - 1, w


def anchor_basis(span):
    """phi[j, i, aj, ai]: the weight of anchor (aj, ai)'s reading in node (i, j)'s displacement."""
    node = lattice(span)
    nx = len(node)
    t = anchor_knots(span)
    u = (node - node[0]) / (node[-1] - node[0])
    phi = np.zeros((nx, nx, ANCH_K, ANCH_K))
    wts = [herm4(t, x) for x in u]
    for j in range(1, nx - 1):
        by, wy = wts[j]
        for i in range(1, nx - 1):
            bx, wx = wts[i]
            for b in range(4):
                for a in range(4):
                    qx, qy = bx + a, by + b
                    if 1 <= qx <= ANCH_K and 1 <= qy <= ANCH_K:
                        phi[j, i, qy - 1, qx - 1] += wy[b] * wx[a]
    return phi


def project(H, mx, my):
    W = H[2, 0] * mx + H[2, 1] * my + H[2, 2]
    return (H[0, 0] * mx + H[0, 1] * my + H[0, 2]) / W, (H[1, 0] * mx + H[1, 1] * my + H[1, 2]) / W


def lens_field(span, H0, W, Hh):
    """L[j, i] = |q|^2 q D at node (i, j)'s place under H0 (the displacement per unit k1)."""
    node = lattice(span)
    X, Y = np.meshgrid(node, node, indexing="xy")
    px, py = project(np.asarray(H0, np.float64).reshape(3, 3), X, Y)
    ax, ay, D = W / 2.0, Hh / 2.0, math.hypot(W / 2.0, Hh / 2.0)
    qx, qy = (px - ax) / D, (py - ay) / D
    r2 = qx * qx + qy * qy
    return np.stack([r2 * qx * D, r2 * qy * D], -1)


def unpack(params):
    p = np.asarray(params, np.float64)
    H = np.append(p[:8], 1.0).reshape(3, 3)
    return H, p[8], p[9:].reshape(ANCH_K, ANCH_K, 2)


def nodes(span, H0, W, Hh, params):
    H, k1, anch = unpack(params)
    return k1 * lens_field(span, H0, W, Hh) + np.einsum("jiba,bac->jic", anchor_basis(span), anch)


def image_sample(img, x, y):
    """Bilinear read at pixel coordinates (centres at half integers) with the clamp, value and gradient (grey/255)."""
    h, w = img.shape
    I = img.astype(np.float64) / 255.0
    xs, ys = x - 0.5, y - 0.5
    xm, ym = w - 1.001, h - 1.001
    inx = (xs > 0) & (xs < xm)
    iny = (ys > 0) & (ys < ym)
    xs, ys = np.clip(xs, 0, xm), np.clip(ys, 0, ym)
    xi, yi = np.floor(xs).astype(np.int64), np.floor(ys).astype(np.int64)
    fx, fy = xs - xi, ys - yi
    a, b = I[yi, xi], I[yi, xi + 1]
    c, d = I[yi + 1, xi], I[yi + 1, xi + 1]
    top, bot = a + (b - a) * fx, c + (d - c) * fx
    val = top * (1 - fy) + bot * fy
    gx = np.where(inx, (1 - fy) * (b - a) + fy * (d - c), 0.0)
    gy = np.where(iny, bot - top, 0.0)
    return val, gx, gy


def loss_grad(img, n, span, H0, params, ref, alpha=1.0, beta=0.0, want_gn=False, spectral=None):
    """Loss and its gradient (NPARAM) in FP64; with want_gn also J^T J and J^T e over the 8 homography parameters.
    spectral = (u, v, mask): the loss on the masked coefficients (u, v) instead (SPEC 6.12's coefficient list)."""
    Hh, W = img.shape
    H, k1, anch = unpack(params)
    node = lattice(span)
    nx = len(node)
    L = lens_field(span, H0, W, Hh)
    phi = anchor_basis(span)
    nd = k1 * L + np.einsum("jiba,bac->jic", phi, anch)
    m = sample_coords(n, span)
    c = cells(m, node)
    u = (m - node[c]) / (node[c + 1] - node[c])
    MX, MY = np.meshgrid(m, m, indexing="xy")
    CI, CJ = np.meshgrid(c, c, indexing="xy")
    U, V = np.meshgrid(u, u, indexing="xy")
    w00, w10, w01, w11 = (1 - U) * (1 - V), U * (1 - V), (1 - U) * V, U * V
    d = (w00[..., None] * nd[CJ, CI] + w10[..., None] * nd[CJ, CI + 1] + w01[..., None] * nd[CJ + 1, CI]
         + w11[..., None] * nd[CJ + 1, CI + 1])
    Wd = H[2, 0] * MX + H[2, 1] * MY + 1.0
    A = H[0, 0] * MX + H[0, 1] * MY + H[0, 2]
    B = H[1, 0] * MX + H[1, 1] * MY + H[1, 2]
    px, py = A / Wd + d[..., 0], B / Wd + d[..., 1]
    val, gx, gy = image_sample(img, px, py)
    e = val - alpha * np.asarray(ref, np.float64) - beta
    loss = float(np.sum(e * e))
    if spectral is not None:
        cu, cv, mk = spectral
        m = np.asarray(mk).astype(bool)
        Fe = np.fft.fft2(e)  # Fe[v, u] = sum e[y, x] exp(-2 pi i (u x + v y) / n)
        Ek = Fe[np.asarray(cv)[m] % n, np.asarray(cu)[m] % n]
        loss = float(2.0 / (n * n) * np.sum(np.abs(Ek) ** 2))
        Z = np.zeros((n, n), complex)
        np.add.at(Z, (np.asarray(cv)[m] % n, np.asarray(cu)[m] % n), Ek)
        # (4 / n^2) Re sum E exp(+2 pi i (u x + v y) / n); ifft2 carries 1 / n^2; e below is half of dL/dw
        e = 2.0 * np.real(np.fft.ifft2(Z))
    z = np.zeros_like(MX)
    jx = [MX / Wd, MY / Wd, 1 / Wd, z, z, z, -A * MX / Wd ** 2, -A * MY / Wd ** 2]
    jy = [z, z, z, MX / Wd, MY / Wd, 1 / Wd, -B * MX / Wd ** 2, -B * MY / Wd ** 2]
    J = np.stack([gx * a + gy * b for a, b in zip(jx, jy)], -1)  # n x n x 8
    grad = np.zeros(NPARAM)
    grad[:8] = 2 * np.einsum("yx,yxk->k", e, J)
    # through the nodes: each sample's (e gx, e gy) spread over its cell's four corners
    G = np.zeros((nx, nx, 2))
    ex, ey = e * gx, e * gy
    for wgt, dj, di in ((w00, 0, 0), (w10, 0, 1), (w01, 1, 0), (w11, 1, 1)):
        np.add.at(G, (CJ + dj, CI + di, 0), 2 * ex * wgt)
        np.add.at(G, (CJ + dj, CI + di, 1), 2 * ey * wgt)
    grad[8] = np.sum(G * L)
    grad[9:] = np.einsum("jiba,jic->bac", phi, G).ravel()
    out = {"loss": loss, "grad": grad, "grid": val, "node_grad": G,
           "stats": (float(val.sum()), float(np.sum(ref)), float(np.sum(val * ref)), float(np.sum(ref * ref)),
                     float(val.size))}
    if want_gn:
        # k1's column: the capture's gradient along the lens field's blend at each sample
        Ls = (w00[..., None] * L[CJ, CI] + w10[..., None] * L[CJ, CI + 1] + w01[..., None] * L[CJ + 1, CI]
              + w11[..., None] * L[CJ + 1, CI + 1])
        J9 = np.concatenate([J, (gx * Ls[..., 0] + gy * Ls[..., 1])[..., None]], -1)
        out["JtJ"] = np.einsum("yxk,yxl->kl", J9, J9)
        out["Jte"] = np.einsum("yx,yxk->k", e, J9)
    return out


def gn_step(img, n, span, H0, params, ref, alpha, beta, damping=1e-6, k1=False):
    """One Gauss-Newton step over the 8 homography parameters (k1 too with k1=True; the anchors held): the normal
    equations (J^T J + damping diag) d = -J^T e solved in FP64."""
    r = loss_grad(img, n, span, H0, params, ref, alpha, beta, want_gn=True)
    m = 9 if k1 else 8
    JtJ, Jte = r["JtJ"][:m, :m], r["Jte"][:m]
    A = JtJ + damping * np.diag(np.diag(JtJ))
    d = np.linalg.solve(A, -Jte)
    p = np.array(params, np.float64)
    p[:m] += d
    return p, r
