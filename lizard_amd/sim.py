"""Synthetic frames for tests and benches: a sender's picture and a camera channel, in torch on the host.

torch's wheel carries its own HIP runtime, separate from the one the native library links, so device memory does not
pass between them: frames made here go to the device through the library's allocator (lizard_amd.DeviceArray).

sender   blocks -> frame slots through the library (lizx_frame_slots: CRC-32C, LDPC, bit map, pilots, whitening;
         liblizard/src/lizard_phy.c:1087-1150) -> the n x n picture of research/SPEC.md 6.5-6.7: QPSK +-1/sqrt 2 per
         axis at each coefficient's (u, v) (lizx_format_coefs), the Hermitian spectrum, the unscaled inverse DFT,
         clipped at 2 sigma (sigma = sqrt(640 subch)) into [0, 1].
paint    one picture sample a paint pixel (LIZARD's rings hold 3 samples and 3 pixels a module), read bilinearly; the
         ring outside the picture is grey 128 (the finder's marks are not part of this path; a whole grey level, so
         blur and rounding leave it in place).
channel  module coordinates -> capture pixels by a homography (rotation, scale, perspective) and a radial term
         k1 |q|^2 q about the image centre, q the offset over the farthest corner's distance (the lens form of
         liblizard/src/acquire.c:1470-1512), then a Gaussian blur and AWGN, rounded to uint8.
The registration a decoder needs comes with each capture: H (module -> pixel, h8 = 1) and the node lattice's
displacements, the radial field at each node.

This is a channel model for measurement inputs, not the reference's sender: the reference resamples the picture into
its drive with a Lanczos-3 kernel and paints anchors and a coded border (lizard_phy.c:970-1236, layout.c), which the
receive path erases or never reads.
"""
import math

import numpy as np

from . import format_blocks, format_coefs, format_info, frame_slots, node_lattice

BORDER = 15  # modules between the symbol's edge and the picture


def picture(version, blocks, cnt=0, device="cpu"):
    """The n x n picture (row y, column x) in [0, 1], float32 on `device`, carrying these blocks."""
    import torch
    info = format_info(version)
    n, subch = info["n"], info["subch"]
    slots = frame_slots(version, blocks, cnt)
    u, v = format_coefs(version)
    s = torch.as_tensor(slots.astype(np.float64), device=device)
    H = torch.complex(1.0 - 2.0 * s[0::2], 1.0 - 2.0 * s[1::2]) / math.sqrt(2.0)
    uu = torch.as_tensor(u, device=device)
    vv = torch.as_tensor(v, device=device)
    Y = torch.zeros(n, n, dtype=torch.complex128, device=device)  # Y[v, u]
    Y[vv % n, uu % n] = H
    Y[(-vv) % n, (-uu) % n] = torch.conj(H)
    p = torch.fft.ifft2(Y).real * (n * n)  # exp(+2 pi i (u x + v y) / n), unscaled
    L = 2.0 * math.sqrt(640.0 * subch)
    return (0.5 + torch.clamp(p, -L, L) / (2.0 * L)).to(torch.float32)


def remodulate(version, blocks, known, cnt=0, blur=0.0):
    """The picture re-modulated from what a receiver knows: the blocks whose CRC passed (known[b]) and the padding
    past the last block (the whitening alone); every other coefficient 0; no clip (the partial sum is not the sent
    picture's); blurred by a Gaussian of `blur` samples, applied as its transfer function on the coefficients (a
    periodic blur). numpy float32, n x n."""
    import torch
    info = format_info(version)
    n, subch = info["n"], info["subch"]
    slots = frame_slots(version, blocks, cnt)
    u, v = format_coefs(version)
    keep = np.ones(subch * 320, np.float64)
    _, s0, ns = format_blocks(version)
    for b in range(len(s0)):
        if not known[b]:
            keep[s0[b] // 2:(s0[b] + ns[b]) // 2] = 0.0
    otf = keep * np.exp(-2.0 * (math.pi * blur / n) ** 2 * (u.astype(np.float64) ** 2 + v.astype(np.float64) ** 2))
    s = torch.as_tensor(slots.astype(np.float64))
    H = torch.complex(1.0 - 2.0 * s[0::2], 1.0 - 2.0 * s[1::2]) / math.sqrt(2.0) * torch.as_tensor(otf)
    Y = torch.zeros(n, n, dtype=torch.complex128)
    uu, vv = torch.as_tensor(u), torch.as_tensor(v)
    Y[vv % n, uu % n] = H
    Y[(-vv) % n, (-uu) % n] = torch.conj(H)
    p = torch.fft.ifft2(Y).real * (n * n)
    return (0.5 + p / (4.0 * math.sqrt(640.0 * subch))).to(torch.float32).numpy()


def known_mask(version, known):
    """The coefficients remodulate() re-modulates (the blocks known[b] and the padding past the last block), one
    byte a coefficient in coefficient order: the mask of the spectral refinement loss."""
    info = format_info(version)
    keep = np.ones(info["subch"] * 320, np.uint8)
    _, s0, ns = format_blocks(version)
    for b in range(len(s0)):
        if not known[b]:
            keep[s0[b] // 2:(s0[b] + ns[b]) // 2] = 0
    return keep


def homography(src, dst):
    """The 3 x 3 projective map (h8 = 1) taking 4 points src to dst (float64 Gauss-Jordan on the 8 x 8 system)."""
    A, b = np.zeros((8, 8)), np.zeros(8)
    for k, ((x, y), (X, Y)) in enumerate(zip(src, dst)):
        A[2 * k] = [x, y, 1, 0, 0, 0, -x * X, -y * X]
        A[2 * k + 1] = [0, 0, 0, x, y, 1, -x * Y, -y * Y]
        b[2 * k], b[2 * k + 1] = X, Y
    h = np.linalg.solve(A, b)
    return np.append(h, 1.0).reshape(3, 3)


class Camera:
    """A frame's geometry and noise: W x Hh pixels; the symbol rotated rot_deg, its side scale x Hh pixels, corners
    pulled by persp (fractions of Hh), its centre shifted by (cx, cy) pixels; radial k1 and k2; blur sigma (pixels);
    AWGN sigma (grey levels)."""

    def __init__(self, W=1280, Hh=720, rot_deg=5.0, scale=0.8, persp=0.02, k1=0.0, blur=0.8, noise=4.0, cx=0.0,
                 cy=0.0, k2=0.0):
        self.W, self.Hh, self.rot_deg, self.scale, self.persp = W, Hh, rot_deg, scale, persp
        self.k1, self.blur, self.noise, self.cx, self.cy = k1, blur, noise, cx, cy
        self.k2 = k2

    @classmethod
    def random(cls, rng, W=1920, Hh=1080, rot=10.0, scale=(0.6, 0.9), persp=0.05, k1=0.05, blur=(0.5, 1.5),
               noise=(2.0, 8.0)):
        return cls(W, Hh, rot_deg=rng.uniform(-rot, rot), scale=rng.uniform(*scale), persp=rng.uniform(0, persp),
                   k1=rng.uniform(-k1, k1), blur=rng.uniform(*blur), noise=rng.uniform(*noise),
                   cx=rng.uniform(-0.02, 0.02) * W, cy=rng.uniform(-0.02, 0.02) * Hh)

    def H(self, span):
        """Module coordinates (0 .. span + 30 on a side) -> capture pixels."""
        side = span + 2 * BORDER
        s = self.scale * self.Hh / side
        th = math.radians(self.rot_deg)
        R = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
        corners = np.array([[0, 0], [side, 0], [side, side], [0, side]], float)
        dst = (corners - side / 2) @ R.T * s + np.array([self.W / 2 + self.cx, self.Hh / 2 + self.cy])
        d = self.persp * self.Hh
        dst[1] += [0.5 * d, -0.5 * d]
        dst[2] -= [0.25 * d, 0.25 * d]
        return homography(corners, dst)

    def axis(self):
        a = np.array([self.W / 2.0, self.Hh / 2.0])
        return a, math.hypot(self.W / 2.0, self.Hh / 2.0)

    def distort(self, p):
        """Undistorted pixel points (..., 2) -> where the lens puts them: p + (k1 |q|^2 + k2 |q|^4) q D,
        q = (p - a) / D."""
        a, D = self.axis()
        q = (p - a) / D
        r2 = np.sum(q * q, -1, keepdims=True)
        return p + (self.k1 * r2 + self.k2 * r2 * r2) * q * D


def nodes_for(cam, span, Hm):
    """The node lattice's displacements (ny x nx x 2, pixels): the lens field at each node's undistorted place."""
    x = node_lattice(span).astype(np.float64)
    M = np.stack(np.meshgrid(x, x, indexing="xy"), -1)  # [j, i] = (node_x[i], node_y[j])
    P = np.concatenate([M, np.ones(M.shape[:2] + (1,))], -1) @ Hm.T
    P = P[..., :2] / P[..., 2:]
    return (cam.distort(P) - P).astype(np.float32)


def capture(pic, span, cam, seed=0):
    """Renders picture pic (n x n torch tensor, any device) through cam. Returns (img uint8 [Hh, W] on pic's device,
    H as 9 float32, nodes float32 [ny, nx, 2])."""
    import torch
    dev = pic.device
    n = pic.shape[0]
    Hm = cam.H(span)
    Hi = torch.as_tensor(np.linalg.inv(Hm), dtype=torch.float64, device=dev)
    Y, X = torch.meshgrid(torch.arange(cam.Hh, device=dev, dtype=torch.float64) + 0.5,
                          torch.arange(cam.W, device=dev, dtype=torch.float64) + 0.5, indexing="ij")
    # undo the lens by fixed-point iteration (a contraction for |k1| well below 1), then the homography
    a, D = cam.axis()
    px, py = X.clone(), Y.clone()
    for _ in range(6):
        qx, qy = (px - a[0]) / D, (py - a[1]) / D
        r2 = qx * qx + qy * qy
        f = cam.k1 * r2 + cam.k2 * r2 * r2
        px = X - f * qx * D
        py = Y - f * qy * D
    w = Hi[2, 0] * px + Hi[2, 1] * py + Hi[2, 2]
    mx = (Hi[0, 0] * px + Hi[0, 1] * py + Hi[0, 2]) / w
    my = (Hi[1, 0] * px + Hi[1, 1] * py + Hi[1, 2]) / w
    # module -> picture sample index (sample t's centre at module 15 + (t + 1/2) span / n)
    tx = (mx - BORDER) * (n / span) - 0.5
    ty = (my - BORDER) * (n / span) - 0.5
    inside = (tx > -0.5) & (tx < n - 0.5) & (ty > -0.5) & (ty < n - 0.5)
    tx = tx.clamp(0, n - 1.000001)
    ty = ty.clamp(0, n - 1.000001)
    x0, y0 = tx.floor().long(), ty.floor().long()
    fx, fy = (tx - x0).float(), (ty - y0).float()
    x1, y1 = (x0 + 1).clamp(max=n - 1), (y0 + 1).clamp(max=n - 1)
    top = pic[y0, x0] * (1 - fx) + pic[y0, x1] * fx
    bot = pic[y1, x0] * (1 - fx) + pic[y1, x1] * fx
    img = torch.where(inside, top * (1 - fy) + bot * fy, torch.full_like(top, 128.0 / 255.0)) * 255.0
    if cam.blur > 0:
        r = max(1, int(math.ceil(3 * cam.blur)))
        k = torch.exp(-0.5 * (torch.arange(-r, r + 1, device=dev, dtype=torch.float32) / cam.blur) ** 2)
        k = k / k.sum()
        t = torch.nn.functional.pad(img[None, None], (r, r, r, r), mode="replicate")
        t = torch.nn.functional.conv2d(t, k.view(1, 1, 1, -1))
        img = torch.nn.functional.conv2d(t, k.view(1, 1, -1, 1))[0, 0]
    if cam.noise > 0:
        g = torch.Generator(device=dev)
        g.manual_seed(int(seed))
        img = img + torch.randn(img.shape, generator=g, device=dev) * cam.noise
    img = torch.clamp(torch.round(img), 0, 255).to(torch.uint8)
    Hn = (Hm / Hm[2, 2]).astype(np.float32).ravel()
    return img, Hn, nodes_for(cam, span, Hm)


def capture_gpu(dec, version, span, d_picture, cam, seed=0):
    """capture() on the GPU (lizx_capture) from a device picture (lizard_amd.DeviceArray, n x n float32): returns
    (img as a DeviceArray of Hh x W uint8, H as 9 float32, nodes float32 [ny, nx, 2]). The same model; the noise
    comes from the library's counter-based generator, not torch's."""
    Hm = cam.H(span)
    img = dec.capture(version, span, d_picture, cam.W, cam.Hh, Hm, k1=cam.k1, blur=cam.blur, noise=cam.noise,
                      seed=seed, k2=cam.k2)
    Hn = (Hm / Hm[2, 2]).astype(np.float32).ravel()
    return img, Hn, nodes_for(cam, span, Hm)


def random_blocks(version, rng):
    """Seeded blocks: a 4-byte id then 472 random bytes each (the encoder adds the CRC)."""
    B = format_info(version)["blocks"]
    out = rng.integers(0, 256, (B, 476)).astype(np.uint8)
    out[:, :4] = np.arange(B, dtype=">u4").view(np.uint8).reshape(B, 4)
    return out.ravel()
