"""ORACLE — test infrastructure only (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it).

The picture side of the receive path restated in FP64 numpy, as the checker for the GPU's float stages:
  modulate      frame slots -> the n x n picture in [0, 1] (research/SPEC.md 6.5-6.7; liblizard/src/lizard_phy.c:
                1133-1200): QPSK +-1/sqrt 2 per axis on the used coefficients, the Hermitian spectrum, the unscaled
                inverse DFT, clipped at 2 sigma, sigma = sqrt(640 subch).
  erase_ranges  the samples erased round the anchors (SPEC 6.13; lizard_phy.c:402-407, acquire.c:41-45)
  finish64      erase fill, detrend, forward DFT, pilot shift and soft values in FP64 (lizard_phy.c:1584-1772,
                :280-329, :1322-1332; SPEC 6.9, 6.10, 7.3, 8.5); the pilot turn through the reference's table index
                (lizard_phy.c:214-218), the angle exact; optionally (an extension the reference lacks) each
                coefficient divided, after the turn, by a known Gaussian blur's OTF exp(-2 pi^2 s^2 (u^2 + v^2) / n^2)
  blur          a picture blurred by a Gaussian of s samples as a periodic picture (its spectrum times the OTF)
Pinned by: numpy's FFT (a library DFT) for the transforms; the closed forms of SPEC 6.10 (the detrend basis is
orthogonal, so a picture that is exactly the shading surface detrends to zero) and 6.13; a known grid shift
recovered by the pilot fit; and the compiled reference (oracle/_ref) on the same pictures (tests/test_soft_*.py).
"""
import numpy as np

TWO_PI = 2.0 * np.pi


def coef_uv(pos, n):
    """(u, v) of each coefficient from its offset in the reference's half-plane layout (SPEC 6.12 P1)."""
    bw = min(n // 2, 64)
    pos = np.asarray(pos, np.int64)
    v = (pos // (n * bw)) * bw + pos % bw
    uw = (pos // bw) % n
    u = np.where(uw < n // 2, uw, uw - n)
    return u, v


def modulate(n, subch, slots, u, v, shift=(0.0, 0.0)):
    """The picture (row y, column x) carrying frame slots (0/1, subch * 640); shift = (dx, dy) moves the picture by a
    fraction of a sample (a grid that sits (dx, dy) off reads coefficient (u, v) turned by 2 pi (u dx + v dy) / n)."""
    S = subch * 320
    H = ((1.0 - 2.0 * slots[0::2]) + 1j * (1.0 - 2.0 * slots[1::2])) / np.sqrt(2.0)
    H = H[:S] * np.exp(1j * TWO_PI * (u * shift[0] + v * shift[1]) / n)
    Y = np.zeros((n, n), complex)  # Y[v, u]
    Y[v % n, u % n] = H
    Y[(-v) % n, (-u) % n] = np.conj(H)
    p = np.real(np.fft.ifft2(Y)) * n * n  # unscaled inverse, exp(+2 pi i (u x + v y) / n)
    L = 2.0 * np.sqrt(640.0 * subch)
    return 0.5 + np.clip(p, -L, L) / (2.0 * L)


def erase_ranges(n, span):
    scale = np.float32(n) / np.float32(span)
    a = span // 64
    lo, hi = [], []
    for i in range(7):
        m0 = int(np.floor(15 + span * (i + 0.5) / 7 - a / 2.0 + 0.5))
        l = int(np.floor(np.float32(m0 - 15) * scale)) - 2
        h = int(np.ceil(np.float32(m0 + a - 15) * scale)) - 1 + 2
        lo.append(max(l, 0))
        hi.append(min(h, n - 1))
    return lo, hi


def basis(n):
    c = (2.0 * np.arange(n) + 1 - n) / n
    q = c * c - np.mean(c * c)
    return c, q


def axis_info(g):
    return np.where(g > 0, (1 - 2.0 ** (-0.3073 * (2 * np.sqrt(np.maximum(g, 0))) ** 1.7870)) ** 1.1064, 0.0)


ALIGN_TURNS = 4096


def turn_index(u, v, n, dx, dy):
    """The pilot turn's table index of each coefficient: round-half-even((u dx + v dy) 4096 / n) in float32, in the
    reference's order of operations, modulo 4096 (liblizard/src/lizard_phy.c:215, ALIGN_TURNS :206)."""
    f = np.float32
    a = (np.asarray(u).astype(f) * f(dx) + np.asarray(v).astype(f) * f(dy)) * (f(ALIGN_TURNS) / f(n))
    return np.rint(a).astype(np.int64) & (ALIGN_TURNS - 1)


def blur(pic, s):
    """The periodic picture blurred by a Gaussian of s samples: every DFT coefficient (u, v) times
    exp(-2 pi^2 s^2 (u^2 + v^2) / n^2), u and v the signed frequencies."""
    n = pic.shape[0]
    f = np.fft.fftfreq(n) * n
    otf = np.exp(-2.0 * np.pi ** 2 * s * s * (f[:, None] ** 2 + f[None, :] ** 2) / (n * n))
    return np.real(np.fft.ifft2(np.fft.fft2(pic) * otf))


def finish64(grid, n, span, u, v, white, pilots, align=None, otf_sigma=0.0):
    """grid: n x n (row y, column x). Returns dict: coef (pos order, detrended, before the turn), align (dx, dy),
    soft (k y per slot, float64, before quantisation), llr (int8), est, rms. The turn by the pilots' shift goes through
    the reference's 4096-step table index (lizard_phy.c:214-218; the angle itself exact); align, where given, is the
    shift the soft values are turned by (the fit's own result is still returned as align)."""
    g = np.array(grid, np.float64)
    c, q = basis(n)
    lo, hi = erase_ranges(n, span)
    insq = np.zeros(n, bool)
    for a, b in zip(lo, hi):
        insq[a:b + 1] = True
    # the shading fit on every eighth sample outside the squares (lizard_phy.c:1605-1624)
    ys, xs = np.meshgrid(np.arange(0, n, 8), np.arange(0, n, 8), indexing="ij")
    keep = ~(insq[ys] & insq[xs])
    B = np.stack([np.ones(ys.shape), c[xs], q[xs], c[ys], q[ys], c[xs] * c[ys]], -1)[keep]
    beta = np.linalg.solve(B.T @ B, B.T @ g[ys, xs][keep])
    Y_, X_ = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    fill = (beta[0] + beta[1] * c[X_] + beta[2] * q[X_] + beta[3] * c[Y_] + beta[4] * q[Y_] + beta[5] * c[X_] * c[Y_])
    sq = insq[Y_] & insq[X_]
    g[sq] = fill[sq]
    # the detrend's projections (SPEC 6.10)
    sx2, sq2 = np.sum(c * c), np.sum(q * q)
    px = np.sum(g * c[None, :]) / (n * sx2)
    py = np.sum(g * c[:, None]) / (n * sx2)
    pxy = np.sum(g * c[None, :] * c[:, None]) / (sx2 * sx2)
    pxx = np.sum(g * q[None, :]) / (n * sq2)
    pyy = np.sum(g * q[:, None]) / (n * sq2)
    F = np.fft.fft2(g)  # F[v, u] = sum g[y, x] exp(-2 pi i (u x + v y) / n)
    X1, X2 = np.fft.fft(c), np.fft.fft(q)
    uw = u % n
    y = F[v, uw] - pxy * X1[uw] * X1[v]
    y = y - np.where(v == 0, n * (px * X1[uw] + pxx * X2[uw]), 0)
    y = y - np.where(uw == 0, n * (py * X1[v] + pyy * X2[v]), 0)
    coef = y.copy()
    # the pilots' grid shift (lizard_phy.c:280-329)
    k = TWO_PI / n
    dx = dy = 0.0
    pu, pv = u[pilots], v[pilots]
    xr = np.where(white[2 * pilots], -1.0, 1.0)
    xi = np.where(white[2 * pilots + 1], -1.0, 1.0)
    ok = True
    for _ in range(2):
        t = y[pilots] * np.exp(-1j * TWO_PI * (pu * dx + pv * dy) / n)
        zr = t.real * xr + t.imag * xi
        zi = t.imag * xr - t.real * xi
        cu, cv = k * zr * pu, k * zr * pv
        A, Bb, C, Pp, Q, YY = (np.sum(cu * cu), np.sum(cu * cv), np.sum(cv * cv), np.sum(cu * zi), np.sum(cv * zi),
                               np.sum(zi * zi))
        det = A * C - Bb * Bb
        if len(pilots) < 8 or not det > 0:
            ok = False
            break
        dx += (Pp * C - Q * Bb) / det
        dy += (A * Q - Bb * Pp) / det
    if ok:
        ex, ey = (Pp * C - Q * Bb) / det, (A * Q - Bb * Pp) / det
        s2 = (YY - ex * Pp - ey * Q) / (len(pilots) - 2)
        chi = (dx * dx * A + 2 * dx * dy * Bb + dy * dy * C) / s2 if s2 > 0 else 0
        if not (chi > 13.8):
            dx = dy = 0.0
    else:
        dx = dy = 0.0
    tx, ty = (dx, dy) if align is None else (float(align[0]), float(align[1]))
    if tx or ty:
        y = y * np.exp(-1j * TWO_PI * turn_index(u, v, n, tx, ty) / ALIGN_TURNS)
    if otf_sigma > 0:
        y = y * np.exp(2.0 * np.pi ** 2 * otf_sigma ** 2 * (u.astype(float) ** 2 + v.astype(float) ** 2) / (n * n))
    # soft values per sub-channel (lizard_phy.c:1322-1332, :1756-1772)
    S = len(u)
    ys = y.reshape(-1, 320)
    e = np.abs(ys) ** 2
    m2, m4 = e.mean(1), (e * e).mean(1)
    a2 = np.sqrt(np.maximum(2 * m2 * m2 - m4, 0))
    nv = np.maximum(m2 - a2, 0.02 * m2)
    kk = 0.7 * 2 * np.sqrt(2 * a2) / nv
    soft = np.empty(2 * S)
    soft[0::2] = (kk[:, None] * ys.real).ravel()
    soft[1::2] = (kk[:, None] * ys.imag).ravel()
    llr = np.rint(8 * np.clip(soft, -10, 10)).astype(np.int8)
    return {"coef": coef, "align": (dx, dy), "soft": soft, "llr": llr, "est": axis_info(a2 / nv),
            "rms": np.sqrt(m2 / 2), "proj": (px, py, pxy, pxx, pyy)}


def to_strips(grid, n):
    """Row-major n x n to the reference's strip layout (strip s of 64 columns at s * n * 64, n rows of 64)."""
    g = np.asarray(grid, np.float32).reshape(n, n // 64, 64).transpose(1, 0, 2)
    return np.ascontiguousarray(g).ravel()
