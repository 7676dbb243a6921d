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
    def __init__(self, version, span=256, seed=0, snr_db=20.0, shift=(0.0, 0.0), ramp=0.05, junk=True, cnt=0):
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

    def fp64(self):
        return picture.finish64(self.grid, self.n, self.span, self.u, self.v, self.white, self.pilots)
