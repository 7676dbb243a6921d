"""lizard_finish_ext with soft values made elsewhere (lizx_finish_llr) on the GPU against the oracle: decline gate,
bit map, LDPC at the PHY's settings, bit pack and CRC-32C, block for block."""
import numpy as np
import pytest

from _libs import oracle, ptr

pytestmark = pytest.mark.gpu


def _frame(version, sigma, seed, est_lo=0.05):
    o = oracle()
    subch = 16 * version
    f = o.or_format_new(o.or_n_for(subch), subch)
    info = np.zeros(6, np.int32)
    o.or_format_info(f, ptr(info))
    B = int(info[2])
    rng = np.random.default_rng(seed)
    blocks = rng.integers(0, 256, B * 476).astype(np.uint8)
    slots = np.zeros(subch * 640, np.uint8)
    o.or_frame_slots(f, ptr(blocks), seed & 3, ptr(slots))
    y = (1.0 - 2.0 * slots) + sigma * rng.standard_normal(slots.size)
    llr = np.clip(np.rint(8 * 2 * y / sigma ** 2), -80, 80).astype(np.int8)
    est = rng.uniform(est_lo, 1.0, subch).astype(np.float32)
    return f, B, blocks, llr, est


@pytest.mark.parametrize("version,sigma", [(4, 0.45), (5, 0.5), (5, 0.7), (20, 0.55), (45, 0.6), (64, 0.5), (64, 0.75)])
def test_finish_llr(decoder_gpu, version, sigma):
    o = oracle()
    f, B, blocks, llr, est = _frame(version, sigma, seed=version * 10 + int(sigma * 100))
    ob, ok, its, be = np.zeros(B * 476, np.uint8), np.zeros(B, np.uint8), np.zeros(B, np.int8), np.zeros(B, np.float32)
    good = o.or_finish_llr(f, ptr(llr), ptr(est), ptr(ob), ptr(ok), ptr(its), ptr(be))
    out = decoder_gpu.finish_llr(version, llr, est)
    assert out["good"] == good
    assert np.array_equal(out["its"], its)
    assert np.array_equal(out["ok"], ok)
    assert np.array_equal(out["est"], be)
    ob = ob.reshape(B, 476)
    for b in range(B):
        if ok[b]:
            assert np.array_equal(out["blocks"][b], ob[b])
            assert np.array_equal(ob[b], blocks.reshape(B, 476)[b])
    o.or_format_free(f)
