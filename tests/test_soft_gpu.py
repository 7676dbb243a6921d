"""The GPU's picture-to-soft-value stages (lizx_finish_grid: anchor erase, detrend, 2D FFT, pilot shift, soft demap,
then LDPC and CRC) against the FP64 restatement (oracle/picture.py) and the compiled reference (oracle/_ref).
Tolerances (north star): coefficients within 1e-5 relative L2 of FP64; LLR floats within 1e-4 relative; int8 soft
values (LLR floats within 1e-4 relative L2 of FP64) differ from the reference's by one, and only at rounding ties
(the GPU's float within 1e-4 of a half-step); the reference's
decline gate, LDPC and CRC on the GPU's own soft values give the GPU's verdicts, iterations and bytes on every frame."""
import numpy as np

import lizard_amd as L
import pytest

from _libs import ptr, ref
from pictures import Frame, assert_finish_equal, picture, ref_finish_soft, ties_against_reference

pytestmark = pytest.mark.gpu

CASES = [  # version, span (ring), grid shift, SNR dB
    (4, 128, (0.0, 0.0), 25), (5, 128, (0.3, -0.2), 20), (5, 256, (0.0, 0.0), 12), (20, 256, (0.15, 0.1), 15),
    (36, 256, (-0.1, 0.05), 18), (45, 512, (0.2, 0.2), 20), (46, 512, (0.0, -0.3), 18), (64, 512, (-0.2, 0.25), 20),
    (8, 128, (0.0, 0.0), 20), (28, 256, (0.05, -0.05), 20)]


@pytest.mark.parametrize("version,span,shift,snr", CASES)
def test_finish_grid(decoder_gpu, version, span, shift, snr):
    F = Frame(version, span, seed=version + span, snr_db=snr, shift=shift)
    g64 = F.fp64()
    out = decoder_gpu.finish_grid(version, span, F.grid)
    c = out["coef"][0::2] + 1j * out["coef"][1::2]
    rel = np.linalg.norm(c - g64["coef"]) / np.linalg.norm(g64["coef"])
    assert rel < 1e-5, rel
    assert abs(out["align"][0] - g64["align"][0]) < 1e-3 and abs(out["align"][1] - g64["align"][1]) < 1e-3
    assert np.allclose(out["sub_est"], g64["est"], rtol=1e-4, atol=1e-7)
    # the float soft values k y within 1e-4 relative L2 of FP64's turned by the same shift (the shift itself is held
    # to FP64's above); their int8 roundings equal FP64's but for one step on a small fraction of slots (the
    # reference's float detrend sums are the algorithm: held to the reference below)
    g64t = F.fp64(align=out["align"]) if any(out["align"]) else g64  # turned by the GPU's own shift
    rel_soft = np.linalg.norm(out["soft"] - g64t["soft"]) / np.linalg.norm(g64t["soft"])
    assert rel_soft < 1e-4, rel_soft
    dl = np.abs(out["llr"].astype(int) - g64["llr"].astype(int))
    assert dl.max() <= 1 and (dl > 0).mean() < 1e-3, ((dl > 0).mean(), dl.max())
    r = ref("scalar")
    if r is None:
        return
    # the integer path on the GPU's own soft values and estimates: the reference's decline gate, LDPC and CRC give
    # the same verdicts, iterations and bytes, whatever the ties
    assert_finish_equal(out, ref_finish_soft(r, version, span, out["llr"], out["sub_est"]))
    B, S = F.blocks_n, F.subch * 320
    fr = r.ref_format_new(F.n, F.subch, span)
    strip = picture.to_strips(F.grid, F.n)
    bl, ok, its, be, bp = (np.zeros(B * 476, np.uint8), np.zeros(B, np.uint8), np.zeros(B, np.int8),
                           np.zeros(B, np.float32), np.zeros(B, np.float32))
    llr, est, rms, al, coef = (np.zeros(2 * S, np.int8), np.zeros(F.subch), np.zeros(F.subch, np.float32),
                               np.zeros(2, np.float32), np.zeros(2 * S, np.float32))
    good = r.ref_finish_grid(fr, ptr(strip), ptr(bl), ptr(ok), ptr(its), ptr(be), ptr(bp), ptr(llr), ptr(est),
                             ptr(rms), ptr(al), ptr(coef))
    r.ref_format_free(fr)
    # the blocks' pilot readings: NaN on the same blocks (no tail), the same values within float rounding
    gp = out["blk_pilot"]
    assert np.array_equal(np.isnan(gp), np.isnan(bp))
    v = ~np.isnan(bp)
    assert np.allclose(gp[v], bp[v], rtol=1e-4, atol=1e-4), np.abs(gp[v] - bp[v]).max()
    cr = coef[0::2] + 1j * coef[1::2]
    assert np.linalg.norm(c - cr) / np.linalg.norm(cr) < 1e-5
    assert np.allclose(out["sub_est"], est, rtol=1e-4, atol=1e-7)
    # against the reference's own soft values: ties only, and where none differs the same verdicts
    tie_ref = ties_against_reference(version, out, dict(llr=llr, coef=coef, align=(float(al[0]), float(al[1]))))
    assert tie_ref < 1e-3, tie_ref
    if tie_ref == 0:
        assert out["good"] == good
        assert np.array_equal(out["its"], its) and np.array_equal(out["ok"], ok)
    for b in range(B):
        if out["ok"][b]:
            assert np.array_equal(out["blocks"][b], F.blocks.reshape(B, 476)[b])


@pytest.mark.parametrize("cnt", [0, 1, 2, 3])
def test_pilot_readings_carry_the_count(decoder_gpu, cnt):
    """On a clean frame each block's pilot tail reads +1 where its bit of the painted count is 0 and -1 where it is 1
    (bit 0 on even blocks, bit 1 on odd; SPEC 7.3); 2/3 and 1/2 blocks have no tail and read NaN."""
    F = Frame(20, 256, seed=9 + cnt, snr_db=40, shift=(0.2, -0.1), ramp=0.0, junk=False, cnt=cnt)
    out = decoder_gpu.finish_grid(20, 256, F.grid)
    rate = L.format_blocks(20)[0]
    for b, pr in enumerate(out["blk_pilot"]):
        if rate[b] in (2, 3):
            assert np.isnan(pr)
        else:
            want = -1.0 if (cnt >> (b & 1)) & 1 else 1.0
            assert abs(pr - want) < 0.15, (b, pr, want)


def test_clean_frame_all_verify(decoder_gpu):
    F = Frame(20, 256, seed=3, snr_db=40, shift=(0, 0), ramp=0.0, junk=False)
    out = decoder_gpu.finish_grid(20, 256, F.grid)
    assert out["good"] == F.blocks_n
    assert np.array_equal(out["blocks"].ravel(), F.blocks)


@pytest.mark.parametrize("version,span,blur", [(5, 128, 0.6), (20, 256, 0.8), (45, 512, 0.5), (64, 512, 0.7)])
def test_otf_equalised_demap_holds_to_fp64(version, span, blur):
    """Equalisation by a known Gaussian blur (an extension; lizx_decoder_otf): on a picture blurred by that Gaussian,
    the estimates within 1e-4 of FP64 equalising the same way (oracle/picture.py finish64 with otf_sigma), the float
    soft values within 1e-4 relative L2 of FP64's, int8 soft values one step from FP64's on a small fraction of
    slots; verified blocks are the sent ones; with the blur set back to 0 the decoder's results are the unequalised
    path's."""
    F = Frame(version, span, seed=3 * version + span, snr_db=24, shift=(0.1, -0.1), blur=blur)
    dec = L.Decoder(0)
    try:
        plain = dec.finish_grid(version, span, F.grid)
        dec.otf(blur)
        out = dec.finish_grid(version, span, F.grid)
        g64 = F.fp64(align=out["align"], otf_sigma=blur)
        assert np.allclose(out["sub_est"], g64["est"], rtol=1e-4, atol=1e-7)
        rel_soft = np.linalg.norm(out["soft"] - g64["soft"]) / np.linalg.norm(g64["soft"])
        assert rel_soft < 1e-4, rel_soft
        dl = np.abs(out["llr"].astype(int) - g64["llr"].astype(int))
        assert dl.max() <= 1 and (dl > 0).mean() < 1e-3, ((dl > 0).mean(), dl.max())
        # the coefficients themselves are not equalised (the debug copy is taken before the turn)
        assert np.array_equal(out["coef"], plain["coef"])
        sent = F.blocks.reshape(F.blocks_n, 476)
        v = out["ok"].astype(bool)
        assert v.sum() > 0 and np.array_equal(out["blocks"][v], sent[v])
        dec.otf(0.0)
        again = dec.finish_grid(version, span, F.grid)
        assert np.array_equal(again["llr"], plain["llr"]) and np.array_equal(again["ok"], plain["ok"])
        with pytest.raises(L.LizError):
            dec.otf(-1.0)
    finally:
        dec.close()
