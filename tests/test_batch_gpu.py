"""Batches (lizx_batch_*): frames of mixed formats through one HIP graph, the same launches without it, frame by
frame, and frame by frame with one LDPC launch a codeword. Every mode must give the single-frame path's bits
(lizx_decode_reg / lizx_finish_grid, themselves held to the compiled reference in test_warp_gpu.py and
test_soft_gpu.py), the graph on every replay, and the reference's own decoder's verdicts and bytes."""
import numpy as np
import pytest

import lizard_amd as L
from lizard_amd import sim
from _libs import ptr, ref
from pictures import Frame, assert_finish_equal, ref_finish_soft, ref_grid_finish, ties_against_reference

pytestmark = pytest.mark.gpu

# (version, span, camera): small, decodable captures of three formats
FRAMES = [
    (5, 128, dict(W=1280, Hh=720, rot_deg=5.0, scale=0.8, persp=0.02, k1=0.0, blur=0.8, noise=4.0)),
    (20, 256, dict(W=1920, Hh=1080, rot_deg=-4.0, scale=0.9, persp=0.02, k1=0.03, blur=0.7, noise=5.0)),
    (5, 128, dict(W=1280, Hh=720, rot_deg=-8.0, scale=0.85, persp=0.03, k1=-0.04, blur=1.2, noise=12.0)),
    (8, 256, dict(W=1280, Hh=720, rot_deg=2.0, scale=0.95, persp=0.01, k1=0.02, blur=0.6, noise=6.0)),
    (20, 256, dict(W=1920, Hh=1080, rot_deg=9.0, scale=0.7, persp=0.04, k1=-0.02, blur=1.4, noise=8.0)),
]


def make_captures(frames, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k, (version, span, cam) in enumerate(frames):
        blocks = sim.random_blocks(version, rng)
        img, H, nodes = sim.capture(sim.picture(version, blocks, cnt=k % 4), span, sim.Camera(**cam), seed=seed + k)
        out.append(dict(version=version, span=span, blocks=blocks, img=img.numpy(), H=H, nodes=nodes))
    return out


def device_frames(caps):
    keep = []
    items = []
    for c in caps:
        d_img, d_nodes = L.DeviceArray.of(c["img"]), L.DeviceArray.of(c["nodes"])
        keep += [d_img, d_nodes]
        items.append(dict(img=d_img, w=c["img"].shape[1], h=c["img"].shape[0], H=c["H"], nodes=d_nodes))
    return L.make_frames(items), keep


def single(decoder, caps):
    outs = [decoder.decode_reg(c["version"], c["span"], c["img"], c["H"], c["nodes"], soft=True) for c in caps]
    return {k: np.concatenate([o[k] for o in outs]) for k in ("ok", "its", "est", "blocks")}, outs


def same(got, want):
    assert np.array_equal(got["its"], want["its"])
    assert np.array_equal(got["ok"], want["ok"])
    assert np.array_equal(got["est"], want["est"])
    v = want["ok"].astype(bool)
    assert np.array_equal(got["blocks"][v], want["blocks"][v])


@pytest.fixture(scope="module")
def caps():
    return make_captures(FRAMES, 11)


def test_batch_modes_match_single_frames(decoder_gpu, caps):
    want, outs = single(decoder_gpu, caps)
    assert want["ok"].sum() > 0.5 * len(want["ok"])  # most of the channel decodes
    b = L.Batch().plan([c["version"] for c in caps], [c["span"] for c in caps])
    assert b.formats == 3 and b.blocks == len(want["ok"])
    frames, keep = device_frames(caps)
    for mode in (L.RUN_GRAPH, L.RUN_STREAM, L.RUN_FRAMES, L.RUN_PLAIN, L.RUN_GRAPH):
        b.run(frames, mode)
        b.sync()
        same(b.blocks_out(), want)
    # soft values of each frame equal the single-frame path's
    for i, c in enumerate(caps):
        s = b.soft(i, c["version"])
        assert np.array_equal(s["llr"], outs[i]["llr"])
        assert np.array_equal(s["sub_est"], outs[i]["sub_est"])
        assert tuple(s["align"]) == tuple(np.float32(outs[i]["align"]))
    b.close()


def test_batch_verified_bytes_are_the_sent_blocks(caps):
    b = L.Batch().plan([c["version"] for c in caps], [c["span"] for c in caps])
    frames, keep = device_frames(caps)
    b.run(frames, L.RUN_GRAPH)
    b.sync()
    got = b.blocks_out()
    sent = np.concatenate([c["blocks"].reshape(-1, 476) for c in caps])
    v = got["ok"].astype(bool)
    assert np.array_equal(got["blocks"][v], sent[v])
    r = ref("scalar")
    if r is not None:  # the reference's own decoder on the same captures and registrations: the same verdicts
        at = 0
        for i, c in enumerate(caps):
            info = L.format_info(c["version"])
            B = info["blocks"]
            f = r.ref_format_new(info["n"], info["subch"], c["span"])
            out, ok = np.zeros(B * 476, np.uint8), np.zeros(B, np.uint8)
            its, be = np.zeros(B, np.int8), np.zeros(B, np.float32)
            r.ref_decode_reg(f, ptr(c["img"]), c["img"].shape[1], c["img"].shape[0], ptr(c["H"]),
                             ptr(np.ascontiguousarray(c["nodes"])), ptr(out), ptr(ok), ptr(its), ptr(be), None, None)
            r.ref_format_free(f)
            assert np.array_equal(got["ok"][at:at + B], ok)
            assert np.array_equal(got["its"][at:at + B], its)
            # and the reference's integer path on the batch's own soft values, whatever the ties
            s = b.soft(i, c["version"])
            assert_finish_equal({k: got[k][at:at + B] for k in ("ok", "its", "blocks")},
                                ref_finish_soft(r, c["version"], c["span"], s["llr"], s["sub_est"]))
            at += B
    b.close()


def test_graph_replays_with_other_work_between(decoder_gpu, caps):
    """The captured graph replayed several times, with eager work of the library's own in between, gives the eager
    result every time; new inputs on the same plan are taken up by the same graph."""
    b = L.Batch().plan([c["version"] for c in caps], [c["span"] for c in caps])
    frames, keep = device_frames(caps)
    b.run(frames, L.RUN_STREAM)
    b.sync()
    eager = b.blocks_out()
    for k in range(3):
        b.run(frames, L.RUN_GRAPH)
        decoder_gpu.decode_reg(caps[k]["version"], caps[k]["span"], caps[k]["img"], caps[k]["H"], caps[k]["nodes"])
        b.sync()
        same(b.blocks_out(), eager)
    other = make_captures(FRAMES, 23)
    want, _ = single(decoder_gpu, other)
    frames2, keep2 = device_frames(other)
    b.run(frames2, L.RUN_GRAPH)
    b.sync()
    same(b.blocks_out(), want)
    b.run(frames, L.RUN_GRAPH)
    b.sync()
    same(b.blocks_out(), eager)
    b.close()


def test_grid_batch_of_mixed_sizes(decoder_gpu):
    """Soft values only, from picture grids of four sizes (n = 384, 768, 1152, 1536) in one batch."""
    rng = np.random.default_rng(5)
    versions = [5, 20, 45, 64, 20, 5]
    spans = [128, 256, 512, 512, 256, 128]
    grids, keep, items = [], [], []
    for v in versions:
        pic = sim.picture(v, sim.random_blocks(v, rng)).numpy()
        g = (pic + rng.standard_normal(pic.shape) * 0.02).astype(np.float32)
        grids.append(g)
        d = L.DeviceArray.of(g)
        keep.append(d)
        items.append(dict(grid=d))
    frames = L.make_frames(items)
    b = L.Batch().plan(versions, spans, source=L.SRC_GRID, stages=L.STAGE_SOFT)
    assert b.formats == 4
    for mode in (L.RUN_GRAPH, L.RUN_FRAMES, L.RUN_GRAPH):
        b.run(frames, mode)
        b.sync()
        for i, (v, s) in enumerate(zip(versions, spans)):
            want = decoder_gpu.finish_grid(v, s, grids[i])
            got = b.soft(i, v)
            assert np.array_equal(got["llr"], want["llr"])
            assert np.array_equal(got["sub_est"], want["sub_est"])
            assert np.array_equal(got["sub_rms"], want["sub_rms"])
    b.close()


def test_grid_batch_of_sixteen_one_format(decoder_gpu):
    """Sixteen grids of one format in one batch: the column pass then runs each frame's columns on one XCD (frames a
    multiple of 8), and every frame's soft values, estimates and rms still equal its own single-frame finish's, graph
    and frame by frame."""
    rng = np.random.default_rng(6)
    versions, spans = [20] * 16, [256] * 16
    grids, keep, items = [], [], []
    for v in versions:
        pic = sim.picture(v, sim.random_blocks(v, rng)).numpy()
        g = (pic + rng.standard_normal(pic.shape) * 0.03).astype(np.float32)
        grids.append(g)
        d = L.DeviceArray.of(g)
        keep.append(d)
        items.append(dict(grid=d))
    frames = L.make_frames(items)
    b = L.Batch().plan(versions, spans, source=L.SRC_GRID, stages=L.STAGE_SOFT)
    for mode in (L.RUN_GRAPH, L.RUN_FRAMES):
        b.run(frames, mode)
        b.sync()
        for i in range(16):
            want = decoder_gpu.finish_grid(20, 256, grids[i])
            got = b.soft(i, 20)
            for k in ("llr", "sub_est", "sub_rms"):
                assert np.array_equal(got[k], want[k]), (mode, i, k)
    b.close()


def test_batch_argument_errors():
    b = L.Batch()
    with pytest.raises(L.LizError):
        b.plan([5], [100])
    with pytest.raises(L.LizError):
        b.plan([3], [128])
    with pytest.raises(L.LizError):
        b.run(L.make_frames([{}]), L.RUN_GRAPH)  # not planned
    b.plan([5], [128])
    with pytest.raises(L.LizError):
        b.run(L.make_frames([{}]), L.RUN_GRAPH)  # no capture
    with pytest.raises(L.LizError):
        b.run(L.make_frames([{}]), 9)
    b.close()


def test_launch_shapes_leave_the_bits_alone(caps):
    """Every launch shape the autotuner may pick gives the same bits; the tuner keeps one of them."""
    b = L.Batch().plan([c["version"] for c in caps], [c["span"] for c in caps])
    frames, keep = device_frames(caps)
    b.run(frames, L.RUN_GRAPH)
    b.sync()
    want = b.blocks_out()
    soft0 = [b.soft(i, c["version"]) for i, c in enumerate(caps)]
    for shape in ([1, 4, 4, 0], [2, 4, 4, 0], [4, 1, 4, 0], [4, 2, 4, 0], [4, 4, 1, 0], [4, 4, 2, 0], [4, 4, 4, 2],
                  [1, 1, 1, 4]):
        b.shape(shape)
        b.run(frames, L.RUN_GRAPH)
        b.sync()
        same(b.blocks_out(), want)
        for i, c in enumerate(caps):
            s = b.soft(i, c["version"])
            assert np.array_equal(s["llr"], soft0[i]["llr"]) and np.array_equal(s["sub_est"], soft0[i]["sub_est"])
    choice = b.tune(frames, reps=2)
    assert choice[0] in (1, 2, 4) and choice[1] in (1, 2, 4) and choice[2] in (1, 2, 4) and choice[3] in (0, 2, 4, 6, 8)
    b.run(frames, L.RUN_GRAPH)
    b.sync()
    same(b.blocks_out(), want)
    with pytest.raises(L.LizError):
        b.shape([3, 4, 4, 0])
    b.close()


@pytest.mark.parametrize("plan", [(4, 4), (16, 16), (4, 16), (16, 8), (8, 4)])
def test_radix_plans_hold_to_fp64(plan):
    """Each radix plan of the row and column FFTs (the 64-point stage across the lanes as R1 x 64 / R1) on grids of
    every size class: estimates within 1e-4 of FP64, int8 soft values one step from FP64's on a small fraction of
    slots at most, the pilot shift within 1e-3 of FP64's."""
    cases = [(5, 128, 1), (20, 256, 2), (45, 512, 3), (64, 512, 4)]  # n = 384, 768, 1152, 1536
    fr = [Frame(v, s, seed=40 + k, snr_db=20.0, shift=(0.1, -0.05)) for v, s, k in cases]
    keep = [L.DeviceArray.of(F.grid) for F in fr]
    frames = L.make_frames([dict(grid=d) for d in keep])
    b = L.Batch().plan([c[0] for c in cases], [c[1] for c in cases], source=L.SRC_GRID, stages=L.STAGE_SOFT)
    b.shape([4, 4, 4, 0, plan[0], plan[1]])
    for mode in (L.RUN_GRAPH, L.RUN_FRAMES):
        b.run(frames, mode)
        b.sync()
        for i, F in enumerate(fr):
            got = b.soft(i, cases[i][0])
            g64 = F.fp64()
            assert np.allclose(got["sub_est"], g64["est"], rtol=1e-4, atol=1e-7)
            assert abs(got["align"][0] - g64["align"][0]) < 1e-3 and abs(got["align"][1] - g64["align"][1]) < 1e-3
            dl = np.abs(got["llr"].astype(int) - g64["llr"].astype(int))
            assert dl.max() <= 1 and (dl > 0).mean() < 1e-3, (plan, i, (dl > 0).mean(), dl.max())
    with pytest.raises(L.LizError):
        b.shape([4, 4, 4, 0, 2, 8])
    b.close()


def test_plan_tuner_keeps_a_decoding_plan(caps):
    """lizx_batch_tune_plans picks shapes and radix plans by timing; the blocks it then verifies are the sent ones."""
    b = L.Batch().plan([c["version"] for c in caps], [c["span"] for c in caps])
    frames, keep = device_frames(caps)
    choice = b.tune_plans(frames, reps=2)
    assert choice[0] in (1, 2, 4) and choice[4] in (4, 8, 16) and choice[5] in (4, 8, 16)
    b.run(frames, L.RUN_GRAPH)
    b.sync()
    got = b.blocks_out()
    sent = np.concatenate([c["blocks"].reshape(-1, 476) for c in caps])
    v = got["ok"].astype(bool)
    assert v.sum() > 0 and np.array_equal(got["blocks"][v], sent[v])
    b.close()


# LIZARD-1024 (n = 1536, 124 blocks) in the 512 ring on 1920x1080 captures: the full-pipeline workload's format
BIG = [(64, 512, dict(W=1920, Hh=1080, rot_deg=rot, scale=sc, persp=pe, k1=k1, blur=bl, noise=nz))
       for rot, sc, pe, k1, bl, nz in ((3.0, 0.97, 0.01, 0.0, 0.5, 2.0), (-6.0, 0.92, 0.03, 0.03, 0.6, 3.0),
                                       (9.0, 0.9, 0.02, -0.04, 0.7, 4.0), (-2.0, 0.95, 0.05, 0.02, 0.5, 6.0),
                                       (5.0, 0.88, 0.0, -0.01, 0.8, 2.0))]


def test_lizard1024_batch_beyond_the_grid_equals_reference(decoder_gpu):
    """Five LIZARD-1024 captures (620 blocks) through the batch graph with the LDPC grid forced to one workgroup a
    CU (256 on an MI355X), so workgroups take blocks of several frames in turn. Each frame: the reference's decline
    gate, LDPC and CRC on the batch's own soft values give the batch's verdicts, iterations and bytes; the soft values
    equal the reference's own decode of the same capture and registration (ref_decode_reg) but at rounding ties of
    the GPU's floats (the single-frame path's, equal to the batch's); verified blocks are the blocks sent; graph,
    frame-by-frame and graph again agree."""
    r = ref("scalar")
    if r is None:
        pytest.skip("oracle/_ref not built")
    caps = make_captures(BIG, 77)
    b = L.Batch().plan([c["version"] for c in caps], [c["span"] for c in caps])
    assert b.blocks == 5 * 124
    b.shape([4, 4, 4, 1])
    frames, keep = device_frames(caps)
    outs = []
    for mode in (L.RUN_GRAPH, L.RUN_FRAMES, L.RUN_GRAPH):
        b.run(frames, mode)
        b.sync()
        outs.append(b.blocks_out())
    for o in outs[1:]:
        same(o, outs[0])
    got = outs[0]
    sent = np.concatenate([c["blocks"].reshape(-1, 476) for c in caps])
    v = got["ok"].astype(bool)
    assert v.sum() > 0 and np.array_equal(got["blocks"][v], sent[v])
    for i, c in enumerate(caps):
        at = slice(124 * i, 124 * (i + 1))
        mine = {k: got[k][at] for k in ("ok", "its", "blocks")}
        s = b.soft(i, c["version"])
        assert_finish_equal(mine, ref_finish_soft(r, 64, 512, s["llr"], s["sub_est"]))
        info = L.format_info(64)
        f = r.ref_format_new(info["n"], info["subch"], 512)
        B = 124
        out, ok = np.zeros(B * 476, np.uint8), np.zeros(B, np.uint8)
        its, be = np.zeros(B, np.int8), np.zeros(B, np.float32)
        llr, al = np.zeros(info["subch"] * 640, np.int8), np.zeros(2, np.float32)
        nodes = np.ascontiguousarray(c["nodes"], np.float32)
        r.ref_decode_reg(f, ptr(c["img"]), c["img"].shape[1], c["img"].shape[0], ptr(c["H"]), ptr(nodes), ptr(out),
                         ptr(ok), ptr(its), ptr(be), ptr(llr), ptr(al))
        grid = np.zeros((info["n"], info["n"]), np.float32)
        r.ref_sample_grid(f, ptr(c["img"]), c["img"].shape[1], c["img"].shape[0], ptr(c["H"]), ptr(nodes), 1,
                          ptr(grid))
        r.ref_format_free(f)
        rg = ref_grid_finish(r, 64, 512, grid)
        assert np.array_equal(rg["llr"], llr)
        one = decoder_gpu.decode_reg(64, 512, c["img"], c["H"], nodes)
        assert np.array_equal(one["llr"], s["llr"])
        tie = ties_against_reference(64, one, rg)
        assert tie < 1e-3, tie
        if tie == 0:
            assert np.array_equal(mine["its"], its) and np.array_equal(mine["ok"], ok)
    b.close()


def test_batch_otf_equals_single_frames():
    """lizx_batch_otf: per-frame blurs equalised in a batch give the single-frame path's soft values under the same
    blur (lizx_decoder_otf), frame by frame; frames with no blur stay unequalised."""
    cases = [(5, 128, 0.6), (20, 256, 0.0), (45, 512, 0.5)]
    fr = [Frame(v, s, seed=70 + k, snr_db=22.0, blur=bl) for k, (v, s, bl) in enumerate(cases)]
    keep = [L.DeviceArray.of(F.grid) for F in fr]
    frames = L.make_frames([dict(grid=d) for d in keep])
    b = L.Batch().plan([c[0] for c in cases], [c[1] for c in cases], source=L.SRC_GRID, stages=L.STAGE_SOFT)
    b.otf([c[2] for c in cases])
    dec = L.Decoder(0)
    try:
        for mode in (L.RUN_GRAPH, L.RUN_FRAMES):
            b.run(frames, mode)
            b.sync()
            for i, (v, s, bl) in enumerate(cases):
                dec.otf(bl)
                want = dec.finish_grid(v, s, fr[i].grid)
                got = b.soft(i, v)
                assert np.array_equal(got["llr"], want["llr"]) and np.array_equal(got["sub_est"], want["sub_est"])
        b.otf(None)
        b.run(frames, L.RUN_GRAPH)
        b.sync()
        dec.otf(0.0)
        want = dec.finish_grid(cases[0][0], cases[0][1], fr[0].grid)
        assert np.array_equal(b.soft(0, cases[0][0])["llr"], want["llr"])
    finally:
        dec.close()
        b.close()
