#!/usr/bin/env python3
"""LIZARD receive hot path on one MI355X: the LDPC-only batched workload.

Workload (one step = one decode of one batch): 4096 codewords with 3840 information bits each (476 random bytes and
their CRC-32C), rates mixed 7/8 : 3/4 : 2/3 : 1/2 = 1597 : 1065 : 614 : 820 (39/26/15/20 %), BPSK on each QPSK axis
over AWGN at Eb/N0 swept 1.5 to 4.0 dB in 0.5 dB steps (one batch a point, steps cycling through the points), LLR =
2 y / sigma^2 quantised to int8 as round-half-even(8 LLR) clamped to +-80, the 7/8 code's 128 known zeros at +127;
layered min-sum, at most 20 iterations with early termination, one persistent kernel per step, CRC-32C checked on
the GPU. Inputs are resident in HBM before the timed region.

Prints one JSON line: LDPC Gbit/s (3840 x codewords / s) as the value, payload MB/s (472 x verified blocks / s),
the plain per-codeword path's rate beside the batched one, the decode kernel's roofline (integer operations against
the vector issue peak, with the LDS and HBM figures beside it) and the reference CPU decoder timed on this machine's
cores, the PHY's own setting beside it ("stall_rule": 30 iterations with the stall rule 9 / 0.95), and the same
decode from host buffers ("host_buffers": the PCIe copies included). More workloads ride along as sub-objects:
  "pipeline"   the full receive path, batched: 32 captured 1920x1080 uint8 frames, each an n = 1536 picture
               (LIZARD-1024, 124 blocks) in the 256 ring through a random homography (rotation +-10 deg, scale 0.6-0.9
               of the frame height, perspective up to 0.05), radial k1 in [-0.05, 0.05], blur sigma 0.5-1.5 px, AWGN
               sigma 2-8 grey levels (made on the GPU: lizx_modulate, lizx_capture); warp -> FFT -> demap -> LDPC -> CRC as one
               HIP graph, against the same launches without the graph, frame by frame, and frame by frame with the
               LDPC launched once a codeword; frames/s and payload MB/s; the reference decoder on the host cores.
  "fft_demap"  64 picture grids with n drawn from {384, 768, 1152, 1536} (each at its top format), the sender's
               picture plus AWGN at 20 dB SNR: erase, detrend, 2D FFT, pilot shift, soft demap as one HIP graph;
               algorithmic HBM GB/s against the 8 TB/s peak; the reference's finish path on the same grids.
  "refine"     256 captured frames of LIZARD-320 whose homography corners are up to 1.5 px off: block-fail before and
               after 10 Gauss-Newton steps over the homography (the hand-written adjoint of the warp, FP64 8x8 solve);
               and the finder's own anchor registration of the same captures, before and after the same steps.
  "small"      one n = 384 LIZARD-80 frame in a 1280x720 capture through liz_decode, against the reference's decoder
               on one core.
  "sender"     the sender's pictures for the FFT+demap mix on the GPU (lizx_modulate): HBM roofline of its two passes,
               the reference's encoder on the host cores beside it; and liz_encoder_paint's whole page (LIZARD-1024)
               against the reference's encode and paint on one core.
  "finder"     16 whole-symbol captures (the pipeline's channel, painted with border and word): lizx_find's hits,
               registration error against the truth and time, the blind liz_decode, the same on 16 host threads, and
               the blind receiver at rate (finder on threads, then one batch graph); the reference's finder and its
               whole blind decode beside them.
  "blake3"     the transfer's chunk check: BLAKE3 chaining values of every 1 MiB chunk of a 1 GiB file in HBM, its
               32-bit vector-issue roofline, the reference's BLAKE3 (AVX2) on the host's threads beside it.
"""
import argparse
import ctypes as C
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import lizard_amd  # noqa: E402

METRIC = "Decoded payload MB/s + LDPC Gbit/s at block-fail<1e-4; batched vs plain, 1 GPU"
MIX = ((6, 1597), (4, 1065), (3, 614), (2, 820))
EBN0 = (1.5, 2.0, 2.5, 3.0, 3.5, 4.0)
MAX_ITER = 20
# MI355X: 256 CUs at 2.4 GHz; LDS 128 bytes a clock per CU (the CDNA3-generation figure; CDNA4 documents doubled LDS
# read bandwidth for some instructions, which would halve the fraction below). HBM3E 8 TB/s.
LDS_PEAK_GBS = 256 * 128 * 2.4
HBM_PEAK_GBS = 8000.0
# Vector integer issue: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz, two 16-bit operations a lane (packed) a clock.
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 * 2 / 1e12
# 32-bit integer vector operations: 256 CUs x 4 SIMDs x 16 lanes a clock at 2.4 GHz (no packed form for 32-bit ints)
VALU32_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12
# Integer operations of the reference's scalar layer kernel per edge per iteration (liblizard/src/ldpc.c:171-191):
# subtract, magnitude, sign, first- and second-minimum updates; magnitude select, sign, add, clamp (two).
OPS_PER_EDGE = 10
# The operating point where block-fail < 1e-4 is shown: Eb/N0 at which args.op_batches x 4096 codewords of the
# workload's rate mix decode with no failure (zero failures in N trials bound the rate below 3 / N at 95 %).
OP_EBN0 = 4.5
# captures of the pipeline leg that the reference decoder reads (its CPU timing and the verdict comparison)
PIPE_REF_FRAMES = 8


def build_batch(ebn0_db, seed):
    """4096 codewords at one Eb/N0: rates in a seeded interleaved order; returns rates, int8 LLRs, offsets, payload
    bytes and the sent codewords' information bits."""
    rng = np.random.default_rng(seed)
    rates = np.concatenate([np.full(c, r, np.uint8) for r, c in MIX])
    rates = rates[rng.permutation(rates.size)]
    dims = {r: lizard_amd.ldpc_dims(r) for r, _ in MIX}
    n = np.array([dims[r]["n"] for r in rates])
    offs = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.int64)
    llr = np.empty(int(n.sum()), np.int8)
    payload = rng.integers(0, 256, (rates.size, 476), dtype=np.uint8)
    info = np.zeros((rates.size, 3840), np.uint8)
    for i in range(rates.size):
        crc = lizard_amd.crc32c(payload[i])
        info[i, :3808] = np.unpackbits(payload[i])
        info[i, 3808:] = (crc >> (31 - np.arange(32))) & 1
    ebn0 = 10.0 ** (ebn0_db / 10.0)
    for r, _ in MIX:
        d = dims[r]
        idx = np.nonzero(rates == r)[0]
        data = np.zeros((idx.size, d["k"]), np.uint8)
        data[:, :3840] = info[idx]
        cw = lizard_amd.ldpc_encode(r, data)
        sigma = np.sqrt(1.0 / (2.0 * (3840.0 / d["nt"]) * ebn0))
        y = (1.0 - 2.0 * cw) + sigma * rng.standard_normal(cw.shape)
        q = np.clip(np.rint(8.0 * (2.0 * y / sigma ** 2).astype(np.float32).astype(np.float64)), -80, 80).astype(np.int8)
        if d["k"] > 3840:
            q[:, 3840:d["k"]] = 127
        for j, i in enumerate(idx):
            llr[offs[i]:offs[i] + d["n"]] = q[j]
    return rates, llr, offs, payload, info


def measured_traffic():
    """HBM bytes a k_ldpc launch moves, from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-kernel
    calibrated); None where that summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_ldpc_traffic_ldpc_batched_4096.txt")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def profile_values(name):
    """The "key: value" lines of a committed profile summary (profiles/name) as floats; {} where it is absent."""
    out = {}
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                try:
                    out[k.strip()] = float(v.split()[0])
                except (ValueError, IndexError):
                    pass
    except OSError:
        pass
    return out


def measured_issue():
    """k_ldpc's vector-instruction issue over the timed launches from the committed rocprofv3 SQ counter summary:
    {"valu_insts_per_codeword_iteration", "valu_busy"}; None where that summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_ldpc_issue_ldpc_batched_4096.txt")
    out = {}
    try:
        for line in open(path):
            for key in ("valu_insts_per_codeword_iteration", "valu_busy"):
                if line.startswith(key + ":"):
                    out[key] = float(line.split(":")[1].split()[0])
    except OSError:
        return None
    return out or None


def host_cpus():
    """The host's processors as this process sees them: all of them, and those it may run on."""
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"visible": os.cpu_count() or 1, "usable": usable}


def host_threads():
    """Threads for a CPU baseline: OMP_NUM_THREADS where set (16 on a one-GPU box: its share of the host), else the
    usable processors, never more than those."""
    usable = host_cpus()["usable"]
    return max(1, min(int(os.environ.get("OMP_NUM_THREADS", "0")) or usable, usable))


def effective_iterations(its):
    its = its.astype(np.int64)
    return np.where(its > 0, its, MAX_ITER)


def cpu_baseline(batches, seconds):
    """The reference decoder (compiled from its sources into oracle/_ref, native x86-64 form) on the host cores: one
    decoder per thread as liblizard/core/cpu/pool.h runs them, over a bounded sample of the same codewords."""
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libref_x86.so")
    kind = "reference"
    if not os.path.exists(ref_so):
        return None
    lib = C.CDLL(ref_so)
    fn = lib.ref_ldpc_decode_many
    fn.restype = C.c_int
    fn.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_float]
    cores = host_threads()
    # work items: 32 codewords of one rate from one Eb/N0 point, round robin over points and rates
    items = []
    for (rates, llr, offs, _, _) in batches:
        for r, _ in MIX:
            n = lizard_amd.ldpc_dims(r)["n"]
            idx = np.nonzero(rates == r)[0]
            for s in range(0, idx.size - 31, 32):
                block = np.concatenate([llr[offs[i]:offs[i] + n] for i in idx[s:s + 32]])
                items.append((r, n, block))
    order = np.random.default_rng(0).permutation(len(items))
    items = [items[i] for i in order]
    lock = threading.Lock()
    state = {"next": 0, "words": 0}
    deadline = [0.0]

    def worker():
        out = np.zeros(32 * 7680, np.uint8)
        its = np.zeros(32, np.int8)
        while True:
            with lock:
                if time.perf_counter() > deadline[0]:
                    return
                r, n, block = items[state["next"] % len(items)]
                state["next"] += 1
            fn(r, 32, block.ctypes.data, out.ctypes.data, its.ctypes.data, MAX_ITER, 0, 0.95)
            with lock:
                state["words"] += 32

    t0 = time.perf_counter()
    deadline[0] = t0 + seconds
    th = [threading.Thread(target=worker) for _ in range(cores)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    wall = time.perf_counter() - t0
    words = state["words"]
    return {"value": round(3840.0 * words / wall / 1e9, 4), "unit": "Gbit/s", "cores": cores, "host_cpus": host_cpus(), "kind": kind,
            "sample": "%d codewords (32-word runs of one rate from all six Eb/N0 batches, shuffled, cycled to the time bound), %.1f s wall on %d "
                      "threads; liblizard ldpc_decode_stall (x86-64 SIMD build), 20 iterations, early termination"
                      % (words, wall, cores)}


def pipeline_frames(count, seed):
    """count captures of LIZARD-1024 in the 256 ring through random cameras, made on the GPU (the sender's picture,
    lizx_modulate, then the camera channel, lizx_capture) and copied to the host, where the reference decoder reads
    them and from where the batch legs upload them."""
    from lizard_amd import sim
    rng = np.random.default_rng(seed)
    dec = lizard_amd.Decoder(0)
    caps = []
    for k in range(count):
        cam = sim.Camera.random(rng)
        blocks = sim.random_blocks(64, rng)
        dpic = dec.modulate(64, blocks, cnt=k % 4)
        dimg, H, nodes = sim.capture_gpu(dec, 64, 512, dpic, cam, seed=seed + k)
        img = dimg.get(np.uint8, cam.W * cam.Hh).reshape(cam.Hh, cam.W)
        dimg.free()
        dpic.free()
        caps.append(dict(img=img, H=H, nodes=np.ascontiguousarray(nodes), blocks=blocks, cam=cam))
    dec.close()
    return caps


def soft_bytes(version):
    """Algorithmic HBM bytes of one frame's soft-value kernels from a grid: row pass n^2 4 read + E n 8 write, column
    pass E n 8 read + 320 subch 8 write (the coefficients in coefficient order), demap 320 subch 8 read + 640 subch
    write (plus sub-channel estimates)."""
    info = lizard_amd.format_info(version)
    n, subch = info["n"], info["subch"]
    u, v = lizard_amd.format_coefs(version)
    E = int(max(np.abs(u).max(), v.max())) + 1
    return {"rows": n * n * 4 + E * n * 8, "cols": E * n * 8 + 320 * subch * 8,
            "demap": 320 * subch * 8 + 640 * subch + subch * 12, "n": n, "E": E}


def run_batch(b, frames, mode, runs):
    b.run(frames, mode)
    b.sync()
    t0 = time.perf_counter()
    for _ in range(runs):
        b.run(frames, mode)
    ms = b.sync()
    return (time.perf_counter() - t0) / runs, ms / runs


def bench_pipeline(args, caps):
    D = lizard_amd.DeviceArray
    F = len(caps)
    keep = []
    items = []
    for c in caps:
        d_img, d_nodes = D.of(c["img"]), D.of(c["nodes"])
        keep += [d_img, d_nodes]
        items.append(dict(img=d_img, w=c["img"].shape[1], h=c["img"].shape[0], H=c["H"], nodes=d_nodes))
    frames = lizard_amd.make_frames(items)
    b = lizard_amd.Batch(0).plan([64] * F, [512] * F)
    # the default launch shape's graph time, then the autotuner's pick (the bits are the same whatever the shape)
    _, default_ms = run_batch(b, frames, lizard_amd.RUN_GRAPH, args.pipe_runs)
    # shapes and the FFTs' radix plans (a radix plan reorders the FFT's float operations; all modes below run the
    # one kept, so their bits are compared as before)
    shape = b.tune_plans(frames, reps=3) if not args.no_tune else [0, 0, 0, 0, 8, 8]
    res = {}
    outs = {}
    modes = args.pipe_modes.split(",")
    for name, mode, runs in (("graph", lizard_amd.RUN_GRAPH, args.pipe_runs),
                             ("stream", lizard_amd.RUN_STREAM, args.pipe_runs),
                             ("frames", lizard_amd.RUN_FRAMES, max(1, args.pipe_runs // 2)),
                             ("plain", lizard_amd.RUN_PLAIN, 1)):
        if name not in modes and name != "graph":
            continue
        wall, ev = run_batch(b, frames, mode, runs)
        out = b.blocks_out()
        outs[name] = out
        res[name] = {"frames_per_s": round(F / wall, 2), "ms_per_batch": round(wall * 1e3, 3),
                     "gpu_ms_per_batch": round(ev, 3), "payload_MBps": round(472.0 * out["good"] / wall / 1e6, 3),
                     "runs": runs}
    same = all(np.array_equal(outs[k]["ok"], outs["graph"]["ok"]) and np.array_equal(outs[k]["its"], outs["graph"]["its"])
               and np.array_equal(outs[k]["blocks"][outs[k]["ok"] > 0], outs["graph"]["blocks"][outs["graph"]["ok"] > 0])
               for k in outs)
    g = outs["graph"]
    # the graph run's soft values of the captures the reference decoder reads (its verdicts are compared there)
    b.run(frames, lizard_amd.RUN_GRAPH)
    b.sync()
    for i in range(min(F, PIPE_REF_FRAMES)):
        caps[i]["gpu_soft"] = b.soft(i, 64)
        caps[i]["gpu_out"] = {k: g[k][124 * i:124 * (i + 1)] for k in ("ok", "its", "blocks")}
    sent = np.concatenate([c["blocks"].reshape(-1, 476) for c in caps])
    right = (g["ok"] > 0) & np.all(g["blocks"] == sent, axis=1)
    out = {
        "workload": "full receive pipeline, batched: %d captured 1920x1080 uint8 frames, n=1536 (LIZARD-1024, 124 "
                    "blocks, 256 ring), random homography (rot +-10 deg, scale 0.6-0.9, persp <= 0.05), k1 in "
                    "[-0.05, 0.05], blur 0.5-1.5 px, AWGN 2-8 grey; warp+erase+row FFT fused, column FFT, pilot "
                    "shift, demap, decline gate, 30-iteration stall-ruled LDPC, CRC-32C" % F,
        "blocks": int(g["ok"].size), "verified": int(g["good"]), "block_fail": round(1.0 - float(right.mean()), 4),
        "verified_bytes_are_sent": bool(right.sum() == g["good"]), "modes_bit_identical": bool(same),
        "modes": res,
        # batched over plain: the graph against frame-by-frame launches (one LDPC launch a frame, the host waiting a
        # frame), the path liz_decode takes once a frame; the target is 50x
        "batched_over_plain": (round(res["graph"]["frames_per_s"] / res["frames"]["frames_per_s"], 2)
                               if "frames" in res else None),
        "batched_over_plain_target": 50.0,
        # an extra: frame by frame with the LDPC launched once a codeword (no product entry decodes that way)
        "graph_over_per_codeword_launches": (round(res["graph"]["frames_per_s"] / res["plain"]["frames_per_s"], 2)
                                             if "plain" in res else None),
        "autotune": {"shape": {"rows_waves": shape[0], "cols_waves": shape[1], "demap_waves": shape[2],
                               "ldpc_wg_per_cu": shape[3] or "max", "rows_radix": "%d x %d" % (shape[4], 64 // shape[4]),
                               "cols_radix": "%d x %d" % (shape[5], 64 // shape[5])},
                     "default_shape_gpu_ms_per_batch": round(default_ms, 3)},
    }
    # the picture stages' HBM traffic and issue at the default shape, from the committed counter summary
    pc = profile_values("pipeline_counters_32x1536.txt")
    if pc:
        stages = {}
        for key, kernel in (("row_pass", "k_rows<WarpSource, 1536, 4>"), ("column_pass", "k_cols<1536, 4>"),
                            ("demap", "k_demap<4>")):
            t, us = pc.get(key + "_traffic_bytes_per_launch"), pc.get(key + "_avg_us")
            if t and us:
                gbs = t / (us * 1e-6) / 1e9
                stages[key] = {"kernel": kernel, "traffic": int(t), "avg_us": us, "achieved": round(gbs, 1),
                               "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)}
        if "row_pass_valu_busy" in pc:
            stages["row_pass"]["valu_busy"] = pc["row_pass_valu_busy"]
        stages["from"] = ("profiles/pipeline_counters_32x1536.txt: rocprofv3 --kernel-trace and --pmc runs of this leg "
                          "(FETCH_SIZE, WRITE_SIZE copy-calibrated; SQ counters), copied, not measured by this run")
        out["stage_rooflines"] = stages
    b.close()
    for d in keep:
        d.free()
    return out


def bench_fft_demap(args):
    from lizard_amd import sim
    rng = np.random.default_rng(77)
    table = {384: (5, 128), 768: (20, 256), 1152: (45, 512), 1536: (64, 512)}
    ns = rng.choice(sorted(table), size=args.fft_frames)
    versions = [table[int(n)][0] for n in ns]
    spans = [table[int(n)][1] for n in ns]
    keep, items, host = [], [], []
    total = 0
    for v, sp in zip(versions, spans):
        pic = sim.picture(v, sim.random_blocks(v, rng)).numpy()
        # AWGN at 20 dB SNR against the picture's own variance
        grid = (pic + rng.standard_normal(pic.shape) * np.sqrt(pic.var() / 100.0)).astype(np.float32)
        host.append(dict(grid=grid, version=v, span=sp))
        d = lizard_amd.DeviceArray.of(grid)
        keep.append(d)
        items.append(dict(grid=d))
        sb = soft_bytes(v)
        total += sb["rows"] + sb["cols"] + sb["demap"]
    frames = lizard_amd.make_frames(items)
    b = lizard_amd.Batch(0).plan(versions, spans, source=lizard_amd.SRC_GRID, stages=lizard_amd.STAGE_SOFT)
    _, default_ms = run_batch(b, frames, lizard_amd.RUN_GRAPH, args.pipe_runs)
    # the tuner over shapes and radix plans; the time of each radix plan alone (default shapes) beside it
    plan_ms = {}
    for r1, r2 in ((8, 8), (4, 4), (16, 16), (4, 16), (16, 4)):
        b.shape([4, 4, 4, 0, r1, r2])
        plan_ms["rows %d x %d, cols %d x %d" % (r1, 64 // r1, r2, 64 // r2)] = round(
            run_batch(b, frames, lizard_amd.RUN_GRAPH, args.pipe_runs)[1], 4)
    b.shape([4, 4, 4, 0, 8, 8])
    shape = b.tune_plans(frames, reps=3)
    wall, ev = run_batch(b, frames, lizard_amd.RUN_GRAPH, args.pipe_runs)
    fwall, fev = run_batch(b, frames, lizard_amd.RUN_FRAMES, max(1, args.pipe_runs // 2))
    gbs = total / (ev * 1e-3) / 1e9
    b.close()
    for d in keep:
        d.free()
    otf = bench_otf(args, rng, table)
    cpu = None if args.no_cpu else cpu_fft_demap(host, args.cpu_seconds)
    return {"cpu_baseline": cpu,
            "graph_over_cpu": round(len(ns) / (ev * 1e-3) / cpu["value"], 1) if cpu and cpu["value"] > 0 else None,
            "workload": "%d picture grids, n drawn from {384, 768, 1152, 1536} (%s), top formats, sender picture + "
                        "AWGN 20 dB; erase, detrend, 2D FFT, pilot shift, soft demap" %
                        (len(ns), ", ".join("%d x %d" % (int((ns == n).sum()), n) for n in sorted(table))),
            "graph_ms_per_batch": round(ev, 3), "frames_per_s": round(len(ns) / wall, 1),
            "autotune": {"shape": {"rows_waves": shape[0], "cols_waves": shape[1], "demap_waves": shape[2],
                                   "rows_radix": "%d x %d" % (shape[4], 64 // shape[4]),
                                   "cols_radix": "%d x %d" % (shape[5], 64 // shape[5])},
                         "default_shape_graph_ms_per_batch": round(default_ms, 3),
                         "graph_ms_by_radix_plan": plan_ms},
            "frame_by_frame_ms_per_batch": round(fev, 3),
            "algorithmic_bytes_per_batch": int(total),
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4),
                         "traffic": profile_values("fft_demap_traffic_64_mixed.txt").get("traffic_bytes_per_batch"),
                         "traffic_from": "profiles/fft_demap_traffic_64_mixed.txt: rocprofv3 --pmc runs of this leg "
                                         "(FETCH_SIZE, WRITE_SIZE, copy-calibrated), copied, not measured by this "
                                         "run"},
            "otf_equalised": otf}


def bench_otf(args, rng, table, frames_n=16):
    """The FFT + demap with equalisation by each frame's known blur (lizx_batch_otf): frames_n grids of the same
    sizes blurred by a Gaussian of sigma in [0.5, 1.0] samples (a periodic blur, through numpy's FFT) plus AWGN at 20
    dB, and the same at sigma 1.0-1.6 and 14 dB; the graph's time with the equalisation against without, and the
    blocks each way verifies (lizx_finish_grid's LDPC and CRC through lizx_decoder_otf)."""
    out = {}
    for name, lo, hi, snr in (("sigma 0.5-1.0, 20 dB", 0.5, 1.0, 20.0), ("sigma 1.0-1.6, 14 dB", 1.0, 1.6, 14.0)):
        out[name] = bench_otf_case(args, rng, table, frames_n, lo, hi, snr)
    return out


def bench_otf_case(args, rng, table, frames_n, lo, hi, snr):
    from lizard_amd import sim
    ns = rng.choice(sorted(table), size=frames_n)
    sig = rng.uniform(lo, hi, frames_n)
    grids, vs, sps, blocks_sent = [], [], [], []
    for n, s_ in zip(ns, sig):
        v, sp = table[int(n)]
        blk = sim.random_blocks(v, rng)
        pic = sim.picture(v, blk).numpy().astype(np.float64)
        f = np.fft.fftfreq(int(n)) * int(n)
        H = np.exp(-2.0 * np.pi ** 2 * s_ * s_ * (f[:, None] ** 2 + f[None, :] ** 2) / (int(n) ** 2))
        pic = np.real(np.fft.ifft2(np.fft.fft2(pic) * H))
        grids.append((pic + rng.standard_normal(pic.shape) * np.sqrt(pic.var() / 10 ** (snr / 10))).astype(np.float32))
        vs.append(v)
        sps.append(sp)
        blocks_sent.append(np.asarray(blk, np.uint8).reshape(-1, 476))
    keep = [lizard_amd.DeviceArray.of(g) for g in grids]
    frames = lizard_amd.make_frames([dict(grid=d) for d in keep])
    b = lizard_amd.Batch(0).plan(vs, sps, source=lizard_amd.SRC_GRID, stages=lizard_amd.STAGE_SOFT)
    _, plain_ms = run_batch(b, frames, lizard_amd.RUN_GRAPH, args.pipe_runs)
    b.otf(sig.astype(np.float32))
    _, eq_ms = run_batch(b, frames, lizard_amd.RUN_GRAPH, args.pipe_runs)
    b.close()
    dec = lizard_amd.Decoder(0)
    verified = {}
    right = True
    for name, on in (("unequalised", False), ("equalised", True)):
        good = 0
        for g, v, sp, s_, sent in zip(grids, vs, sps, sig, blocks_sent):
            dec.otf(float(s_) if on else 0.0)
            out = dec.finish_grid(v, sp, g, soft=False)
            ok = out["ok"].astype(bool)
            right &= bool(np.array_equal(out["blocks"][ok], sent[ok]))
            good += int(ok.sum())
        verified[name] = good
    dec.close()
    for d in keep:
        d.free()
    return {"workload": "%d grids of the mix's sizes blurred by a Gaussian of sigma %.1f-%.1f samples, AWGN %g dB" %
                        (frames_n, lo, hi, snr),
            "graph_ms_per_batch": {"unequalised": round(plain_ms, 4), "equalised": round(eq_ms, 4)},
            "blocks": int(sum(x.shape[0] for x in blocks_sent)), "verified": verified,
            "verified_bytes_are_sent": right}


def cpu_fft_demap(grids, seconds):
    """The reference's finish path on the same picture grids (x86-64 build; erase, detrend, FFT, pilot shift, soft
    values, then its LDPC, which at 20 dB converges in one iteration), one format instance per size and thread; grids
    re-laid into its strip layout beforehand (picture.to_strips), copied in per call (the finish transforms in
    place)."""
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libref_x86.so")
    if not os.path.exists(ref_so):
        return None
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import picture as oracle_picture
    lib = C.CDLL(ref_so)
    lib.ref_format_new.restype = C.c_void_p
    lib.ref_format_new.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.ref_format_free.argtypes = [C.c_void_p]
    fn = lib.ref_finish_grid
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p] * 12
    strips = []
    for g in grids:
        info = lizard_amd.format_info(g["version"])
        strips.append(dict(s=oracle_picture.to_strips(g["grid"], info["n"]), n=info["n"], subch=info["subch"],
                           span=g["span"], blocks=info["blocks"]))
    cores = host_threads()
    lock = threading.Lock()
    state = {"next": 0, "frames": 0, "good": 0, "blocks": 0}
    deadline = [0.0]

    def worker():
        forms = {}
        work = np.empty(max(g["s"].size for g in strips), np.float32)
        out = np.zeros(124 * 476, np.uint8)
        ok, its, be = np.zeros(124, np.uint8), np.zeros(124, np.int8), np.zeros(124, np.float32)
        while True:
            with lock:
                if time.perf_counter() > deadline[0]:
                    break
                g = strips[state["next"] % len(strips)]
                state["next"] += 1
            key = (g["n"], g["subch"], g["span"])
            if key not in forms:
                forms[key] = lib.ref_format_new(*key)
            w = work[:g["s"].size]
            w[:] = g["s"]
            good = fn(forms[key], w.ctypes.data, out.ctypes.data, ok.ctypes.data, its.ctypes.data, be.ctypes.data,
                      None, None, None, None, None, None)
            with lock:
                state["frames"] += 1
                state["good"] += good
                state["blocks"] += g["blocks"]
        for f in forms.values():
            lib.ref_format_free(f)

    t0 = time.perf_counter()
    deadline[0] = t0 + seconds
    th = [threading.Thread(target=worker) for _ in range(cores)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    wall = time.perf_counter() - t0
    return {"value": round(state["frames"] / wall, 2), "unit": "frames/s", "cores": cores, "host_cpus": host_cpus(), "kind": "reference",
            "blocks_verified": round(state["good"] / max(state["blocks"], 1), 4),
            "sample": "%d grids (the same %d cycled), %.1f s wall on %d threads; liblizard lizard_finish on supplied "
                      "grids (x86-64 SIMD build), which also runs its LDPC (one iteration at 20 dB)" %
                      (state["frames"], len(strips), wall, cores)}


def bench_sender(args, dec):
    """The sender's pictures on the GPU (lizx_modulate) for the FFT+demap leg's mix of sizes: the bit side (CRC,
    LDPC encoding, bit map, whitening), QPSK, the Hermitian spectrum's inverse 2D DFT and the clip, all on the GPU
    after the blocks' upload; HBM roofline over the GPU time; the reference's own encoder (lizard_encode) on the host
    cores."""
    rng = np.random.default_rng(78)
    table = {384: 5, 768: 20, 1152: 45, 1536: 64}
    ns = rng.choice(sorted(table), size=args.fft_frames)
    groups = {}
    for n in ns:
        groups.setdefault(table[int(n)], 0)
        groups[table[int(n)]] += 1
    work, total = [], 0
    for v, k in sorted(groups.items()):
        info = lizard_amd.format_info(v)
        u, vv = lizard_amd.format_coefs(v)
        E = int(max(np.abs(u).max(), vv.max())) + 1
        n = info["n"]
        blocks = rng.integers(0, 256, (k, info["blocks"] * 476)).astype(np.uint8)
        out = lizard_amd.DeviceArray(k * n * n * 4)
        work.append((v, blocks, out))
        # blocks in, slots out and back in, the column pass's E n complex out and back in, the picture out
        total += k * (info["blocks"] * 476 + 2 * info["subch"] * 640 + 2 * E * n * 8 + n * n * 4)
    for v, blocks, out in work:
        dec.modulate(v, blocks, out=out)
    runs = max(1, args.pipe_runs)
    t0 = time.perf_counter()
    kms = 0.0
    for _ in range(runs):
        for v, blocks, out in work:
            kms += dec.modulate(v, blocks, out=out, timed=True)[1]
    wall = (time.perf_counter() - t0) / runs
    kms /= runs
    for _, _, out in work:
        out.free()
    gbs = total / (kms * 1e-3) / 1e9
    cpu = None if args.no_cpu else cpu_sender(work, args.cpu_seconds)
    # the drop-in: liz_encoder_paint, blocks in and the whole GREY8 page out in host memory, frame by frame
    enc = lizard_amd.Encoder(64, 2, 60)
    g = enc.geometry
    eb = rng.integers(0, 256, (8, g["frame_blocks"] * 476)).astype(np.uint8)
    enc.paint(eb[0], 0)
    t0 = time.perf_counter()
    reps = 24
    for i in range(reps):
        enc.paint(eb[i % 8], i)
    enc_ms = (time.perf_counter() - t0) * 1e3 / reps
    enc.close()
    encoder = {"workload": "liz_encoder_paint, LIZARD-1024 in the 256 ring (a %d x %d GREY8 page), blocks from and the "
                           "page to host memory, frame by frame" % (g["side"], g["side"]),
               "ms_per_frame": round(enc_ms, 3)}
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libref_x86.so")
    if os.path.exists(ref_so) and not args.no_cpu:
        lib = C.CDLL(ref_so)
        lib.ref_format_new.restype = C.c_void_p
        lib.ref_format_new.argtypes = [C.c_int, C.c_int, C.c_int]
        lib.ref_format_free.argtypes = [C.c_void_p]
        lib.ref_set_fps.argtypes = [C.c_void_p, C.c_int]
        lib.ref_encode.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        lib.ref_paint_grey.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        f = lib.ref_format_new(1536, 1024, 512)
        lib.ref_set_fps(f, 60)
        px = g["side"] - 4 * g["pxm"]
        drive = np.zeros(px * px, np.float32)
        page = np.zeros(g["side"] * g["side"], np.uint8)
        t0 = time.perf_counter()
        n_ref = 0
        while time.perf_counter() - t0 < 2.0:
            lib.ref_encode(f, eb[n_ref % 8].ctypes.data, n_ref & 3, drive.ctypes.data, None)
            lib.ref_paint_grey(f, drive.ctypes.data, page.ctypes.data)
            n_ref += 1
        ref_ms = (time.perf_counter() - t0) * 1e3 / n_ref
        lib.ref_format_free(f)
        encoder["cpu_baseline"] = {"value": round(ref_ms, 3), "unit": "ms/frame", "cores": 1, "kind": "reference",
                                   "sample": "%d frames, 2 s, one thread: lizard_encode + lizard_paint_grey (x86-64 "
                                             "SIMD build)" % n_ref}
    return {"workload": "%d sender pictures, n drawn from {384, 768, 1152, 1536} (%s), top formats, random blocks" %
                        (len(ns), ", ".join("%d x %d" % (int((ns == n).sum()), n) for n in sorted(table))),
            "encoder": encoder,
            "gpu_ms_per_batch": round(kms, 3), "gpu_frames_per_s": round(len(ns) / (kms * 1e-3), 1),
            "frames_per_s_end_to_end": round(len(ns) / wall, 1),
            "algorithmic_bytes_per_batch": int(total),
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4)},
            "cpu_baseline": cpu}


def cpu_sender(work, seconds):
    """The reference's encoder (lizard_encode, x86-64 build) on the same blocks, one format instance per size and
    thread: its bit side, inverse FFT and picture, and also the Lanczos-3 resample into its drive."""
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libref_x86.so")
    if not os.path.exists(ref_so):
        return None
    lib = C.CDLL(ref_so)
    lib.ref_format_new.restype = C.c_void_p
    lib.ref_format_new.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.ref_format_free.argtypes = [C.c_void_p]
    lib.ref_format_info.argtypes = [C.c_void_p, C.c_void_p]
    fn = lib.ref_encode
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    span = {5: 128, 20: 256, 45: 512, 64: 512}
    items = [(v, blocks[i]) for v, blocks, _ in work for i in range(blocks.shape[0])]
    cores = host_threads()
    lock = threading.Lock()
    state = {"next": 0, "frames": 0}
    deadline = [0.0]

    def worker():
        forms = {}
        while True:
            with lock:
                if time.perf_counter() > deadline[0]:
                    break
                v, bl = items[state["next"] % len(items)]
                state["next"] += 1
            if v not in forms:
                info = lizard_amd.format_info(v)
                f = lib.ref_format_new(info["n"], info["subch"], span[v])
                fi = np.zeros(10, np.int32)
                lib.ref_format_info(f, fi.ctypes.data)
                forms[v] = (f, np.zeros(int(fi[8]) ** 2, np.float32))
            f, drive = forms[v]
            fn(f, bl.ctypes.data, 0, drive.ctypes.data, None)
            with lock:
                state["frames"] += 1
        for f, _ in forms.values():
            lib.ref_format_free(f)

    t0 = time.perf_counter()
    deadline[0] = t0 + seconds
    th = [threading.Thread(target=worker) for _ in range(cores)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    wall = time.perf_counter() - t0
    return {"value": round(state["frames"] / wall, 2), "unit": "frames/s", "cores": cores, "host_cpus": host_cpus(), "kind": "reference",
            "sample": "%d frames (the same %d cycled), %.1f s wall on %d threads; liblizard lizard_encode (x86-64 SIMD "
                      "build): bit side, inverse FFT, clipped picture, and the Lanczos-3 resample into its drive" %
                      (state["frames"], len(items), wall, cores)}


def refine_frames(count, seed, k2=0.0):
    """count captures of LIZARD-320 (n = 768) in the 128 ring, the whole symbol as a page shows it (border, word,
    anchors) through random cameras (sender and channel on the GPU), each with a registration whose homography corners
    are up to 1.5 px off (the lens term known); with k2, a second radial term drawn in [-k2, k2] (which the registration
    model's H and k1 cannot hold)."""
    from lizard_amd import sim
    rng = np.random.default_rng(seed)
    dec = lizard_amd.Decoder(0)
    side = 256 + 30
    cs = np.array([[0, 0], [side, 0], [side, side], [0, side]], float)
    out = []
    for k in range(count):
        cam = sim.Camera.random(rng, scale=(0.75, 0.9))
        kk2 = float(rng.uniform(-k2, k2)) if k2 else 0.0
        blocks = sim.random_blocks(20, rng)
        dpic = dec.modulate(20, blocks)
        dimg = dec.capture(20, 256, dpic, cam.W, cam.Hh, cam.H(256), k1=cam.k1, blur=cam.blur, noise=cam.noise,
                           seed=seed + k, fps=30, k2=kk2)
        img = dimg.get(np.uint8, cam.W * cam.Hh).reshape(cam.Hh, cam.W)
        dimg.free()
        dpic.free()
        Ht = cam.H(256)
        Ht = Ht / Ht[2, 2]
        w = cs @ Ht[:, :2].T + Ht[:, 2]
        H0 = sim.homography(cs, w[:, :2] / w[:, 2:] + rng.uniform(-1.5, 1.5, (4, 2)))
        scale = np.sqrt(abs(np.linalg.det(Ht[:2, :2]))) * 256 / 768  # capture pixels a picture sample
        out.append(dict(img=img, H0=H0, k1=cam.k1, k2=kk2, blocks=blocks, blur=np.sqrt(cam.blur ** 2 + 1 / 3) / scale))
    dec.close()
    return out


def refine_two_terms(dec, count, B):
    """Captures behind a lens of two radial terms (k1 in [-0.05, 0.05], k2 in [-0.06, 0.06]) registered by H (corners
    1.5 px off) and k1 alone, then three decision-directed rounds of the spectral refinement over H and k1, and over
    H, k1 and the anchors' offsets (which carry what k2 adds); block-fail after each round."""
    from lizard_amd import sim
    frames = refine_frames(count, 919, k2=0.06)
    D = lizard_amd.DeviceArray
    sent = np.concatenate([fr["blocks"].reshape(-1, 476) for fr in frames])
    keep = []

    def decode(ps):
        items = []
        for fr, p in zip(frames, ps):
            nodes = lizard_amd.reg_nodes(256, fr["H0"].astype(np.float32), fr["img"].shape[1], fr["img"].shape[0], p)
            d_img = fr.get("d_img") or D.of(fr["img"])
            fr["d_img"] = d_img
            d_nodes = D.of(nodes)
            keep.append(d_nodes)
            items.append(dict(img=d_img, w=fr["img"].shape[1], h=fr["img"].shape[0],
                              H=np.append(p[:8], 1.0).reshape(3, 3), nodes=d_nodes))
        b = lizard_amd.Batch(0).plan([20] * len(frames), [256] * len(frames))
        b.run(lizard_amd.make_frames(items), lizard_amd.RUN_GRAPH)
        b.sync()
        out = b.blocks_out()
        b.close()
        return out

    def fail(out):
        return round(1.0 - float(((out["ok"] > 0) & np.all(out["blocks"] == sent, axis=1)).mean()), 5)

    p0 = [lizard_amd.reg_params(fr["H0"], fr["k1"]) for fr in frames]
    out0 = decode(p0)
    res = {"frames": len(frames), "block_fail_before": fail(out0)}
    for anchors, name in ((False, "H_and_k1"), (True, "H_k1_and_anchors")):
        ps = [p.copy() for p in p0]
        out, path = out0, []
        t0 = time.perf_counter()
        for rnd in range(3):
            ok = out["ok"].reshape(len(frames), B).astype(bool)
            for k, fr in enumerate(frames):
                ref_pic = sim.remodulate(20, fr["blocks"], ok[k], 0, fr["blur"])
                ps[k], _ = dec.reg_refine_spectral(20, 256, fr["img"], fr["H0"], ps[k], ref_pic,
                                                   sim.known_mask(20, ok[k]), steps=8, anchors=anchors)
            out = decode(ps)
            path.append(fail(out))
        res["block_fail_after_" + name] = path
        res["ms_per_frame_round_" + name] = round((time.perf_counter() - t0) / (3 * len(frames)) * 1e3, 3)
    for fr in frames:
        if fr.get("d_img"):
            fr["d_img"].free()
    for d in keep:
        d.free()
    res["what"] = ("the registration knows H (corners 1.5 px off) and the camera's k1 but not its k2; decision-directed "
                   "rounds of lizx_reg_refine_spectral (8 steps) over H and k1, or over H, k1 and the 49 anchors' "
                   "offsets; block-fail after each round")
    return res


def bench_small(args):
    """One n = 384 capture (LIZARD-80, 64 ring) in a 1280x720 frame: homography (5 deg, 0.02 perspective), blur 0.8
    px, AWGN 4 grey; the full decode through liz_decode (held registration, host buffers in and out) against the
    reference's decoder on one host core, both on the same capture and registration."""
    from lizard_amd import sim
    rng = np.random.default_rng(384)
    cam = sim.Camera(1280, 720, rot_deg=5.0, scale=0.8, persp=0.02, k1=0.0, blur=0.8, noise=4.0)
    blocks = sim.random_blocks(5, rng)
    img, H, nodes = sim.capture(sim.picture(5, blocks), 128, cam, seed=1)
    img = img.numpy()
    nodes = np.ascontiguousarray(nodes)
    dec = lizard_amd.Decoder(384)
    if args.small_eager:
        dec.graphs(False)
    dec.hold(5, 128, H, nodes)
    k, got, out, _ = dec.decode(img)

    def steady(call, warm_s=0.3, run_s=0.5):
        """ms a call in a tight loop once clocks have settled: warm_s of calls untimed, then run_s timed."""
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < warm_s:
            call()
        t0, n = time.perf_counter(), 0
        while time.perf_counter() - t0 < run_s:
            call()
            n += 1
        return (time.perf_counter() - t0) * 1e3 / n, n

    # the C-ABI call with its arguments made once (as the reference's decoder is timed below), and the Python wrapper
    ver = np.zeros(dec.max_blocks() * 476, np.uint8)
    held, dout = C.c_int(0), lizard_amd._Decoded()
    L = lizard_amd.lib()
    abi_args = (dec.h, img.ctypes.data_as(C.c_void_p), 1280, 720, img.strides[0], 1, C.byref(held),
                ver.ctypes.data_as(C.c_void_p), C.byref(dout))
    gpu_ms, n_abi = steady(lambda: L.liz_decode(*abi_args))
    wrap_ms, _ = steady(lambda: dec.decode(img), warm_s=0.05, run_s=0.3)
    same = int(dout.verified) == int(k) and np.array_equal(ver[:k * 476], got.ravel())
    dec.close()
    res = {"workload": "one n=384 LIZARD-80 picture (9 blocks, 64 ring) in a 1280x720 capture: homography 5 deg, "
                       "persp 0.02, blur 0.8 px, AWGN 4 grey; liz_decode through the held registration",
           "verified": int(k), "blocks": 9, "bytes_are_sent": bool(np.array_equal(got.ravel(), blocks[:k * 476])),
           "gpu_ms_per_frame": round(gpu_ms, 4), "wrapper_ms_per_frame": round(wrap_ms, 4),
           "abi_equals_wrapper": bool(same), "graph": not args.small_eager,
           "timing": "%d liz_decode calls through the C ABI in 0.5 s after 0.3 s of the same (host frame in, blocks "
                     "out, a stream sync each); wrapper: lizard_amd.Decoder.decode, the same loop" % n_abi,
           "pilot_r": [round(out.pilot_r[0], 3), round(out.pilot_r[1], 3)]}
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libref_x86.so")
    if os.path.exists(ref_so) and not args.no_cpu:
        lib = C.CDLL(ref_so)
        lib.ref_format_new.restype = C.c_void_p
        lib.ref_format_new.argtypes = [C.c_int, C.c_int, C.c_int]
        lib.ref_format_free.argtypes = [C.c_void_p]
        fn = lib.ref_decode_reg
        fn.restype = C.c_int
        fn.argtypes = [C.c_void_p] * 2 + [C.c_int] * 2 + [C.c_void_p] * 8
        f = lib.ref_format_new(384, 80, 128)
        o = np.zeros(9 * 476, np.uint8)
        ok, its, be = np.zeros(9, np.uint8), np.zeros(9, np.int8), np.zeros(9, np.float32)
        call = lambda: fn(f, img.ctypes.data, 1280, 720, H.ctypes.data, nodes.ctypes.data, o.ctypes.data,  # noqa: E731
                          ok.ctypes.data, its.ctypes.data, be.ctypes.data, None, None)
        for _ in range(3):
            call()
        t0 = time.perf_counter()
        n = 0
        while time.perf_counter() - t0 < 2.0:
            good = call()
            n += 1
        cpu_ms = (time.perf_counter() - t0) * 1e3 / n
        lib.ref_format_free(f)
        res["cpu_baseline"] = {"value": round(cpu_ms, 3), "unit": "ms/frame", "cores": 1, "kind": "reference",
                               "verified": int(good),
                               "sample": "%d decodes of the same capture, 2 s, one thread; liblizard lizard_finish "
                                         "through the held registration (x86-64 SIMD build)" % n}
    return res


def finder_frames(count, seed):
    """The full-pipeline workload's channel (1920x1080, LIZARD-1024 in the 256 ring: rotation +-10 deg, scale 0.6-0.9
    of the height, perspective <= 0.05, k1 +-0.05, blur 0.5-1.5 px, AWGN 2-8 grey) on the whole symbol as a page
    shows it (border, word, anchors, guard; lizx_capture_symbol), pictures and captures made on the GPU; with each
    capture's true registration."""
    from lizard_amd import sim
    rng = np.random.default_rng(seed)
    dec = lizard_amd.Decoder(0)
    version, span = 64, 512
    out = []
    for i in range(count):
        blocks = sim.random_blocks(version, rng)
        cam = sim.Camera(1920, 1080, rot_deg=float(rng.uniform(-10, 10)), scale=float(rng.uniform(0.6, 0.9)),
                         persp=float(rng.uniform(0, 0.05)), k1=float(rng.uniform(-0.05, 0.05)),
                         blur=float(rng.uniform(0.5, 1.5)), noise=float(rng.uniform(2, 8)))
        Hm = cam.H(span)
        dpic = dec.modulate(version, blocks)
        dimg = dec.capture(version, span, dpic, cam.W, cam.Hh, Hm, k1=cam.k1, blur=cam.blur, noise=cam.noise, seed=i,
                           fps=60)
        img = dimg.get(np.uint8, cam.W * cam.Hh).reshape(cam.Hh, cam.W)
        dimg.free()
        dpic.free()
        out.append(dict(img=img, H=(Hm / Hm[2, 2]).astype(np.float32).ravel(), nodes=sim.nodes_for(cam, span, Hm),
                        k1=cam.k1, blocks=blocks))
    dec.close()
    return out, version, span


def bench_finder(args):
    """The finder (lizx_find) and the blind decode (liz_decode with nothing held: find, word, decode) on whole-symbol
    captures, against the true registration's decode and the reference's own finder (lizard_acquire) on the host."""
    frames, version, span = finder_frames(16, 1234)
    node_x = lizard_amd.node_lattice(span)
    X, Y = np.meshgrid(node_x, node_x)
    P = np.stack([X.ravel(), Y.ravel(), np.ones(X.size)])

    def place(Hm, nd):
        q = np.asarray(Hm, np.float64).reshape(3, 3) @ P
        return q[:2] / q[2] + np.asarray(nd, np.float64).reshape(-1, 2).T

    dec = lizard_amd.Decoder(0)
    for f in frames[:2]:
        dec.find(f["img"])
    found, errs, k1err, stages = 0, [], [], []
    t0 = time.perf_counter()
    res = [dec.find(f["img"]) for f in frames]
    find_ms = (time.perf_counter() - t0) * 1e3 / len(frames)
    for f, r in zip(frames, res):
        if r is None or r["span"] != span:
            continue
        found += 1
        errs.append(float(np.hypot(*(place(r["H"], r["nodes"]) - place(f["H"], f["nodes"]))).mean()))
        k1err.append(abs(r["k1"] - f["k1"]))
        stages.append(r["ms"])
    # blind decode against the decode through the true registration
    blind = lizard_amd.Decoder(0)
    t0 = time.perf_counter()
    kb = 0
    for f in frames:
        k, _, out, _ = blind.decode(f["img"])
        kb += k
    blind_ms = (time.perf_counter() - t0) * 1e3 / len(frames)
    blind.close()
    held = lizard_amd.Decoder(0)
    t0 = time.perf_counter()
    kt = 0
    for f in frames:
        held.hold(version, span, f["H"], f["nodes"])
        k, _, _, _ = held.decode(f["img"])
        kt += k
    held_ms = (time.perf_counter() - t0) * 1e3 / len(frames)
    held.close()
    dec.close()
    # the same blind decodes on host threads, one decoder (stream, buffers) a thread: the finder's geometry is host work
    import concurrent.futures as cf
    nthreads = min(16, os.cpu_count() or 1)
    decs = [lizard_amd.Decoder(0) for _ in range(nthreads)]
    rounds = 4
    jobs = [(i % nthreads, frames[i % len(frames)]) for i in range(rounds * len(frames))]
    locks = [threading.Lock() for _ in range(nthreads)]

    def blind_job(j):
        k, f = j
        with locks[k]:
            return decs[k].decode(f["img"])[0]

    with cf.ThreadPoolExecutor(nthreads) as ex:
        list(ex.map(blind_job, jobs[:nthreads]))
        t0 = time.perf_counter()
        kv = sum(ex.map(blind_job, [(i % nthreads, jobs[i][1]) for i in range(len(jobs))]))
        tw = time.perf_counter() - t0
    threaded = {"threads": nthreads, "frames_per_s": round(len(jobs) / tw, 1), "verified": int(kv),
                "frames": len(jobs)}

    # the blind receiver at rate: the finder and the word on host threads (a decoder a thread; lizx_find,
    # lizx_read_word), then one batch graph over all the frames through what was found (lizx_batch_*). The captures
    # are in device memory already (a camera's frames arriving there); the plan is made once, the stream's format fixed.
    F = 32
    bf = [frames[i % len(frames)] for i in range(F)]
    D = lizard_amd.DeviceArray
    d_imgs = [D.of(f["img"]) for f in bf]
    nx = len(node_x)
    d_nodes = [D(nx * nx * 8) for _ in range(F)]
    batch = lizard_amd.Batch(0).plan([version] * F, [span] * F)

    finder = lizard_amd.Decoder(0)
    bimgs = [f["img"] for f in bf]

    def find_word(i):
        k = i % nthreads
        with locks[k]:
            r = decs[k].find(bimgs[i])
        if r is None:
            return None
        return r, lizard_amd.read_word(r["span"], bimgs[i], r["H"], r["nodes"])

    def blind_batch(native):
        with cf.ThreadPoolExecutor(nthreads) as ex:
            if native:  # one lizx_find_batch over the frames, each frame's word read by the thread that found it
                regs = finder.find_batch(bimgs, words=True)
                found = [None if r is None or r["version"] is None else (r, (r["version"],)) for r in regs]
            else:  # lizx_find and the word on the threads, a decoder each
                found = list(ex.map(find_word, range(F)))
        items = []
        for i, fw in enumerate(found):
            assert fw is not None and fw[1] is not None and fw[0]["span"] == span and fw[1][0] == version
            d_nodes[i].put(fw[0]["nodes"])
            items.append(dict(img=d_imgs[i], w=1920, h=1080, H=fw[0]["H"], nodes=d_nodes[i]))
        batch.run(lizard_amd.make_frames(items), lizard_amd.RUN_GRAPH)
        batch.sync()
        return batch.blocks_out()

    sent = np.concatenate([f["blocks"].reshape(-1, 476) for f in bf])
    blind_pipe = {"frames": F, "threads": nthreads}
    for native, key in ((False, "find_on_threads"), (True, "find_batch")):
        blind_batch(native)
        t0 = time.perf_counter()
        reps = 5
        for _ in range(reps):
            got = blind_batch(native)
        wb = (time.perf_counter() - t0) / reps
        right = (got["ok"] > 0) & np.all(got["blocks"] == sent, axis=1)
        blind_pipe[key] = {"frames_per_s": round(F / wb, 1), "ms_per_batch": round(wb * 1e3, 3),
                           "verified": int(got["good"]), "verified_bytes_are_sent": bool(right.sum() == got["good"])}
    blind_pipe["frames_per_s"] = max(blind_pipe["find_on_threads"]["frames_per_s"],
                                     blind_pipe["find_batch"]["frames_per_s"])
    blind_pipe["what"] = ("the frames' registrations found (find_on_threads: lizx_find and lizx_read_word on host "
                          "threads, a decoder each; find_batch: one lizx_find_batch, its frames and their words on the "
                          "library's own threads and streams), then one batch graph "
                          "(lizx_batch) through them")
    batch.close()
    finder.close()
    for d in d_imgs + d_nodes:
        d.free()
    for d in decs:
        d.close()
    st = np.mean(np.array(stages), axis=0) if stages else np.zeros(6)
    out = {"workload": "16 captures 1920x1080 of LIZARD-1024 (n=1536, 256 ring) painted whole (border, word, anchors, "
                       "guard) on the GPU: rotation +-10 deg, scale 0.6-0.9, perspective <= 0.05, k1 +-0.05, blur "
                       "0.5-1.5 px, AWGN 2-8 grey",
           "found": found, "frames": len(frames), "find_ms_per_frame": round(find_ms, 3),
           "find_stage_ms": {k: round(float(v), 3) for k, v in zip(("upload", "vote_peaks", "sides", "probes",
                                                                       "rings", "refine"), st)},
           "node_err_px_mean": round(float(np.mean(errs)), 4) if errs else None,
           "node_err_px_worst_frame": round(float(np.max(errs)), 4) if errs else None,
           "k1_err_max": round(float(np.max(k1err)), 5) if k1err else None,
           "blind_decode_ms_per_frame": round(blind_ms, 3), "blind_verified": int(kb),
           "held_truth_decode_ms_per_frame": round(held_ms, 3), "held_truth_verified": int(kt),
           "blind_decode_threaded": threaded,
           "blind_batch": blind_pipe,
           "blocks": 124 * len(frames)}
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libref_x86.so")
    if os.path.exists(ref_so) and not args.no_cpu:
        lib = C.CDLL(ref_so)
        lib.ref_format_new.restype = C.c_void_p
        lib.ref_format_new.argtypes = [C.c_int, C.c_int, C.c_int]
        lib.ref_format_free.argtypes = [C.c_void_p]
        fn = lib.ref_acquire
        fn.restype = C.c_int
        fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int] + [C.c_void_p] * 4
        fmt = lib.ref_format_new(1536, 1024, span)
        nx = len(node_x)
        H = np.zeros(9, np.float32)
        nodes = np.zeros(2 * nx * nx, np.float32)
        quad = np.zeros(8, np.float32)
        o = np.zeros(1, np.int32)
        rfound, rerr = 0, []
        t0 = time.perf_counter()
        for f in frames:
            img = f["img"]
            if fn(fmt, img.ctypes.data, img.shape[1], img.shape[0], H.ctypes.data, nodes.ctypes.data,
                  quad.ctypes.data, o.ctypes.data):
                rfound += 1
                rerr.append(float(np.hypot(*(place(H, nodes) - place(f["H"], f["nodes"]))).mean()))
        cpu_ms = (time.perf_counter() - t0) * 1e3 / len(frames)
        # the reference's whole blind decode of a frame: its finder, then lizard_finish through what it found
        dr = lib.ref_decode_reg
        dr.restype = C.c_int
        dr.argtypes = [C.c_void_p] * 2 + [C.c_int] * 2 + [C.c_void_p] * 8
        blk = np.zeros(124 * 476, np.uint8)
        ok, its, be = np.zeros(124, np.uint8), np.zeros(124, np.int8), np.zeros(124, np.float32)
        rver = 0
        t0 = time.perf_counter()
        for f in frames:
            img = f["img"]
            if fn(fmt, img.ctypes.data, img.shape[1], img.shape[0], H.ctypes.data, nodes.ctypes.data,
                  quad.ctypes.data, o.ctypes.data):
                rver += dr(fmt, img.ctypes.data, img.shape[1], img.shape[0], H.ctypes.data, nodes.ctypes.data,
                           blk.ctypes.data, ok.ctypes.data, its.ctypes.data, be.ctypes.data, None, None)
        blind_cpu_ms = (time.perf_counter() - t0) * 1e3 / len(frames)
        lib.ref_format_free(fmt)
        out["reference_blind_decode"] = {"ms_per_frame": round(blind_cpu_ms, 3), "verified": int(rver), "cores": 1,
                                         "what": "lizard_acquire then lizard_finish through its registration, one "
                                                 "thread (x86-64 SIMD build)"}

        # and on the same host threads, one format instance (its own workspace) a thread
        def ref_job(k):
            fm = lib.ref_format_new(1536, 1024, span)
            Hk, nk, qk, ok_ = (np.zeros(9, np.float32), np.zeros(2 * nx * nx, np.float32), np.zeros(8, np.float32),
                               np.zeros(1, np.int32))
            bk, okk, itk, bek = (np.zeros(124 * 476, np.uint8), np.zeros(124, np.uint8), np.zeros(124, np.int8),
                                 np.zeros(124, np.float32))
            done = 0
            for i in range(k, len(jobs), nthreads):
                img = jobs[i][1]["img"]
                if fn(fm, img.ctypes.data, img.shape[1], img.shape[0], Hk.ctypes.data, nk.ctypes.data,
                      qk.ctypes.data, ok_.ctypes.data):
                    dr(fm, img.ctypes.data, img.shape[1], img.shape[0], Hk.ctypes.data, nk.ctypes.data,
                       bk.ctypes.data, okk.ctypes.data, itk.ctypes.data, bek.ctypes.data, None, None)
                done += 1
            lib.ref_format_free(fm)
            return done

        with cf.ThreadPoolExecutor(nthreads) as ex:
            t0 = time.perf_counter()
            nref = sum(ex.map(ref_job, range(nthreads)))
            tr = time.perf_counter() - t0
        out["reference_blind_decode"]["threaded_frames_per_s"] = round(nref / tr, 1)
        out["reference_blind_decode"]["threads"] = nthreads
        out["blind_decode_over_reference"] = round(blind_cpu_ms / blind_ms, 2)
        out["cpu_baseline"] = {"value": round(cpu_ms, 3), "unit": "ms/frame", "cores": 1, "kind": "reference",
                               "found": rfound,
                               "node_err_px_mean": round(float(np.mean(rerr)), 4) if rerr else None,
                               "sample": "the same 16 captures, one thread: liblizard lizard_acquire (the thin frame's "
                                         "finder, corner nodes, anchors, border mesh; x86-64 SIMD build)"}
    return out


B3_OPS_PER_COMPRESSION = 7 * 8 * 14 + 8  # 7 rounds of 8 G (6 adds, 4 xors, 4 rotates), the final 8 xors


def bench_blake3(args):
    """The transfer's chunk check on the GPU: BLAKE3 chaining values of every 1 MiB chunk of a 1 GiB file resident in
    HBM (lizx_b3_chunk_cvs), and the whole file's hash; the reference's BLAKE3 (xfer_chunk_cv over the vendored
    official C with its AVX2 kernel) on the host's cores beside it."""
    length, k = 1 << 30, 20
    rng = np.random.default_rng(33)
    data = rng.integers(0, 256, length, dtype=np.uint8)
    d = lizard_amd.DeviceArray.of(data)
    chunks = length >> k
    dc = lizard_amd.DeviceArray(32 * chunks)
    dec = lizard_amd.Decoder(0)
    for _ in range(2):
        dec.b3_chunk_cvs(d, length, k, d_cvs=dc)
    ms = []
    for _ in range(5):
        _, t = dec.b3_chunk_cvs(d, length, k, d_cvs=dc, timed=True)
        ms.append(t)
    kms = float(np.median(ms))
    cvs = dc.get(np.uint8, 32 * chunks).reshape(chunks, 32)
    root, hms = dec.b3_hash_device(d, length, timed=True)
    same_root = lizard_amd.b3_root(cvs) == root
    dec.close()
    dc.free()
    d.free()
    compressions = length // 64 + (length // 1024 - chunks)  # leaves' blocks, then parents under each chunk
    ops = compressions * B3_OPS_PER_COMPRESSION
    out = {"workload": "BLAKE3 chaining values of every 1 MiB chunk of a 1 GiB file resident in HBM (the transfer's "
                       "chunk check, xfer_chunk_cv over the file), then the file's root",
           "GBps": round(length / (kms * 1e-3) / 1e9, 2), "kernel_ms": round(kms, 4),
           "hash_kernel_ms": round(hms, 4), "root_from_chunks_equals_hash": bool(same_root),
           "roofline": {"bound": "valu", "achieved": round(ops / (kms * 1e-3) / 1e12, 2), "peak": VALU32_PEAK_TOPS,
                        "unit": "TOP/s", "frac": round(ops / (kms * 1e-3) / 1e12 / VALU32_PEAK_TOPS, 4),
                        "ops_per_compression": B3_OPS_PER_COMPRESSION, "compressions": compressions,
                        "hbm_GBps": round(length / (kms * 1e-3) / 1e9, 1), "hbm_peak_GBps": HBM_PEAK_GBS}}
    xfer_so = os.path.join(ROOT, "oracle", "_ref", "libref_xfer.so")
    if os.path.exists(xfer_so) and not args.no_cpu:
        import concurrent.futures as cf
        lib = C.CDLL(xfer_so)
        fn = lib.xfer_chunk_cv
        fn.restype = C.c_int
        fn.argtypes = [C.c_void_p, C.c_size_t, C.c_uint32, C.c_int, C.c_void_p]
        threads = min(16, os.cpu_count() or 1)
        sample = min(chunks, 16 * threads)  # chunks of the same file
        outb = np.zeros((sample, 32), np.uint8)

        def work(i):
            fn(data.ctypes.data + (i << k), 1 << k, i, k, outb[i].ctypes.data)

        with cf.ThreadPoolExecutor(threads) as ex:
            list(ex.map(work, range(threads)))
            t0 = time.perf_counter()
            list(ex.map(work, range(sample)))
            wall = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(sample * (1 << k) / wall / 1e9, 3), "unit": "GB/s", "cores": threads,
                               "kind": "reference", "same_values": bool(np.array_equal(outb, cvs[:sample])),
                               "sample": "%d of the file's 1 MiB chunks on %d threads (ctypes releases the GIL): "
                                         "xfer_chunk_cv over the vendored BLAKE3 C with its AVX2 kernel" % (sample,
                                                                                                          threads)}
    return out


def bench_refine(args):
    """Decision-directed registration refinement: decode through the perturbed registrations (one batch), re-modulate
    the CRC-passed blocks, 10 Gauss-Newton steps a frame on the GPU, decode again."""
    from lizard_amd import sim
    t0 = time.perf_counter()
    frames = refine_frames(args.refine_frames, 777)
    gen_s = time.perf_counter() - t0
    F = len(frames)
    D = lizard_amd.DeviceArray
    dec = lizard_amd.Decoder(0)
    keep = []

    def decode(Hs, node_list=None):
        items = []
        for k, (fr, H) in enumerate(zip(frames, Hs)):
            if node_list is None:
                p = lizard_amd.reg_params(fr["H0"], fr["k1"])
                nodes = lizard_amd.reg_nodes(256, fr["H0"].astype(np.float32), fr["img"].shape[1], fr["img"].shape[0],
                                             p)
            else:
                nodes = node_list[k]
            d_img = fr.get("d_img") or D.of(fr["img"])
            fr["d_img"] = d_img
            d_nodes = D.of(nodes)
            keep.append(d_nodes)
            items.append(dict(img=d_img, w=fr["img"].shape[1], h=fr["img"].shape[0], H=H, nodes=d_nodes))
        b = lizard_amd.Batch(0).plan([20] * F, [256] * F)
        b.run(lizard_amd.make_frames(items), lizard_amd.RUN_GRAPH)
        b.sync()
        out = b.blocks_out()
        b.close()
        return out

    before = decode([fr["H0"] for fr in frames])
    B = lizard_amd.format_info(20)["blocks"]
    sent = np.concatenate([fr["blocks"].reshape(-1, 476) for fr in frames])
    ok0 = before["ok"].reshape(F, B).astype(bool)
    t0 = time.perf_counter()
    Hs, traces = [], []
    ref_s = 0.0
    for k, fr in enumerate(frames):
        r0 = time.perf_counter()
        ref_pic = sim.remodulate(20, fr["blocks"], ok0[k], 0, fr["blur"])
        ref_s += time.perf_counter() - r0
        p = lizard_amd.reg_params(fr["H0"], fr["k1"])
        p1, tr = dec.reg_refine(20, 256, fr["img"], fr["H0"], p, ref_pic, steps=10)
        Hs.append(np.append(p1[:8], 1.0).reshape(3, 3))
        traces.append(tr)
    refine_s = time.perf_counter() - t0 - ref_s
    after = decode(Hs)
    right0 = (before["ok"] > 0) & np.all(before["blocks"] == sent, axis=1)
    right1 = (after["ok"] > 0) & np.all(after["blocks"] == sent, axis=1)
    tr = np.array(traces)
    # the lens unknown: the same perturbed homographies with no lens in the registration (k1 = 0), then the steps over
    # H alone and over H and k1 together (12 each), on the re-modulation of what the lens-less registration verified
    def nodes_k1(fr, k1):
        p = lizard_amd.reg_params(fr["H0"], k1)
        return lizard_amd.reg_nodes(256, fr["H0"].astype(np.float32), fr["img"].shape[1], fr["img"].shape[0], p)

    nl_nodes = [nodes_k1(fr, 0.0) for fr in frames]
    nl0 = decode([fr["H0"] for fr in frames], nl_nodes)

    def fail(out):
        return round(1.0 - float(((out["ok"] > 0) & np.all(out["blocks"] == sent, axis=1)).mean()), 5)

    lens_unknown = {"block_fail_before": fail(nl0)}
    spec_s = 0.0
    for free_k1, name, spectral in ((False, "H", False), (True, "H_and_k1", False), (True, "H_and_k1_spectral", True)):
        # decision-directed rounds: each refines from where the last left off, on the re-modulation of what the last
        # decode verified (the spectral loss: on those blocks' and the padding's coefficients alone)
        ps = [lizard_amd.reg_params(fr["H0"], 0.0) for fr in frames]
        out, path = nl0, []
        for rnd in range(3):
            ok = out["ok"].reshape(F, B).astype(bool)
            for k, fr in enumerate(frames):
                ref_pic = sim.remodulate(20, fr["blocks"], ok[k], 0, fr["blur"])
                if spectral:
                    t1 = time.perf_counter()
                    ps[k], _ = dec.reg_refine_spectral(20, 256, fr["img"], fr["H0"], ps[k], ref_pic,
                                                       sim.known_mask(20, ok[k]), steps=8)
                    spec_s += time.perf_counter() - t1
                else:
                    ps[k], _ = dec.reg_refine(20, 256, fr["img"], fr["H0"], ps[k], ref_pic, steps=8, k1=free_k1)
            out = decode([np.append(p[:8], 1.0).reshape(3, 3) for p in ps],
                         [nodes_k1(fr, float(p[8])) for fr, p in zip(frames, ps)])
            path.append(fail(out))
        lens_unknown["block_fail_after_" + name] = path
        if free_k1:
            err = [abs(float(p[8]) - fr["k1"]) for fr, p in zip(frames, ps)]
            sfx = "_spectral" if spectral else ""
            lens_unknown["k1_abs_err_median" + sfx] = round(float(np.median(err)), 5)
            lens_unknown["k1_abs_err_max" + sfx] = round(float(np.max(err)), 5)
    lens_unknown["spectral_ms_per_frame_round"] = round(spec_s / (3 * F) * 1e3, 3)
    lens_unknown["what"] = ("the registration with no lens (k1 = 0; the cameras' k1 in [-0.05, 0.05]) and the corners "
                            "1.5 px off, then three decision-directed rounds (re-modulate what the last decode "
                            "verified, 8 Gauss-Newton steps, decode) over H alone (lizx_reg_refine) or over H and k1 "
                            "together (lizx_reg_refine_k1), or over H and k1 on the loss restricted to the verified "
                            "blocks' coefficients through the FFT's adjoint (lizx_reg_refine_spectral, 8 steps with "
                            "a line search); block-fail after each round")
    two_terms = refine_two_terms(dec, min(F, 64), B)
    # the plain anchor registration (the finder on each capture, from the anchors outward: over its own border nodes,
    # and the reference's chain with the border refined by its templates) and the same Gauss-Newton steps from it, the
    # nodes held, on the re-modulation of what it verified
    anchor_legs = {}
    for mode, name in ((1, "anchor_registration"), (2, "anchor_registration_reference_chain")):
        dec.anchors(mode)
        Ha, Na, lost = [], [], 0
        for fr in frames:
            f = dec.find(fr["img"])
            if f is None or f["span"] != 256:
                lost += 1
                Ha.append(fr["H0"])
                Na.append(np.zeros((len(lizard_amd.node_lattice(256)),) * 2 + (2,), np.float32))
                continue
            Ha.append(np.asarray(f["H"], np.float64).reshape(3, 3))
            Na.append(f["nodes"])
        anch = decode(Ha, Na)
        ok2 = anch["ok"].reshape(F, B).astype(bool)
        Hr = []
        for k, fr in enumerate(frames):
            ref_pic = sim.remodulate(20, fr["blocks"], ok2[k], 0, fr["blur"])
            H1, _ = dec.reg_refine_nodes(20, 256, fr["img"], Ha[k], Na[k], ref_pic, steps=10)
            Hr.append(H1.reshape(3, 3).astype(np.float64))
        anch_ref = decode(Hr, Na)
        right2 = (anch["ok"] > 0) & np.all(anch["blocks"] == sent, axis=1)
        right3 = (anch_ref["ok"] > 0) & np.all(anch_ref["blocks"] == sent, axis=1)
        anchor_legs[name] = {"block_fail": round(1.0 - float(right2.mean()), 5),
                             "block_fail_after_refinement": round(1.0 - float(right3.mean()), 5),
                             "frames_not_found": lost}
    dec.anchors(1)
    anchor_legs["note"] = ("the finder's own registration of each capture (border fit, 7 x 7 anchors registered "
                           "outward, interior filled from the border and the anchors; the reference chain's border "
                           "from the nodes' templates), then 10 Gauss-Newton steps over H from it with the nodes "
                           "held, on the re-modulation of the blocks it verified")
    dec.close()
    for d in keep:
        d.free()
    for fr in frames:
        fr["d_img"].free()
    return {"workload": "%d captured 1920x1080 frames of LIZARD-320 (n=768, 39 blocks, 128 ring), random camera "
                        "(rot +-10 deg, scale 0.75-0.9, persp <= 0.05, k1 in [-0.05, 0.05], blur 0.5-1.5 px, AWGN "
                        "2-8 grey); initial homography corners up to 1.5 px off; loss = L2 of the warped capture "
                        "against the blurred re-modulation of the CRC-passed blocks; 10 Gauss-Newton steps over the "
                        "homography on the GPU (FP64 8x8 solve)" % F,
            "blocks": int(B * F), "block_fail_before": round(1.0 - float(right0.mean()), 5),
            "block_fail_after": round(1.0 - float(right1.mean()), 5),
            "frames_improved": int(np.sum(right1.reshape(F, B).sum(1) > right0.reshape(F, B).sum(1))),
            "frames_worse": int(np.sum(right1.reshape(F, B).sum(1) < right0.reshape(F, B).sum(1))),
            "loss_ratio_median": round(float(np.median(tr[:, -1] / tr[:, 0])), 4),
            "from_anchor_registration": anchor_legs,
            "lens_unknown": lens_unknown,
            "lens_two_terms": two_terms,
            "refine_ms_per_frame": round(refine_s * 1e3 / F, 3),
            "generation_s": round(gen_s, 1)}


def cpu_pipeline(caps_host, seconds):
    """The reference decoder's finish path (sampler, erase, FFT, soft values, LDPC, CRC; x86-64 build) through the
    same registrations, one format instance per thread."""
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libref_x86.so")
    if not os.path.exists(ref_so):
        return None
    lib = C.CDLL(ref_so)
    lib.ref_format_new.restype = C.c_void_p
    lib.ref_format_new.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.ref_format_free.argtypes = [C.c_void_p]
    fn = lib.ref_decode_reg
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p] * 2 + [C.c_int] * 2 + [C.c_void_p] * 8
    cores = host_threads()
    # untimed, first: the reference's verdicts on each capture against the GPU graph's, the soft values' rounding
    # ties, and the reference's integer path (lizard_finish_ext) on the GPU's own soft values and estimates
    lib.ref_finish_llr.restype = C.c_int
    lib.ref_finish_llr.argtypes = [C.c_void_p] * 7
    f = lib.ref_format_new(1536, 1024, 512)
    cmp = {"frames": 0, "verdicts_equal": 0, "blocks_differing": 0, "soft_values": 0, "soft_differing": 0,
           "finish_on_gpu_soft_equal": 0}
    for c in caps_host:
        if "gpu_soft" not in c:
            continue
        out = np.zeros(124 * 476, np.uint8)
        ok, its, be = np.zeros(124, np.uint8), np.zeros(124, np.int8), np.zeros(124, np.float32)
        llr = np.zeros(1024 * 640, np.int8)
        fn(f, c["img"].ctypes.data, c["img"].shape[1], c["img"].shape[0], c["H"].ctypes.data, c["nodes"].ctypes.data,
           out.ctypes.data, ok.ctypes.data, its.ctypes.data, be.ctypes.data, llr.ctypes.data, None)
        g = c["gpu_out"]
        same = np.array_equal(g["ok"], ok) and np.array_equal(g["its"], its)
        cmp["frames"] += 1
        cmp["verdicts_equal"] += int(same)
        cmp["blocks_differing"] += int(np.sum((g["ok"] != ok) | (g["its"] != its)))
        cmp["soft_values"] += llr.size
        cmp["soft_differing"] += int(np.sum(c["gpu_soft"]["llr"] != llr))
        est = np.ascontiguousarray(c["gpu_soft"]["sub_est"], np.float32)
        lib.ref_finish_llr(f, c["gpu_soft"]["llr"].ctypes.data, est.ctypes.data, out.ctypes.data, ok.ctypes.data,
                           its.ctypes.data, be.ctypes.data)
        v = ok > 0
        cmp["finish_on_gpu_soft_equal"] += int(np.array_equal(g["ok"], ok) and np.array_equal(g["its"], its) and
                                               np.array_equal(g["blocks"][v], out.reshape(124, 476)[v]))
    lib.ref_format_free(f)
    lock = threading.Lock()
    state = {"next": 0, "frames": 0, "good": 0}
    deadline = [0.0]

    def worker():
        f = lib.ref_format_new(1536, 1024, 512)
        out = np.zeros(124 * 476, np.uint8)
        ok, its, be = np.zeros(124, np.uint8), np.zeros(124, np.int8), np.zeros(124, np.float32)
        while True:
            with lock:
                if time.perf_counter() > deadline[0]:
                    break
                c = caps_host[state["next"] % len(caps_host)]
                state["next"] += 1
            good = fn(f, c["img"].ctypes.data, c["img"].shape[1], c["img"].shape[0], c["H"].ctypes.data,
                      c["nodes"].ctypes.data, out.ctypes.data, ok.ctypes.data, its.ctypes.data, be.ctypes.data, None,
                      None)
            with lock:
                state["frames"] += 1
                state["good"] += good
        lib.ref_format_free(f)

    t0 = time.perf_counter()
    deadline[0] = t0 + seconds
    th = [threading.Thread(target=worker) for _ in range(cores)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    wall = time.perf_counter() - t0
    return {"value": round(472.0 * state["good"] / wall / 1e6, 4), "unit": "MB/s payload", "cores": cores,
            "host_cpus": host_cpus(), "kind": "reference", "frames_per_s": round(state["frames"] / wall, 2),
            "verdicts_equal_reference": {
                "frames_equal": "%d of %d" % (cmp["verdicts_equal"], cmp["frames"]),
                "blocks_differing": cmp["blocks_differing"],
                "soft_tie_fraction": round(cmp["soft_differing"] / max(1, cmp["soft_values"]), 8),
                "reference_finish_on_gpu_soft_values_equal": "%d of %d" % (cmp["finish_on_gpu_soft_equal"],
                                                                          cmp["frames"])},
            "sample": "%d frames (the same %d captures cycled), %.1f s wall on %d threads; liblizard lizard_finish "
                      "through the held registration (x86-64 SIMD build)" % (state["frames"], len(caps_host), wall,
                                                                            cores)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--pipe-frames", type=int, default=32)
    ap.add_argument("--pipe-runs", type=int, default=10)
    ap.add_argument("--fft-frames", type=int, default=64)
    ap.add_argument("--no-pipeline", action="store_true")
    ap.add_argument("--refine-frames", type=int, default=256)
    ap.add_argument("--pipe-modes", default="graph,stream,frames,plain",
                    help="pipeline run modes to time (graph always): graph, stream, frames, plain")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuners (default launch shapes)")
    ap.add_argument("--small-eager", action="store_true",
                    help="small: the decoder's kernels launched one by one instead of replayed as a graph (the same "
                         "kernels and bits; for kernel traces)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="with --only: a 1 GiB device copy after the leg, whose counters calibrate a --pmc run")
    ap.add_argument("--op-batches", type=int, default=8,
                    help="4096-codeword batches at the block-fail < 1e-4 operating point")
    ap.add_argument("--only", choices=("ldpc", "pipeline", "fft_demap", "refine", "small", "sender", "finder",
                                       "blake3"), default=None,
                    help="run one workload alone (for per-kernel profiles); the JSON line then carries only it")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only; the path shards by frame and codeword with no exchange between "
              "GPUs, and no figure at several GPUs is produced", file=sys.stderr)
        sys.exit(2)
    if lizard_amd.device_count() < 1:
        print("bench.py: no GPU", file=sys.stderr)
        sys.exit(1)

    if args.only in ("pipeline", "fft_demap", "refine", "small", "sender", "finder", "blake3"):
        out = (bench_pipeline(args, pipeline_frames(args.pipe_frames, 4242)) if args.only == "pipeline"
               else bench_fft_demap(args) if args.only == "fft_demap" else bench_refine(args) if args.only == "refine"
               else bench_sender(args, lizard_amd.Decoder(0)) if args.only == "sender"
               else bench_finder(args) if args.only == "finder" else bench_blake3(args) if args.only == "blake3"
               else bench_small(args))
        if args.copy_probe:
            nb = 1 << 30
            a, b = lizard_amd.DeviceArray(nb), lizard_amd.DeviceArray(nb)
            out["copy_probe_ms"] = round(lizard_amd.copy_probe(b, a, nb), 4)
            a.free(), b.free()
        print(json.dumps({"only": args.only, args.only: out}))
        return
    if args.only == "ldpc":
        args.no_pipeline = True

    batches = [build_batch(db, seed=1000 + i) for i, db in enumerate(EBN0)]
    count = batches[0][0].size
    D = lizard_amd.DeviceArray
    dev = []
    for rates, llr, offs, _, _ in batches:
        dev.append((D.of(rates), D.of(llr), D.of(offs)))
    span = max(b[1].size for b in batches)
    d_bits = D(span)
    d_its = D(count)
    d_bytes = D(count * 476)
    d_ok = D(count)
    dec = lizard_amd.LdpcDecoder(0)

    def step(s, plain=False, bits=False):
        d_rt, d_llr, d_off = dev[s % len(dev)]
        dec.decode_device(count, d_rt, d_llr, d_off, d_bits if bits else None, d_its, max_iter=MAX_ITER, stall_it=0,
                          crc=True, plain=plain, d_bytes=d_bytes, d_crc_ok=d_ok)

    # verification pass (untimed): per point, iterations, CRC verdicts and decisions against what was sent
    points = []
    its_all = []
    verified_all = []
    for p in range(len(EBN0)):
        step(p, bits=True)
        dec.sync()
        its = d_its.get(np.int8, count)
        ok = d_ok.get(np.uint8, count)
        bits = d_bits.get(np.uint8, span)
        rates, llr, offs, payload, info = batches[p]
        right = np.array([bool(ok[i]) and np.array_equal(bits[offs[i]:offs[i] + 3840], info[i]) for i in range(count)])
        fails = {}
        for r, c in MIX:
            m = rates == r
            fails[{6: "7/8", 4: "3/4", 3: "2/3", 2: "1/2"}[r]] = round(float(1.0 - right[m].mean()), 6)
        points.append({"ebn0_db": EBN0[p], "block_fail": round(float(1.0 - right.mean()), 6), "block_fail_by_rate": fails,
                       "its_mean": round(float(effective_iterations(its).mean()), 3), "verified": int(ok.sum())})
        its_all.append(its)
        verified_all.append(int(ok.sum()))

    # the PHY's own setting beside it, before the timed steps (so those are the workload's last launches): 30
    # iterations with the stall rule (9, 0.95), CRC verdicts per point
    def step30(s):
        d_rt, d_llr, d_off = dev[s % len(dev)]
        dec.decode_device(count, d_rt, d_llr, d_off, None, d_its, max_iter=30, stall_it=9, stall_ratio=0.95, crc=True,
                          d_bytes=d_bytes, d_crc_ok=d_ok)

    ok30, outcome30 = [], []
    for p in range(len(EBN0)):
        step30(p)
        dec.sync()
        ok30.append(int(d_ok.get(np.uint8, count).sum()))
        it = d_its.get(np.int8, count).astype(np.int64)
        conv = it[it > 0]
        outcome30.append({"its_mean_converged": round(float(conv.mean()), 3) if conv.size else None,
                          "capped": int((it == -1).sum()), "stalled": int((it == -3).sum())})
    for s in range(2):
        step30(s)
    dec.sync()
    t0 = time.perf_counter()
    for s in range(args.steps):
        step30(s)
    dec.sync()
    w30 = time.perf_counter() - t0
    stall = {"max_iter": 30, "stall_it": 9, "stall_ratio": 0.95,
             "Gbps": round(3840.0 * count * args.steps / w30 / 1e9, 3),
             "verified_by_point": ok30, "by_point": outcome30}

    # the same decode from host buffers (lizx_ldpc_decode: the LLRs in and the decisions out over PCIe each call),
    # before the timed region (whose launches are the last of the run), one call a point; never the headline value
    dec.decode(batches[0][0], batches[0][1], batches[0][2], max_iter=MAX_ITER, stall_it=0)
    t0 = time.perf_counter()
    for p in range(len(EBN0)):
        dec.decode(batches[p][0], batches[p][1], batches[p][2], max_iter=MAX_ITER, stall_it=0)
    wh = time.perf_counter() - t0
    host_buffers = {"Gbps": round(3840.0 * count * len(EBN0) / wh / 1e9, 3),
                    "ms_per_step": round(wh / len(EBN0) * 1e3, 4),
                    "what": "lizx_ldpc_decode with host buffers: each call copies the LLRs in and the decisions out"}

    # the operating point where block-fail < 1e-4 is shown (before the timed steps): op_batches x 4096 codewords of
    # the same rate mix at OP_EBN0, every decision checked against the bits sent, then the batches decoded back to
    # back and timed
    op = None
    if args.op_batches > 0:
        opb = [build_batch(OP_EBN0, seed=5000 + i) for i in range(args.op_batches)]
        d_op = [(D.of(r), D.of(l), D.of(o)) for r, l, o, _, _ in opb]
        op_fail = op_ver = 0
        op_its = []
        for k, (rates, llr, offs, payload, info) in enumerate(opb):
            d_rt, d_llr, d_off = d_op[k]
            dec.decode_device(count, d_rt, d_llr, d_off, d_bits, d_its, max_iter=MAX_ITER, stall_it=0, crc=True,
                              d_bytes=d_bytes, d_crc_ok=d_ok)
            dec.sync()
            ok = d_ok.get(np.uint8, count)
            bits = d_bits.get(np.uint8, span)
            right = np.array([bool(ok[i]) and np.array_equal(bits[offs[i]:offs[i] + 3840], info[i])
                              for i in range(count)])
            op_fail += int(count - right.sum())
            op_ver += int(ok.sum())
            op_its.append(effective_iterations(d_its.get(np.int8, count)))
        for k in range(len(opb)):
            d_rt, d_llr, d_off = d_op[k]
            dec.decode_device(count, d_rt, d_llr, d_off, None, d_its, max_iter=MAX_ITER, stall_it=0, crc=True,
                              d_bytes=d_bytes, d_crc_ok=d_ok)
        dec.sync()
        t0 = time.perf_counter()
        for k in range(len(opb)):
            d_rt, d_llr, d_off = d_op[k]
            dec.decode_device(count, d_rt, d_llr, d_off, None, d_its, max_iter=MAX_ITER, stall_it=0, crc=True,
                              d_bytes=d_bytes, d_crc_ok=d_ok)
        op_ms = dec.sync()
        op_wall = time.perf_counter() - t0
        words = count * len(opb)
        op = {"ebn0_db": OP_EBN0, "codewords": words, "block_failures": op_fail,
              "block_fail": op_fail / words,
              "block_fail_upper_95": round(3.0 / words, 7) if op_fail == 0 else None,
              "below_1e-4": bool(op_fail == 0 and 3.0 / words < 1e-4),
              "its_mean": round(float(np.concatenate(op_its).mean()), 3),
              "Gbps": round(3840.0 * words / op_wall / 1e9, 3),
              "payload_MBps": round(472.0 * op_ver / op_wall / 1e6, 2),
              "gpu_ms_per_batch": round(op_ms / len(opb), 4),
              "what": "the workload's rate mix at one Eb/N0, every decision checked against the bits sent; zero "
                      "failures in N codewords bound the rate below 3/N (95 %)"}
        for t in d_op:
            for a in t:
                a.free()

    # the autotuner picks the persistent grid's size on the first Eb/N0 point's batch (the bits do not depend on it)
    tuned_per_cu = dec.tune(count, dev[0][0], dev[0][1], dev[0][2], d_its, max_iter=MAX_ITER, stall_it=0, reps=3)
    for s in range(args.warmup):
        step(s)
    dec.sync()
    t0 = time.perf_counter()
    for s in range(args.steps):
        step(s)
    ev_ms = dec.sync()
    wall = time.perf_counter() - t0

    codewords = count * args.steps
    verified = sum(verified_all[s % len(EBN0)] for s in range(args.steps))
    gbps = 3840.0 * codewords / wall / 1e9

    payload_mbps = 472.0 * verified / wall / 1e6

    # rooflines of the decode kernel over the iterations actually run, per launch: integer operations (the bound:
    # the messages live in registers, so the kernel issues vector instructions rather than moving LDS bytes) and the
    # algorithmic LDS bytes of the reference layout
    edge_its, lds_per_launch = [], []
    for p in range(len(EBN0)):
        rates = batches[p][0]
        edges = np.array([lizard_amd.ldpc_dims(int(r))["edges"] for r in rates], np.int64)
        eff = effective_iterations(its_all[p])
        edge_its.append(float((edges * eff).sum()))
        lds_per_launch.append(float((6 * edges * eff).sum()))
    lds_bytes = sum(lds_per_launch[s % len(EBN0)] for s in range(args.steps)) / args.steps
    ops = OPS_PER_EDGE * sum(edge_its[s % len(EBN0)] for s in range(args.steps)) / args.steps
    launch_ms = ev_ms / args.steps
    lds_gbs = lds_bytes / (launch_ms * 1e-3) / 1e9
    tops = ops / (launch_ms * 1e-3) / 1e12
    hbm_bytes = float(sum(b[1].size for b in batches) / len(batches) + count * (476 + 2))

    plain = None
    if not args.no_plain:
        t0 = time.perf_counter()
        for s in range(len(EBN0)):
            step(s, plain=True)
        dec.sync()
        pw = time.perf_counter() - t0
        plain = 3840.0 * count * len(EBN0) / pw / 1e9

    # the practical HBM copy rate (the build's own stream-copy probe), for reference
    nb = 1 << 30
    a, b = D(nb), D(nb)
    lizard_amd.copy_probe(b, a, nb)
    copy_ms = min(lizard_amd.copy_probe(b, a, nb) for _ in range(3))
    copy_gbs = 2.0 * nb / (copy_ms * 1e-3) / 1e9
    a.free(), b.free()

    cpu = None if args.no_cpu else cpu_baseline(batches, args.cpu_seconds)

    pipe = fft = refine = small = sender = finder = blake3 = None
    if not args.no_pipeline:
        small = bench_small(args)
        finder = bench_finder(args)
        blake3 = bench_blake3(args)
        sdec = lizard_amd.Decoder(0)
        sender = bench_sender(args, sdec)
        sdec.close()
        refine = bench_refine(args)
        caps = pipeline_frames(args.pipe_frames, 4242)
        pipe = bench_pipeline(args, caps)
        fft = bench_fft_demap(args)
        if not args.no_cpu:
            host = caps[:8]  # the reference decoder cycles over the first captures of the same workload
            pipe["cpu_baseline"] = cpu_pipeline(host, args.cpu_seconds)
            if pipe["cpu_baseline"] and pipe["cpu_baseline"]["value"] > 0:
                pipe["graph_over_cpu"] = round(pipe["modes"]["graph"]["payload_MBps"] / pipe["cpu_baseline"]["value"], 2)

    line = {
        "metric": METRIC, "value": round(gbps, 3), "unit": "Gbit/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int8 messages / int16 posteriors", "data": "synthetic",
        "config": {"workload": "LDPC only, batched: 4096 codewords x 3840 info bits, rates 7/8:3/4:2/3:1/2 = "
                               "1597:1065:614:820, BPSK/AWGN Eb/N0 1.5-4.0 dB (one batch a point, steps cycling), "
                               "int8 LLR = rhe(8 * 2y/sigma^2) clamped +-80, layered min-sum max 20 it, early "
                               "termination, one persistent kernel, CRC-32C on the GPU",
                   "codewords": count, "ebn0_db": list(EBN0), "max_iter": MAX_ITER},
        "payload_MBps": round(payload_mbps, 2),
        "operating_point": op,
        "autotune": {"ldpc_wg_per_cu": tuned_per_cu or "max"},
        "stall_rule": stall,
        "host_buffers": host_buffers,
        # an extra, labelled as such: the same batches with one LDPC launch a codeword (no product entry decodes so)
        "per_codeword_launches_Gbps": round(plain, 4) if plain else None,
        "batched_over_per_codeword_launches": round(gbps / plain, 2) if plain else None,
        "points": points,
        "roofline": {"bound": "valu", "kernel": "k_ldpc", "achieved": round(tops, 2), "peak": VALU_PEAK_TOPS,
                     "unit": "TOP/s", "frac": round(tops / VALU_PEAK_TOPS, 4), "traffic": measured_traffic(),
                     "traffic_from": "profiles/k_ldpc_traffic_ldpc_batched_4096.txt: two rocprofv3 --pmc runs of "
                                     "this leg (FETCH_SIZE, WRITE_SIZE), copied, not measured by this run",
                     "issue_from": "profiles/k_ldpc_issue_ldpc_batched_4096.txt: one rocprofv3 --pmc run of SQ "
                                   "counters, copied, not measured by this run",
                     "launch_ms": round(launch_ms, 4), "ops_per_launch": ops, "ops_per_edge": OPS_PER_EDGE,
                     "issue": measured_issue(),
                     "lds": {"achieved": round(lds_gbs, 1), "peak": LDS_PEAK_GBS, "unit": "GB/s",
                             "frac": round(lds_gbs / LDS_PEAK_GBS, 4), "bytes_per_launch": lds_bytes},
                     "iterations_per_s": round(sum(float(effective_iterations(its_all[s % len(EBN0)]).sum())
                                                   for s in range(args.steps)) / (ev_ms * 1e-3), 1),
                     "hbm": {"achieved": round(hbm_bytes / (launch_ms * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS,
                             "copy_probe_GBps": round(copy_gbs, 1), "algorithmic_bytes_per_launch": hbm_bytes}},
        "cpu_baseline": cpu,
        "pipeline": pipe,
        "fft_demap": fft,
        "refine": refine,
        "small": small,
        "sender": sender,
        "finder": finder,
        "blake3": blake3,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
