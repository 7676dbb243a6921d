// Batches of frames: many captures (or picture grids) of mixed formats decoded by one launch sequence, captured once
// as a HIP graph and replayed. The reference batches whole frames per receiver (liblizard/include/lizard.h:285-287,
// liblizard/core/rx/receiver.h:77-95) and decodes each on its own CPU worker; here the frames of one format share
// each stage's launch (frames along the grid's y axis) and every frame's blocks share one persistent LDPC launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <vector>

#include "common.hpp"
#include "format.hpp"
#include "ldpc.hpp"
#include "spectrum.hpp"

using namespace lz;

namespace {

// The frames of one format, in the batch's internal order (contiguous), with their work areas.
struct Group {
  std::unique_ptr<FramePlan> plan;
  std::vector<int> frames;  // caller indices
  int first = 0;            // the group's first internal index
  FrameBufs fb;             // X, coefficients, row sums, shading, projections, shift, rms for its frames
  SoftBufs view;            // fb's pointers with llr and est pointing into the batch-wide arrays
};

}  // namespace

struct lizx_batch {
  std::unique_ptr<LdpcEngine> eng;
  int frames = 0, source = 0, stages = 0, blocks = 0;
  std::vector<std::unique_ptr<Group>> groups;
  std::vector<int> group_of, slot_of;  // caller frame -> group, index within the group
  std::vector<int64_t> llr_off, est_off, blk_off;  // caller frame -> offsets (soft bytes, estimates, blocks)
  int64_t llr_total = 0, est_total = 0;
  // device
  DevBuf wframes, grids, llr, est, tasks, frame_tasks, its, blk_est, bytes, crc_ok;
  std::vector<int> frame_task0;  // caller frame -> first task of its own list in frame_tasks
  std::vector<WarpFrame> host_frames, sent_frames;  // this run's descriptors, and those on the device
  std::vector<const float *> host_grids, sent_grids;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint8_t> out_bytes;
  // launch shapes (the autotuner's choices): waves a workgroup of the row FFT, column FFT and demap; the persistent
  // LDPC grid's workgroups a CU (0: as many as fit)
  LaunchShape shape{};
  int ldpc_per_cu = 0;
  // one more stream a format beyond the first, so the formats' soft-value launches overlap (default priority)
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;

  ~lizx_batch() {
    drop_graph();
    for (auto e : join) (void)hipEventDestroy(e);
    for (auto st : side) (void)hipStreamDestroy(st);
    if (fork) (void)hipEventDestroy(fork);
  }
  void ensure_streams(int count) {
    if (!fork) hip_check(hipEventCreateWithFlags(&fork, hipEventDisableTiming), "hipEventCreate");
    while (static_cast<int>(side.size()) < count) {
      hipStream_t st = nullptr;
      hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
      side.push_back(st);
      hipEvent_t e = nullptr;
      hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
      join.push_back(e);
    }
  }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }

  LdpcParams ldpc_params() const {
    LdpcEngine &e = *eng;
    LdpcParams p{};
    p.llr = llr.as<int8_t>();
    p.its = its.as<int8_t>();
    p.blk_est = blk_est.as<float>();
    p.bytes = bytes.as<uint8_t>();
    p.crc_ok = crc_ok.as<uint8_t>();
    p.white = e.white.as<uint8_t>();
    for (int ci = 0; ci < NCODES; ci++) p.s2p[ci] = e.s2p[ci].as<int16_t>();
    p.est = est.as<double>();
    p.crc_tab = e.crc_tab.as<uint32_t>();
    p.crc0 = e.crc0;
    p.max_iter = 30;  // lizard_phy.c:1824 with LIZARD_STALL_IT, LIZARD_STALL_RATIO (:81-82)
    p.stall_it = 9;
    p.stall_ratio = 0.95f;
    p.mode = MODE_FRAME;
    p.flags = 1;
    return p;
  }

  // The soft-value launches of frames [i0, i0 + count) of group g (internal indices within the group) on stream s.
  void soft(const Group &g, int i0, int count, hipStream_t s) const {
    const FramePlan &p = *g.plan;
    const SoftBufs v = g.view.at(p, i0);
    if (source == LIZX_SRC_GRID)
      launch_soft_grids(p, v, grids.as<const float *>() + g.first + i0, count, s, shape);
    else
      launch_soft_warp(p, v, wframes.as<WarpFrame>() + g.first + i0, count, s, shape);
  }

  // The whole batch's launches, in stream order (what the graph holds).
  void enqueue_batch(bool timed) const {
    const hipStream_t s = eng->stream;
    if (groups.size() == 1) {
      soft(*groups[0], 0, static_cast<int>(groups[0]->frames.size()), s);
    } else {
      // fork: each format on its own stream, joined back before the LDPC (a graph captures the same dependencies);
      // the groups with the most work (frames x n^2) launched first, the longest branch starting soonest
      std::vector<size_t> order(groups.size());
      for (size_t g = 0; g < order.size(); g++) order[g] = g;
      const auto work = [&](size_t g) {
        return static_cast<double>(groups[g]->frames.size()) * groups[g]->plan->n * groups[g]->plan->n;
      };
      std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return work(a) > work(b); });
      hip_check(hipEventRecord(fork, s), "hipEventRecord");
      for (const size_t g : order) {
        hip_check(hipStreamWaitEvent(side[g], fork, 0), "hipStreamWaitEvent");
        soft(*groups[g], 0, static_cast<int>(groups[g]->frames.size()), side[g]);
        hip_check(hipEventRecord(join[g], side[g]), "hipEventRecord");
        hip_check(hipStreamWaitEvent(s, join[g], 0), "hipStreamWaitEvent");
      }
    }
    if (stages == LIZX_STAGE_BLOCKS) {
      LdpcParams p = ldpc_params();
      p.tasks = tasks.as<LdpcTask>();
      p.count = blocks;
      eng->launch(p, false, timed, false, ldpc_per_cu);
    }
  }

  // One frame's launches (caller index f).
  void enqueue_frame(int f, bool plain) const {
    const Group &g = *groups[group_of[f]];
    soft(g, slot_of[f], 1, eng->stream);
    if (stages == LIZX_STAGE_BLOCKS) {
      LdpcParams p = ldpc_params();
      p.tasks = frame_tasks.as<LdpcTask>() + frame_task0[f];
      p.count = frame_task0[f + 1] - frame_task0[f];
      eng->launch(p, plain, true, false, ldpc_per_cu);
    }
  }
};

extern "C" {

LIZ_API lizx_batch *lizx_batch_new(int device) {
  return guard_new<lizx_batch>([&] {
    const int dev = current_device(device);
    auto b = std::make_unique<lizx_batch>();
    b->eng = std::make_unique<LdpcEngine>(dev);
    return b.release();
  });
}

LIZ_API void lizx_batch_free(lizx_batch *b) {
  if (b) {
    (void)hipSetDevice(b->eng->device);
    delete b;
  }
}

LIZ_API int lizx_batch_plan(lizx_batch *b, int frames, const int *version, const int *span, int source, int stages) {
  return guard([&] {
    need(b && frames > 0 && version && span, "a batch, frames and their formats");
    need(source == LIZX_SRC_CAPTURE || source == LIZX_SRC_GRID, "a source (LIZX_SRC_CAPTURE or LIZX_SRC_GRID)");
    need(stages == LIZX_STAGE_SOFT || stages == LIZX_STAGE_BLOCKS, "stages (LIZX_STAGE_SOFT or LIZX_STAGE_BLOCKS)");
    hip_check(hipSetDevice(b->eng->device), "hipSetDevice");
    b->eng->wait("the previous runs");
    b->drop_graph();
    // the batch is unplanned until this call succeeds (frames = 0 rejects every run)
    b->frames = 0;
    b->blocks = 0;
    b->groups.clear();
    b->sent_frames.clear();
    b->sent_grids.clear();
    b->source = source;
    b->stages = stages;
    b->group_of.assign(frames, -1);
    b->slot_of.assign(frames, 0);
    // group the frames by format, in order of first appearance
    for (int f = 0; f < frames; f++) {
      if (!format_for(version[f])) throw Error(LIZ_E_FORMAT, "no format at that version (4 to 64)");
      if (span[f] != 128 && span[f] != 256 && span[f] != 512) throw Error(LIZ_E_ARG, "a span of 128, 256 or 512");
      int g = 0;
      for (; g < static_cast<int>(b->groups.size()); g++)
        if (b->groups[g]->plan->version == version[f] && b->groups[g]->plan->span == span[f]) break;
      if (g == static_cast<int>(b->groups.size())) {
        auto grp = std::make_unique<Group>();
        grp->plan = make_plan(version[f], span[f]);
        b->groups.push_back(std::move(grp));
      }
      b->group_of[f] = g;
      b->slot_of[f] = static_cast<int>(b->groups[g]->frames.size());
      b->groups[g]->frames.push_back(f);
    }
    // offsets: soft values and estimates in internal order (a group's frames contiguous), blocks in caller order
    b->llr_off.assign(frames, 0);
    b->est_off.assign(frames, 0);
    b->blk_off.assign(frames + 1, 0);
    int64_t lo = 0, eo = 0;
    int first = 0;
    for (auto &g : b->groups) {
      g->first = first;
      first += static_cast<int>(g->frames.size());
      for (int f : g->frames) {
        b->llr_off[f] = lo;
        b->est_off[f] = eo;
        lo += static_cast<int64_t>(g->plan->subch) * 2 * SUB;
        eo += g->plan->subch;
      }
    }
    b->llr_total = lo;
    b->est_total = eo;
    for (int f = 0; f < frames; f++) b->blk_off[f + 1] = b->blk_off[f] + b->groups[b->group_of[f]]->plan->fmt->blocks;
    b->blocks = static_cast<int>(b->blk_off[frames]);
    // device areas
    b->llr.reserve(lo);
    b->est.reserve(eo * sizeof(double));
    for (auto &g : b->groups) {
      const int cnt = static_cast<int>(g->frames.size());
      g->fb.reserve(*g->plan, cnt, false);
      g->view = g->fb.view(false);
      g->view.blk_pilot = nullptr;  // a batch reports blocks, not pilot readings
      g->view.llr = b->llr.as<int8_t>() + b->llr_off[g->frames[0]];
      g->view.est = b->est.as<double>() + b->est_off[g->frames[0]];
    }
    if (b->groups.size() > 1) b->ensure_streams(static_cast<int>(b->groups.size()));
    b->wframes.reserve(frames * sizeof(WarpFrame));
    b->grids.reserve(frames * sizeof(float *));
    // the LDPC tasks: all blocks sorted by code (1/2 first) for the batch, and per frame for the frame modes
    if (stages == LIZX_STAGE_BLOCKS) {
      b->eng->ensure_frame_tables(static_cast<int>(format_for(64)->white.size()));
      std::vector<LdpcTask> all, per;
      b->frame_task0.assign(frames + 1, 0);
      for (int f = 0; f < frames; f++) {
        const Format *fm = b->groups[b->group_of[f]]->plan->fmt;
        std::vector<LdpcTask> mine;
        for (int k = 0; k < fm->blocks; k++)
          mine.push_back(LdpcTask{b->llr_off[f] + fm->blk_slot[k], -1, fm->blk_code[k], fm->blk_slot[k],
                                  static_cast<int32_t>(b->est_off[f]), static_cast<int32_t>(b->blk_off[f] + k)});
        std::stable_sort(mine.begin(), mine.end(), [](const LdpcTask &x, const LdpcTask &y) { return x.code > y.code; });
        all.insert(all.end(), mine.begin(), mine.end());
        per.insert(per.end(), mine.begin(), mine.end());
        b->frame_task0[f + 1] = static_cast<int>(per.size());
      }
      std::stable_sort(all.begin(), all.end(), [](const LdpcTask &x, const LdpcTask &y) { return x.code > y.code; });
      const size_t tb = all.size() * sizeof(LdpcTask);
      b->tasks.reserve(tb);
      b->frame_tasks.reserve(tb);
      copy_wait(b->tasks.p, all.data(), tb, hipMemcpyHostToDevice, b->eng->stream, "tasks");
      copy_wait(b->frame_tasks.p, per.data(), tb, hipMemcpyHostToDevice, b->eng->stream, "tasks");
      b->its.reserve(b->blocks);
      b->blk_est.reserve(b->blocks * sizeof(float));
      b->bytes.reserve(static_cast<size_t>(b->blocks) * BLOCK_BYTES);
      b->crc_ok.reserve(b->blocks);
    }
    b->frames = frames;
    return 0;
  });
}

LIZ_API int lizx_batch_dims(const lizx_batch *b, int *dims) {
  return guard([&] {
    need(b && dims, "a batch and dims");
    dims[0] = b->frames;
    dims[1] = b->blocks;
    dims[2] = static_cast<int>(b->llr_total);
    dims[3] = static_cast<int>(b->groups.size());
    return 0;
  });
}

LIZ_API int lizx_batch_run(lizx_batch *b, const lizx_frame *frames, int mode) {
  return guard([&] {
    need(b && b->frames > 0 && frames, "a planned batch and its frames");
    need(mode >= LIZX_RUN_GRAPH && mode <= LIZX_RUN_PLAIN, "a run mode (LIZX_RUN_*)");
    LdpcEngine &e = *b->eng;
    hip_check(hipSetDevice(e.device), "hipSetDevice");
    // the frames' descriptors, in internal order, to fixed device addresses the launches (and the graph) read
    b->host_frames.assign(b->frames, WarpFrame{});
    b->host_grids.assign(b->frames, nullptr);
    for (int f = 0; f < b->frames; f++) {
      const Group &g = *b->groups[b->group_of[f]];
      const int at = g.first + b->slot_of[f];
      const lizx_frame &in = frames[f];
      if (b->source == LIZX_SRC_GRID) {
        need(in.grid != nullptr, "a grid for every frame");
        b->host_grids[at] = in.grid;
      } else {
        need(in.img && in.nodes && in.w >= 2 && in.h >= 2, "a capture and nodes for every frame");
        need(warp_frame_fits(in.h, in.w), "captures under 2^31 pixels with rows under 2^24 pixels");
        WarpFrame w{};
        std::memcpy(w.H, in.H, sizeof w.H);
        w.nodes = reinterpret_cast<const float2 *>(in.nodes);
        w.img = in.img;
        w.w = in.w;
        w.h = in.h;
        w.stride = in.w;
        b->host_frames[at] = w;
      }
    }
    // the descriptors go to the device only when they changed: the stream is drained first (a queued run may still
    // read them), so a batch replayed on the same inputs keeps the GPU fed back to back
    const bool grid = b->source == LIZX_SRC_GRID;
    const bool same = grid ? b->host_grids == b->sent_grids
                           : (b->sent_frames.size() == b->host_frames.size() &&
                              std::memcmp(b->sent_frames.data(), b->host_frames.data(),
                                          b->frames * sizeof(WarpFrame)) == 0);
    if (!same) {
      e.wait("the previous runs");
      if (grid) {
        copy_wait(b->grids.p, b->host_grids.data(), b->frames * sizeof(float *), hipMemcpyHostToDevice, b->eng->stream,
                  "grids");
        b->sent_grids = b->host_grids;
      } else {
        copy_wait(b->wframes.p, b->host_frames.data(), b->frames * sizeof(WarpFrame), hipMemcpyHostToDevice,
                  b->eng->stream, "frames");
        b->sent_frames = b->host_frames;
      }
    }
    if (!e.armed) {
      hip_check(hipEventRecord(e.ev0, e.stream), "hipEventRecord");
      e.armed = true;
    }
    if (mode == LIZX_RUN_GRAPH) {
      if (!b->exec) {
        hip_check(hipStreamBeginCapture(e.stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
        try {
          b->enqueue_batch(false);
        } catch (...) {
          hipGraph_t g = nullptr;
          (void)hipStreamEndCapture(e.stream, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        hip_check(hipStreamEndCapture(e.stream, &b->graph), "hipStreamEndCapture");
        hip_check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
      }
      hip_check(hipGraphLaunch(b->exec, e.stream), "hipGraphLaunch");
    } else if (mode == LIZX_RUN_STREAM) {
      b->enqueue_batch(false);
    } else {
      for (int f = 0; f < b->frames; f++) {
        b->enqueue_frame(f, mode == LIZX_RUN_PLAIN);
        e.wait("a frame");  // a decoder called once a frame returns here
      }
    }
    hip_check(hipEventRecord(e.ev1, e.stream), "hipEventRecord");
    return 0;
  });
}

// The autotuner: coordinate descent over the launch shapes (and, with plans, the FFTs' radix plans), each candidate
// timed on the batch's own frames; the best is kept and the graph re-captured with it. Shapes leave the results'
// bits unchanged; a radix plan orders the FFT's float operations differently.
static void tune(lizx_batch *b, const lizx_frame *frames, int reps, bool plans, int *choice) {
  LdpcEngine &e = *b->eng;
  hip_check(hipSetDevice(e.device), "hipSetDevice");
  auto time_it = [&]() {
    // one run to settle, then reps timed ones, on the stream (no graph: the shape changes between candidates)
    if (lizx_batch_run(b, frames, LIZX_RUN_STREAM)) throw Error(liz_last_error_code(), liz_last_error());
    e.sync();
    for (int r = 0; r < reps; r++)
      if (lizx_batch_run(b, frames, LIZX_RUN_STREAM)) throw Error(liz_last_error_code(), liz_last_error());
    e.sync();
    return e.last_ms;
  };
  auto descend = [&](int *knob, std::initializer_list<int> values) {
    int best = *knob;
    float best_ms = time_it();
    for (int v : values) {
      if (v == best) continue;
      const int keep = *knob;
      *knob = v;
      const float ms = time_it();
      if (ms < best_ms) {
        best_ms = ms;
        best = v;
      }
      *knob = keep;
    }
    *knob = best;
  };
  descend(&b->shape.rows, {1, 2, 4});
  descend(&b->shape.cols, {1, 2, 4});
  descend(&b->shape.demap, {1, 2, 4});
  if (plans) {
    descend(&b->shape.rows_radix, {4, 8, 16});
    descend(&b->shape.cols_radix, {4, 8, 16});
  }
  if (b->stages == LIZX_STAGE_BLOCKS) {
    int *pc = &b->ldpc_per_cu;
    if (e.wg_per_cu >= 8) descend(pc, {0, 2, 4, 6, 8});
    else if (e.wg_per_cu >= 6) descend(pc, {0, 2, 4, 6});
    else if (e.wg_per_cu >= 4) descend(pc, {0, 2, 4});
    else descend(pc, {0, 2});
  }
  b->drop_graph();
  if (choice) {
    choice[0] = b->shape.rows;
    choice[1] = b->shape.cols;
    choice[2] = b->shape.demap;
    choice[3] = b->ldpc_per_cu;
    if (plans) {
      choice[4] = b->shape.rows_radix;
      choice[5] = b->shape.cols_radix;
    }
  }
}

LIZ_API int lizx_batch_tune(lizx_batch *b, const lizx_frame *frames, int reps, int *choice) {
  return guard([&] {
    need(b && b->frames > 0 && frames && reps > 0 && reps <= 100, "a planned batch, its frames and 1 to 100 repeats");
    tune(b, frames, reps, false, choice);
    return 0;
  });
}

LIZ_API int lizx_batch_tune_plans(lizx_batch *b, const lizx_frame *frames, int reps, int *choice) {
  return guard([&] {
    need(b && b->frames > 0 && frames && reps > 0 && reps <= 100, "a planned batch, its frames and 1 to 100 repeats");
    tune(b, frames, reps, true, choice);
    return 0;
  });
}

LIZ_API int lizx_batch_otf(lizx_batch *b, const float *sigma) {
  return guard([&] {
    need(b && b->frames > 0, "a planned batch");
    hip_check(hipSetDevice(b->eng->device), "hipSetDevice");
    for (auto &g : b->groups) {
      std::vector<float> sg(g->frames.size(), 0.0f);
      for (size_t i = 0; sigma && i < sg.size(); i++) sg[i] = sigma[g->frames[i]];
      g->fb.set_otf(sg.data(), static_cast<int>(sg.size()), b->eng->stream);
      g->view.otf = g->fb.has_otf ? g->fb.otf.as<float>() : nullptr;
    }
    b->drop_graph();  // the kernels' arguments change
    return 0;
  });
}

LIZ_API int lizx_batch_shape_plans(lizx_batch *b, const int *choice) {
  return guard([&] {
    need(b && choice, "a batch and a shape");
    for (int k = 4; k < 6; k++) need(choice[k] == 4 || choice[k] == 8 || choice[k] == 16, "radix plans 4, 8, 16");
    if (lizx_batch_shape(b, choice)) throw Error(liz_last_error_code(), liz_last_error());
    b->shape.rows_radix = choice[4];
    b->shape.cols_radix = choice[5];
    return 0;
  });
}

LIZ_API int lizx_batch_shape(lizx_batch *b, const int *choice) {
  return guard([&] {
    need(b && choice, "a batch and a shape");
    for (int k = 0; k < 3; k++) need(choice[k] == 1 || choice[k] == 2 || choice[k] == 4, "waves per workgroup 1, 2, 4");
    need(choice[3] >= 0 && choice[3] <= 64, "LDPC workgroups a CU, 0 for as many as fit");
    b->shape.rows = choice[0];
    b->shape.cols = choice[1];
    b->shape.demap = choice[2];
    b->ldpc_per_cu = choice[3];
    b->drop_graph();
    return 0;
  });
}

LIZ_API int lizx_batch_sync(lizx_batch *b, float *ms) {
  return guard([&] {
    need(b != nullptr, "a batch");
    hip_check(hipSetDevice(b->eng->device), "hipSetDevice");
    b->eng->sync();
    if (ms) *ms = b->eng->last_ms;
    return 0;
  });
}

LIZ_API int lizx_batch_blocks(lizx_batch *b, uint8_t *blocks, uint8_t *ok, int8_t *its, float *blk_est) {
  return guard([&] {
    need(b && b->frames > 0 && b->stages == LIZX_STAGE_BLOCKS, "a batch planned with LIZX_STAGE_BLOCKS");
    need(blocks && ok && its && blk_est, "output arrays");
    hip_check(hipSetDevice(b->eng->device), "hipSetDevice");
    b->eng->wait("the runs");
    const int B = b->blocks;
    copy_wait(its, b->its.p, B, hipMemcpyDeviceToHost, b->eng->stream, "iterations");
    copy_wait(blk_est, b->blk_est.p, B * sizeof(float), hipMemcpyDeviceToHost, b->eng->stream, "estimates");
    copy_wait(ok, b->crc_ok.p, B, hipMemcpyDeviceToHost, b->eng->stream, "verdicts");
    b->out_bytes.resize(static_cast<size_t>(B) * BLOCK_BYTES);
    copy_wait(b->out_bytes.data(), b->bytes.p, b->out_bytes.size(), hipMemcpyDeviceToHost, b->eng->stream, "blocks");
    int good = 0;
    for (int k = 0; k < B; k++) {
      if (its[k] == -2) ok[k] = 0;
      if (ok[k]) {
        std::memcpy(blocks + static_cast<size_t>(k) * BLOCK_BYTES, b->out_bytes.data() + static_cast<size_t>(k) * BLOCK_BYTES,
                    BLOCK_BYTES);
        good++;
      }
    }
    return good;
  });
}

LIZ_API int lizx_batch_soft(lizx_batch *b, int frame, int8_t *llr, double *sub_est, float *sub_rms, float *align) {
  return guard([&] {
    need(b && frame >= 0 && frame < b->frames, "a batch and a frame of it");
    hip_check(hipSetDevice(b->eng->device), "hipSetDevice");
    b->eng->wait("the runs");
    const Group &g = *b->groups[b->group_of[frame]];
    const FramePlan &p = *g.plan;
    const SoftBufs v = g.view.at(p, b->slot_of[frame]);
    const size_t S = static_cast<size_t>(p.subch) * SUB;
    if (llr) copy_wait(llr, v.llr, 2 * S, hipMemcpyDeviceToHost, b->eng->stream, "soft values");
    if (sub_est) copy_wait(sub_est, v.est, p.subch * sizeof(double), hipMemcpyDeviceToHost, b->eng->stream, "estimates");
    if (sub_rms) copy_wait(sub_rms, v.rms, p.subch * sizeof(float), hipMemcpyDeviceToHost, b->eng->stream, "rms");
    if (align) copy_wait(align, v.align, 2 * sizeof(float), hipMemcpyDeviceToHost, b->eng->stream, "shift");
    return 0;
  });
}

}  // extern "C"
