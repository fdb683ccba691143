// autotune.cpp -- the traversal autotuner, a client of the public C ABI.
//
// For one representative frame it times the trace stages (primary trace,
// proposal network, finish + shadow trace; HIP events, median of `repeats`
// frames after one warm-up) over LDS stack depth {8, 16, 32} x rays per
// work item {1, 2, 4}, keeps the fastest, then tries the refill threshold
// {8, 16, 32, 48} there, then the tail mode {1, 3, 4} (the cooperative tail;
// the split tail from 16 or 32 rays left), and for J3DG_LEARNED frames then
// sweeps the bound margin over {0, 0.005, 0.01, 0.02, 0.05, 0.1, 0.2, 0.4,
// 0.8, 1.6, 3.2}.  BVH width {4, 8}: the same frame given twice, over width-4
// and width-8 uploads of its meshes (the caller owns the uploads), each tuned,
// the faster kept.  Batches (j3dg_render_batch) are tuned on the time of their
// one graph launch over width x depth x rays per work item, then the tail
// mode.  The result cache lives with the caller.
#include <algorithm>
#include <cstring>
#include <vector>

#include "j3dg.h"

namespace {

int time_frame(j3dg_ctx* ctx, const j3dg_frame_desc& fd, uint32_t target, uint32_t repeats, float* ms) {
  int rc = j3dg_render(ctx, &fd, target, J3DG_STAGE_TRACE, nullptr, nullptr);
  if (rc) return rc;
  if ((rc = j3dg_synchronize(ctx)) || (rc = j3dg_reset_timings(ctx))) return rc;
  for (uint32_t i = 0; i < repeats; ++i)
    if ((rc = j3dg_render(ctx, &fd, target, J3DG_STAGE_TRACE, nullptr, nullptr))) return rc;
  std::vector<float> tr(repeats), ml(repeats), sh(repeats), tot(repeats);
  uint32_t n = 0;
  if ((rc = j3dg_get_timings(ctx, tr.data(), nullptr, ml.data(), sh.data(), repeats, &n))) return rc;
  if (n == 0) return J3DG_ERR_INTERNAL;
  for (uint32_t i = 0; i < n; ++i) tot[i] = tr[i] + ml[i] + sh[i];
  std::nth_element(tot.begin(), tot.begin() + n / 2, tot.begin() + n);
  *ms = tot[n / 2];
  return J3DG_OK;
}

// Median time of one graph launch of the batch (trace + shade) after a warm-up
// launch, which also captures the graph for the current tuning.
int time_batch(j3dg_ctx* ctx, const j3dg_frame_desc* frames, const uint32_t* targets, uint32_t n, uint32_t repeats,
               float* ms) {
  int rc = j3dg_render_batch(ctx, frames, targets, n);
  if (rc) return rc;
  if ((rc = j3dg_synchronize(ctx)) || (rc = j3dg_reset_timings(ctx))) return rc;
  for (uint32_t i = 0; i < repeats; ++i)
    if ((rc = j3dg_render_batch(ctx, frames, targets, n))) return rc;
  std::vector<float> tr(repeats);
  uint32_t got = 0;
  if ((rc = j3dg_get_timings(ctx, tr.data(), nullptr, nullptr, nullptr, repeats, &got))) return rc;
  if (got == 0) return J3DG_ERR_INTERNAL;
  std::nth_element(tr.begin(), tr.begin() + got / 2, tr.begin() + got);
  *ms = tr[got / 2];
  return J3DG_OK;
}

// The BVH width of a frame's meshes (that of its first object; 0 without objects).
uint32_t frame_width(j3dg_ctx* ctx, const j3dg_frame_desc& fd) {
  if (fd.n_objects == 0) return 0;
  const uint32_t slot = fd.object_list ? fd.object_list[0] : fd.objects[0];
  j3dg_mesh_info info;
  return j3dg_get_mesh_info(ctx, slot, &info) == J3DG_OK ? info.width : 0;
}

}  // namespace

extern "C" int j3dg_autotune_widths(j3dg_ctx* ctx, const j3dg_frame_desc* frame_w4, const j3dg_frame_desc* frame_w8,
                                    uint32_t target, uint32_t repeats, j3dg_autotune_report* rep) {
  if (!ctx || !rep || (!frame_w4 && !frame_w8)) return J3DG_ERR_INVALID_ARG;
  if (!frame_w4) return j3dg_autotune(ctx, frame_w8, target, repeats, rep);
  if (!frame_w8) return j3dg_autotune(ctx, frame_w4, target, repeats, rep);
  j3dg_tuning base;
  int rc = j3dg_get_tuning(ctx, &base);
  if (rc) return rc;
  j3dg_autotune_report r4, r8;
  if ((rc = j3dg_autotune(ctx, frame_w4, target, repeats, &r4))) return rc;
  if ((rc = j3dg_set_tuning(ctx, &base))) return rc;
  if ((rc = j3dg_autotune(ctx, frame_w8, target, repeats, &r8))) return rc;
  *rep = r8.best_ms <= r4.best_ms ? r8 : r4;
  rep->n_tried = r4.n_tried + r8.n_tried;
  return j3dg_set_tuning(ctx, &rep->best);
}

extern "C" int j3dg_autotune_batch(j3dg_ctx* ctx, const j3dg_frame_desc* frames_w4, const j3dg_frame_desc* frames_w8,
                                   const uint32_t* targets, uint32_t n, uint32_t repeats, j3dg_autotune_report* rep) {
  if (!ctx || !rep || !targets || n == 0 || (!frames_w4 && !frames_w8)) return J3DG_ERR_INVALID_ARG;
  if (repeats == 0) repeats = 3;
  if (repeats > 256) return J3DG_ERR_INVALID_ARG;
  std::memset(rep, 0, sizeof(*rep));
  j3dg_tuning base;
  int rc = j3dg_get_tuning(ctx, &base);
  if (rc) return rc;
  const j3dg_frame_desc* sets[2] = {frames_w4, frames_w8};
  rc = time_batch(ctx, frames_w8 ? frames_w8 : frames_w4, targets, n, repeats, &rep->baseline_ms);
  if (rc) return rc;
  j3dg_tuning best = base;
  float best_ms = -1.f;
  uint32_t best_width = 0;
  const uint32_t depths[3] = {8, 16, 32}, rpis[3] = {1, 2, 4};
  for (const j3dg_frame_desc* fs : sets) {
    if (!fs) continue;
    for (uint32_t d : depths)
      for (uint32_t r : rpis) {
        j3dg_tuning t = base;
        t.stack_depth = d;
        t.rays_per_item = r;
        if ((rc = j3dg_set_tuning(ctx, &t))) return rc;
        float ms = 0.f;
        if ((rc = time_batch(ctx, fs, targets, n, repeats, &ms))) {
          j3dg_set_tuning(ctx, &base);
          return rc;
        }
        rep->n_tried++;
        if (best_ms < 0.f || ms < best_ms) best_ms = ms, best = t, best_width = frame_width(ctx, fs[0]);
      }
  }
  // the tail mode at the best setting (over the width that setting came from)
  const j3dg_frame_desc* bfs = best_width == 4 && frames_w4 ? frames_w4 : frames_w8 ? frames_w8 : frames_w4;
  const uint32_t tails[3] = {1, 3, 4};
  const j3dg_tuning at = best;
  for (uint32_t m : tails) {
    if (m == at.coop_tail) continue;
    j3dg_tuning t = at;
    t.coop_tail = m;
    if ((rc = j3dg_set_tuning(ctx, &t))) return rc;
    float ms = 0.f;
    if ((rc = time_batch(ctx, bfs, targets, n, repeats, &ms))) {
      j3dg_set_tuning(ctx, &base);
      return rc;
    }
    rep->n_tried++;
    if (ms < best_ms) best_ms = ms, best = t;
  }
  if ((rc = j3dg_set_tuning(ctx, &best))) return rc;
  rep->best = best;
  rep->best_ms = best_ms;
  rep->best_margin = 0.f;
  rep->width = best_width;
  return J3DG_OK;
}

extern "C" int j3dg_autotune(j3dg_ctx* ctx, const j3dg_frame_desc* frame, uint32_t target, uint32_t repeats,
                             j3dg_autotune_report* rep) {
  if (!ctx || !frame || !rep) return J3DG_ERR_INVALID_ARG;
  if (repeats == 0) repeats = 5;
  if (repeats > 256) return J3DG_ERR_INVALID_ARG;
  std::memset(rep, 0, sizeof(*rep));
  j3dg_tuning base;
  int rc = j3dg_get_tuning(ctx, &base);
  if (rc) return rc;
  rc = time_frame(ctx, *frame, target, repeats, &rep->baseline_ms);
  if (rc) return rc;
  j3dg_tuning best = base;
  float best_ms = rep->baseline_ms;
  const uint32_t depths[3] = {8, 16, 32}, rpis[3] = {1, 2, 4};
  for (uint32_t d : depths)
    for (uint32_t r : rpis) {
      j3dg_tuning t = base;
      t.stack_depth = d;
      t.rays_per_item = r;
      if ((rc = j3dg_set_tuning(ctx, &t))) return rc;
      float ms = 0.f;
      if ((rc = time_frame(ctx, *frame, target, repeats, &ms))) {
        j3dg_set_tuning(ctx, &base);
        return rc;
      }
      rep->n_tried++;
      if (ms < best_ms) best_ms = ms, best = t;
    }
  const uint32_t refills[4] = {8, 16, 32, 48};
  const j3dg_tuning at = best;
  for (uint32_t f : refills) {
    if (f == at.refill_min) continue;
    j3dg_tuning t = at;
    t.refill_min = f;
    if ((rc = j3dg_set_tuning(ctx, &t))) return rc;
    float ms = 0.f;
    if ((rc = time_frame(ctx, *frame, target, repeats, &ms))) {
      j3dg_set_tuning(ctx, &base);
      return rc;
    }
    rep->n_tried++;
    if (ms < best_ms) best_ms = ms, best = t;
  }
  const uint32_t tails[3] = {1, 3, 4};
  const j3dg_tuning at2 = best;
  for (uint32_t m : tails) {
    if (m == at2.coop_tail) continue;
    j3dg_tuning t = at2;
    t.coop_tail = m;
    if ((rc = j3dg_set_tuning(ctx, &t))) return rc;
    float ms = 0.f;
    if ((rc = time_frame(ctx, *frame, target, repeats, &ms))) {
      j3dg_set_tuning(ctx, &base);
      return rc;
    }
    rep->n_tried++;
    if (ms < best_ms) best_ms = ms, best = t;
  }
  if ((rc = j3dg_set_tuning(ctx, &best))) return rc;
  rep->best = best;
  rep->best_ms = best_ms;
  rep->best_margin = frame->margin;
  rep->width = frame_width(ctx, *frame);
  if (frame->flags & J3DG_LEARNED) {
    const float margins[11] = {0.f, 0.005f, 0.01f, 0.02f, 0.05f, 0.1f, 0.2f, 0.4f, 0.8f, 1.6f, 3.2f};
    float bm = best_ms;
    for (float m : margins) {
      j3dg_frame_desc fd = *frame;
      fd.margin = m;
      float ms = 0.f;
      if ((rc = time_frame(ctx, fd, target, repeats, &ms))) return rc;
      rep->n_tried++;
      if (ms < bm) bm = ms, rep->best_margin = m;
    }
    rep->best_ms = bm;
  }
  return J3DG_OK;
}
