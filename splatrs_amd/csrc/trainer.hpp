// The trainer object behind splat_trainer_* (shared by trainer.hip and learn.hip).
#pragma once

#include <deque>
#include <memory>

#include "internal.hpp"

using splat::DevBuf;
using splat::ProjBuffers;
using splat::BinBuffers;
using splat::StepDyn;
using splat::StepStatus;
using splat::StepReport;

struct splat_scorer;
struct LearnState;

struct TrainerView {
  splat_camera cam;
  DevBuf rgb;
};

// Everything a captured step bakes in: sizes, flags and device buffers.  Two steps with
// byte-identical plans can replay the same graph (every buffer in the plan is reserved to
// at least the sizes in the plan before the replay).
struct StepPlan {
  uint64_t n, stride, entry_cap;
  uint32_t width, height;
  int32_t sh_degree, antialias, ordered, accumulate, absgrad, pad0;
  float near_plane, lambda, bg[3], strong_threshold;
  int32_t fwd_batch, fwd_persistent, fwd_per_cu, bwd_batch, bwd_persistent, bwd_per_cu;
  int32_t bwd_rows, pad1;
  float *params, *m, *v, *grad_accum;
  uint32_t *denom, *max_radii;
  float* max_grad;   // extra statistics for the learned scorer (null when not tracked)
  uint32_t* strong;
  ProjBuffers pb;
  BinBuffers bb;
  float *image, *final_t, *dimage, *loss_scratch, *slots, *loss_one;
  uint32_t *n_contrib, *small, *big_list, *scan_sums, *counter;
  unsigned long long* total64;
  StepDyn* dyn;
  StepStatus* status;
  StepReport* reports;
};

struct PendingStep {
  uint64_t seq;
  uint32_t view;
  splat_step_params p;
  uint64_t adam_step;
  float* loss_out;
  hipEvent_t done;
};

struct splat_trainer {
  splat_ctx* ctx = nullptr;
  int device = 0;
  size_t n = 0, cap = 0, stride = 0;
  uint64_t adam_step = 0;
  DevBuf params, m, v;                  // planes [59][stride]
  DevBuf grad_accum, denom, max_radii;  // statistics, capacity cap
  std::vector<std::unique_ptr<TrainerView>> views;
  std::map<std::string, DevBuf> store;  // per-view buffers (projection, binning, images)
  size_t entry_cap = 0;
  float* loss_target = nullptr;  // when set, each step writes its loss value here (device)
  // graph-captured steps
  bool use_graph = true;
  uint64_t seq = 0;
  std::deque<PendingStep> pending;
  std::vector<std::pair<StepPlan, hipGraphExec_t>> graphs;
  std::vector<hipEvent_t> events;
  StepReport* reports_host = nullptr;  // pinned, mapped; written by the step's last kernel
  StepReport* reports_dev = nullptr;
  unsigned long long last_total = 0;
  splat::RasterVariant fwd_shape, bwd_shape;  // compositing launch shapes (splat_trainer_tune)
  uint64_t graph_captures = 0, graph_replays = 0, step_aborts = 0;
  // learned densification (learn.hip): extra statistics and lineage planes, the attached
  // scorer and its sample store
  bool track = false;                  // keep max_grad / strong / birth / origin
  float strong_threshold = 0.0002f;
  DevBuf max_grad, strong, birth, origin;
  splat_scorer* scorer = nullptr;
  std::unique_ptr<LearnState> learn;
  splat_trainer();  // defined where LearnState is complete (learn.hip)
  ~splat_trainer();
};


namespace splat {
// trainer.hip
void trainer_drain(splat_trainer* t);
// densify_and_prune on the resident model with optional external decisions; carries the
// lineage planes when they are tracked (sample_of: logged decision id per row, or null).
void trainer_densify(splat_trainer* t, const splat_densify_config& cfg, float extent,
                     uint64_t seed, const uint8_t* decide, const uint32_t* sample_of,
                     uint32_t iteration, splat_densify_report* report);
// Mean PSNR (src/loss.rs:175-190 against the RGB8 targets) over the first `views` views.
float trainer_eval_psnr(splat_trainer* t, uint32_t views, int32_t sh_degree, const float bg[3]);
// Each row's share of the listed views' images: the sum of alpha T over their pixels (the
// red colour gradient of a backward pass under dL/dimage = (1, 0, 0), black background),
// into contrib[n].
void trainer_image_share(splat_trainer* t, const uint32_t* views, uint32_t count,
                         int32_t sh_degree, float* contrib);
// learn.hip
void learn_track_alloc(splat_trainer* t);
void learn_zero_stats(splat_trainer* t);
// One densification event of splat_train with a scorer attached (log / learned decisions,
// outcome labels of older events, online fit).
void learn_densify(splat_trainer* t, const splat_densify_config& cfg, float extent,
                   uint64_t seed, uint32_t iteration, int32_t sh_degree, const float bg[3],
                   uint32_t last_reset, splat_densify_report* report);
}  // namespace splat
