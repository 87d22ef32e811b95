// Device-resident training (src/train.rs:196-303): the model, its gradient, both Adam
// moments, the densification statistics and every view's image stay in HBM; a step runs
// project -> bin -> composite -> loss -> backward -> statistics -> Adam on the context
// stream.
//
// Two ways to run a step:
//   * graph (default): the step's launches are captured once into a HIP graph (keyed by
//     every buffer and size it bakes in) and replayed each iteration.  The per-iteration
//     values go through a small device record written just before the replay.  The entry
//     arrays have a preallocated capacity; a step that needs more sets a sticky abort flag
//     (nothing of the model changes), and the host, which checks each step's report two
//     steps behind, grows the arrays and replays the aborted steps in order.  No host
//     synchronisation sits inside a step.
//   * eager: the same kernels launched one by one, reading the entry count back to size
//     the entry arrays (used for per-kernel event timing).
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <memory>
#include <new>

#include "internal.hpp"
#include "trainer.hpp"

using namespace splat;

namespace {

void drain(splat_trainer* t);

// After a failed call, finish the steps in flight; if that fails too, forget them.
void settle(splat_trainer* t) {
  try {
    drain(t);
  } catch (...) {
    t->pending.clear();
  }
}

template <class F>
int tguard(splat_trainer* t, F&& f) {
  if (!t || !t->ctx) return SPLAT_ERR_INVALID;
  splat_ctx* ctx = t->ctx;
  try {
    SPLAT_HIP(hipSetDevice(ctx->device));
    f();
    ctx->err.clear();
    return SPLAT_OK;
  } catch (const Error& e) {
    settle(t);
    ctx->err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    settle(t);
    ctx->err = "host allocation failed";
    return SPLAT_ERR_OOM;
  } catch (const std::exception& e) {
    settle(t);
    ctx->err = e.what();
    return SPLAT_ERR_INVALID;
  }
}

void alloc_model(splat_trainer* t, size_t cap) {
  t->cap = cap;
  t->stride = plane_stride(cap);
  const size_t bytes = t->stride * kParams * 4;
  t->params.reserve(bytes);
  t->m.reserve(bytes);
  t->v.reserve(bytes);
  t->grad_accum.reserve((cap + 1) * 4);
  t->denom.reserve((cap + 1) * 4);
  t->max_radii.reserve((cap + 1) * 4);
}

void zero_stats(splat_trainer* t) {
  splat_ctx* ctx = t->ctx;
  fill_u32(ctx, t->grad_accum.as<uint32_t>(), t->n, 0u);  // +0.0f has all bits clear
  fill_u32(ctx, t->denom.as<uint32_t>(), t->n, 0u);
  fill_u32(ctx, t->max_radii.as<uint32_t>(), t->n, 0u);
  if (t->track) learn_zero_stats(t);
}

ProjBuffers proj_buffers(splat_trainer* t) {
  ProjBuffers pb;
  const size_t c = t->cap + 1;
  pb.rec = (float4*)t->store["rec"].reserve(c * 64);
  pb.depth = (float*)t->store["depth"].reserve(c * 4);
  pb.radius = (uint32_t*)t->store["radius"].reserve(c * 4);
  pb.rect = (uint2*)t->store["rect"].reserve(c * 8);
  pb.ntiles = (uint32_t*)t->store["ntiles"].reserve(c * 4);
  return pb;
}

// Rows with a positive radius, added to *count: the grid strides over the rows and each
// workgroup adds its total once (a few hundred atomics on the counter, not one per wave).
constexpr uint32_t kCountBlocks = 256;
__global__ void __launch_bounds__(256) count_visible_kernel(const uint32_t* __restrict__ radius,
                                                            uint32_t n, uint32_t* count) {
  __shared__ uint32_t s_cnt[4];
  uint32_t c = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    c += (uint32_t)__popcll(__ballot(radius[i] > 0));  // the wave's count, in every lane
  if ((threadIdx.x & 63u) == 0) s_cnt[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t total = (s_cnt[0] + s_cnt[1]) + (s_cnt[2] + s_cnt[3]);
    if (total) atomicAdd(count, total);
  }
}

// Per-block sums of squared errors (binary64) between a clamped image and RGB8 targets.
__global__ void __launch_bounds__(256) sq_err_kernel(const float* __restrict__ img,
                                                     const uint8_t* __restrict__ rgb, size_t n,
                                                     double* __restrict__ partial) {
  __shared__ double sh[4];
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double d = (double)rclamp(img[i], 0.0f, 1.0f) -
                     (double)rclamp((float)rgb[i] / 255.0f, 0.0f, 1.0f);
    acc += d * d;
  }
  for (int k = 32; k >= 1; k >>= 1) acc += __shfl_xor(acc, k, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// Before each replay of a graph-captured step (kernel arguments are frozen in a graph):
// the step's per-iteration values, and the counters its binning accumulates into cleared --
// the 64-bit entry total (2 words), the ntc words of tile counter sets and the queued-tile
// counts small[1..3].
__global__ void set_dyn_kernel(StepDyn d, StepDyn* __restrict__ out, uint32_t* __restrict__ total64,
                               uint32_t* __restrict__ tile_count, uint32_t ntc,
                               uint32_t* __restrict__ small) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *out = d;
  if (i < 2) total64[i] = 0u;
  if (i < 3) small[1 + i] = 0u;
  if (i < ntc) tile_count[i] = 0u;
}

}  // namespace

namespace {

constexpr uint32_t kReportSlots = 64;
constexpr size_t kStepLag = 2;     // steps in flight before the host checks the oldest
constexpr size_t kMaxGraphs = 8;   // cached step graphs per trainer

float adam_c(float beta, uint64_t step) { return 1.0f - powi(beta, (int32_t)step); }

splat_proj_opts step_opts(const splat_step_params& p) {
  splat_proj_opts o;
  o.near_plane = p.near_plane;
  o.eps2d = 0.3f;
  o.antialias = p.antialias;
  o.sh_degree = p.sh_degree;
  return o;
}

// Reserves every buffer of a graph-captured step and returns its plan.
StepPlan plan_step(splat_trainer* t, const splat_camera& cam, const splat_step_params& p) {
  StepPlan P;
  std::memset(&P, 0, sizeof(P));
  auto& st = t->store;
  const uint32_t W = cam.width, H = cam.height;
  const size_t npx = (size_t)W * H;
  const uint32_t tx = (W + kTile - 1) / kTile, ty = (H + kTile - 1) / kTile;
  const size_t T = (size_t)tx * ty;
  const size_t n = t->n;
  if (t->entry_cap == 0) t->entry_cap = std::max<size_t>(4 * n, 1 << 16);
  const size_t ec = t->entry_cap;
  P.n = n;
  P.stride = t->stride;
  P.entry_cap = ec;
  P.width = W;
  P.height = H;
  P.sh_degree = p.sh_degree;
  P.antialias = p.antialias;
  P.ordered = p.ordered_backward != 0;
  P.accumulate = p.accumulate_stats != 0;
  P.absgrad = p.absgrad != 0;
  P.near_plane = p.near_plane;
  P.lambda = p.lambda_dssim;
  for (int c = 0; c < 3; c++) P.bg[c] = p.background[c];
  P.params = t->params.as<float>();
  P.m = t->m.as<float>();
  P.v = t->v.as<float>();
  P.grad_accum = t->grad_accum.as<float>();
  P.denom = t->denom.as<uint32_t>();
  P.max_radii = t->max_radii.as<uint32_t>();
  if (t->track) {
    P.max_grad = t->max_grad.as<float>();
    P.strong = t->strong.as<uint32_t>();
    P.strong_threshold = t->strong_threshold;
  }
  P.fwd_batch = t->fwd_shape.fwd_batch;
  P.fwd_persistent = t->fwd_shape.persistent;
  P.fwd_per_cu = (int32_t)t->fwd_shape.per_cu;
  P.bwd_batch = t->bwd_shape.bwd_batch;
  P.bwd_persistent = t->bwd_shape.persistent;
  P.bwd_per_cu = (int32_t)t->bwd_shape.per_cu;
  P.bwd_rows = t->bwd_shape.rows ? 1 : 0;
  P.counter = (uint32_t*)st["raster_counter"].reserve(64);
  P.pb = proj_buffers(t);
  BinBuffers& bb = P.bb;
  bb.tiles_x = tx;
  bb.tiles_y = ty;
  bb.entry_offset = (uint32_t*)st["entry_offset"].reserve((t->cap + 1) * 4);
  bb.tile_count = (uint32_t*)st["tile_count"].reserve(kTileSub * T * 4);
  bb.tile_sub = (uint32_t*)st["tile_sub"].reserve((kTileSub + 1) * T * 4);
  bb.block_marks = (uint32_t*)st["block_marks"].reserve((ec / 256 + 2) * 4);
  bb.tile_start = (uint32_t*)st["tile_start"].reserve((T + 1) * 4);
  bb.tile_order = (uint32_t*)st["tile_order"].reserve(T * 4 + 4);
  bb.entry_gaussian = (uint32_t*)st["entry_gaussian"].reserve(ec * 4);
  bb.entry_tile = (uint32_t*)st["entry_tile"].reserve(ec * 4);
  bb.entry_slot = (uint32_t*)st["entry_slot"].reserve(ec * 4);
  bb.bucket = (uint4*)st["bucket"].reserve(ec * 16);
  bb.sorted_entry = (uint32_t*)st["sorted_entry"].reserve(ec * 4);
  bb.sorted_gauss = (uint32_t*)st["sorted_gauss"].reserve(ec * 4);
  P.small = (uint32_t*)st["bin_small"].reserve(64);
  P.total64 = (unsigned long long*)st["bin_total64"].reserve(8);
  P.big_list = (uint32_t*)st["bin_big_list"].reserve(T * 4);
  P.scan_sums = (uint32_t*)st["step_scan_sums"].reserve(scan_temp_bytes(std::max(t->cap + 1, T)));
  P.image = (float*)st["image"].reserve(npx * 12 + 4);
  P.final_t = (float*)st["final_t"].reserve(npx * 4 + 4);
  P.n_contrib = (uint32_t*)st["n_contrib"].reserve(npx * 4 + 4);
  P.dimage = (float*)st["dimage"].reserve(npx * 12 + 4);
  P.loss_scratch = (float*)st["loss_scratch"].reserve(loss_scratch_floats(W, H) * 4);
  P.slots = (float*)st["slots"].reserve((ec + 1) * kSlots * 4);
  P.loss_one = (float*)st["step_loss"].reserve(4);
  P.dyn = (StepDyn*)st["step_dyn"].reserve(sizeof(StepDyn));
  P.status = (StepStatus*)st["step_status"].reserve(sizeof(StepStatus));
  P.reports = t->reports_dev;
  return P;
}

// The launches of one step, reading per-iteration values from P.dyn.
void enqueue_step(splat_trainer* t, const StepPlan& P, const uint8_t* any_target) {
  splat_ctx* ctx = t->ctx;
  const size_t n = P.n;
  splat_camera cam0;
  std::memset(&cam0, 0, sizeof(cam0));
  cam0.width = P.width;
  cam0.height = P.height;
  splat_proj_opts opts;
  opts.near_plane = P.near_plane;
  opts.eps2d = 0.3f;
  opts.antialias = P.antialias;
  opts.sh_degree = P.sh_degree;
  project_forward(ctx, P.params, n, P.stride, cam0, opts, P.pb, P.dyn);
  bin_build_device(ctx, P.pb, n, P.bb, P.entry_cap, P.small, P.total64, P.big_list, P.scan_sums,
                   P.status);
  RasterVariant fv, bv;
  fv.fwd_batch = P.fwd_batch;
  fv.persistent = P.fwd_persistent != 0;
  fv.per_cu = (unsigned)P.fwd_per_cu;
  bv.bwd_batch = P.bwd_batch;
  bv.persistent = P.bwd_persistent != 0;
  bv.per_cu = (unsigned)P.bwd_per_cu;
  bv.rows = P.bwd_rows != 0;
  raster_forward(ctx, P.pb, P.bb, P.width, P.height, P.bg, P.image, P.final_t, P.n_contrib, fv,
                 nullptr, P.counter);
  // the loss pass also writes the loss value and the step's report
  loss_forward_backward(ctx, P.image, nullptr, any_target, P.width, P.height, P.lambda, true,
                        P.dimage, P.loss_scratch, nullptr, P.dyn, P.reports, P.status);
  raster_backward(ctx, P.pb, P.bb, P.width, P.height, P.bg, P.final_t, P.n_contrib, P.dimage,
                  P.ordered != 0, P.slots, P.accumulate && P.absgrad, bv, nullptr, P.counter);
  StatsArgs sa{};
  if (P.accumulate) {
    sa.radius = P.pb.radius;
    sa.sx = (float)P.width * 0.5f;
    sa.sy = (float)P.height * 0.5f;
    sa.absgrad = P.absgrad;
    sa.grad_accum = P.grad_accum;
    sa.denom = P.denom;
    sa.max_radii = P.max_radii;
    sa.max_grad = P.max_grad;
    sa.strong = P.strong;
    sa.strong_threshold = P.strong_threshold;
  }
  splat_adam_state ast{0.9f, 0.999f, 1e-15f, 0, 0};
  splat_learning_rates lr0;
  std::memset(&lr0, 0, sizeof(lr0));
  grad_adam(ctx, P.params, P.m, P.v, n, P.stride, cam0, opts, P.bb.entry_offset, P.slots, sa,
            1.0f, 1.0f, ast, lr0, P.dyn, P.status);
}

hipGraphExec_t graph_for(splat_trainer* t, const StepPlan& P, const uint8_t* any_target) {
  for (auto& g : t->graphs)
    if (std::memcmp(&g.first, &P, sizeof(P)) == 0) return g.second;
  splat_ctx* ctx = t->ctx;
  const bool timing = ctx->timing;
  ctx->timing = false;
  SPLAT_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
  try {
    enqueue_step(t, P, any_target);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(ctx->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    ctx->timing = timing;
    throw;
  }
  ctx->timing = timing;
  hipGraph_t g = nullptr;
  SPLAT_HIP(hipStreamEndCapture(ctx->stream, &g));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  SPLAT_HIP(e);
  if (t->graphs.size() >= kMaxGraphs) {
    (void)hipGraphExecDestroy(t->graphs.front().second);
    t->graphs.erase(t->graphs.begin());
  }
  t->graphs.emplace_back(P, exec);
  t->graph_captures++;
  return exec;
}

void ensure_reports(splat_trainer* t) {
  if (t->reports_host) return;
  void* h = nullptr;
  SPLAT_HIP(hipHostMalloc(&h, kReportSlots * sizeof(StepReport),
                          hipHostMallocMapped | hipHostMallocCoherent));
  std::memset(h, 0xff, kReportSlots * sizeof(StepReport));
  t->reports_host = (StepReport*)h;
  void* d = nullptr;
  SPLAT_HIP(hipHostGetDevicePointer(&d, h, 0));
  t->reports_dev = (StepReport*)d;
}

void clear_abort(splat_trainer* t) {
  StepStatus* st = (StepStatus*)t->store["step_status"].reserve(sizeof(StepStatus));
  fill_u32(t->ctx, reinterpret_cast<uint32_t*>(st), sizeof(StepStatus) / 4, 0u);
}

void launch_graph_step(splat_trainer* t, uint32_t view, const splat_step_params& p,
                       uint64_t adam_step, float* loss_out) {
  splat_ctx* ctx = t->ctx;
  ensure_reports(t);
  const bool fresh = !t->store.count("step_status");
  const TrainerView& tv = *t->views[view];
  const StepPlan P = plan_step(t, tv.cam, p);
  if (fresh) clear_abort(t);
  hipGraphExec_t exec = graph_for(t, P, tv.rgb.as<uint8_t>());
  StepDyn d;
  std::memset(&d, 0, sizeof(d));
  d.cam = camera_dev(tv.cam);
  d.target = tv.rgb.as<uint8_t>();
  d.loss_out = loss_out ? loss_out : P.loss_one;
  d.lr = p.lr;
  d.c1 = adam_c(0.9f, adam_step);
  d.c2 = adam_c(0.999f, adam_step);
  d.seq = (uint32_t)t->seq;
  d.report_slot = (uint32_t)(t->seq % kReportSlots);
  const uint32_t ntc = kTileSub * P.bb.tiles_x * P.bb.tiles_y;
  launch(ctx, "step_args", set_dyn_kernel, dim3(blocks_for(std::max<uint32_t>(ntc, 3u), 256)),
         dim3(256), 0, d, P.dyn, reinterpret_cast<uint32_t*>(P.total64), P.bb.tile_count, ntc,
         P.small);
  SPLAT_HIP(hipGraphLaunch(exec, ctx->stream));
  t->graph_replays++;
  hipEvent_t ev;
  if (t->events.empty()) {
    SPLAT_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  } else {
    ev = t->events.back();
    t->events.pop_back();
  }
  SPLAT_HIP(hipEventRecord(ev, ctx->stream));
  t->pending.push_back(PendingStep{t->seq, view, p, adam_step, loss_out, ev});
  t->seq++;
}

// Waits for the oldest step in flight and checks its report; an aborted step (entry
// arrays too small) grows the arrays and replays it and every later step in order.
void verify_front(splat_trainer* t) {
  PendingStep pe = t->pending.front();
  SPLAT_HIP(hipEventSynchronize(pe.done));
  const volatile StepReport* r = t->reports_host + pe.seq % kReportSlots;
  const uint32_t seq = r->seq, abort = r->abort;
  const unsigned long long total = r->total;
  if (seq != (uint32_t)pe.seq) fail(SPLAT_ERR_HIP, "step report missing");
  if (!abort) {
    t->last_total = total;
    t->events.push_back(pe.done);
    t->pending.pop_front();
    return;
  }
  t->step_aborts++;
  SPLAT_HIP(hipStreamSynchronize(t->ctx->stream));
  SPLAT_REQUIRE(total < 0xffffffffull, "too many tile entries");
  t->entry_cap = std::max<size_t>(t->entry_cap, (size_t)(total + total / 4 + 1024));
  clear_abort(t);
  std::deque<PendingStep> redo;
  redo.swap(t->pending);
  for (auto& q : redo) {
    t->events.push_back(q.done);
    launch_graph_step(t, q.view, q.p, q.adam_step, q.loss_out);
  }
}

void drain(splat_trainer* t) {
  while (!t->pending.empty()) verify_front(t);
}

}  // namespace

namespace splat {

void trainer_drain(splat_trainer* t) { drain(t); }

// Per-view sums of squared clamped differences (src/loss.rs:175-190, in binary64) of the
// rendered views against their RGB8 targets (View::to_real, src/dataset.rs:23-25).
std::vector<double> trainer_view_sse(splat_trainer* t, const uint32_t* views, uint32_t count,
                                     int32_t sh_degree, int32_t antialias, const float bg[3]) {
  drain(t);
  splat_ctx* ctx = t->ctx;
  constexpr unsigned kBlocks = 512;
  double* partial = (double*)t->store["psnr_partial"].reserve(kBlocks * 8 * (size_t)count + 8);
  for (uint32_t q = 0; q < count; q++) {
    const TrainerView& tv = *t->views[views[q]];
    const uint32_t W = tv.cam.width, H = tv.cam.height;
    const size_t npx = (size_t)W * H;
    splat_proj_opts opts{0.2f, 0.3f, antialias, sh_degree};
    ProjBuffers pb = proj_buffers(t);
    project_forward(ctx, t->params.as<float>(), t->n, t->stride, tv.cam, opts, pb);
    BinBuffers bb{};
    bin_build(ctx, pb, t->n, (W + kTile - 1) / kTile, (H + kTile - 1) / kTile, bb,
              &t->entry_cap, t->store, nullptr);
    float* img = (float*)t->store["image"].reserve(npx * 12 + 4);
    float* final_t = (float*)t->store["final_t"].reserve(npx * 4 + 4);
    uint32_t* n_contrib = (uint32_t*)t->store["n_contrib"].reserve(npx * 4 + 4);
    raster_forward(ctx, pb, bb, W, H, bg, img, final_t, n_contrib);
    launch(ctx, "sq_err", sq_err_kernel, dim3(kBlocks), dim3(256), 0, (const float*)img,
           tv.rgb.as<uint8_t>(), npx * 3, partial + (size_t)q * kBlocks);
  }
  std::vector<double> h((size_t)kBlocks * count), sse(count, 0.0);
  if (count) {
    SPLAT_HIP(hipMemcpyAsync(h.data(), partial, h.size() * 8, hipMemcpyDeviceToHost,
                             ctx->stream));
    SPLAT_HIP(hipStreamSynchronize(ctx->stream));
  }
  for (uint32_t q = 0; q < count; q++)
    for (unsigned b = 0; b < kBlocks; b++) sse[q] += h[(size_t)q * kBlocks + b];
  return sse;
}

// dL/dimage of the contribution pass: 1 in the red channel of every pixel.
__global__ void __launch_bounds__(256) unit_red_kernel(float* __restrict__ d, size_t npx) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npx) return;
  d[3 * p] = 1.0f;
  d[3 * p + 1] = 0.0f;
  d[3 * p + 2] = 0.0f;
}

// contrib[i] (+)= the red colour gradient of row i, which under dL/dimage = (1, 0, 0) is
// the row's sum of alpha T over the view's pixels.
__global__ void __launch_bounds__(256) add_share_kernel(const splat_proj_grad* __restrict__ g,
                                                        uint32_t n, int first,
                                                        float* __restrict__ contrib) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float c = g[i].color[0];
  contrib[i] = first ? c : contrib[i] + c;
}

void trainer_image_share(splat_trainer* t, const uint32_t* views, uint32_t count,
                         int32_t sh_degree, float* contrib) {
  drain(t);
  splat_ctx* ctx = t->ctx;
  const size_t n = t->n;
  if (n == 0) return;
  const float black[3] = {0.0f, 0.0f, 0.0f};
  splat_proj_grad* grads = (splat_proj_grad*)t->store["share_grads"].reserve((n + 1) * 36);
  for (uint32_t q = 0; q < count; q++) {
    const TrainerView& tv = *t->views[views[q]];
    const uint32_t W = tv.cam.width, H = tv.cam.height;
    const size_t npx = (size_t)W * H;
    splat_proj_opts opts{0.2f, 0.3f, 0, sh_degree};
    ProjBuffers pb = proj_buffers(t);
    project_forward(ctx, t->params.as<float>(), n, t->stride, tv.cam, opts, pb);
    BinBuffers bb{};
    const size_t E = bin_build(ctx, pb, n, (W + kTile - 1) / kTile, (H + kTile - 1) / kTile, bb,
                               &t->entry_cap, t->store, nullptr);
    float* img = (float*)t->store["image"].reserve(npx * 12 + 4);
    float* final_t = (float*)t->store["final_t"].reserve(npx * 4 + 4);
    uint32_t* n_contrib = (uint32_t*)t->store["n_contrib"].reserve(npx * 4 + 4);
    float* dimage = (float*)t->store["share_dimage"].reserve(npx * 12 + 4);
    float* slots = (float*)t->store["share_slots"].reserve((E + 1) * kSlots * 4);
    raster_forward(ctx, pb, bb, W, H, black, img, final_t, n_contrib);
    launch(ctx, "share_dimage", unit_red_kernel, dim3(blocks_for(npx, 256)), dim3(256), 0, dimage,
           npx);
    raster_backward(ctx, pb, bb, W, H, black, final_t, n_contrib, dimage, false, slots, false);
    reduce_slots(ctx, bb.entry_offset, slots, n, grads, nullptr);
    launch(ctx, "share_add", add_share_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
           (const splat_proj_grad*)grads, (uint32_t)n, q == 0 ? 1 : 0, contrib);
  }
}

float trainer_eval_psnr(splat_trainer* t, uint32_t views, int32_t sh_degree, const float bg[3]) {
  views = std::min<uint32_t>(views, (uint32_t)t->views.size());
  if (views == 0) return 0.0f;
  std::vector<uint32_t> list(views);
  for (uint32_t v = 0; v < views; v++) list[v] = v;
  const std::vector<double> sse = trainer_view_sse(t, list.data(), views, sh_degree, 0, bg);
  double tot = 0;
  for (uint32_t v = 0; v < views; v++) {
    const TrainerView& tv = *t->views[v];
    const double mse = sse[v] / (double)((size_t)tv.cam.width * tv.cam.height * 3);
    tot += mse <= 0 ? 100.0 : -10.0 * std::log10(mse);
  }
  return (float)(tot / views);
}

void trainer_densify(splat_trainer* t, const splat_densify_config& cfg, float extent,
                     uint64_t seed, const uint8_t* decide, const uint32_t* sample_of,
                     uint32_t iteration, splat_densify_report* report) {
  drain(t);
  splat_ctx* ctx = t->ctx;
  const size_t n = t->n;
  *report = splat_densify_report{0, 0, 0};
  if (n == 0) return;
  const size_t children = cfg.split_children;
  const size_t cap2 = std::max(t->cap, n * (1 + std::max<size_t>(children, 1)));
  const size_t stride2 = plane_stride(cap2);
  // Output planes: the previous call's input planes, kept for reuse (no allocation once
  // they are large enough); swapped with the trainer's below.
  DevBuf &po = t->store["densify_params"], &mo = t->store["densify_m"],
         &vo = t->store["densify_v"], &bo = t->store["densify_birth"],
         &oo = t->store["densify_origin"];
  po.reserve(stride2 * kParams * 4);
  mo.reserve(stride2 * kParams * 4);
  vo.reserve(stride2 * kParams * 4);
  DensifyLineage ln{};
  if (t->track) {
    ln.birth_in = t->birth.as<uint32_t>();
    ln.origin_in = t->origin.as<uint32_t>();
    ln.sample_of = sample_of;
    ln.birth_out = (uint32_t*)bo.reserve((cap2 + 1) * 4);
    ln.origin_out = (uint32_t*)oo.reserve((cap2 + 1) * 4);
    ln.iteration = iteration;
  }
  const size_t k = densify_planes(ctx, t->params.as<float>(), t->m.as<float>(), t->v.as<float>(),
                                  n, t->stride, t->grad_accum.as<float>(), t->denom.as<uint32_t>(),
                                  t->max_radii.as<uint32_t>(), cfg, extent, nullptr, 0, seed,
                                  po.as<float>(), mo.as<float>(), vo.as<float>(), cap2, stride2,
                                  report, decide, t->track ? &ln : nullptr);
  SPLAT_REQUIRE(k != (size_t)-1 && k <= cap2, "densify output exceeds its buffer");
  SPLAT_HIP(hipStreamSynchronize(ctx->stream));
  std::swap(t->params.ptr, po.ptr);
  std::swap(t->params.bytes, po.bytes);
  std::swap(t->m.ptr, mo.ptr);
  std::swap(t->m.bytes, mo.bytes);
  std::swap(t->v.ptr, vo.ptr);
  std::swap(t->v.bytes, vo.bytes);
  if (t->track) {
    std::swap(t->birth.ptr, bo.ptr);
    std::swap(t->birth.bytes, bo.bytes);
    std::swap(t->origin.ptr, oo.ptr);
    std::swap(t->origin.bytes, oo.bytes);
    t->max_grad.reserve((cap2 + 1) * 4);
    t->strong.reserve((cap2 + 1) * 4);
  }
  t->cap = cap2;
  t->stride = stride2;
  t->grad_accum.reserve((cap2 + 1) * 4);
  t->denom.reserve((cap2 + 1) * 4);
  t->max_radii.reserve((cap2 + 1) * 4);
  t->n = k;
  zero_stats(t);  // src/densify.rs:219
  SPLAT_HIP(hipStreamSynchronize(ctx->stream));
}

}  // namespace splat

extern "C" {

int splat_trainer_create(splat_ctx* ctx, const float* gaussians, uint64_t n, uint64_t capacity,
                         splat_trainer** out) {
  if (!ctx || !out) return SPLAT_ERR_INVALID;
  *out = nullptr;
  splat_trainer* t = new (std::nothrow) splat_trainer();
  if (!t) return SPLAT_ERR_OOM;
  t->ctx = ctx;
  t->device = ctx->device;
  const int st = tguard(t, [&] {
    SPLAT_REQUIRE(n == 0 || gaussians, "null argument");
    t->n = n;
    alloc_model(t, std::max<size_t>(std::max<size_t>(capacity, n), 64));
    if (n) {
      float* staging = (float*)t->store["aos"].reserve(n * kParams * 4);
      SPLAT_HIP(hipMemcpyAsync(staging, gaussians, n * kParams * 4, hipMemcpyHostToDevice,
                               ctx->stream));
      aos_to_planes(ctx, staging, n, t->params.as<float>(), t->stride);
    }
    fill_u32(ctx, t->m.as<uint32_t>(), t->stride * kParams, 0u);
    fill_u32(ctx, t->v.as<uint32_t>(), t->stride * kParams, 0u);
    zero_stats(t);
    SPLAT_HIP(hipStreamSynchronize(ctx->stream));
    t->store.erase("aos");
  });
  if (st != SPLAT_OK) {
    delete t;
    return st;
  }
  *out = t;
  return SPLAT_OK;
}

int splat_trainer_destroy(splat_trainer* t) {
  if (!t) return SPLAT_OK;
  // The context may already be gone: use only the device id captured at creation.
  (void)hipSetDevice(t->device);
  (void)hipDeviceSynchronize();
  for (auto& g : t->graphs) (void)hipGraphExecDestroy(g.second);
  for (auto& pe : t->pending) (void)hipEventDestroy(pe.done);
  for (hipEvent_t e : t->events) (void)hipEventDestroy(e);
  if (t->reports_host) (void)hipHostFree(t->reports_host);
  delete t;
  return SPLAT_OK;
}

int splat_trainer_add_view(splat_trainer* t, const splat_camera* cam, const uint8_t* rgb,
                           uint32_t* index) {
  return tguard(t, [&] {
    SPLAT_REQUIRE(cam && rgb, "null argument");
    SPLAT_REQUIRE(cam->width > 0 && cam->height > 0 && cam->width <= 65535 * kTile &&
                      cam->height <= 65535 * kTile,
                  "bad image size");
    auto v = std::make_unique<TrainerView>();
    v->cam = *cam;
    const size_t bytes = (size_t)cam->width * cam->height * 3;
    v->rgb.reserve(bytes);
    SPLAT_HIP(hipMemcpyAsync(v->rgb.ptr, rgb, bytes, hipMemcpyHostToDevice, t->ctx->stream));
    SPLAT_HIP(hipStreamSynchronize(t->ctx->stream));
    t->views.push_back(std::move(v));
    if (index) *index = (uint32_t)(t->views.size() - 1);
  });
}

int splat_trainer_add_view_downscaled(splat_trainer* t, const splat_camera* cam,
                                      const uint8_t* rgb, uint32_t width, uint32_t height,
                                      uint32_t downscale, splat_camera* used,
                                      uint32_t* index) {
  return tguard(t, [&] {
    SPLAT_REQUIRE(cam && rgb, "null argument");
    if (downscale == 0) fail(SPLAT_ERR_INVALID, "downscale must be at least 1");
    // src/dataset.rs:159-184: the camera matched to the decoded image, then both scaled
    splat_camera matched;
    if (splat_camera_match_image_size(cam, width, height, &matched) != SPLAT_OK)
      fail(SPLAT_ERR_INVALID, "image of " + std::to_string(width) + "x" +
                                  std::to_string(height) + " does not match camera of " +
                                  std::to_string(cam->width) + "x" + std::to_string(cam->height));
    uint32_t w = 0, h = 0;
    (void)splat_downscaled_size(matched.width, matched.height, downscale, &w, &h);
    SPLAT_REQUIRE(w <= 65535 * kTile && h <= 65535 * kTile, "bad image size");
    splat_camera c = matched;  // Camera::resized, src/camera.rs:48-62
    const float sx = (float)w / (float)matched.width, sy = (float)h / (float)matched.height;
    c.width = w;
    c.height = h;
    c.fx = matched.fx * sx;
    c.fy = matched.fy * sy;
    c.cx = matched.cx * sx;
    c.cy = matched.cy * sy;
    splat_ctx* ctx = t->ctx;
    drain(t);
    auto v = std::make_unique<TrainerView>();
    v->cam = c;
    v->rgb.reserve((size_t)w * h * 3);
    const size_t in_bytes = (size_t)width * height * 3;
    if (downscale > 1) {
      uint8_t* staged = (uint8_t*)t->store["view_upload"].reserve(in_bytes + 4);
      SPLAT_HIP(hipMemcpyAsync(staged, rgb, in_bytes, hipMemcpyHostToDevice, ctx->stream));
      resize_rgb8_device(ctx, staged, width, height, w, h, v->rgb.as<uint8_t>());
    } else {
      SPLAT_HIP(hipMemcpyAsync(v->rgb.ptr, rgb, in_bytes, hipMemcpyHostToDevice, ctx->stream));
    }
    SPLAT_HIP(hipStreamSynchronize(ctx->stream));
    t->views.push_back(std::move(v));
    if (used) *used = c;
    if (index) *index = (uint32_t)(t->views.size() - 1);
  });
}

int splat_trainer_view(splat_trainer* t, uint32_t view, splat_camera* cam, uint8_t* rgb,
                       uint64_t capacity) {
  return tguard(t, [&] {
    SPLAT_REQUIRE(view < t->views.size(), "bad view index");
    const TrainerView& tv = *t->views[view];
    if (cam) *cam = tv.cam;
    if (!rgb) return;
    const size_t bytes = (size_t)tv.cam.width * tv.cam.height * 3;
    if (capacity < bytes) fail(SPLAT_ERR_CAPACITY, "output capacity too small");
    drain(t);
    SPLAT_HIP(hipMemcpyAsync(rgb, tv.rgb.ptr, bytes, hipMemcpyDeviceToHost, t->ctx->stream));
    SPLAT_HIP(hipStreamSynchronize(t->ctx->stream));
  });
}

int splat_trainer_step(splat_trainer* t, uint32_t view, const splat_step_params* p,
                       splat_step_result* r) {
  return tguard(t, [&] {
    SPLAT_REQUIRE(p && view < t->views.size(), "bad view index");
    SPLAT_REQUIRE(p->sh_degree >= 0 && p->sh_degree <= 3, "sh_degree must be 0..3");
    splat_ctx* ctx = t->ctx;
    if (t->use_graph) {
      t->adam_step += 1;
      launch_graph_step(t, view, *p, t->adam_step, t->loss_target);
      while (t->pending.size() > kStepLag) verify_front(t);
      if (!r) return;
      drain(t);
      r->num_entries = t->last_total;
      r->num_visible = 0;
      r->loss = 0.0f;
      if (p->want_loss) {
        const float* lp = t->loss_target ? t->loss_target : (const float*)t->store["step_loss"].ptr;
        uint32_t* cnt = (uint32_t*)t->store["visible_count"].reserve(4);
        fill_u32(ctx, cnt, 1, 0u);
        launch(ctx, "count_visible", count_visible_kernel,
               dim3(std::min(blocks_for(t->n, 256), kCountBlocks)), dim3(256), 0,
               (const uint32_t*)t->store["radius"].ptr, (uint32_t)t->n, cnt);
        uint32_t vis = 0;
        float loss = 0.0f;
        SPLAT_HIP(hipMemcpyAsync(&vis, cnt, 4, hipMemcpyDeviceToHost, ctx->stream));
        SPLAT_HIP(hipMemcpyAsync(&loss, lp, 4, hipMemcpyDeviceToHost, ctx->stream));
        SPLAT_HIP(hipStreamSynchronize(ctx->stream));
        r->num_visible = vis;
        r->loss = loss;
      }
      return;
    }
    drain(t);
    const TrainerView& tv = *t->views[view];
    const uint32_t W = tv.cam.width, H = tv.cam.height;
    const size_t npx = (size_t)W * H;
    const uint32_t tx = (W + kTile - 1) / kTile, ty = (H + kTile - 1) / kTile;
    const size_t n = t->n;
    splat_proj_opts opts;
    opts.near_plane = p->near_plane;
    opts.eps2d = 0.3f;
    opts.antialias = p->antialias;
    opts.sh_degree = p->sh_degree;

    // forward (src/train.rs:211-221)
    ProjBuffers pb = proj_buffers(t);
    project_forward(ctx, t->params.as<float>(), n, t->stride, tv.cam, opts, pb);
    BinBuffers bb{};
    const size_t E = bin_build(ctx, pb, n, tx, ty, bb, &t->entry_cap, t->store, nullptr);
    float* image = (float*)t->store["image"].reserve(npx * 12 + 4);
    float* final_t = (float*)t->store["final_t"].reserve(npx * 4 + 4);
    uint32_t* n_contrib = (uint32_t*)t->store["n_contrib"].reserve(npx * 4 + 4);
    uint32_t* counter = (uint32_t*)t->store["raster_counter"].reserve(64);
    raster_forward(ctx, pb, bb, W, H, p->background, image, final_t, n_contrib, t->fwd_shape,
                   nullptr, counter);

    // loss (src/train.rs:221-222)
    float* dimage = (float*)t->store["dimage"].reserve(npx * 12 + 4);
    float* scratch = (float*)t->store["loss_scratch"].reserve(loss_scratch_floats(W, H) * 4);
    double parts[2] = {0, 0};
    loss_forward_backward(ctx, image, nullptr, tv.rgb.as<uint8_t>(), W, H, p->lambda_dssim, true,
                          dimage, scratch, p->want_loss ? parts : nullptr);
    if (t->loss_target) loss_value_to(ctx, scratch, W, H, p->lambda_dssim, t->loss_target);

    // backward (src/train.rs:225-235)
    float* slots = (float*)t->store["slots"].reserve((E + 1) * kSlots * 4);
    raster_backward(ctx, pb, bb, W, H, p->background, final_t, n_contrib, dimage,
                    p->ordered_backward != 0, slots, p->accumulate_stats && p->absgrad,
                    t->bwd_shape, nullptr, counter);
    // slot reduce, statistics (src/train.rs:228-235) and project backward in one pass over
    // the Gaussians, then the optimiser step (src/train.rs:238-247, src/adam.rs:77-90)
    StatsArgs sa{};
    if (p->accumulate_stats) {
      sa.radius = pb.radius;
      sa.sx = (float)W * 0.5f;
      sa.sy = (float)H * 0.5f;
      sa.absgrad = p->absgrad;
      sa.grad_accum = t->grad_accum.as<float>();
      sa.denom = t->denom.as<uint32_t>();
      sa.max_radii = t->max_radii.as<uint32_t>();
      if (t->track) {
        sa.max_grad = t->max_grad.as<float>();
        sa.strong = t->strong.as<uint32_t>();
        sa.strong_threshold = t->strong_threshold;
      }
    }
    t->adam_step += 1;
    splat_adam_state st;
    st.beta1 = 0.9f;
    st.beta2 = 0.999f;
    st.eps = 1e-15f;
    st.step = t->adam_step;
    const float c1 = 1.0f - powi(st.beta1, (int32_t)st.step);
    const float c2 = 1.0f - powi(st.beta2, (int32_t)st.step);
    grad_adam(ctx, t->params.as<float>(), t->m.as<float>(), t->v.as<float>(), n, t->stride,
              tv.cam, opts, bb.entry_offset, slots, sa, c1, c2, st, p->lr);

    if (r) {
      r->num_entries = E;
      r->num_visible = 0;
      r->loss = 0.0f;
      if (p->want_loss) {
        uint32_t* cnt = (uint32_t*)t->store["visible_count"].reserve(4);
        fill_u32(ctx, cnt, 1, 0u);
        launch(ctx, "count_visible", count_visible_kernel,
               dim3(std::min(blocks_for(n, 256), kCountBlocks)), dim3(256), 0,
               (const uint32_t*)pb.radius, (uint32_t)n, cnt);
        uint32_t vis = 0;
        SPLAT_HIP(hipMemcpyAsync(&vis, cnt, 4, hipMemcpyDeviceToHost, ctx->stream));
        SPLAT_HIP(hipStreamSynchronize(ctx->stream));
        r->num_visible = vis;
        const float inv_n = 1.0f / (float)(npx * 3);
        const float l1 = (float)parts[1] * inv_n;
        const float ssim = (float)parts[0] / (float)(npx * 3);
        r->loss = (1.0f - p->lambda_dssim) * l1 + p->lambda_dssim * (1.0f - ssim);
      }
    }
  });
}

int splat_trainer_densify(splat_trainer* t, const splat_densify_config* cfg, float extent,
                          uint64_t seed, splat_densify_report* report) {
  return tguard(t, [&] {
    SPLAT_REQUIRE(cfg && report, "null argument");
    trainer_densify(t, *cfg, extent, seed, nullptr, nullptr, 0, report);
  });
}

int splat_trainer_reset_opacity(splat_trainer* t, float max_alpha) {
  return tguard(t, [&] {
    drain(t);
    const float limit = logf(max_alpha / (1.0f - max_alpha));  // inverse_sigmoid
    reset_opacity_planes(t->ctx, t->params.as<float>(), t->m.as<float>(), t->v.as<float>(), t->n,
                         t->stride, limit);
  });
}

int splat_trainer_render(splat_trainer* t, uint32_t view, int32_t sh_degree, int32_t antialias,
                         const float background[3], float* image) {
  return tguard(t, [&] {
    SPLAT_REQUIRE(view < t->views.size() && background && image, "bad argument");
    SPLAT_REQUIRE(sh_degree >= 0 && sh_degree <= 3, "sh_degree must be 0..3");
    drain(t);
    splat_ctx* ctx = t->ctx;
    const splat_camera& cam = t->views[view]->cam;
    const uint32_t W = cam.width, H = cam.height;
    const size_t npx = (size_t)W * H;
    splat_proj_opts opts{0.2f, 0.3f, antialias, sh_degree};
    ProjBuffers pb = proj_buffers(t);
    project_forward(ctx, t->params.as<float>(), t->n, t->stride, cam, opts, pb);
    BinBuffers bb{};
    bin_build(ctx, pb, t->n, (W + kTile - 1) / kTile, (H + kTile - 1) / kTile, bb, &t->entry_cap,
              t->store, nullptr);
    float* img = (float*)t->store["image"].reserve(npx * 12 + 4);
    float* final_t = (float*)t->store["final_t"].reserve(npx * 4 + 4);
    uint32_t* n_contrib = (uint32_t*)t->store["n_contrib"].reserve(npx * 4 + 4);
    raster_forward(ctx, pb, bb, W, H, background, img, final_t, n_contrib);
    SPLAT_HIP(hipMemcpyAsync(image, img, npx * 12, hipMemcpyDeviceToHost, ctx->stream));
    SPLAT_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int splat_trainer_set_graph(splat_trainer* t, int32_t enabled) {
  return tguard(t, [&] {
    drain(t);
    t->use_graph = enabled != 0;
  });
}

int splat_trainer_tune(splat_trainer* t, uint32_t view, const splat_step_params* p,
                       splat_tune_result* out) {
  return tguard(t, [&] {
    SPLAT_REQUIRE(p && out && view < t->views.size(), "bad argument");
    SPLAT_REQUIRE(p->sh_degree >= 0 && p->sh_degree <= 3, "sh_degree must be 0..3");
    drain(t);
    splat_ctx* ctx = t->ctx;
    const TrainerView& tv = *t->views[view];
    const uint32_t W = tv.cam.width, H = tv.cam.height;
    const size_t npx = (size_t)W * H;
    // the forward half of a step (the model is not changed) fills the tuning buffers
    ProjBuffers pb = proj_buffers(t);
    project_forward(ctx, t->params.as<float>(), t->n, t->stride, tv.cam, step_opts(*p), pb);
    BinBuffers bb{};
    const size_t E = bin_build(ctx, pb, t->n, (W + kTile - 1) / kTile, (H + kTile - 1) / kTile,
                               bb, &t->entry_cap, t->store, nullptr);
    float* image = (float*)t->store["image"].reserve(npx * 12 + 4);
    float* final_t = (float*)t->store["final_t"].reserve(npx * 4 + 4);
    uint32_t* n_contrib = (uint32_t*)t->store["n_contrib"].reserve(npx * 4 + 4);
    float* dimage = (float*)t->store["dimage"].reserve(npx * 12 + 4);
    float* scratch = (float*)t->store["loss_scratch"].reserve(loss_scratch_floats(W, H) * 4);
    float* slots = (float*)t->store["slots"].reserve((std::max(E, t->entry_cap) + 1) * kSlots * 4);
    uint32_t* counter = (uint32_t*)t->store["raster_counter"].reserve(64);
    raster_forward(ctx, pb, bb, W, H, p->background, image, final_t, n_contrib);
    loss_forward_backward(ctx, image, nullptr, tv.rgb.as<uint8_t>(), W, H, p->lambda_dssim, true,
                          dimage, scratch, nullptr);
    const RasterTune tr = tune_raster(ctx, pb, bb, W, H, p->background, image, final_t,
                                      n_contrib, dimage, p->ordered_backward != 0, slots,
                                      p->accumulate_stats && p->absgrad, nullptr, counter);
    t->fwd_shape = tr.fwd;
    t->bwd_shape = tr.bwd;
    fill_tune_result(tr, out);
  });
}

int splat_trainer_pair_counts(splat_trainer* t, uint32_t view, const splat_step_params* p,
                              uint64_t counts[9]) {
  return tguard(t, [&] {
    SPLAT_REQUIRE(p && counts && view < t->views.size(), "bad argument");
    SPLAT_REQUIRE(p->sh_degree >= 0 && p->sh_degree <= 3, "sh_degree must be 0..3");
    drain(t);
    splat_ctx* ctx = t->ctx;
    const TrainerView& tv = *t->views[view];
    const uint32_t W = tv.cam.width, H = tv.cam.height;
    const size_t npx = (size_t)W * H;
    // the compositing half of a step; the model is not changed
    ProjBuffers pb = proj_buffers(t);
    project_forward(ctx, t->params.as<float>(), t->n, t->stride, tv.cam, step_opts(*p), pb);
    BinBuffers bb{};
    const size_t E = bin_build(ctx, pb, t->n, (W + kTile - 1) / kTile, (H + kTile - 1) / kTile,
                               bb, &t->entry_cap, t->store, nullptr);
    float* image = (float*)t->store["image"].reserve(npx * 12 + 4);
    float* final_t = (float*)t->store["final_t"].reserve(npx * 4 + 4);
    uint32_t* n_contrib = (uint32_t*)t->store["n_contrib"].reserve(npx * 4 + 4);
    float* dimage = (float*)t->store["dimage"].reserve(npx * 12 + 4);
    float* scratch = (float*)t->store["loss_scratch"].reserve(loss_scratch_floats(W, H) * 4);
    float* slots = (float*)t->store["slots"].reserve((std::max(E, t->entry_cap) + 1) * kSlots * 4);
    auto* pairs = (unsigned long long*)t->store["pair_counts"].reserve(16 * 8);
    raster_forward_counted(ctx, pb, bb, W, H, p->background, image, final_t, n_contrib, pairs);
    loss_forward_backward(ctx, image, nullptr, tv.rgb.as<uint8_t>(), W, H, p->lambda_dssim, true,
                          dimage, scratch, nullptr);
    raster_backward_counted(ctx, pb, bb, W, H, p->background, final_t, n_contrib, dimage, slots,
                            pairs);
    SPLAT_HIP(hipMemcpyAsync(counts, pairs, 9 * 8, hipMemcpyDeviceToHost, ctx->stream));
    SPLAT_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int splat_trainer_set_shapes(splat_trainer* t, const splat_tune_result* shapes) {
  return tguard(t, [&] {
    SPLAT_REQUIRE(shapes, "null argument");
    SPLAT_REQUIRE((shapes->fwd_batch == 32 || shapes->fwd_batch == 64 ||
                   shapes->fwd_batch == 128 || shapes->fwd_batch == 256) &&
                      (shapes->bwd_batch == 16 || shapes->bwd_batch == 32),
                  "unsupported batch shape");
    drain(t);
    t->fwd_shape.fwd_batch = shapes->fwd_batch;
    t->fwd_shape.persistent = shapes->fwd_persistent != 0;
    t->fwd_shape.per_cu = (unsigned)std::max(1, shapes->fwd_per_cu);
    t->bwd_shape.bwd_batch = shapes->bwd_batch;
    t->bwd_shape.persistent = shapes->bwd_persistent != 0;
    t->bwd_shape.per_cu = (unsigned)std::max(1, shapes->bwd_per_cu);
    t->bwd_shape.rows = shapes->bwd_row_lists != 0;
  });
}

int splat_trainer_sync(splat_trainer* t) {
  return tguard(t, [&] {
    drain(t);
    SPLAT_HIP(hipStreamSynchronize(t->ctx->stream));
  });
}

int splat_trainer_reserve_entries(splat_trainer* t, uint64_t entries) {
  return tguard(t, [&] {
    SPLAT_REQUIRE(entries > 0 && entries < 0xffffffffull, "bad entry capacity");
    drain(t);
    t->entry_cap = (size_t)entries;
  });
}

int splat_trainer_graph_stats(splat_trainer* t, uint64_t* captures, uint64_t* replays,
                              uint64_t* aborts) {
  if (!t) return SPLAT_ERR_INVALID;
  if (captures) *captures = t->graph_captures;
  if (replays) *replays = t->graph_replays;
  if (aborts) *aborts = t->step_aborts;
  return SPLAT_OK;
}

int splat_trainer_info(splat_trainer* t, uint64_t* n, uint64_t* adam_step) {
  if (!t) return SPLAT_ERR_INVALID;
  if (n) *n = t->n;
  if (adam_step) *adam_step = t->adam_step;
  return SPLAT_OK;
}

int splat_trainer_download(splat_trainer* t, float* gaussians, float* m, float* v,
                           float* grad_accum, uint32_t* denom, uint32_t* max_radii,
                           uint64_t capacity) {
  return tguard(t, [&] {
    SPLAT_REQUIRE(capacity >= t->n, "output capacity too small");
    drain(t);
    splat_ctx* ctx = t->ctx;
    const size_t n = t->n;
    if (n == 0) return;
    float* staging = (float*)t->store["aos_out"].reserve(n * kParams * 4);
    const std::pair<float*, const DevBuf*> arrays[3] = {{gaussians, &t->params}, {m, &t->m},
                                                        {v, &t->v}};
    for (auto& a : arrays) {
      if (!a.first) continue;
      planes_to_aos(ctx, a.second->as<float>(), n, t->stride, staging);
      SPLAT_HIP(hipMemcpyAsync(a.first, staging, n * kParams * 4, hipMemcpyDeviceToHost,
                               ctx->stream));
      SPLAT_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (grad_accum)
      SPLAT_HIP(hipMemcpyAsync(grad_accum, t->grad_accum.ptr, n * 4, hipMemcpyDeviceToHost,
                               ctx->stream));
    if (denom)
      SPLAT_HIP(hipMemcpyAsync(denom, t->denom.ptr, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (max_radii)
      SPLAT_HIP(hipMemcpyAsync(max_radii, t->max_radii.ptr, n * 4, hipMemcpyDeviceToHost,
                               ctx->stream));
    SPLAT_HIP(hipStreamSynchronize(ctx->stream));
    t->store.erase("aos_out");
  });
}

int splat_trainer_evaluate(splat_trainer* t, const uint32_t* views, uint32_t count,
                           int32_t sh_degree, int32_t antialias, const float* background,
                           float* psnr, float* mean) {
  return tguard(t, [&] {
    SPLAT_REQUIRE(count == 0 || views, "null argument");
    SPLAT_REQUIRE(sh_degree >= 0 && sh_degree <= 3, "sh_degree must be 0..3");
    for (uint32_t q = 0; q < count; q++)
      SPLAT_REQUIRE(views[q] < t->views.size(), "view index out of range");
    const float zero[3] = {0.0f, 0.0f, 0.0f};
    const std::vector<double> sse =
        trainer_view_sse(t, views, count, sh_degree, antialias, background ? background : zero);
    // evaluate, src/train.rs:309-321: binary32 PSNR per view, summed in order
    float total = 0.0f;
    for (uint32_t q = 0; q < count; q++) {
      const TrainerView& tv = *t->views[views[q]];
      const double mse = sse[q] / (double)std::max<size_t>((size_t)tv.cam.width * tv.cam.height * 3, 1);
      const float p = mse <= 0.0 ? INFINITY : (float)(-10.0 * std::log10(mse));
      if (psnr) psnr[q] = p;
      total += p;
    }
    if (mean) *mean = count ? total / (float)count : NAN;
  });
}

int splat_trainer_restore(splat_trainer* t, const float* gaussians, const float* m,
                          const float* v, const float* grad_accum, const uint32_t* denom,
                          const uint32_t* max_radii, uint64_t n, uint64_t adam_step) {
  return tguard(t, [&] {
    SPLAT_REQUIRE(n == 0 || gaussians, "null argument");
    drain(t);
    splat_ctx* ctx = t->ctx;
    if (n > t->cap) alloc_model(t, n);
    t->n = n;
    t->adam_step = adam_step;
    if (n) {
      float* staging = (float*)t->store["aos"].reserve(n * kParams * 4);
      const std::pair<const float*, DevBuf*> arrays[3] = {{gaussians, &t->params}, {m, &t->m},
                                                          {v, &t->v}};
      for (auto& a : arrays) {
        if (a.first) {
          SPLAT_HIP(hipMemcpyAsync(staging, a.first, n * kParams * 4, hipMemcpyHostToDevice,
                                   ctx->stream));
          aos_to_planes(ctx, staging, n, a.second->as<float>(), t->stride);
          SPLAT_HIP(hipStreamSynchronize(ctx->stream));
        } else {
          fill_u32(ctx, a.second->as<uint32_t>(), t->stride * kParams, 0u);
        }
      }
    }
    zero_stats(t);
    if (t->track) learn_track_alloc(t);  // sized for the capacity; no lineage
    if (n && grad_accum)
      SPLAT_HIP(hipMemcpyAsync(t->grad_accum.ptr, grad_accum, n * 4, hipMemcpyHostToDevice,
                               ctx->stream));
    if (n && denom)
      SPLAT_HIP(hipMemcpyAsync(t->denom.ptr, denom, n * 4, hipMemcpyHostToDevice, ctx->stream));
    if (n && max_radii)
      SPLAT_HIP(hipMemcpyAsync(t->max_radii.ptr, max_radii, n * 4, hipMemcpyHostToDevice,
                               ctx->stream));
    SPLAT_HIP(hipStreamSynchronize(ctx->stream));
    t->store.erase("aos");
  });
}

}  // extern "C"

// ---------------------------------------------------------------- train loop -------

namespace {

uint64_t sm64_next(uint64_t& state) {  // SplitMix64
  uint64_t z = (state += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Per-epoch shuffled view order popped from the back (src/train.rs:203-207).  The
// reference shuffles with ChaCha8; this draws a Fisher-Yates shuffle from SplitMix64.
struct ViewOrder {
  uint64_t state;
  uint32_t n;
  std::vector<uint32_t> order;
  ViewOrder(uint64_t seed, uint32_t num) : state(seed), n(num) {}
  uint32_t next() {
    if (order.empty()) {
      order.resize(n);
      for (uint32_t i = 0; i < n; i++) order[i] = i;
      for (uint32_t i = n; i-- > 1;) {
        const uint32_t j = (uint32_t)(sm64_next(state) % (uint64_t)(i + 1));
        std::swap(order[i], order[j]);
      }
    }
    const uint32_t v = order.back();
    order.pop_back();
    return v;
  }
};

// TrainConfig::position_lr, src/train.rs:132-137 (glibc logf/expf, as the reference).
float position_lr(const splat_train_config& c, uint32_t it, float extent) {
  float t = (float)it / (float)std::max<uint32_t>(c.position_lr_max_steps, 1u);
  t = rclamp(t, 0.0f, 1.0f);
  const float init = c.position_lr_init * extent;
  const float fin = c.position_lr_final * extent;
  return expf(logf(init) * (1.0f - t) + logf(fin) * t);
}

// psnr, src/loss.rs:175-190, against RGB8 (View::to_real).
float psnr_u8(const std::vector<float>& img, const std::vector<uint8_t>& gt) {
  double mse = 0;
  for (size_t i = 0; i < img.size(); i++) {
    const double d = (double)rclamp(img[i], 0.0f, 1.0f) -
                     (double)rclamp((float)gt[i] / 255.0f, 0.0f, 1.0f);
    mse += d * d;
  }
  mse /= (double)std::max<size_t>(img.size(), 1);
  return mse <= 0 ? INFINITY : (float)(-10.0 * std::log10(mse));
}

}  // namespace

extern "C" {

void splat_train_config_default(splat_train_config* c) {  // src/train.rs:76-109
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->iterations = 30000;
  c->position_lr_init = 0.00016f;
  c->position_lr_final = 0.0000016f;
  c->position_lr_max_steps = 30000;
  c->feature_lr = 0.0025f;
  c->opacity_lr = 0.025f;
  c->scaling_lr = 0.005f;
  c->rotation_lr = 0.001f;
  c->lambda_dssim = 0.2f;
  c->densify = 1;
  c->densify_from_iter = 500;
  c->densify_until_iter = 15000;
  c->densification_interval = 100;
  c->opacity_reset_interval = 3000;
  c->densify_grad_threshold = 0.0002f;
  c->percent_dense = 0.01f;
  c->min_opacity = 0.005f;
  c->sh_degree = 3;
  c->sh_degree_interval = 1000;
  c->near_plane = 0.2f;
  c->seed = 0;
  c->log_interval = 100;
}

uint32_t splat_train_view_at(uint64_t seed, uint32_t num_views, uint32_t it) {
  if (num_views == 0 || it == 0) return 0;
  ViewOrder o(seed, num_views);
  uint32_t v = 0;
  for (uint32_t i = 1; i <= it; i++) v = o.next();
  return v;
}

int splat_train(splat_trainer* t, float extent, const splat_train_config* cfg,
                splat_progress_fn on_progress, splat_save_fn on_save, void* user) {
  if (!t || !cfg) return SPLAT_ERR_INVALID;
  splat_ctx* ctx = t->ctx;
  // src/train.rs:169-183: validation before anything runs or saves
  if (t->views.empty()) {
    ctx->err = "no training views";
    return SPLAT_ERR_INVALID;
  }
  if (cfg->iterations == 0) {
    ctx->err = "iterations must be at least 1";
    return SPLAT_ERR_INVALID;
  }
  if (cfg->sh_degree < 0 || cfg->sh_degree > 3) {
    ctx->err = "sh_degree exceeds the maximum of 3";
    return SPLAT_ERR_INVALID;
  }
  const uint32_t first_it = std::max<uint32_t>(cfg->start_iteration, 1u);
  if (first_it > cfg->iterations) {
    ctx->err = "start_iteration is past the last iteration";
    return SPLAT_ERR_INVALID;
  }
  if (first_it > 1 && t->learn) {
    ctx->err = "a run with an attached scorer cannot resume";
    return SPLAT_ERR_INVALID;
  }
  int status = SPLAT_OK;
  const int st = tguard(t, [&] {
    const uint32_t log_iv = cfg->log_interval;
    // Kept in the trainer's store: steps still in flight write into it.
    float* hist = (float*)t->store["loss_hist"].reserve((size_t)(log_iv ? log_iv : 1) * 4 + 4);
    ViewOrder order(cfg->seed, (uint32_t)t->views.size());
    t->strong_threshold = cfg->densify_grad_threshold;
    uint32_t last_reset = 0;
    const float bg = cfg->white_background ? 1.0f : 0.0f;
    int32_t sh = 0;  // model.sh_degree starts at 0 (src/train.rs:186-187)
    const auto sh_steps = [&](uint32_t it) {
      return cfg->sh_degree_interval > 0 && it % cfg->sh_degree_interval == 0 &&
             sh < cfg->sh_degree;
    };
    const auto resets_at = [&](uint32_t it) {
      return cfg->densify && it < cfg->densify_until_iter && cfg->opacity_reset_interval > 0 &&
             (it % cfg->opacity_reset_interval == 0 ||
              (cfg->white_background && it == cfg->densify_from_iter));
    };
    // A resumed run replays the schedule state of iterations 1 .. first_it - 1.
    for (uint32_t it = 1; it < first_it; it++) {
      if (sh_steps(it)) sh++;
      order.next();
      if (resets_at(it)) last_reset = it;
    }
    float ema = 0.0f;
    uint64_t dc = 0, ds = 0, dp = 0;
    uint32_t since_log = 0;
    auto timer = std::chrono::steady_clock::now();
    std::vector<float> losses;
    for (uint32_t it = first_it; it <= cfg->iterations; it++) {
      if (sh_steps(it)) sh++;
      const uint32_t view = order.next();
      const bool in_window = cfg->densify && it < cfg->densify_until_iter;
      splat_step_params p;
      std::memset(&p, 0, sizeof(p));
      p.sh_degree = sh;
      p.accumulate_stats = in_window;
      p.absgrad = cfg->absgrad;
      p.ordered_backward = cfg->ordered_backward;
      p.antialias = cfg->antialias;
      p.near_plane = cfg->near_plane;
      p.lambda_dssim = cfg->lambda_dssim;
      p.background[0] = p.background[1] = p.background[2] = bg;
      p.lr.mean = position_lr(*cfg, it, extent);
      p.lr.log_scale = cfg->scaling_lr;
      p.lr.quat = cfg->rotation_lr;
      p.lr.opacity = cfg->opacity_lr;
      p.lr.sh_dc = cfg->feature_lr;
      p.lr.sh_rest = cfg->feature_lr / 20.0f;
      t->loss_target = log_iv ? hist + since_log % log_iv : nullptr;
      const int s1 = splat_trainer_step(t, view, &p, nullptr);
      t->loss_target = nullptr;
      if (s1 != SPLAT_OK) fail(s1, ctx->err);
      since_log++;
      if (in_window) {  // density control, src/train.rs:249-271
        if (it > cfg->densify_from_iter && cfg->densification_interval > 0 &&
            it % cfg->densification_interval == 0) {
          splat_densify_config dcfg;
          dcfg.grad_threshold = cfg->densify_grad_threshold;
          dcfg.percent_dense = cfg->percent_dense;
          dcfg.min_opacity = cfg->min_opacity;
          dcfg.prune_big_world = cfg->opacity_reset_interval > 0 && it > cfg->opacity_reset_interval;
          dcfg.max_world_size_frac = 0.1f;
          dcfg.has_max_screen_size = cfg->has_max_screen_size;
          dcfg.max_screen_size = cfg->max_screen_size;
          dcfg.split_children = 2;
          uint64_t s = cfg->seed ^ (0x5851f42d4c957f2dull * (uint64_t)it);
          splat_densify_report rep;
          if (t->learn) {
            learn_densify(t, dcfg, extent, sm64_next(s), it, sh, p.background, last_reset,
                          &rep);
          } else {
            const int s2 = splat_trainer_densify(t, &dcfg, extent, sm64_next(s), &rep);
            if (s2 != SPLAT_OK) fail(s2, ctx->err);
          }
          dc += rep.cloned;
          ds += rep.split;
          dp += rep.pruned;
        }
        if (resets_at(it)) {
          const int s3 = splat_trainer_reset_opacity(t, 0.01f);
          if (s3 != SPLAT_OK) fail(s3, ctx->err);
          last_reset = it;
        }
      }
      const bool log_now = log_iv > 0 && (it % log_iv == 0 || it == cfg->iterations);
      if (log_now) {
        // EMA over every iteration's loss (src/train.rs:274-278), read back in one copy.
        losses.resize(since_log);
        drain(t);
        SPLAT_HIP(hipMemcpyAsync(losses.data(), hist, since_log * 4, hipMemcpyDeviceToHost,
                                 ctx->stream));
        SPLAT_HIP(hipStreamSynchronize(ctx->stream));
        const uint32_t first = it - since_log + 1;
        for (uint32_t k = 0; k < since_log; k++)
          ema = (first + k == first_it) ? losses[k] : 0.4f * losses[k] + 0.6f * ema;
        if (log_now && on_progress) {
          const auto now = std::chrono::steady_clock::now();
          const double ms = std::chrono::duration<double, std::milli>(now - timer).count();
          const TrainerView& tv = *t->views[view];
          const size_t npx = (size_t)tv.cam.width * tv.cam.height;
          std::vector<float> img(npx * 3);
          std::vector<uint8_t> gt(npx * 3);
          SPLAT_HIP(hipMemcpyAsync(img.data(), t->store["image"].ptr, npx * 12,
                                   hipMemcpyDeviceToHost, ctx->stream));
          SPLAT_HIP(hipMemcpyAsync(gt.data(), tv.rgb.ptr, npx * 3, hipMemcpyDeviceToHost,
                                   ctx->stream));
          SPLAT_HIP(hipStreamSynchronize(ctx->stream));
          splat_progress pr;
          pr.iteration = it;
          pr.loss = losses[since_log - 1];
          pr.ema_loss = ema;
          pr.psnr = psnr_u8(img, gt);
          pr.num_gaussians = t->n;
          pr.ms_per_iter = ms / std::max<uint32_t>(since_log, 1);
          pr.cloned = dc;
          pr.split = ds;
          pr.pruned = dp;
          on_progress(&pr, user);
          dc = ds = dp = 0;
          timer = std::chrono::steady_clock::now();
        }
        since_log = 0;
      }
      bool save = false;
      for (uint32_t k = 0; k < cfg->num_save_iterations; k++)
        if (cfg->save_iterations && cfg->save_iterations[k] == it) save = true;
      if (save && it != cfg->iterations && on_save) {
        drain(t);
        SPLAT_HIP(hipStreamSynchronize(ctx->stream));
        if (on_save(it, t, user) != 0) {
          status = SPLAT_ERR_IO;
          return;
        }
      }
    }
    drain(t);
    SPLAT_HIP(hipStreamSynchronize(ctx->stream));
    if (on_save && on_save(cfg->iterations, t, user) != 0) status = SPLAT_ERR_IO;
  });
  t->loss_target = nullptr;
  if (st != SPLAT_OK) return st;
  if (status != SPLAT_OK) ctx->err = "on_save failed";
  return status;
}

}  // extern "C"
