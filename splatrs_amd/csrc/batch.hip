// Batched training: many independent scenes and many views of uneven resolution trained in
// one step, each stage one launch over every (scene, view) pair, the whole step captured
// in a HIP graph (no counterpart in the reference, which trains one view of one scene per
// iteration, src/train.rs:211-247).
//
// Semantics: a step renders every listed view of its scene, computes the reference's
// photometric loss and gradient per view, sums each scene's parameter gradients over its
// views in list order, and applies one Adam step (src/adam.rs:77-90) to every scene with
// the scene's own position learning rate.  A batch of one view of one scene is the
// reference's iteration.
//
// Layout: all scenes share one set of parameter planes; scene s owns rows [off_s, off_s +
// n_s) (off_s a multiple of 64).  Projection rows, tiles and pixels of the pairs are laid
// end to end (BatchDev).  Launch shapes of the compositing kernels are chosen by
// splat_batch_tune, which times the candidates on the batch itself.
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "internal.hpp"

using namespace splat;

namespace {

constexpr uint32_t kMaxScenes = 64;

struct BatchScene {
  size_t n = 0, off = 0;
};

struct BatchView {
  uint32_t scene;
  splat_camera cam;
  DevBuf rgb;
};

// Per-step values written before each replay: Adam corrections and rates, and the
// scenes' position learning rates (BatchDev::lr_mean points at lr_mean).
struct BatchDyn {
  StepDyn step;
  float lr_mean[kMaxScenes];
};

__global__ void set_batch_dyn_kernel(BatchDyn d, BatchDyn* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = d;
}

// Everything a captured batched step bakes in.
struct BatchKey {
  std::vector<uint32_t> views;
  int32_t sh_degree, ordered, antialias;
  float near_plane, lambda, bg[3];
  RasterVariant fwd, bwd;
  size_t entry_cap;
  void* params;
  bool operator==(const BatchKey& o) const {
    return views == o.views && sh_degree == o.sh_degree && ordered == o.ordered &&
           antialias == o.antialias && near_plane == o.near_plane && lambda == o.lambda &&
           bg[0] == o.bg[0] && bg[1] == o.bg[1] && bg[2] == o.bg[2] &&
           fwd.fwd_batch == o.fwd.fwd_batch && fwd.persistent == o.fwd.persistent &&
           fwd.per_cu == o.fwd.per_cu && bwd.bwd_batch == o.bwd.bwd_batch &&
           bwd.persistent == o.bwd.persistent && bwd.per_cu == o.bwd.per_cu &&
           bwd.rows == o.bwd.rows &&
           entry_cap == o.entry_cap && params == o.params;
  }
};

}  // namespace

struct splat_batch {
  splat_ctx* ctx = nullptr;
  int device = 0;
  std::vector<BatchScene> scenes;
  std::vector<std::unique_ptr<BatchView>> views;
  size_t rows = 0, stride = 0;
  uint64_t adam_step = 0;
  bool started = false;
  DevBuf params, m, v;
  std::map<std::string, DevBuf> store;
  size_t entry_cap = 0;
  RasterVariant fwd, bwd;  // launch shapes (splat_batch_tune)
  bool tuned = false;
  // current step layout
  std::vector<uint32_t> plan_views;
  BatchDev bd{};
  size_t work = 0, tiles = 0, pixels = 0;
  unsigned loss_blocks_total = 0;
  // captured step
  BatchKey key;
  hipGraphExec_t exec = nullptr;
  uint64_t captures = 0, replays = 0, aborts = 0;
};

namespace {

template <class F>
int bguard(splat_batch* b, F&& f) {
  if (!b || !b->ctx) return SPLAT_ERR_INVALID;
  return ctx_guard(b->ctx, std::forward<F>(f));
}

// Host tables of a view list -> device (BatchDev with device pointers).
void build_plan(splat_batch* b, const std::vector<uint32_t>& views) {
  const uint32_t P = (uint32_t)views.size(), S = (uint32_t)b->scenes.size();
  std::vector<uint32_t> pair_scene(P), work_base(P + 1), tile_base(P + 1), pix_base(P + 1),
      tiles_x(P), width(P), height(P), scene_off(S + 1), scene_n(S), spb(S + 1), spairs(P),
      lbb(P + 1);
  std::vector<CameraDev> cams(P);
  std::vector<uint64_t> targets(P);
  size_t work = 0, tiles = 0, pix = 0, lblocks = 0;
  for (uint32_t k = 0; k < P; k++) {
    const BatchView& v = *b->views[views[k]];
    const uint32_t W = v.cam.width, H = v.cam.height;
    pair_scene[k] = v.scene;
    work_base[k] = (uint32_t)work;
    tile_base[k] = (uint32_t)tiles;
    pix_base[k] = (uint32_t)pix;
    lbb[k] = (uint32_t)lblocks;
    tiles_x[k] = (W + kTile - 1) / kTile;
    SPLAT_REQUIRE((size_t)tiles_x[k] * ((H + kTile - 1) / kTile) < kMaxViewTiles,
                  "view too large for radix binning");
    width[k] = W;
    height[k] = H;
    cams[k] = camera_dev(v.cam);
    targets[k] = (uint64_t)(uintptr_t)v.rgb.ptr;
    work += b->scenes[v.scene].n;
    tiles += (size_t)tiles_x[k] * ((H + kTile - 1) / kTile);
    pix += (size_t)W * H;
    lblocks += loss_blocks(W, H);
  }
  SPLAT_REQUIRE(work < 0xffffffffull && pix < 0xffffffffull / 3 && tiles < 0xffffffffull,
                "batch too large");
  work_base[P] = (uint32_t)work;
  tile_base[P] = (uint32_t)tiles;
  pix_base[P] = (uint32_t)pix;
  lbb[P] = (uint32_t)lblocks;
  uint32_t q = 0;
  for (uint32_t s = 0; s < S; s++) {
    scene_off[s] = (uint32_t)b->scenes[s].off;
    scene_n[s] = (uint32_t)b->scenes[s].n;
    spb[s] = q;
    for (uint32_t k = 0; k < P; k++)
      if (pair_scene[k] == s) spairs[q++] = k;
  }
  spb[S] = q;
  scene_off[S] = (uint32_t)b->rows;
  // one device table buffer
  std::vector<uint8_t> blob;
  auto put = [&](const void* src, size_t bytes) {
    const size_t at = (blob.size() + 15) & ~size_t(15);
    blob.resize(at + bytes);
    std::memcpy(blob.data() + at, src, bytes);
    return at;
  };
  const size_t o_ps = put(pair_scene.data(), P * 4), o_wb = put(work_base.data(), (P + 1) * 4),
               o_tb = put(tile_base.data(), (P + 1) * 4), o_pb = put(pix_base.data(), (P + 1) * 4),
               o_tx = put(tiles_x.data(), P * 4), o_w = put(width.data(), P * 4),
               o_h = put(height.data(), P * 4), o_c = put(cams.data(), P * sizeof(CameraDev)),
               o_so = put(scene_off.data(), (S + 1) * 4), o_sn = put(scene_n.data(), S * 4),
               o_spb = put(spb.data(), (S + 1) * 4), o_sp = put(spairs.data(), P * 4),
               o_t = put(targets.data(), P * 8), o_lb = put(lbb.data(), (P + 1) * 4);
  uint8_t* d = (uint8_t*)b->store["tables"].reserve(blob.size());
  SPLAT_HIP(hipMemcpyAsync(d, blob.data(), blob.size(), hipMemcpyHostToDevice, b->ctx->stream));
  SPLAT_HIP(hipStreamSynchronize(b->ctx->stream));
  BatchDyn* dyn = (BatchDyn*)b->store["dyn"].reserve(sizeof(BatchDyn));
  BatchDev bd;
  bd.npairs = P;
  bd.nscenes = S;
  bd.pair_scene = (const uint32_t*)(d + o_ps);
  bd.work_base = (const uint32_t*)(d + o_wb);
  bd.tile_base = (const uint32_t*)(d + o_tb);
  bd.pix_base = (const uint32_t*)(d + o_pb);
  bd.tiles_x = (const uint32_t*)(d + o_tx);
  bd.width = (const uint32_t*)(d + o_w);
  bd.height = (const uint32_t*)(d + o_h);
  bd.cams = (const CameraDev*)(d + o_c);
  bd.scene_off = (const uint32_t*)(d + o_so);
  bd.scene_n = (const uint32_t*)(d + o_sn);
  bd.scene_pair_begin = (const uint32_t*)(d + o_spb);
  bd.scene_pairs = (const uint32_t*)(d + o_sp);
  bd.targets = (const uint64_t*)(d + o_t);
  bd.loss_block_base = (const uint32_t*)(d + o_lb);
  bd.lr_mean = dyn->lr_mean;
  b->bd = bd;
  b->plan_views = views;
  b->work = work;
  b->tiles = tiles;
  b->pixels = pix;
  b->loss_blocks_total = (unsigned)lblocks;
}

struct BatchBufs {
  ProjBuffers pb;
  BinBuffers bb;
  uint32_t *small, *big_list, *scan_sums, *counter, *n_contrib;
  unsigned long long* total64;
  float *image, *final_t, *dimage, *scratch, *slots, *losses;
  splat_proj_grad* pgrad;
  StepStatus* status;
  BatchDyn* dyn;
  RadixScratch rs;
};

BatchBufs reserve_bufs(splat_batch* b) {
  auto& st = b->store;
  BatchBufs x;
  const size_t w = b->work + 1, T = b->tiles, px = b->pixels;
  if (b->entry_cap == 0) b->entry_cap = std::max<size_t>(4 * b->work, 1 << 16);
  const size_t ec = b->entry_cap;
  x.pb.rec = (float4*)st["rec"].reserve(w * 64);
  x.pb.depth = (float*)st["depth"].reserve(w * 4);
  x.pb.radius = (uint32_t*)st["radius"].reserve(w * 4);
  x.pb.rect = (uint2*)st["rect"].reserve(w * 8);
  x.pb.ntiles = (uint32_t*)st["ntiles"].reserve(w * 4);
  x.bb.tiles_x = (uint32_t)T;
  x.bb.tiles_y = 1;
  x.bb.entry_offset = (uint32_t*)st["entry_offset"].reserve(w * 4);
  // Batched steps bin by radix sorts (dense scenes put thousands of entries in a tile):
  // no per-tile counters, buckets or per-tile sorts.
  x.bb.tile_count = x.bb.tile_sub = x.bb.entry_tile = x.bb.entry_slot = nullptr;
  x.bb.bucket = nullptr;
  x.bb.block_marks = (uint32_t*)st["block_marks"].reserve((b->entry_cap / 256 + 2) * 4);
  x.bb.tile_start = (uint32_t*)st["tile_start"].reserve(T * 4 + 8);
  // Tiles run in index order: neighbouring tiles of one view share Gaussian records in L2,
  // which the persistent backward loses (+18%) when the largest tiles go first.
  x.bb.tile_order = nullptr;
  x.bb.entry_gaussian = nullptr;  // not needed: the tile sort carries each entry's Gaussian
  x.bb.sorted_entry = (uint32_t*)st["sorted_entry"].reserve(ec * 4);
  x.bb.sorted_gauss = (uint32_t*)st["sorted_gauss"].reserve(ec * 4);
  x.small = (uint32_t*)st["small"].reserve(64);
  x.total64 = (unsigned long long*)st["total64"].reserve(8);
  x.big_list = nullptr;
  x.rs = radix_scratch(st, w, ec);
  x.scan_sums = (uint32_t*)st["scan_sums"].reserve(
      scan_temp_bytes(std::max({w, T, radix_count_words(std::max(w, ec))})));
  x.counter = (uint32_t*)st["counter"].reserve(64);
  x.image = (float*)st["image"].reserve(px * 12 + 4);
  x.final_t = (float*)st["final_t"].reserve(px * 4 + 4);
  x.n_contrib = (uint32_t*)st["n_contrib"].reserve(px * 4 + 4);
  x.dimage = (float*)st["dimage"].reserve(px * 12 + 4);
  x.scratch = (float*)st["loss_scratch"].reserve((px * 9 + 2 * (size_t)b->loss_blocks_total + 64) * 4);
  x.slots = (float*)st["slots"].reserve((ec + 1) * kSlots * 4);
  x.pgrad = (splat_proj_grad*)st["pgrad"].reserve(w * sizeof(splat_proj_grad));
  x.losses = (float*)st["losses"].reserve(b->plan_views.size() * 4 + 4);
  x.status = (StepStatus*)st["status"].reserve(sizeof(StepStatus));
  x.dyn = (BatchDyn*)st["dyn"].reserve(sizeof(BatchDyn));
  return x;
}

// The launches of one batched step.
void enqueue(splat_batch* b, const BatchBufs& x, const splat_batch_step_params& p,
             const RasterVariant& fwd, const RasterVariant& bwd) {
  splat_ctx* ctx = b->ctx;
  const BatchDev& bd = b->bd;
  project_forward_batch(ctx, b->params.as<float>(), b->stride, bd, b->rows, p.near_plane,
                        p.antialias, p.sh_degree, x.pb);
  bin_build_radix_device(ctx, x.pb, b->work, x.bb, b->entry_cap, x.rs, x.small, x.total64,
                         x.scan_sums, x.status, &bd);
  raster_forward(ctx, x.pb, x.bb, 0, 0, p.background, x.image, x.final_t, x.n_contrib, fwd, &bd,
                 x.counter);
  loss_forward_backward_batch(ctx, x.image, bd, b->loss_blocks_total, b->pixels * 3,
                              p.lambda_dssim, x.dimage, x.scratch, x.losses);
  raster_backward(ctx, x.pb, x.bb, 0, 0, p.background, x.final_t, x.n_contrib, x.dimage,
                  p.ordered_backward != 0, x.slots, false, bwd, &bd, x.counter);
  reduce_slots(ctx, x.bb.entry_offset, x.slots, b->work, x.pgrad, nullptr, x.status);
  splat_adam_state ast{0.9f, 0.999f, 1e-15f, 0, 0};
  grad_adam_batch(ctx, b->params.as<float>(), b->m.as<float>(), b->v.as<float>(), b->stride, bd,
                  b->rows, p.near_plane, p.antialias, p.sh_degree, x.bb.entry_offset, x.pgrad, ast,
                  &x.dyn->step, x.status);
}

void start(splat_batch* b) {
  if (b->started) return;
  SPLAT_REQUIRE(!b->scenes.empty(), "no scenes");
  b->started = true;
}

void clear_status(splat_batch* b, StepStatus* st) {
  fill_u32(b->ctx, reinterpret_cast<uint32_t*>(st), sizeof(StepStatus) / 4, 0u);
}

StepStatus read_status(splat_batch* b, const StepStatus* st) {
  StepStatus h;
  SPLAT_HIP(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, b->ctx->stream));
  SPLAT_HIP(hipStreamSynchronize(b->ctx->stream));
  return h;
}

void grow_entries(splat_batch* b, unsigned long long total) {
  SPLAT_REQUIRE(total < 0xffffffffull, "too many tile entries");
  b->entry_cap = std::max<size_t>(b->entry_cap, (size_t)(total + total / 4 + 1024));
  b->aborts++;
}

void write_dyn(splat_batch* b, const splat_batch_step_params& p, const float* lr_mean,
               BatchDyn* dst) {
  BatchDyn d;
  std::memset(&d, 0, sizeof(d));
  d.step.lr = p.lr;
  d.step.c1 = 1.0f - powi(0.9f, (int32_t)b->adam_step);
  d.step.c2 = 1.0f - powi(0.999f, (int32_t)b->adam_step);
  for (size_t s = 0; s < b->scenes.size(); s++) d.lr_mean[s] = lr_mean[s];
  launch(b->ctx, "step_args", set_batch_dyn_kernel, dim3(1), dim3(64), 0, d, dst);
}

std::vector<uint32_t> view_list(splat_batch* b, const uint32_t* views, uint32_t nviews) {
  SPLAT_REQUIRE(views && nviews > 0, "no views");
  std::vector<uint32_t> v(views, views + nviews);
  std::vector<uint8_t> seen(b->views.size(), 0);
  for (uint32_t x : v) {
    SPLAT_REQUIRE(x < b->views.size(), "bad view index");
    SPLAT_REQUIRE(!seen[x], "a view appears twice in one step");
    seen[x] = 1;
  }
  return v;
}

void check_params(const splat_batch_step_params* p) {
  SPLAT_REQUIRE(p, "null argument");
  SPLAT_REQUIRE(p->sh_degree >= 0 && p->sh_degree <= 3, "sh_degree must be 0..3");
}

}  // namespace

extern "C" {

int splat_batch_create(splat_ctx* ctx, splat_batch** out) {
  if (!ctx || !out) return SPLAT_ERR_INVALID;
  *out = nullptr;
  splat_batch* b = new (std::nothrow) splat_batch();
  if (!b) return SPLAT_ERR_OOM;
  b->ctx = ctx;
  b->device = ctx->device;
  *out = b;
  return SPLAT_OK;
}

int splat_batch_destroy(splat_batch* b) {
  if (!b) return SPLAT_OK;
  (void)hipSetDevice(b->device);
  (void)hipDeviceSynchronize();
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  delete b;
  return SPLAT_OK;
}

int splat_batch_add_scene(splat_batch* b, const float* gaussians, uint64_t n, uint32_t* scene) {
  return bguard(b, [&] {
    SPLAT_REQUIRE(!b->started, "scenes must be added before the first step");
    SPLAT_REQUIRE(n > 0 && gaussians, "a scene needs Gaussians");
    SPLAT_REQUIRE(b->scenes.size() < kMaxScenes, "too many scenes");
    splat_ctx* ctx = b->ctx;
    BatchScene sc;
    sc.n = n;
    sc.off = b->rows;
    const size_t rows = b->rows + ((n + 63) & ~size_t(63));
    const size_t stride = plane_stride(rows);
    // grow the planes, keeping the scenes already there
    DevBuf np, nm, nv;
    np.reserve(stride * kParams * 4);
    nm.reserve(stride * kParams * 4);
    nv.reserve(stride * kParams * 4);
    fill_u32(ctx, np.as<uint32_t>(), stride * kParams, 0u);
    fill_u32(ctx, nm.as<uint32_t>(), stride * kParams, 0u);
    fill_u32(ctx, nv.as<uint32_t>(), stride * kParams, 0u);
    if (b->rows) {
      for (int j = 0; j < kParams; j++)
        SPLAT_HIP(hipMemcpyAsync(np.as<float>() + (size_t)j * stride,
                                 b->params.as<float>() + (size_t)j * b->stride, b->rows * 4,
                                 hipMemcpyDeviceToDevice, ctx->stream));
    }
    float* staging = (float*)b->store["aos"].reserve(n * kParams * 4);
    SPLAT_HIP(hipMemcpyAsync(staging, gaussians, n * kParams * 4, hipMemcpyHostToDevice,
                             ctx->stream));
    aos_to_planes(ctx, staging, n, np.as<float>() + sc.off, stride);
    SPLAT_HIP(hipStreamSynchronize(ctx->stream));
    std::swap(b->params.ptr, np.ptr);
    std::swap(b->params.bytes, np.bytes);
    std::swap(b->m.ptr, nm.ptr);
    std::swap(b->m.bytes, nm.bytes);
    std::swap(b->v.ptr, nv.ptr);
    std::swap(b->v.bytes, nv.bytes);
    b->rows = rows;
    b->stride = stride;
    b->store.erase("aos");
    b->scenes.push_back(sc);
    if (scene) *scene = (uint32_t)(b->scenes.size() - 1);
  });
}

int splat_batch_add_view(splat_batch* b, uint32_t scene, const splat_camera* cam,
                         const uint8_t* rgb, uint32_t* view) {
  return bguard(b, [&] {
    SPLAT_REQUIRE(cam && rgb, "null argument");
    SPLAT_REQUIRE(scene < b->scenes.size(), "bad scene index");
    SPLAT_REQUIRE(cam->width > 0 && cam->height > 0 && cam->width <= 65535 * kTile &&
                      cam->height <= 65535 * kTile,
                  "bad image size");
    auto v = std::make_unique<BatchView>();
    v->scene = scene;
    v->cam = *cam;
    const size_t bytes = (size_t)cam->width * cam->height * 3;
    v->rgb.reserve(bytes);
    SPLAT_HIP(hipMemcpyAsync(v->rgb.ptr, rgb, bytes, hipMemcpyHostToDevice, b->ctx->stream));
    SPLAT_HIP(hipStreamSynchronize(b->ctx->stream));
    b->views.push_back(std::move(v));
    if (view) *view = (uint32_t)(b->views.size() - 1);
  });
}

int splat_batch_step(splat_batch* b, const uint32_t* views, uint32_t nviews,
                     const splat_batch_step_params* p, const float* lr_mean, float* losses) {
  return bguard(b, [&] {
    check_params(p);
    SPLAT_REQUIRE(lr_mean, "null argument");
    start(b);
    splat_ctx* ctx = b->ctx;
    const std::vector<uint32_t> list = view_list(b, views, nviews);
    if (list != b->plan_views) build_plan(b, list);
    b->adam_step++;
    for (;;) {
      const BatchBufs x = reserve_bufs(b);
      BatchKey key;
      key.views = list;
      key.sh_degree = p->sh_degree;
      key.ordered = p->ordered_backward;
      key.antialias = p->antialias;
      key.near_plane = p->near_plane;
      key.lambda = p->lambda_dssim;
      for (int c = 0; c < 3; c++) key.bg[c] = p->background[c];
      key.fwd = b->fwd;
      key.bwd = b->bwd;
      key.entry_cap = b->entry_cap;
      key.params = b->params.ptr;
      if (!b->exec || !(key == b->key)) {
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
        const bool timing = ctx->timing;
        ctx->timing = false;
        SPLAT_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
        try {
          enqueue(b, x, *p, b->fwd, b->bwd);
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
        const hipError_t e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        SPLAT_HIP(e);
        b->key = key;
        b->captures++;
      }
      clear_status(b, x.status);
      write_dyn(b, *p, lr_mean, x.dyn);
      SPLAT_HIP(hipGraphLaunch(b->exec, ctx->stream));
      b->replays++;
      const StepStatus st = read_status(b, x.status);
      if (!st.abort) {
        if (losses) {
          SPLAT_HIP(hipMemcpyAsync(losses, x.losses, list.size() * 4, hipMemcpyDeviceToHost,
                                   ctx->stream));
          SPLAT_HIP(hipStreamSynchronize(ctx->stream));
        }
        return;
      }
      grow_entries(b, st.total);  // nothing of the model changed: run the step again
    }
  });
}

int splat_batch_tune(splat_batch* b, const uint32_t* views, uint32_t nviews,
                     const splat_batch_step_params* p, splat_tune_result* out) {
  return bguard(b, [&] {
    check_params(p);
    SPLAT_REQUIRE(out, "null argument");
    start(b);
    splat_ctx* ctx = b->ctx;
    const std::vector<uint32_t> list = view_list(b, views, nviews);
    if (list != b->plan_views) build_plan(b, list);
    // One eager pass of the forward half fills the buffers the candidates run on; the
    // model is left untouched (no Adam).
    BatchBufs x;
    for (;;) {
      x = reserve_bufs(b);
      clear_status(b, x.status);
      project_forward_batch(ctx, b->params.as<float>(), b->stride, b->bd, b->rows, p->near_plane,
                            p->antialias, p->sh_degree, x.pb);
      bin_build_radix_device(ctx, x.pb, b->work, x.bb, b->entry_cap, x.rs, x.small, x.total64,
                             x.scan_sums, x.status, &b->bd);
      const StepStatus st = read_status(b, x.status);
      if (!st.abort) break;
      grow_entries(b, st.total);
    }
    RasterVariant def;
    raster_forward(ctx, x.pb, x.bb, 0, 0, p->background, x.image, x.final_t, x.n_contrib, def,
                   &b->bd, x.counter);
    loss_forward_backward_batch(ctx, x.image, b->bd, b->loss_blocks_total, b->pixels * 3,
                                p->lambda_dssim, x.dimage, x.scratch, nullptr);
    const RasterTune tr = tune_raster(ctx, x.pb, x.bb, 0, 0, p->background, x.image, x.final_t,
                                      x.n_contrib, x.dimage, p->ordered_backward != 0, x.slots,
                                      false, &b->bd, x.counter);
    b->fwd = tr.fwd;
    b->bwd = tr.bwd;
    b->tuned = true;
    fill_tune_result(tr, out);
  });
}

int splat_batch_set_shapes(splat_batch* b, const splat_tune_result* shapes) {
  return bguard(b, [&] {
    SPLAT_REQUIRE(shapes, "null argument");
    auto ok_f = [](int v) { return v == 32 || v == 64 || v == 128 || v == 256; };
    SPLAT_REQUIRE(ok_f(shapes->fwd_batch) && (shapes->bwd_batch == 16 || shapes->bwd_batch == 32),
                  "unsupported batch shape");
    b->fwd.fwd_batch = shapes->fwd_batch;
    b->fwd.persistent = shapes->fwd_persistent != 0;
    b->fwd.per_cu = (unsigned)std::max(1, shapes->fwd_per_cu);
    b->bwd.bwd_batch = shapes->bwd_batch;
    b->bwd.persistent = shapes->bwd_persistent != 0;
    b->bwd.per_cu = (unsigned)std::max(1, shapes->bwd_per_cu);
    b->bwd.rows = shapes->bwd_row_lists != 0;
  });
}

int splat_batch_download(splat_batch* b, uint32_t scene, float* gaussians, float* m, float* v,
                         uint64_t capacity) {
  return bguard(b, [&] {
    SPLAT_REQUIRE(scene < b->scenes.size(), "bad scene index");
    const BatchScene& sc = b->scenes[scene];
    SPLAT_REQUIRE(capacity >= sc.n, "output capacity too small");
    splat_ctx* ctx = b->ctx;
    float* staging = (float*)b->store["aos_out"].reserve(sc.n * kParams * 4);
    const std::pair<float*, const DevBuf*> arrays[3] = {{gaussians, &b->params}, {m, &b->m},
                                                        {v, &b->v}};
    for (auto& a : arrays) {
      if (!a.first) continue;
      planes_to_aos(ctx, a.second->as<float>() + sc.off, sc.n, b->stride, staging);
      SPLAT_HIP(hipMemcpyAsync(a.first, staging, sc.n * kParams * 4, hipMemcpyDeviceToHost,
                               ctx->stream));
      SPLAT_HIP(hipStreamSynchronize(ctx->stream));
    }
    b->store.erase("aos_out");
  });
}

int splat_batch_info(splat_batch* b, uint64_t* rows, uint64_t* adam_step, uint64_t* captures,
                     uint64_t* replays, uint64_t* aborts) {
  if (!b) return SPLAT_ERR_INVALID;
  if (rows) *rows = b->rows;
  if (adam_step) *adam_step = b->adam_step;
  if (captures) *captures = b->captures;
  if (replays) *replays = b->replays;
  if (aborts) *aborts = b->aborts;
  return SPLAT_OK;
}

}  // extern "C"
