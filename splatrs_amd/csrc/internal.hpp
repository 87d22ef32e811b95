// Internal runtime of libsplat: context, errors, timed kernel launches, device buffers and
// the HBM layouts shared by the kernel files.
//
// HBM layouts (one GPU, FP32):
//   params / grads / Adam moments: "planes" -- parameter j of Gaussian i at [j * S + i],
//     j in the reference's flat order (0-2 mean, 3-5 log_scale, 6-9 quat, 10 opacity,
//     11 + 3k + c SH), S = plane stride (>= N, multiple of 64).  Every kernel then reads
//     and writes whole 256-byte wave rows.
//   projection: one 64-byte rasterizer record per Gaussian ({mx, my, A, B},
//     {C, opacity, sigma_max, half_x}, {half_y, r, g, b}, unused), depth, radius, tile
//     rectangle (uint2 {x0 | y0 << 16, x1 | y1 << 16}) and the tile-entry count.
//   binning: entry arrays (tile, slot, gaussian), per-tile counts/starts, and the sorted
//     per-tile lists (entry id, gaussian id) that the compositing kernels read.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <map>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/splat.h"
#include "fmath.hpp"

namespace splat {

constexpr int kTile = 16;
constexpr int kParams = 59;
constexpr int kSlots = 11;
constexpr float kAlphaMin = 1.0f / 255.0f;  // src/rasterize.rs:53
constexpr float kAlphaMax = 0.99f;          // src/rasterize.rs:55
constexpr float kTMin = 1e-4f;              // src/rasterize.rs:57

struct Error {
  int code;
  std::string msg;
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error{code, msg}; }

#define SPLAT_HIP(call)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      ::splat::fail(e_ == hipErrorOutOfMemory ? SPLAT_ERR_OOM : SPLAT_ERR_HIP,           \
                    std::string(#call) + ": " + hipGetErrorString(e_));                  \
  } while (0)

#define SPLAT_REQUIRE(cond, msg)                                                         \
  do {                                                                                   \
    if (!(cond)) ::splat::fail(SPLAT_ERR_PRECONDITION, msg);                             \
  } while (0)

// Device allocation that grows on demand (contents are not preserved on growth).
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (ptr) (void)hipFree(ptr);
  }
  void* reserve(size_t b) {
    if (b > bytes) {
      if (ptr) {
        SPLAT_HIP(hipDeviceSynchronize());  // work in flight may still use the old buffer
        SPLAT_HIP(hipFree(ptr));
      }
      ptr = nullptr;
      bytes = 0;
      size_t want = b < 256 ? 256 : b;
      SPLAT_HIP(hipMalloc(&ptr, want));
      bytes = want;
    }
    return ptr;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

struct KernelStat {
  uint64_t launches = 0;
  double ms = 0;
};

struct PendingTiming {
  const char* name;
  hipEvent_t a, b;
};

}  // namespace splat

struct splat_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool timing = false;
  int bin_method = 0;  // splat_bin: 0 per-tile counting + sorts, 1 radix sorts
  std::vector<splat::PendingTiming> pending;
  std::vector<hipEvent_t> event_pool;
  std::map<std::string, splat::KernelStat> kstats;
  std::map<std::string, splat::DevBuf> ws;  // named scratch buffers of the host tier

  template <class T>
  T* buf(const char* name, size_t count) {
    return static_cast<T*>(ws[name].reserve(count * sizeof(T)));
  }
  hipEvent_t event() {
    if (!event_pool.empty()) {
      hipEvent_t e = event_pool.back();
      event_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    SPLAT_HIP(hipEventCreate(&e));
    return e;
  }
  // Collect finished timing events (synchronises the stream).
  void drain_timing() {
    if (pending.empty()) return;
    SPLAT_HIP(hipStreamSynchronize(stream));
    for (auto& p : pending) {
      float ms = 0;
      SPLAT_HIP(hipEventElapsedTime(&ms, p.a, p.b));
      auto& s = kstats[p.name];
      s.launches++;
      s.ms += ms;
      event_pool.push_back(p.a);
      event_pool.push_back(p.b);
    }
    pending.clear();
  }
};

namespace splat {

// Launch `kernel` on the context stream, bracketed by events when timing is on.
template <class K, class... Args>
inline void launch(splat_ctx* ctx, const char* name, K kernel, dim3 grid, dim3 block,
                   size_t shmem, Args... args) {
  if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
  hipEvent_t a = nullptr;
  if (ctx->timing) {
    a = ctx->event();
    SPLAT_HIP(hipEventRecord(a, ctx->stream));
  }
  hipLaunchKernelGGL(kernel, grid, block, shmem, ctx->stream, args...);
  SPLAT_HIP(hipGetLastError());
  if (ctx->timing) {
    hipEvent_t b = ctx->event();
    SPLAT_HIP(hipEventRecord(b, ctx->stream));
    ctx->pending.push_back({name, a, b});
    if (ctx->pending.size() > 4096) ctx->drain_timing();
  }
}

// Runs f for a C-ABI entry point: HIP device selection, errors -> status + message.
template <class F>
int ctx_guard(splat_ctx* ctx, F&& f) {
  if (!ctx) return SPLAT_ERR_INVALID;
  try {
    SPLAT_HIP(hipSetDevice(ctx->device));
    f();
    ctx->err.clear();
    return SPLAT_OK;
  } catch (const Error& e) {
    ctx->err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    ctx->err = "host allocation failed";
    return SPLAT_ERR_OOM;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return SPLAT_ERR_INVALID;
  }
}

inline unsigned blocks_for(size_t n, unsigned threads) {
  return (unsigned)((n + threads - 1) / threads);
}

inline size_t plane_stride(size_t n) { return (n + 63) & ~size_t(63); }

// ---- device-side views -------------------------------------------------------------

struct CameraDev {  // src/camera.rs:13-32
  uint32_t width, height;
  float fx, fy, cx, cy;
  float rot[9];
  float trans[3];
};

inline CameraDev camera_dev(const splat_camera& c) {
  CameraDev d;
  d.width = c.width;
  d.height = c.height;
  d.fx = c.fx;
  d.fy = c.fy;
  d.cx = c.cx;
  d.cy = c.cy;
  for (int i = 0; i < 9; i++) d.rot[i] = c.rot[i];
  for (int i = 0; i < 3; i++) d.trans[i] = c.trans[i];
  return d;
}

// Per-step arguments of a graph-captured train step.  Kernel arguments are frozen when a
// graph is instantiated, so the values that change from one iteration to the next (view,
// target image, learning rates, Adam bias corrections, loss destination) are written into
// this device record by a small kernel launched before every replay of the graph.
struct StepDyn {
  CameraDev cam;
  const uint8_t* target;  // the view's RGB8 image
  float* loss_out;        // loss value destination (device), or null
  splat_learning_rates lr;
  float c1, c2;           // Adam bias corrections 1 - beta^step
  uint32_t seq;           // step sequence number (host bookkeeping)
  uint32_t report_slot;   // slot of the host-visible step report
};

// Device status of a graph-captured step.  The tile-entry arrays are preallocated with a
// capacity; when a step needs more entries, `abort` is set and stays set, so that step and
// every later one leave the model untouched until the host grows the arrays, clears the
// flag and replays them in order.
struct StepStatus {
  unsigned long long total;  // tile entries of the current step (64-bit)
  uint32_t entries;          // the same, clamped to 32 bits
  uint32_t abort;            // sticky
};

// Host-visible record a step writes at its end (pinned, mapped memory).
struct StepReport {
  uint32_t seq;
  uint32_t abort;
  unsigned long long total;
};

// Where a train step's loss pass writes its loss value and report (reports null: not
// written): the pass's per-block partial sums, their count, the number of loss terms.
struct StepLossReport {
  const float2* partial;
  uint32_t nb;
  float n;
  const StepStatus* status;
  StepReport* reports;
};

// Tables of a batched step: many (scene, view) pairs composited in one launch per stage.
// Pair v renders scene pair_scene[v] through cams[v]; its projections occupy
// [work_base[v], work_base[v+1]) of the projection buffers (one per scene Gaussian), its
// tiles [tile_base[v], tile_base[v+1]) of the global tile space and its pixels
// [pix_base[v], pix_base[v+1]) of the image buffers.  Scene s owns parameter rows
// [scene_off[s], scene_off[s] + scene_n[s]) of the shared planes; its pairs are
// scene_pairs[scene_pair_begin[s] .. scene_pair_begin[s+1]) in step order.
struct BatchDev {
  uint32_t npairs, nscenes;
  const uint32_t* pair_scene;
  const uint32_t* work_base;
  const uint32_t* tile_base;
  const uint32_t* pix_base;
  const uint32_t* tiles_x;
  const uint32_t* width;
  const uint32_t* height;
  const CameraDev* cams;
  const uint32_t* scene_off;
  const uint32_t* scene_n;
  const uint32_t* scene_pair_begin;
  const uint32_t* scene_pairs;
  const uint64_t* targets;         // RGB8 target image of each pair (device addresses)
  const uint32_t* loss_block_base; // loss-kernel blocks of each pair (prefix)
  const float* lr_mean;            // per-scene position learning rate of the step
};

// Largest k with base[k] <= x, for base[0] = 0 <= x (base has n + 1 entries).
__device__ __forceinline__ uint32_t find_segment(const uint32_t* base, uint32_t n, uint32_t x) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (base[mid] <= x)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Inclusive scan of one value per thread across the workgroup (THREADS lanes, sh holds
// THREADS / 64 + 1 words); returns the inclusive prefix and writes the workgroup total to
// *total.
template <int THREADS>
__device__ __forceinline__ uint32_t block_inclusive_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  if (lane == 63) sh[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int w = 0; w < THREADS / 64; w++) {
      const uint32_t t = sh[w];
      sh[w] = run;
      run += t;
    }
    sh[THREADS / 64] = run;
  }
  __syncthreads();
  v += sh[wave];
  *total = sh[THREADS / 64];
  __syncthreads();
  return v;
}

// The logical block of workgroup w of a launch of nb workgroups: workgroup w runs on die
// w % 8 (each die has its own L2), and die x takes the contiguous range of logical blocks
// [x s + min(x, r), ...) of s + (x < r) blocks (s = nb / 8, r = nb % 8) in order, so that
// neighbouring logical blocks, which share or complete each other's cache lines, meet in
// one L2.  A bijection of [0, nb): die x runs exactly s + (x < r) workgroups.
__device__ __forceinline__ uint32_t die_block(uint32_t w, uint32_t nb) {
  const uint32_t x = w % 8u, i = w / 8u, s = nb / 8u, r = nb % 8u;
  return x * s + min(x, r) + i;
}

// Learning rate of parameter plane j (LearningRates, src/adam.rs:11-26).
__device__ __forceinline__ float plane_lr(int j, const splat_learning_rates& lr) {
  if (j < 3) return lr.mean;
  if (j < 6) return lr.log_scale;
  if (j < 10) return lr.quat;
  if (j == 10) return lr.opacity;
  if (j < 14) return lr.sh_dc;
  return lr.sh_rest;
}

// Streaming 16-byte loads and stores with the non-temporal hint: data touched once per
// pass (Adam's moment planes) should not displace reusable lines from the caches.
typedef float float4_vec __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 load_stream(const float* p) {
  const float4_vec t = __builtin_nontemporal_load(reinterpret_cast<const float4_vec*>(p));
  return make_float4(t.x, t.y, t.z, t.w);
}
__device__ __forceinline__ void store_stream(float* p, float4 x) {
  const float4_vec t = {x.x, x.y, x.z, x.w};
  __builtin_nontemporal_store(t, reinterpret_cast<float4_vec*>(p));
}

// Adam's update of one scalar, src/adam.rs:44-61.
__device__ __forceinline__ void adam_update(float& p, float g, float& m, float& v, float lr,
                                            float b1, float b2, float c1, float c2, float eps) {
  m = b1 * m + (1.0f - b1) * g;
  v = b2 * v + (1.0f - b2) * g * g;
  const float m_hat = m / c1;
  const float v_hat = v / c2;
  p -= lr * m_hat / (sqrtf(v_hat) + eps);
}

// Projection outputs for one view (device pointers, capacity n).
struct ProjBuffers {
  // One 64-byte rasterizer record per Gaussian (a single cache line per gather):
  // rec[4i] = {mx, my, A, B}, rec[4i+1] = {C, opacity, sigma_max, half_x} (half_x < 0
  // marks an empty footprint), rec[4i+2] = {half_y, r, g, b}, rec[4i+3] unused.
  float4* rec;
  float* depth;
  uint32_t* radius;
  uint2* rect;      // {x0 | y0 << 16, x1 | y1 << 16}, tile_max exclusive
  uint32_t* ntiles; // entries of each Gaussian (0 when culled)
};

// Binning of one view.
constexpr uint32_t kTileSub = 8;  // counter sets per tile in the binning's counting pass

struct BinBuffers {
  uint32_t* entry_offset;    // n + 1 (exclusive scan of ntiles)
  uint32_t* entry_gaussian;  // E
  uint32_t* entry_tile;      // E
  uint32_t* entry_slot;      // E (position within the tile, from the counting pass)
  uint32_t* tile_count;      // kTileSub * T: per-tile counters, one set per keygen workgroup
                             // index mod kTileSub, laid out [kTileSub][T]
  uint32_t* tile_sub;        // (kTileSub + 1) * T: each counter's offset in its tile
                             // [kTileSub][T], then the tile totals [T]
  uint32_t* tile_start;      // T + 1
  uint32_t* block_marks;     // >= capacity / 256 + 2: Gaussian of each keygen workgroup's
                             // first entry (a workgroup handles 256 entries)
  uint4* bucket;             // E records {entry, depth bits, Gaussian, 0} (the 64-bit key
                             // depth bits << 32 | entry in .x/.y), grouped by tile
  uint32_t* sorted_entry;    // E: entry ids in (tile, depth, entry) order
  uint32_t* sorted_gauss;    // E: their Gaussian ids
  uint32_t tiles_x, tiles_y;
  uint32_t* tile_order = nullptr;  // T: compositing dispatch order (null: tile index order)
};

// ---- kernel launchers (one per kernel file) -------------------------------------------

// project.hip
// Kernels that take `const StepDyn* dyn` read the per-step values from it when it is not
// null (graph-captured steps) and from their other arguments otherwise.  Kernels that take
// `const StepStatus* status` do nothing once its abort flag is set.
void project_forward(splat_ctx* ctx, const float* params, size_t n, size_t stride,
                     const splat_camera& cam, const splat_proj_opts& opts, ProjBuffers pb,
                     const StepDyn* dyn = nullptr);
void proj_pack(splat_ctx* ctx, ProjBuffers pb, size_t n, splat_proj* out);
void proj_unpack(splat_ctx* ctx, const splat_proj* in, size_t n, uint32_t tiles_x,
                 uint32_t tiles_y, ProjBuffers pb, uint32_t* bad_flag);
void project_backward(splat_ctx* ctx, const float* params, size_t n, size_t stride,
                      const splat_camera& cam, const splat_proj_opts& opts,
                      const splat_proj_grad* dproj, float* grads, const StepDyn* dyn = nullptr);
void aos_to_planes(splat_ctx* ctx, const float* aos, size_t n, float* planes, size_t stride);
// Densification statistics updated inside a fused kernel (grad_accum null: none).
struct StatsArgs {
  const uint32_t* radius;
  float sx, sy;  // width / 2, height / 2
  int absgrad;
  float* grad_accum;
  uint32_t* denom;
  uint32_t* max_radii;
  float* max_grad;  // optional extra statistics of the learned scorer
  uint32_t* strong;
  float strong_threshold;
};
// Fused backward tail and optimizer of a train step: slot reduce + statistics + project
// backward + Adam in one pass (no gradient planes); c1, c2, lr come from dyn when it is
// given.
void grad_adam(splat_ctx* ctx, float* params, float* m, float* v, size_t n, size_t stride,
               const splat_camera& cam, const splat_proj_opts& opts, const uint32_t* entry_offset,
               const float* slots, const StatsArgs& sa, float c1, float c2,
               const splat_adam_state& st, const splat_learning_rates& lr,
               const StepDyn* dyn = nullptr, const StepStatus* status = nullptr);
// Batched: projections of every (scene, view) pair into the work rows (one lane per scene
// row, `rows` plane rows in total, each read once for all of its scene's views), and the backward
// tail + Adam of every scene Gaussian (total `rows` plane rows, a multiple of 64): the
// projection backward of each pair's summed slots (dproj, one per work row) summed over
// the scene's views in step order, then Adam with the scene's position learning rate;
// c1, c2, lr come from dyn.
void project_forward_batch(splat_ctx* ctx, const float* params, size_t stride,
                           const BatchDev& bd, size_t rows, float near_plane, int antialias,
                           int degree, ProjBuffers pb);
void grad_adam_batch(splat_ctx* ctx, float* params, float* m, float* v, size_t stride,
                     const BatchDev& bd, size_t rows, float near_plane, int antialias, int degree,
                     const uint32_t* entry_offset, const splat_proj_grad* dproj,
                     const splat_adam_state& st, const StepDyn* dyn, const StepStatus* status);
void planes_to_aos(splat_ctx* ctx, const float* planes, size_t n, size_t stride, float* aos);

// scan.hip
// `sums` (scan_temp_bytes(n) bytes) defaults to a context scratch buffer; graph-captured
// code passes its own so the captured pointer stays valid.  `total64` (optional, zeroed by
// the caller) receives the exact 64-bit sum of the input.
// Extra work of the last pass of the entry-offset scan of a graph-captured step: the
// capacity guard (total64 against `capacity`, into `status`) and the keygen marks.
struct ScanTail {
  uint32_t* marks;
  uint32_t marks_cap, epb;
  unsigned long long capacity;
  const unsigned long long* total64;
  StepStatus* status;
};
// Returns whether `tail` was done (only by the multi-pass scan, n > scan_single's limit;
// the caller does it otherwise).
bool exclusive_scan_u32(splat_ctx* ctx, const uint32_t* in, uint32_t* out, size_t n,
                        uint32_t* total_dev, uint32_t* sums = nullptr,
                        unsigned long long* total64 = nullptr, const ScanTail* tail = nullptr);
size_t scan_temp_bytes(size_t n);
size_t scan_max_items();  // largest n exclusive_scan_u32 takes
// Fill kernel (used instead of hipMemset so the same code can sit inside a HIP graph).
void fill_u32(splat_ctx* ctx, uint32_t* p, size_t n, uint32_t v);

// binning.hip
// Builds the sorted per-tile lists.  Returns E (synchronises once to read it).
size_t bin_build(splat_ctx* ctx, ProjBuffers pb, size_t n, uint32_t tiles_x, uint32_t tiles_y,
                 BinBuffers& bb, size_t* capacity_io, std::map<std::string, DevBuf>& store,
                 uint64_t* keys_out);
// The same without a host synchronisation: the entry arrays of `bb` (capacity `capacity`)
// and its per-Gaussian / per-tile arrays must already be allocated.  With `bd`, the rows
// are the projections of a batched step and bb.tiles_x * bb.tiles_y is the global tile
// count (bb.tiles_y = 1).  The entry count goes
// to `status`; a count above the capacity sets the abort flag and the entry kernels exit.
// The caller clears, before it, the words the pass accumulates into: total64 (2 words),
// bb.tile_count (kTileSub x tiles) and small[1..3].
void bin_build_device(splat_ctx* ctx, ProjBuffers pb, size_t n, const BinBuffers& bb,
                      size_t capacity, uint32_t* small, unsigned long long* total64,
                      uint32_t* big_list, uint32_t* scan_sums, StepStatus* status,
                      const BatchDev* bd = nullptr);
void bin_guard(splat_ctx* ctx, const unsigned long long* total64, size_t capacity,
               StepStatus* status);

// Scratch of the radix binning: (key, value) ping-pong pairs over the rows (depth order)
// and over the entries (tile order), the rows' offsets in depth order and the radix
// histograms.
struct RadixScratch {
  uint32_t *rk[2], *rv[2];  // rows (n)
  uint32_t *ek[2], *ev[2], *eg[2];  // entries (capacity): tile, entry id, Gaussian
  uint32_t* soff;           // n + 1
  uint4* rows;              // n + 1 row records of the entry generator, depth order
  uint4* recs;              // n + 1 the same records in row order
  uint32_t* counts;         // radix_count_words(max(n, capacity))
};
RadixScratch radix_scratch(std::map<std::string, DevBuf>& store, size_t n, size_t capacity);
constexpr size_t kMaxViewTiles = size_t(1) << 22;  // tiles of one view for radix binning
// Tile binning by two stable radix sorts: the rows by depth, then their entries (emitted in
// depth order, no per-tile counters) by tile.  Each Gaussian has at most one entry per
// tile and entry ids grow with the Gaussian index, so the result is the (tile, depth,
// entry) order of bin_build_device, built from bb's entry_offset, block_marks, tile_start,
// sorted_entry and sorted_gauss (and entry_gaussian when bb.entry_gaussian is set).  With
// `status` (graph steps) the entry count is read on the device; without it `e_host` is the
// count.  keys_out (optional) receives the sorted (tile << 32 | depth bits) keys.
void bin_build_radix_device(splat_ctx* ctx, ProjBuffers pb, size_t n, const BinBuffers& bb,
                            size_t capacity, const RadixScratch& rs, uint32_t* small,
                            unsigned long long* total64, uint32_t* scan_sums, StepStatus* status,
                            const BatchDev* bd, uint64_t* keys_out = nullptr,
                            uint32_t e_host = 0, bool offsets_done = false);

// radix.hip
// Stable LSD radix sort of (key, value) u32 pairs by key bits [0, key_bits), up to 9 bits
// per pass, ping-ponging between (k0, v0) and (k1, v1) (and a second value between w0 and
// w1 when given); returns the index (0 or 1) of the buffers that hold the sorted keys.
// The values end there too, unless v_final (w_final) is given: the last pass writes them
// there (key_bits > 0).  The item count is *n_dev when n_dev is given (grids cover
// `capacity`), else n_host.  counts: radix_count_words(capacity) words; scan_sums:
// scan_temp_bytes(radix_count_words(capacity)).
size_t radix_count_words(size_t capacity);
int radix_sort_pairs(splat_ctx* ctx, uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1,
                     size_t capacity, uint32_t n_host, const uint32_t* n_dev, uint32_t key_bits,
                     uint32_t* counts, uint32_t* scan_sums, const StepStatus* status,
                     uint32_t* w0 = nullptr, uint32_t* w1 = nullptr,
                     uint32_t* v_final = nullptr, uint32_t* w_final = nullptr);

// raster.hip
// Launch shape of the compositing kernels (results do not depend on it): entries per LDS
// batch of the forward (32, 64, 128, 256) and of the tree-reduced backward (16, 32),
// whether a fixed grid of per_cu workgroups per compute unit takes tiles from a counter,
// and whether each 16-lane row of a backward wave walks its own list of entries (rows)
// or the wave walks the entries that reach any of its rows.
struct RasterVariant {
  int fwd_batch = 64;
  int bwd_batch = 32;
  bool persistent = false;
  unsigned per_cu = 4;
  bool rows = true;
};
// With `bd` the tiles are the global tile space of a batched step (bb.tiles_x * bb.tiles_y
// tiles; width/height unused).  `counter` (one word) is needed by persistent launches.
void raster_forward(splat_ctx* ctx, const ProjBuffers& pb, const BinBuffers& bb,
                    uint32_t width, uint32_t height, const float bg[3], float* image,
                    float* final_t, uint32_t* n_contrib, const RasterVariant& rv = {},
                    const BatchDev* bd = nullptr, uint32_t* counter = nullptr);
void raster_backward(splat_ctx* ctx, const ProjBuffers& pb, const BinBuffers& bb,
                     uint32_t width, uint32_t height, const float bg[3], const float* final_t,
                     const uint32_t* n_contrib, const float* dimage, bool ordered,
                     float* slots, bool need_abs, const RasterVariant& rv = {},
                     const BatchDev* bd = nullptr, uint32_t* counter = nullptr);
// The forward and tree-reduced backward compositing of one view with pair counters
// (9 words at `pairs`; the forward clears them): forward visits, active visits, evaluated,
// passing and composited pairs; backward visits, active visits, evaluated and passing
// pairs.  They write what the uncounted kernels write.
void raster_forward_counted(splat_ctx* ctx, const ProjBuffers& pb, const BinBuffers& bb,
                            uint32_t width, uint32_t height, const float bg[3], float* image,
                            float* final_t, uint32_t* n_contrib, unsigned long long* pairs);
void raster_backward_counted(splat_ctx* ctx, const ProjBuffers& pb, const BinBuffers& bb,
                             uint32_t width, uint32_t height, const float bg[3],
                             const float* final_t, const uint32_t* n_contrib,
                             const float* dimage, float* slots, unsigned long long* pairs);
void reduce_slots(splat_ctx* ctx, const uint32_t* entry_offset, const float* slots, size_t n,
                  splat_proj_grad* grads, float* abs_grads, const StepStatus* status = nullptr);
// Autotuner of the compositing launch shapes: times every candidate (best of 3, HIP events)
// on buffers holding a binned view (forward) and its dL/dimage (backward).
struct RasterTune {
  RasterVariant fwd, bwd;
  float fwd_ms = 0, bwd_ms = 0, default_fwd_ms = 0, default_bwd_ms = 0;
  uint32_t candidates = 0;
};
RasterTune tune_raster(splat_ctx* ctx, const ProjBuffers& pb, const BinBuffers& bb,
                       uint32_t width, uint32_t height, const float bg[3], float* image,
                       float* final_t, uint32_t* n_contrib, const float* dimage, bool ordered,
                       float* slots, bool need_abs, const BatchDev* bd, uint32_t* counter);
void fill_tune_result(const RasterTune& t, splat_tune_result* out);

// dataset.hip: image::imageops::resize(.., Lanczos3) of a device RGB8 image into dst
// (out_w x out_h), as dataset::load downscales views; synchronises.
void resize_rgb8_device(splat_ctx* ctx, const uint8_t* src, uint32_t w, uint32_t h,
                        uint32_t out_w, uint32_t out_h, uint8_t* dst);

// loss.hip
struct LossOut {
  float* dimage;     // width*height*3
  float* partials;   // per-block (ssim_sum, l1_sum) pairs
  int num_partials;
};
// With `reports` (a graph-captured train step, `dyn` and `status` given): the loss value
// (to dyn->loss_out when set) and the step's report (reports[dyn->report_slot]) are
// written by the gradient pass's first workgroup.
void loss_forward_backward(splat_ctx* ctx, const float* pred, const float* target_f32,
                           const uint8_t* target_u8, uint32_t width, uint32_t height,
                           float lambda, bool want_l1, float* dimage, float* scratch,
                           double* loss_parts_host, const StepDyn* dyn = nullptr,
                           StepReport* reports = nullptr, const StepStatus* status = nullptr);
size_t loss_scratch_floats(uint32_t width, uint32_t height);
hipError_t loss_launch_bare(hipStream_t stream, const float* pred, const float* target,
                            uint32_t width, uint32_t height, float lambda, float* dimage,
                            float* loss_out, float* scratch);
// Resolves this device's address of the no-allocation scratch array (once per device).
int noalloc_setup(int device);
// Blocks of the loss kernels for one image (the batched grid is the sum over pairs).
unsigned loss_blocks(uint32_t width, uint32_t height);
// Batched: loss and dL/dimage of every pair in two launches; scratch holds 3 maps of
// total_values floats (pixels x 3 over all pairs) and 2 floats per block; losses (optional,
// device) gets one value per pair.
void loss_forward_backward_batch(splat_ctx* ctx, const float* pred, const BatchDev& bd,
                                 unsigned total_blocks, size_t total_values, float lambda,
                                 float* dimage, float* scratch, float* losses);
// The loss value of the last loss_forward_backward on `scratch`, into device memory
// (`out_dev`, or dyn->loss_out when `dyn` is given).
void loss_value_to(splat_ctx* ctx, const float* scratch, uint32_t width, uint32_t height,
                   float lambda, float* out_dev, const StepDyn* dyn = nullptr);
void blur_image(splat_ctx* ctx, const float* src, uint32_t width, uint32_t height,
                uint32_t channels, float* dst);
void gaussian_kernel_host(float k[11]);

// optim.hip
void adam_planes(splat_ctx* ctx, float* params, const float* grads, float* m, float* v,
                 size_t n, size_t stride, float c1, float c2, const splat_adam_state& st,
                 const splat_learning_rates& lr, const StepDyn* dyn = nullptr,
                 const StepStatus* status = nullptr);
struct DensifyParams {
  float grad_threshold, size_limit, min_opacity, world_limit, shrink;
  int prune_big_world, has_max_screen;
  uint32_t max_screen, children;
};
DensifyParams densify_params(const splat_densify_config& cfg, float extent);

// Densification decisions of one Gaussian as bits: clone, split, prune the row, prune a
// copy of the row (the copy is not subject to the screen-size rule).
constexpr uint32_t kDecClone = 1, kDecSplit = 2, kDecPrune = 4, kDecPruneCopy = 8;
// Not a decision: the row's image share was below the learned prune head's bar when logged
// (ignored by densification).
constexpr uint32_t kDecLowShare = 16;
constexpr uint32_t kNoSample = 0xffffffffu;

// Per-Gaussian lineage carried through densification (the learned scorer's age feature and
// outcome labels): birth iteration and the logged decision it descends from.
struct DensifyLineage {
  const uint32_t* birth_in;
  const uint32_t* origin_in;
  const uint32_t* sample_of;  // logged decision id of each row at this event, or kNoSample
  uint32_t* birth_out;
  uint32_t* origin_out;
  uint32_t iteration;
};

// The reference's decisions (DecisionBits) for every Gaussian.
void densify_decide(splat_ctx* ctx, const float* params, size_t n, size_t stride,
                    const float* grad_accum, const uint32_t* denom, const uint32_t* max_radii,
                    const splat_densify_config& cfg, float extent, uint8_t* out);
// Counter-based standard normal number i of stream `seed`: SplitMix64 of (seed, 2i) and
// (seed, 2i + 1) give two uniform doubles, Box-Muller as the reference's normal_sample
// (src/densify.rs:116-121, u1 kept >= 1e-12).
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ float normal_draw(uint64_t seed, uint64_t i) {
  const uint64_t a = splitmix64(seed ^ splitmix64(2 * i));
  const uint64_t b = splitmix64(seed ^ splitmix64(2 * i + 1));
  double u1 = (double)(a >> 11) * 0x1.0p-53;
  const double u2 = (double)(b >> 11) * 0x1.0p-53;
  u1 = u1 > 1e-12 ? u1 : 1e-12;
  return (float)(sqrt(-2.0 * log(u1)) * cos(2.0 * 3.141592653589793 * u2));
}

// `decide` (optional) replaces the reference's clone/split/prune decisions; `lineage`
// (optional) carries birth/origin planes.  The split children's offsets use the standard
// normals `normals_dev` (3 per child, in split-row order), or, when it is null, the
// draws normal_draw(normals_seed, i) for the same indices i.
size_t densify_planes(splat_ctx* ctx, const float* params, const float* m, const float* v,
                      size_t n, size_t stride, const float* grad_accum, const uint32_t* denom,
                      const uint32_t* max_radii, const splat_densify_config& cfg, float extent,
                      const float* normals_dev, size_t normals_len, uint64_t normals_seed,
                      float* po, float* mo, float* vo, size_t out_capacity, size_t out_stride,
                      splat_densify_report* report, const uint8_t* decide = nullptr,
                      const DensifyLineage* lineage = nullptr);
void reset_opacity_planes(splat_ctx* ctx, float* params, float* m, float* v, size_t n,
                          size_t stride, float limit);
void stats_accumulate_dev(splat_ctx* ctx, const uint32_t* radius,
                          const splat_proj_grad* grads, const float* abs_grads, size_t n,
                          uint32_t width, uint32_t height, float* grad_accum, uint32_t* denom,
                          uint32_t* max_radii, const StepStatus* status = nullptr,
                          float* max_grad = nullptr, uint32_t* strong = nullptr,
                          float strong_threshold = 0.0f);

}  // namespace splat
