// Tile binning (src/rasterize.rs:60-128) as a two-level sort.
//
// The reference sorts (tile << 32 | depth bits, entry) pairs.  Keys are unique (the entry
// index breaks ties), so any correct sort yields its order.  Here:
//   1. entry_offset = exclusive scan of each Gaussian's tile count (projection wrote it);
//   2. keygen (one lane per entry): Gaussian by binary search in entry_offset, tile from
//      the rectangle, and an atomic per-tile counter that gives the entry a slot;
//   3. tile_start = exclusive scan of the per-tile counts;
//   4. scatter (depth bits << 32 | entry) into its tile's bucket;
//   5. sort each bucket by that 64-bit value: bitonic in LDS (tiles up to 1024 entries in
//      256-lane workgroups; up to 16384 in 1024-lane workgroups with 128 KB of LDS) and
//      in global memory beyond.  This equals the reference's (key, entry) order.
#include "internal.hpp"

namespace splat {
namespace {

constexpr int kSmallCap = 1024;   // entries per tile sorted by the 256-lane kernel
constexpr int kBigCap = 16384;    // entries per tile sorted in LDS by the 1024-lane kernel

// Last g in [lo, hi) with off[g] <= e, given off[lo] <= e < off[hi] (Gaussians with zero
// entries are skipped: they share their offset with the next Gaussian).
template <class Off>
__device__ __forceinline__ uint32_t search_offsets(Off off, uint32_t lo, uint32_t hi, uint32_t e) {
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off(mid) <= e)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

constexpr uint32_t kKeygenSpan = 2048;  // Gaussian offsets staged per keygen workgroup

// Gaussian of entry e for a workgroup of consecutive entries (call from every lane; e may
// be past e_total).  Two lanes search entry_offset for the workgroup's first and last
// entry; the offsets in between are staged in LDS and each lane searches there.  Spans of
// more than kKeygenSpan Gaussians (long runs of culled ones) search global memory.
__device__ __forceinline__ uint32_t entry_gaussian_of(uint32_t e, uint32_t e_total,
                                                      const uint32_t* __restrict__ entry_offset,
                                                      uint32_t n, uint32_t* s_off, uint32_t* s_g) {
  const uint32_t eb = blockIdx.x * blockDim.x;
  if (eb >= e_total) return 0u;  // uniform: no entry of this workgroup exists
  const auto global_off = [&](uint32_t g) { return entry_offset[g]; };
  if (threadIdx.x < 2) {
    const uint32_t key = threadIdx.x == 0 ? eb : min(eb + blockDim.x, e_total) - 1;
    s_g[threadIdx.x] = search_offsets(global_off, 0u, n, key);
  }
  __syncthreads();
  const uint32_t g0 = s_g[0], g1 = s_g[1];
  const bool staged = g1 - g0 + 1 <= kKeygenSpan;
  if (staged)
    for (uint32_t q = threadIdx.x; q <= g1 - g0 + 1; q += blockDim.x) s_off[q] = entry_offset[g0 + q];
  __syncthreads();
  if (e >= e_total) return 0u;
  if (staged)
    return g0 + search_offsets([&](uint32_t q) { return s_off[q]; }, 0u, g1 - g0 + 1, e);
  return search_offsets(global_off, g0, g1 + 1, e);
}

__global__ void __launch_bounds__(256) keygen_kernel(const uint32_t* __restrict__ entry_offset,
                                                     uint32_t n, uint32_t e_total,
                                                     const uint2* __restrict__ rect,
                                                     uint32_t tiles_x,
                                                     uint32_t* __restrict__ entry_gaussian,
                                                     uint32_t* __restrict__ entry_tile,
                                                     uint32_t* __restrict__ entry_slot,
                                                     uint32_t* __restrict__ tile_count,
                                                     const StepStatus* __restrict__ status) {
  __shared__ uint32_t s_off[kKeygenSpan + 1];
  __shared__ uint32_t s_g[2];
  if (status) {
    if (status->abort) return;
    e_total = status->entries;
  }
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t g = entry_gaussian_of(e, e_total, entry_offset, n, s_off, s_g);
  if (e >= e_total) return;
  const uint32_t local = e - entry_offset[g];
  const uint2 rc = rect[g];
  const uint32_t x0 = rc.x & 0xffffu, y0 = rc.x >> 16, x1 = rc.y & 0xffffu;
  const uint32_t w = x1 - x0;
  const uint32_t ty = y0 + local / w, tx = x0 + local % w;
  const uint32_t tile = ty * tiles_x + tx;
  entry_gaussian[e] = g;
  entry_tile[e] = tile;
  entry_slot[e] = atomicAdd(&tile_count[tile], 1u);
}

// keygen over the global tile space of a batched step: the row's pair gives its tile grid
// and tile offset.
__global__ void __launch_bounds__(256) keygen_batch_kernel(const uint32_t* __restrict__ entry_offset,
                                                           uint32_t n, BatchDev bd,
                                                           const uint2* __restrict__ rect,
                                                           uint32_t* __restrict__ entry_gaussian,
                                                           uint32_t* __restrict__ entry_tile,
                                                           uint32_t* __restrict__ entry_slot,
                                                           uint32_t* __restrict__ tile_count,
                                                           const StepStatus* __restrict__ status) {
  __shared__ uint32_t s_off[kKeygenSpan + 1];
  __shared__ uint32_t s_g[2];
  if (status->abort) return;
  const uint32_t e_total = status->entries;
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t g = entry_gaussian_of(e, e_total, entry_offset, n, s_off, s_g);
  if (e >= e_total) return;
  const uint32_t v = find_segment(bd.work_base, bd.npairs, g);
  const uint32_t local = e - entry_offset[g];
  const uint2 rc = rect[g];
  const uint32_t x0 = rc.x & 0xffffu, y0 = rc.x >> 16, x1 = rc.y & 0xffffu;
  const uint32_t w = x1 - x0;
  const uint32_t ty = y0 + local / w, tx = x0 + local % w;
  const uint32_t tile = bd.tile_base[v] + ty * bd.tiles_x[v] + tx;
  entry_gaussian[e] = g;
  entry_tile[e] = tile;
  entry_slot[e] = atomicAdd(&tile_count[tile], 1u);
}

__global__ void __launch_bounds__(256) scatter_kernel(uint32_t e_total,
                                                      const uint32_t* __restrict__ entry_gaussian,
                                                      const uint32_t* __restrict__ entry_tile,
                                                      const uint32_t* __restrict__ entry_slot,
                                                      const uint32_t* __restrict__ tile_start,
                                                      const float* __restrict__ depth,
                                                      uint64_t* __restrict__ bucket,
                                                      const StepStatus* __restrict__ status) {
  if (status) {
    if (status->abort) return;
    e_total = status->entries;
  }
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= e_total) return;
  const uint32_t pos = tile_start[entry_tile[e]] + entry_slot[e];
  bucket[pos] = ((uint64_t)f2bits(depth[entry_gaussian[e]]) << 32) | e;
}

// Entry count of the step and the capacity check of a graph-captured step.
__global__ void guard_kernel(const unsigned long long* __restrict__ total64,
                             unsigned long long capacity, StepStatus* __restrict__ status) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const unsigned long long t = *total64;
  status->total = t;
  status->entries = t > 0xffffffffull ? 0xffffffffu : (uint32_t)t;
  if (t > capacity) status->abort = 1u;
}

// One bitonic step over `n` keys (virtual +inf padding beyond n).  Pairs p in [0, half).
template <class Load, class Store>
__device__ __forceinline__ void bitonic_pairs(uint32_t n, uint32_t half, uint32_t size,
                                              uint32_t stride, bool flip, Load ld, Store st) {
  for (uint32_t p = threadIdx.x; p < half; p += blockDim.x) {
    uint32_t i, j;
    if (flip) {
      const uint32_t h = size >> 1;
      const uint32_t blk = p / h, off = p % h;
      i = blk * size + off;
      j = blk * size + size - 1 - off;
    } else {
      i = (p / stride) * 2 * stride + p % stride;
      j = i + stride;
    }
    if (j < n) {
      const uint64_t a = ld(i), b = ld(j);
      if (a > b) {
        st(i, b);
        st(j, a);
      }
    }
  }
}

__device__ __forceinline__ uint32_t next_pow2(uint32_t v) {
  uint32_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

template <class Load, class Store>
__device__ void bitonic_sort(uint32_t n, Load ld, Store st) {
  const uint32_t N = next_pow2(n);
  const uint32_t half = N >> 1;
  for (uint32_t size = 2; size <= N; size <<= 1) {
    bitonic_pairs(n, half, size, 0, true, ld, st);
    __syncthreads();
    for (uint32_t stride = size >> 2; stride >= 1; stride >>= 1) {
      bitonic_pairs(n, half, size, stride, false, ld, st);
      __syncthreads();
    }
  }
}

__device__ __forceinline__ void write_sorted(uint32_t pos, uint64_t v, uint32_t tile,
                                             const uint32_t* __restrict__ entry_gaussian,
                                             uint32_t* __restrict__ sorted_entry,
                                             uint32_t* __restrict__ sorted_gauss,
                                             uint64_t* __restrict__ keys_out) {
  const uint32_t entry = (uint32_t)v;
  sorted_entry[pos] = entry;
  sorted_gauss[pos] = entry_gaussian[entry];
  if (keys_out) keys_out[pos] = ((uint64_t)tile << 32) | (v >> 32);
}

// Tiles with at most kSmallCap entries; larger tiles are queued for the big kernel.
__global__ void __launch_bounds__(256) sort_small_kernel(uint32_t num_tiles,
                                                         const uint32_t* __restrict__ tile_start,
                                                         uint64_t* __restrict__ bucket,
                                                         const uint32_t* __restrict__ entry_gaussian,
                                                         uint32_t* __restrict__ sorted_entry,
                                                         uint32_t* __restrict__ sorted_gauss,
                                                         uint64_t* __restrict__ keys_out,
                                                         uint32_t* __restrict__ big_list,
                                                         uint32_t* __restrict__ big_count) {
  __shared__ uint64_t sh[kSmallCap];
  const uint32_t tile = blockIdx.x;
  const uint32_t s = tile_start[tile], n = tile_start[tile + 1] - s;
  if (n == 0) return;
  if (n > kSmallCap) {
    if (threadIdx.x == 0) big_list[atomicAdd(big_count, 1u)] = tile;
    return;
  }
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) sh[i] = bucket[s + i];
  __syncthreads();
  bitonic_sort(n, [&](uint32_t i) { return sh[i]; }, [&](uint32_t i, uint64_t v) { sh[i] = v; });
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
    write_sorted(s + i, sh[i], tile, entry_gaussian, sorted_entry, sorted_gauss, keys_out);
}

// Queued tiles: LDS bitonic up to kBigCap entries, global-memory bitonic beyond.  The
// grid strides over the queue; every workgroup exits once the queue is exhausted.
__global__ void __launch_bounds__(1024) sort_big_kernel(const uint32_t* __restrict__ tile_start,
                                                        uint64_t* __restrict__ bucket,
                                                        const uint32_t* __restrict__ entry_gaussian,
                                                        uint32_t* __restrict__ sorted_entry,
                                                        uint32_t* __restrict__ sorted_gauss,
                                                        uint64_t* __restrict__ keys_out,
                                                        const uint32_t* __restrict__ big_list,
                                                        const uint32_t* __restrict__ big_count) {
  __shared__ uint64_t sh[kBigCap];
  const uint32_t count = *big_count;
  for (uint32_t q = blockIdx.x; q < count; q += gridDim.x) {
    const uint32_t tile = big_list[q];
    const uint32_t s = tile_start[tile], n = tile_start[tile + 1] - s;
    if (n <= kBigCap) {
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) sh[i] = bucket[s + i];
      __syncthreads();
      bitonic_sort(n, [&](uint32_t i) { return sh[i]; },
                   [&](uint32_t i, uint64_t v) { sh[i] = v; });
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        write_sorted(s + i, sh[i], tile, entry_gaussian, sorted_entry, sorted_gauss, keys_out);
    } else {
      uint64_t* seg = bucket + s;
      bitonic_sort(n, [&](uint32_t i) { return seg[i]; },
                   [&](uint32_t i, uint64_t v) { seg[i] = v; });
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        write_sorted(s + i, seg[i], tile, entry_gaussian, sorted_entry, sorted_gauss, keys_out);
    }
    __syncthreads();
  }
}

}  // namespace

size_t bin_build(splat_ctx* ctx, ProjBuffers pb, size_t n, uint32_t tiles_x, uint32_t tiles_y,
                 BinBuffers& bb, size_t* capacity_io, std::map<std::string, DevBuf>& store,
                 uint64_t* keys_out) {
  const size_t T = (size_t)tiles_x * tiles_y;
  bb.tiles_x = tiles_x;
  bb.tiles_y = tiles_y;
  bb.entry_offset = (uint32_t*)store["entry_offset"].reserve((n + 1) * 4);
  bb.tile_count = (uint32_t*)store["tile_count"].reserve(T * 4);
  bb.tile_start = (uint32_t*)store["tile_start"].reserve((T + 1) * 4);
  uint32_t* small = (uint32_t*)store["bin_small"].reserve(64);  // [0] E, [1] big count, [2] T
  unsigned long long* total64 = (unsigned long long*)store["bin_total64"].reserve(8);
  uint32_t* big_list = (uint32_t*)store["bin_big_list"].reserve(T * 4);

  // 1. entry offsets; the 64-bit total guards the u32 offsets (src/rasterize.rs:74)
  fill_u32(ctx, reinterpret_cast<uint32_t*>(total64), 2, 0u);
  exclusive_scan_u32(ctx, pb.ntiles, bb.entry_offset, n, small, nullptr, total64);
  unsigned long long e64 = 0;
  SPLAT_HIP(hipMemcpyAsync(&e64, total64, 8, hipMemcpyDeviceToHost, ctx->stream));
  SPLAT_HIP(hipStreamSynchronize(ctx->stream));
  SPLAT_REQUIRE(e64 < 0xffffffffull, "too many tile entries");
  const size_t E = (size_t)e64;

  if (*capacity_io < E || !store.count("entry_gaussian")) {
    const size_t cap = E + E / 4 + 1024;
    *capacity_io = cap;
    store["entry_gaussian"].reserve(cap * 4);
    store["entry_tile"].reserve(cap * 4);
    store["entry_slot"].reserve(cap * 4);
    store["bucket"].reserve(cap * 8);
    store["sorted_entry"].reserve(cap * 4);
    store["sorted_gauss"].reserve(cap * 4);
  }
  bb.entry_gaussian = store["entry_gaussian"].as<uint32_t>();
  bb.entry_tile = store["entry_tile"].as<uint32_t>();
  bb.entry_slot = store["entry_slot"].as<uint32_t>();
  bb.bucket = store["bucket"].as<uint64_t>();
  bb.sorted_entry = store["sorted_entry"].as<uint32_t>();
  bb.sorted_gauss = store["sorted_gauss"].as<uint32_t>();

  // 2. keygen + per-tile counts
  fill_u32(ctx, bb.tile_count, T, 0);
  fill_u32(ctx, small + 1, 1, 0);
  launch(ctx, "bin_keygen", keygen_kernel, dim3(blocks_for(E, 256)), dim3(256), 0,
         (const uint32_t*)bb.entry_offset, (uint32_t)n, (uint32_t)E, (const uint2*)pb.rect,
         tiles_x, bb.entry_gaussian, bb.entry_tile, bb.entry_slot, bb.tile_count,
         (const StepStatus*)nullptr);
  // 3. tile starts
  exclusive_scan_u32(ctx, bb.tile_count, bb.tile_start, T, small + 2);
  // 4. scatter into buckets
  launch(ctx, "bin_scatter", scatter_kernel, dim3(blocks_for(E, 256)), dim3(256), 0,
         (uint32_t)E, (const uint32_t*)bb.entry_gaussian, (const uint32_t*)bb.entry_tile,
         (const uint32_t*)bb.entry_slot, (const uint32_t*)bb.tile_start,
         (const float*)pb.depth, bb.bucket, (const StepStatus*)nullptr);
  // 5. per-tile sorts
  launch(ctx, "bin_sort_small", sort_small_kernel, dim3((unsigned)T), dim3(256), 0, (uint32_t)T,
         (const uint32_t*)bb.tile_start, bb.bucket, (const uint32_t*)bb.entry_gaussian,
         bb.sorted_entry, bb.sorted_gauss, keys_out, big_list, small + 1);
  launch(ctx, "bin_sort_big", sort_big_kernel, dim3(256), dim3(1024), 0,
         (const uint32_t*)bb.tile_start, bb.bucket, (const uint32_t*)bb.entry_gaussian,
         bb.sorted_entry, bb.sorted_gauss, keys_out, (const uint32_t*)big_list,
         (const uint32_t*)(small + 1));
  return E;
}

void bin_guard(splat_ctx* ctx, const unsigned long long* total64, size_t capacity,
               StepStatus* status) {
  launch(ctx, "bin_guard", guard_kernel, dim3(1), dim3(64), 0, total64,
         (unsigned long long)capacity, status);
}

void bin_build_device(splat_ctx* ctx, ProjBuffers pb, size_t n, const BinBuffers& bb,
                      size_t capacity, uint32_t* small, unsigned long long* total64,
                      uint32_t* big_list, uint32_t* scan_sums, StepStatus* status,
                      const BatchDev* bd) {
  const size_t T = (size_t)bb.tiles_x * bb.tiles_y;
  fill_u32(ctx, reinterpret_cast<uint32_t*>(total64), 2, 0u);
  exclusive_scan_u32(ctx, pb.ntiles, bb.entry_offset, n, small, scan_sums, total64);
  bin_guard(ctx, total64, capacity, status);
  fill_u32(ctx, bb.tile_count, T, 0);
  fill_u32(ctx, small + 1, 1, 0);
  // The entry kernels cover the capacity and read the step's count from `status`.
  if (bd)
    launch(ctx, "bin_keygen", keygen_batch_kernel, dim3(blocks_for(capacity, 256)), dim3(256), 0,
           (const uint32_t*)bb.entry_offset, (uint32_t)n, *bd, (const uint2*)pb.rect,
           bb.entry_gaussian, bb.entry_tile, bb.entry_slot, bb.tile_count,
           (const StepStatus*)status);
  else
    launch(ctx, "bin_keygen", keygen_kernel, dim3(blocks_for(capacity, 256)), dim3(256), 0,
           (const uint32_t*)bb.entry_offset, (uint32_t)n, 0u, (const uint2*)pb.rect, bb.tiles_x,
           bb.entry_gaussian, bb.entry_tile, bb.entry_slot, bb.tile_count,
           (const StepStatus*)status);
  exclusive_scan_u32(ctx, bb.tile_count, bb.tile_start, T, small + 2, scan_sums);
  launch(ctx, "bin_scatter", scatter_kernel, dim3(blocks_for(capacity, 256)), dim3(256), 0, 0u,
         (const uint32_t*)bb.entry_gaussian, (const uint32_t*)bb.entry_tile,
         (const uint32_t*)bb.entry_slot, (const uint32_t*)bb.tile_start, (const float*)pb.depth,
         bb.bucket, (const StepStatus*)status);
  launch(ctx, "bin_sort_small", sort_small_kernel, dim3((unsigned)T), dim3(256), 0, (uint32_t)T,
         (const uint32_t*)bb.tile_start, bb.bucket, (const uint32_t*)bb.entry_gaussian,
         bb.sorted_entry, bb.sorted_gauss, (uint64_t*)nullptr, big_list, small + 1);
  launch(ctx, "bin_sort_big", sort_big_kernel, dim3(256), dim3(1024), 0,
         (const uint32_t*)bb.tile_start, bb.bucket, (const uint32_t*)bb.entry_gaussian,
         bb.sorted_entry, bb.sorted_gauss, (uint64_t*)nullptr, (const uint32_t*)big_list,
         (const uint32_t*)(small + 1));
}

}  // namespace splat
