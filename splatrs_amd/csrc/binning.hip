// Tile binning (src/rasterize.rs:60-128) as a two-level sort.
//
// The reference sorts (tile << 32 | depth bits, entry) pairs.  Keys are unique (the entry
// index breaks ties), so any correct sort yields its order.  Here:
//   1. entry_offset = exclusive scan of each Gaussian's tile count (projection wrote it);
//   2. keygen (one lane per entry): Gaussian by binary search in entry_offset, tile from
//      the rectangle, and an atomic counter that gives the entry a slot.  Each tile has
//      kSub counters, one per workgroup index mod kSub (workgroups are dealt to the XCDs
//      in turn, so a counter set stays in one XCD's cache lines), laid out [kSub][T];
//   3. per tile, the counters' offsets within the tile and its total; tile_start =
//      exclusive scan of the totals;
//   4. scatter one 16-byte record per entry into its tile's bucket: the key (depth bits
//      << 32 | entry) and the entry's Gaussian;
//   5. sort each bucket by that 64-bit value, the Gaussian carried along: tiles of up to
//      1024 entries by one wave each (a bitonic network in registers, lanes exchanging
//      by DPP and the gfx950 row swaps), up to 16384 by a bitonic sort in LDS (1024-lane workgroups,
//      128 KB), in global memory beyond.  This equals the reference's (key, entry) order.
#include "internal.hpp"

namespace splat {
namespace {

constexpr int kWaveCap = 512;     // entries per tile sorted by one wave (8 per lane)
constexpr int kMidCap = 1024;     // ... by one wave of the second kernel (16 per lane)
constexpr uint32_t kKeygenThreads = 256;
constexpr uint32_t kKeygenPerLane = 1;
constexpr uint32_t kKeygenEntries = kKeygenThreads * kKeygenPerLane;  // per keygen workgroup
// (More entries per lane do not help: the step's 2.1M returning counter atomics bound the
// pass, not the per-workgroup offset lookup.)

// Counter set of the keygen workgroup that handles entry e.
__device__ __forceinline__ uint32_t entry_sub(uint32_t e) { return (e / kKeygenEntries) % kTileSub; }
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

constexpr uint32_t kKeygenSpan = 1024;  // Gaussians (offsets, rectangles) staged per keygen workgroup

// marks[b] = the Gaussian that holds entry b * EPB (the first entry of keygen workgroup b,
// EPB entries per workgroup), written by that Gaussian.
template <uint32_t EPB = kKeygenEntries>
__global__ void __launch_bounds__(256) block_marks_kernel(const uint32_t* __restrict__ entry_offset,
                                                          uint32_t n, uint32_t* __restrict__ marks,
                                                          const StepStatus* __restrict__ status) {
  if (status && status->abort) return;
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const uint32_t lo = entry_offset[g], hi = entry_offset[g + 1];
  for (uint32_t b = (lo + EPB - 1) / EPB; b * EPB < hi; b++) marks[b] = g;
}

// The Gaussians of one keygen workgroup's entries [b * kKeygenEntries, ...): they lie in
// [g0, ghi) with entry_offset[g0] <= every entry < entry_offset[ghi].  Their offsets are
// staged in LDS (call from every lane) unless the span exceeds SPAN Gaussians (long
// runs of culled ones); then lookups search global memory.
struct KeygenSpan {
  uint32_t g0, ghi;
  bool staged;
};

// (With `rect`, the span's rectangles are staged in s_rect as well.)  SPAN: the staging
// capacity in Gaussians.
template <uint32_t EPB = kKeygenEntries, uint32_t SPAN = kKeygenSpan>
__device__ __forceinline__ KeygenSpan keygen_span(uint32_t e_total,
                                                  const uint32_t* __restrict__ entry_offset,
                                                  uint32_t n, const uint32_t* __restrict__ marks,
                                                  uint32_t* s_off,
                                                  const uint2* __restrict__ rect = nullptr,
                                                  uint2* s_rect = nullptr) {
  const uint32_t b = blockIdx.x;
  const uint32_t nb = (e_total + EPB - 1) / EPB;
  KeygenSpan sp;
  sp.g0 = marks[b];
  sp.ghi = b + 1 < nb ? marks[b + 1] + 1 : n;
  sp.staged = sp.ghi - sp.g0 <= SPAN;
  if (sp.staged)
    for (uint32_t q = threadIdx.x; q <= sp.ghi - sp.g0; q += blockDim.x) {
      s_off[q] = entry_offset[sp.g0 + q];
      if (rect && q < sp.ghi - sp.g0) s_rect[q] = rect[sp.g0 + q];
    }
  __syncthreads();
  return sp;
}

// (x / w, x % w) for 0 < w < 2^16 and x < 2^22 (a rectangle of a view's tile grid, which
// has fewer than 2^22 tiles): the estimate from the reciprocal (relative error below
// 2^-22) is the quotient or one less, and one correction step makes it exact.
__device__ __forceinline__ uint2 divmod_small(uint32_t x, uint32_t w) {
  uint32_t q = (uint32_t)((float)x * __builtin_amdgcn_rcpf((float)w));
  int32_t r = (int32_t)(x - q * w);
  if (r < 0) {
    q--;
    r += (int32_t)w;
  } else if (r >= (int32_t)w) {
    q++;
    r -= (int32_t)w;
  }
  return make_uint2(q, (uint32_t)r);
}

// Tile of an entry within its rectangle of width w: (local / w, local % w), exact for any
// local (the reciprocal form covers every rectangle of fewer than 2^22 tiles).
__device__ __forceinline__ uint2 rect_divmod(uint32_t local, uint32_t w) {
  if (local < (1u << 22)) return divmod_small(local, w);
  return make_uint2(local / w, local % w);
}

__device__ __forceinline__ uint32_t keygen_find(const KeygenSpan& sp, uint32_t e,
                                                const uint32_t* __restrict__ entry_offset,
                                                const uint32_t* s_off) {
  if (sp.staged)
    return sp.g0 + search_offsets([&](uint32_t q) { return s_off[q]; }, 0u, sp.ghi - sp.g0, e);
  return search_offsets([&](uint32_t g) { return entry_offset[g]; }, sp.g0, sp.ghi, e);
}

// The span's rectangles are staged beside the offsets, so that after the search an entry
// needs no dependent global load before its counter atomic.
__global__ void __launch_bounds__(256) keygen_kernel(const uint32_t* __restrict__ entry_offset,
                                                     uint32_t n, uint32_t e_total,
                                                     const uint2* __restrict__ rect,
                                                     uint32_t tiles_x,
                                                     uint32_t* __restrict__ entry_gaussian,
                                                     uint32_t* __restrict__ entry_tile,
                                                     uint32_t* __restrict__ entry_slot,
                                                     uint32_t* __restrict__ tile_count,
                                                     uint32_t num_tiles,
                                                     const uint32_t* __restrict__ marks,
                                                     const StepStatus* __restrict__ status) {
  __shared__ uint32_t s_off[kKeygenSpan + 1];
  __shared__ uint2 s_rect[kKeygenSpan];
  if (status) {
    if (status->abort) return;
    e_total = status->entries;
  }
  const uint32_t eb = blockIdx.x * kKeygenEntries;
  if (eb >= e_total) return;  // uniform: no entry of this workgroup exists
  const KeygenSpan sp = keygen_span(e_total, entry_offset, n, marks, s_off, rect, s_rect);
#pragma unroll
  for (uint32_t k = 0; k < kKeygenPerLane; k++) {
    const uint32_t e = eb + k * kKeygenThreads + threadIdx.x;
    if (e >= e_total) break;
    const uint32_t g = keygen_find(sp, e, entry_offset, s_off);
    const uint32_t local = e - (sp.staged ? s_off[g - sp.g0] : entry_offset[g]);
    const uint2 rc = sp.staged ? s_rect[g - sp.g0] : rect[g];
    const uint32_t x0 = rc.x & 0xffffu, y0 = rc.x >> 16, x1 = rc.y & 0xffffu;
    const uint32_t w = x1 - x0;
    const uint2 qr = rect_divmod(local, w);
    const uint32_t ty = y0 + qr.x, tx = x0 + qr.y;
    const uint32_t tile = ty * tiles_x + tx;
    entry_gaussian[e] = g;
    entry_tile[e] = tile;
    entry_slot[e] = atomicAdd(&tile_count[entry_sub(e) * num_tiles + tile], 1u);
  }
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
                                                           uint32_t num_tiles,
                                                           const uint32_t* __restrict__ marks,
                                                           const StepStatus* __restrict__ status) {
  __shared__ uint32_t s_off[kKeygenSpan + 1];
  if (status->abort) return;
  const uint32_t e_total = status->entries;
  const uint32_t eb = blockIdx.x * kKeygenEntries;
  if (eb >= e_total) return;  // uniform: no entry of this workgroup exists
  const KeygenSpan sp = keygen_span(e_total, entry_offset, n, marks, s_off);
#pragma unroll
  for (uint32_t k = 0; k < kKeygenPerLane; k++) {
    const uint32_t e = eb + k * kKeygenThreads + threadIdx.x;
    if (e >= e_total) break;
    const uint32_t g = keygen_find(sp, e, entry_offset, s_off);
    const uint32_t v = find_segment(bd.work_base, bd.npairs, g);
    const uint32_t local = e - entry_offset[g];
    const uint2 rc = rect[g];
    const uint32_t x0 = rc.x & 0xffffu, y0 = rc.x >> 16, x1 = rc.y & 0xffffu;
    const uint32_t w = x1 - x0;
    const uint2 qr = rect_divmod(local, w);
    const uint32_t ty = y0 + qr.x, tx = x0 + qr.y;
    const uint32_t tile = bd.tile_base[v] + ty * bd.tiles_x[v] + tx;
    entry_gaussian[e] = g;
    entry_tile[e] = tile;
    entry_slot[e] = atomicAdd(&tile_count[entry_sub(e) * num_tiles + tile], 1u);
  }
}

// Per tile: each counter set's offset within the tile and the tile's total (the input of
// the tile_start scan).
__global__ void __launch_bounds__(256) tile_totals_kernel(uint32_t num_tiles,
                                                          const uint32_t* __restrict__ tile_count,
                                                          uint32_t* __restrict__ tile_sub,
                                                          uint32_t* __restrict__ tile_total) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= num_tiles) return;
  uint32_t acc = 0;
#pragma unroll
  for (uint32_t k = 0; k < kTileSub; k++) {
    tile_sub[k * num_tiles + t] = acc;
    acc += tile_count[k * num_tiles + t];
  }
  tile_total[t] = acc;
}

__global__ void __launch_bounds__(256) scatter_kernel(uint32_t e_total,
                                                      const uint32_t* __restrict__ entry_gaussian,
                                                      const uint32_t* __restrict__ entry_tile,
                                                      const uint32_t* __restrict__ entry_slot,
                                                      const uint32_t* __restrict__ tile_start,
                                                      const uint32_t* __restrict__ tile_sub,
                                                      uint32_t num_tiles,
                                                      const float* __restrict__ depth,
                                                      uint4* __restrict__ bucket,
                                                      const StepStatus* __restrict__ status) {
  if (status) {
    if (status->abort) return;
    e_total = status->entries;
  }
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= e_total) return;
  const uint32_t tile = entry_tile[e];
  const uint32_t g = entry_gaussian[e];
  const uint32_t pos = tile_start[tile] + tile_sub[entry_sub(e) * num_tiles + tile] + entry_slot[e];
  bucket[pos] = make_uint4(e, f2bits(depth[g]), g, 0u);  // key low word, high word, Gaussian
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
// `size` (>= 2) and `stride` are powers of two: shifts and masks index the pairs.
template <class Load, class Store>
__device__ __forceinline__ void bitonic_pairs(uint32_t n, uint32_t half, uint32_t size,
                                              uint32_t stride, bool flip, Load ld, Store st) {
  const uint32_t lh = (uint32_t)__builtin_ctz(size) - 1u;  // log2(size / 2)
  const uint32_t ls = stride ? (uint32_t)__builtin_ctz(stride) : 0u;
  for (uint32_t p = threadIdx.x; p < half; p += blockDim.x) {
    uint32_t i, j;
    if (flip) {
      const uint32_t blk = p >> lh, off = p & ((1u << lh) - 1u);
      i = (blk << (lh + 1u)) + off;
      j = (blk << (lh + 1u)) + size - 1 - off;
    } else {
      i = ((p >> ls) << (ls + 1u)) + (p & (stride - 1u));
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

// ---- per-tile sorts in one wave ----------------------------------------------------
// A tile of at most 64 K entries is sorted by one wave holding K (key, Gaussian) pairs per
// lane in registers: a bitonic network over N = 64 K positions (padding keys all ones),
// position i = lane K + k.  Compare-exchanges of positions less than K apart stay inside
// the lane; the others pair lane l with lane l ^ (stride / K), exchanged by ds_bpermute,
// each lane keeping the smaller or the larger of its pair.  No LDS array, no barrier.

// The value of lane ^ LX for LX = 1, 2, 4, 8: DPP within 16-lane rows (quad permutes, the
// half-row mirror, the rotation by 8).
template <uint32_t LX>
__device__ __forceinline__ uint32_t from_lane_xor(uint32_t v) {
  if constexpr (LX == 1) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, true);  // [1,0,3,2]
  } else if constexpr (LX == 2) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, true);  // [2,3,0,1]
  } else if constexpr (LX == 4) {  // quad reversal, then the half-row mirror
    const int r = __builtin_amdgcn_update_dpp(0, (int)v, 0x1B, 0xF, 0xF, true);
    return (uint32_t)__builtin_amdgcn_update_dpp(0, r, 0x141, 0xF, 0xF, true);
  } else {
    static_assert(LX == 8, "lane distance");
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, true);  // row_ror 8
  }
}

// For LX = 16, 32, the gfx950 swaps: permlane16_swap exchanges the odd rows of its first
// operand with the even rows of its second, permlane32_swap the upper half of the first
// with the lower half of the second.  Given the same value twice, the first result holds
// in every lane the value of the lower lane of its pair (lane & ~LX), the second the
// value of the upper one.
template <uint32_t LX>
__device__ __forceinline__ uint2 pair_values(uint32_t v) {
  if constexpr (LX == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return make_uint2(r[0], r[1]);
  } else {
    static_assert(LX == 32, "lane distance");
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return make_uint2(r[0], r[1]);
  }
}

// Compare-exchange of the in-lane pairs (k, k | S) at network size `size`.
template <int K, int S>
__device__ __forceinline__ void wave_sort_in_lane(uint64_t (&key)[K], uint32_t (&gid)[K],
                                                  uint32_t lane, uint32_t size) {
#pragma unroll
  for (int k = 0; k < K; k++) {
    if (k & S) continue;
    const int k2 = k | S;
    // keys are distinct except the all-ones padding, whose order does not matter, so
    // "a > b when ascending, a < b when descending" is (a > b) == up
    const bool up = ((lane * K + k) & size) == 0;
    const uint64_t a = key[k], b = key[k2];
    const bool sw = (a > b) == up;
    key[k] = sw ? b : a;
    key[k2] = sw ? a : b;
    const uint32_t ga = gid[k], gb = gid[k2];
    gid[k] = sw ? gb : ga;
    gid[k2] = sw ? ga : gb;
  }
}

// Compare-exchange with lane ^ LX (stride = LX K >= K): the pair's direction and this
// lane's side depend on the lane only (size > stride >= K).  Keys are distinct except the
// all-ones padding, whose order does not matter, so "take the smaller when keeping the
// minimum, else the larger" is one comparison equal to keep_min.
template <int K, uint32_t LX>
__device__ __forceinline__ void wave_sort_across(uint64_t (&key)[K], uint32_t (&gid)[K],
                                                 uint32_t lane, uint32_t size) {
  const bool up = ((lane * K) & size) == 0;
  const bool lower = (lane & LX) == 0;
  const bool keep_min = lower == up;
#pragma unroll
  for (int k = 0; k < K; k++) {
    if constexpr (LX <= 8) {
      const uint32_t lo = from_lane_xor<LX>((uint32_t)key[k]);
      const uint32_t hi = from_lane_xor<LX>((uint32_t)(key[k] >> 32));
      const uint32_t pg = from_lane_xor<LX>(gid[k]);
      const uint64_t pk = ((uint64_t)hi << 32) | lo;
      const bool take = (pk < key[k]) == keep_min;
      key[k] = take ? pk : key[k];
      gid[k] = take ? pg : gid[k];
    } else {
      const uint2 lo = pair_values<LX>((uint32_t)key[k]);
      const uint2 hi = pair_values<LX>((uint32_t)(key[k] >> 32));
      const uint2 g = pair_values<LX>(gid[k]);
      const uint64_t a = ((uint64_t)hi.x << 32) | lo.x, b = ((uint64_t)hi.y << 32) | lo.y;
      const bool take_a = (a < b) == keep_min;  // a: the lower lane's key, b: the upper's
      key[k] = take_a ? a : b;
      gid[k] = take_a ? g.x : g.y;
    }
  }
}

// The network's stages from (SIZE, STRIDE) on, unrolled so that every register index is
// a constant.
template <int K, uint32_t SIZE, uint32_t STRIDE>
__device__ __forceinline__ void wave_sort_stages(uint64_t (&key)[K], uint32_t (&gid)[K],
                                                 uint32_t lane) {
  if constexpr (STRIDE >= (uint32_t)K)
    wave_sort_across<K, STRIDE / K>(key, gid, lane, SIZE);
  else
    wave_sort_in_lane<K, (int)STRIDE>(key, gid, lane, SIZE);
  if constexpr (STRIDE > 1)
    wave_sort_stages<K, SIZE, STRIDE / 2>(key, gid, lane);
  else if constexpr (SIZE < 64u * K)
    wave_sort_stages<K, SIZE * 2, SIZE>(key, gid, lane);
}

template <int K>
__device__ __forceinline__ void wave_sort_tile(uint32_t lane, uint32_t tile, uint32_t s,
                                               uint32_t n, const uint4* __restrict__ bucket,
                                               uint32_t* __restrict__ sorted_entry,
                                               uint32_t* __restrict__ sorted_gauss,
                                               uint64_t* __restrict__ keys_out) {
  uint64_t key[K];
  uint32_t gid[K];
  // any assignment of the inputs to positions sorts; this one reads coalesced
#pragma unroll
  for (int k = 0; k < K; k++) {
    const uint32_t p = (uint32_t)k * 64u + lane;
    const uint4 r = p < n ? bucket[s + p] : make_uint4(~0u, ~0u, 0u, 0u);
    key[k] = ((uint64_t)r.y << 32) | r.x;
    gid[k] = r.z;
  }
  wave_sort_stages<K, 2, 1>(key, gid, lane);
#pragma unroll
  for (int k = 0; k < K; k++) {
    const uint32_t r = lane * K + k;
    if (r < n) {
      sorted_entry[s + r] = (uint32_t)key[k];
      sorted_gauss[s + r] = gid[k];
      if (keys_out) keys_out[s + r] = ((uint64_t)tile << 32) | (key[k] >> 32);
    }
  }
}

// One wave per tile (four per workgroup).  Larger tiles are queued in `queue` (T words):
// up to kMidCap entries from the end (count in *mid_count), beyond from the start (count
// in *big_count); a tile is queued at most once, so the two never meet.
__global__ void __launch_bounds__(256) sort_tiles_kernel(uint32_t num_tiles,
                                                         const uint32_t* __restrict__ tile_start,
                                                         const uint4* __restrict__ bucket,
                                                         uint32_t* __restrict__ sorted_entry,
                                                         uint32_t* __restrict__ sorted_gauss,
                                                         uint64_t* __restrict__ keys_out,
                                                         uint32_t* __restrict__ queue,
                                                         uint32_t* __restrict__ big_count,
                                                         uint32_t* __restrict__ mid_count) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t tile = blockIdx.x * 4u + (threadIdx.x >> 6);
  if (tile >= num_tiles) return;
  const uint32_t s = tile_start[tile], n = tile_start[tile + 1] - s;
  if (n == 0) return;
  if (n > (uint32_t)kWaveCap) {
    if (lane == 0) {
      if (n > (uint32_t)kMidCap)
        queue[atomicAdd(big_count, 1u)] = tile;
      else
        queue[num_tiles - 1 - atomicAdd(mid_count, 1u)] = tile;
    }
    return;
  }
  if (n <= 64)
    wave_sort_tile<1>(lane, tile, s, n, bucket, sorted_entry, sorted_gauss, keys_out);
  else if (n <= 128)
    wave_sort_tile<2>(lane, tile, s, n, bucket, sorted_entry, sorted_gauss, keys_out);
  else if (n <= 256)
    wave_sort_tile<4>(lane, tile, s, n, bucket, sorted_entry, sorted_gauss, keys_out);
  else
    wave_sort_tile<8>(lane, tile, s, n, bucket, sorted_entry, sorted_gauss, keys_out);
}

// The queued tiles of kWaveCap < n <= kMidCap entries, one wave each (the grid's waves
// stride over the queue; a wave leaves when it is exhausted).
__global__ void __launch_bounds__(256) sort_mid_kernel(uint32_t num_tiles,
                                                       const uint32_t* __restrict__ tile_start,
                                                       const uint4* __restrict__ bucket,
                                                       uint32_t* __restrict__ sorted_entry,
                                                       uint32_t* __restrict__ sorted_gauss,
                                                       uint64_t* __restrict__ keys_out,
                                                       const uint32_t* __restrict__ queue,
                                                       const uint32_t* __restrict__ mid_count) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t count = *mid_count;
  for (uint32_t q = blockIdx.x * 4u + (threadIdx.x >> 6); q < count; q += gridDim.x * 4u) {
    const uint32_t tile = queue[num_tiles - 1 - q];
    const uint32_t s = tile_start[tile], n = tile_start[tile + 1] - s;
    wave_sort_tile<16>(lane, tile, s, n, bucket, sorted_entry, sorted_gauss, keys_out);
  }
}

// Queued tiles: LDS bitonic up to kBigCap entries, global-memory bitonic beyond.  The
// grid strides over the queue; every workgroup exits once the queue is exhausted.
__global__ void __launch_bounds__(1024) sort_big_kernel(const uint32_t* __restrict__ tile_start,
                                                        uint4* __restrict__ bucket,
                                                        const uint32_t* __restrict__ entry_gaussian,
                                                        uint32_t* __restrict__ sorted_entry,
                                                        uint32_t* __restrict__ sorted_gauss,
                                                        uint64_t* __restrict__ keys_out,
                                                        const uint32_t* __restrict__ big_list,
                                                        const uint32_t* __restrict__ big_count) {
  __shared__ uint64_t sh[kBigCap];
  const uint32_t count = *big_count;
  // the records' keys (their Gaussians come from entry_gaussian)
  const auto rec_key = [&](uint32_t i) {
    const uint2 k = *reinterpret_cast<const uint2*>(&bucket[i]);
    return ((uint64_t)k.y << 32) | k.x;
  };
  for (uint32_t q = blockIdx.x; q < count; q += gridDim.x) {
    const uint32_t tile = big_list[q];
    const uint32_t s = tile_start[tile], n = tile_start[tile + 1] - s;
    if (n <= kBigCap) {
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) sh[i] = rec_key(s + i);
      __syncthreads();
      bitonic_sort(n, [&](uint32_t i) { return sh[i]; },
                   [&](uint32_t i, uint64_t v) { sh[i] = v; });
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        write_sorted(s + i, sh[i], tile, entry_gaussian, sorted_entry, sorted_gauss, keys_out);
    } else {
      bitonic_sort(n, [&](uint32_t i) { return rec_key(s + i); },
                   [&](uint32_t i, uint64_t v) {
                     *reinterpret_cast<uint2*>(&bucket[s + i]) =
                         make_uint2((uint32_t)v, (uint32_t)(v >> 32));
                   });
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        write_sorted(s + i, rec_key(s + i), tile, entry_gaussian, sorted_entry, sorted_gauss,
                     keys_out);
    }
    __syncthreads();
  }
}

// ---- radix binning ---------------------------------------------------------------------

constexpr uint32_t kStagedPairs = 1024;  // views whose work-row table is staged in LDS

// Depth keys of the rows (rows without entries sort last), their indices, and each row's
// record for keygen_ordered, in row order (coalesced): {first tile (row-major over the
// row's rectangle, src/rasterize.rs:86-99), rectangle width | tile-grid stride << 16,
// first entry id, row}.  Rows without entries get no record.
template <bool BATCH>
__global__ void __launch_bounds__(256) depth_keys_kernel(
    const float* __restrict__ depth, const uint32_t* __restrict__ ntiles,
    const uint2* __restrict__ rect, const uint32_t* __restrict__ entry_offset, uint32_t n,
    uint32_t tiles_x, BatchDev bd, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
    uint4* __restrict__ recs, const StepStatus* __restrict__ status) {
  __shared__ uint32_t s_wb[BATCH ? kStagedPairs + 1 : 1];  // the views' first work rows
  if (status && status->abort) return;  // uniform
  const bool staged = BATCH && bd.npairs <= kStagedPairs;
  if (staged) {
    for (uint32_t q = threadIdx.x; q <= bd.npairs; q += blockDim.x) s_wb[q] = bd.work_base[q];
    __syncthreads();
  }
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const uint32_t c = ntiles[g];
  keys[g] = c ? f2bits(depth[g]) : 0xffffffffu;  // visible depths are finite > 0
  vals[g] = g;
  if (c == 0) return;
  const uint2 rc = rect[g];
  const uint32_t x0 = rc.x & 0xffffu, y0 = rc.x >> 16, x1 = rc.y & 0xffffu;
  uint32_t first, stride;
  if (BATCH) {
    const uint32_t v = find_segment(staged ? s_wb : bd.work_base, bd.npairs, g);
    stride = bd.tiles_x[v];
    first = bd.tile_base[v] + y0 * stride + x0;
  } else {
    stride = tiles_x;
    first = y0 * stride + x0;
  }
  recs[g] = make_uint4(first, (x1 - x0) | (stride << 16), entry_offset[g], g);
}

// Entry counts and records of the rows in depth order.
__global__ void __launch_bounds__(256) ordered_rows_kernel(
    const uint32_t* __restrict__ order, const uint32_t* __restrict__ ntiles,
    const uint4* __restrict__ recs, uint32_t n, uint32_t* __restrict__ cnt,
    uint4* __restrict__ rows, const StepStatus* __restrict__ status) {
  if (status && status->abort) return;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t g = order[i];
  const uint32_t c = ntiles[g];
  cnt[i] = c;
  if (c) rows[i] = recs[g];
}

constexpr uint32_t kOrderedPerLane = 8;
constexpr uint32_t kOrderedEntries = kKeygenThreads * kOrderedPerLane;  // per workgroup
constexpr uint32_t kOrderedSpan = 2048;  // rows whose offsets are staged per workgroup

// Position j of the depth-ordered entry list: the row (soff, marks: its offsets in depth
// order), whose record gives the entry's tile (the local index row-major over the
// rectangle) and the reference's entry id (the row's Gaussian-order offset + the local
// index).
__global__ void __launch_bounds__(256) keygen_ordered_kernel(
    const uint32_t* __restrict__ soff, const uint4* __restrict__ rows, uint32_t n,
    uint32_t e_total, const uint32_t* __restrict__ marks, uint32_t* __restrict__ tkeys,
    uint32_t* __restrict__ tvals, uint32_t* __restrict__ tgauss,
    uint32_t* __restrict__ entry_gaussian, const StepStatus* __restrict__ status) {
  __shared__ uint32_t s_off[kOrderedSpan + 1];
  if (status) {
    if (status->abort) return;
    e_total = status->entries;
  }
  const uint32_t eb = blockIdx.x * kOrderedEntries;
  if (eb >= e_total) return;  // uniform: no entry of this workgroup exists
  const KeygenSpan sp = keygen_span<kOrderedEntries, kOrderedSpan>(e_total, soff, n, marks, s_off);
#pragma unroll
  for (uint32_t u = 0; u < kOrderedPerLane; u++) {
    const uint32_t j = eb + u * kKeygenThreads + threadIdx.x;
    if (j >= e_total) break;
    const uint32_t i = keygen_find(sp, j, soff, s_off);
    const uint4 row = rows[i];
    const uint32_t local = j - (sp.staged ? s_off[i - sp.g0] : soff[i]);
    const uint2 qr = divmod_small(local, row.y & 0xffffu);
    const uint32_t e = row.z + local;
    tkeys[j] = row.x + qr.x * (row.y >> 16) + qr.y;
    tvals[j] = e;
    tgauss[j] = row.w;
    if (entry_gaussian) entry_gaussian[e] = row.w;
  }
}

// The tile-sorted entries: tile ranges, (optional) keys and, when tvals is given, the
// sorted lists (otherwise the sort wrote them in place).  Position j starts every tile in
// (tile[j - 1], tile[j]]; the last position also ends the tiles after its own, so each
// tile_start word is written exactly once when there are entries.  Each lane takes
// kFinalizePerLane consecutive positions (one 16-byte load of their keys when all exist).
constexpr uint32_t kFinalizePerLane = 4;

__global__ void __launch_bounds__(256) radix_finalize_kernel(
    const uint32_t* __restrict__ tkeys, const uint32_t* __restrict__ tvals,
    const uint32_t* __restrict__ tgauss, uint32_t e_total, uint32_t num_tiles,
    const float* __restrict__ depth, uint32_t* __restrict__ sorted_entry,
    uint32_t* __restrict__ sorted_gauss, uint32_t* __restrict__ tile_start,
    uint64_t* __restrict__ keys_out, const StepStatus* __restrict__ status) {
  if (status) {
    if (status->abort) return;
    e_total = status->entries;
  }
  const uint32_t j0 = (blockIdx.x * blockDim.x + threadIdx.x) * kFinalizePerLane;
  if (j0 >= e_total) return;
  uint32_t tk[kFinalizePerLane];
  if (j0 + kFinalizePerLane <= e_total) {  // j0 is a multiple of 4: an aligned 16-byte load
    const uint4 q = *reinterpret_cast<const uint4*>(tkeys + j0);
    tk[0] = q.x;
    tk[1] = q.y;
    tk[2] = q.z;
    tk[3] = q.w;
  } else {
#pragma unroll
    for (uint32_t k = 0; k < kFinalizePerLane; k++)
      tk[k] = j0 + k < e_total ? tkeys[j0 + k] : 0u;
  }
  uint32_t u0 = j0 == 0 ? 0u : tkeys[j0 - 1] + 1u;  // first tile position j0 starts
#pragma unroll
  for (uint32_t k = 0; k < kFinalizePerLane; k++) {
    const uint32_t j = j0 + k;
    if (j >= e_total) break;
    const uint32_t t = tk[k];
    if (tvals) {
      sorted_entry[j] = tvals[j];
      sorted_gauss[j] = tgauss[j];
    }
    if (keys_out) {
      const uint32_t g = tvals ? tgauss[j] : sorted_gauss[j];
      keys_out[j] = ((uint64_t)t << 32) | f2bits(depth[g]);
    }
    for (uint32_t u = u0; u <= t; u++) tile_start[u] = j;
    u0 = t + 1;
    if (j + 1 == e_total)
      for (uint32_t u = t + 1; u <= num_tiles; u++) tile_start[u] = e_total;
  }
}

// Per-counter offsets and tile totals, then tile_start = exclusive scan of the totals.
void tile_starts(splat_ctx* ctx, const BinBuffers& bb, size_t T, uint32_t* total_dev,
                 uint32_t* scan_sums) {
  uint32_t* totals = bb.tile_sub + kTileSub * T;
  launch(ctx, "bin_tile_totals", tile_totals_kernel, dim3(blocks_for(T, 256)), dim3(256), 0,
         (uint32_t)T, (const uint32_t*)bb.tile_count, bb.tile_sub, totals);
  exclusive_scan_u32(ctx, totals, bb.tile_start, T, total_dev, scan_sums);
}

// Dispatch order of the compositing kernels.  Workgroup b of a launch runs on die
// (accelerator complex) b % 8, each with its own L2, so the order gives die x a contiguous
// band of tiles -- neighbouring tiles share Gaussian records -- as the blocks x, x + 8,
// x + 16, ...: band x holds the ceil((T - x) / 8) tiles after the bands before it, and its
// tiles go out by decreasing entry count (kOrderBuckets buckets relative to the largest
// tile, order within a bucket unspecified) so the largest start first and the smallest
// fill the end of the launch.  Bands fix each die's share of the work; below kBandTiles
// tiles (a few per compute unit) that share is too coarse to balance, and the tiles go
// out heaviest first over the whole view instead (one band).  One workgroup; the result is
// a permutation of [0, T) whatever the counts.
constexpr uint32_t kOrderBuckets = 64;
constexpr uint32_t kOrderBands = 8;
constexpr uint32_t kBandTiles = 4096;
constexpr uint32_t kPlanThreads = 1024;
constexpr uint32_t kPlanItems = 16;
constexpr uint32_t kPlanTiles = kPlanThreads * kPlanItems;  // tiles whose counts fit in LDS

// LDS index of tile q in a count array padded by one word per 32 (a thread reading its 16
// consecutive tiles then spreads over the banks).
__device__ __forceinline__ uint32_t plan_slot(uint32_t q) { return q + (q >> 5); }

struct OrderShared {
  uint32_t smax;
  uint32_t hist[kOrderBands][kOrderBuckets];
  uint32_t cursor[kOrderBands][kOrderBuckets];
};

// The dispatch order from the tiles' entry counts (count(q), any source) and their
// maximum m_lane reduced here; every thread of the 1024-lane workgroup calls it.
template <class Count>
__device__ __forceinline__ void order_from_counts(uint32_t T, uint32_t m_lane, Count count,
                                                  OrderShared& o, uint32_t* __restrict__ order) {
  const uint32_t t = threadIdx.x;
  const uint32_t nb = T >= kBandTiles ? kOrderBands : 1u;
  if (t == 0) o.smax = 0;
  for (uint32_t q = t; q < kOrderBands * kOrderBuckets; q += blockDim.x)
    o.hist[q / kOrderBuckets][q % kOrderBuckets] = 0;
  __syncthreads();
  uint32_t m = m_lane;
  for (int d = 32; d >= 1; d >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, d, 64));
  if ((t & 63u) == 0) atomicMax(&o.smax, m);
  __syncthreads();
  const uint32_t mx = o.smax;
  const float scale = (float)kOrderBuckets / ((float)mx + 1.0f);
  const auto bucket = [&](uint32_t q) {  // 0: the largest tiles (any rounding is harmless)
    return min((uint32_t)((float)(mx - count(q)) * scale), kOrderBuckets - 1);
  };
  // bands 0 .. r-1 hold s + 1 tiles, the others s (s = T / nb, r = T % nb)
  const uint32_t bs = T / nb, br = T % nb, big = br * (bs + 1);
  const auto band = [&](uint32_t q) {
    return q < big ? q / (bs + 1) : br + (bs ? (q - big) / bs : 0u);
  };
  for (uint32_t q = t; q < T; q += blockDim.x) atomicAdd(&o.hist[band(q)][bucket(q)], 1u);
  __syncthreads();
  if (t < kOrderBands) {
    uint32_t run = 0;
    for (uint32_t b = 0; b < kOrderBuckets; b++) {
      o.cursor[t][b] = run;
      run += o.hist[t][b];
    }
  }
  __syncthreads();
  for (uint32_t q = t; q < T; q += blockDim.x) {
    const uint32_t x = band(q);
    const uint32_t p = atomicAdd(&o.cursor[x][bucket(q)], 1u);  // rank within the band
    order[p * nb + x] = q;
  }
}

// The order from tile_start (views binned by radix sorts).  Counts of up to kPlanTiles
// tiles are staged in LDS first, so the three passes over them read no global memory.
__global__ void __launch_bounds__(1024) tile_order_kernel(const uint32_t* __restrict__ tile_start,
                                                          uint32_t T, uint32_t* __restrict__ order) {
  __shared__ uint32_t cnt[kPlanTiles + kPlanTiles / 32];
  __shared__ OrderShared o;
  const bool staged = T <= kPlanTiles;
  uint32_t m = 0;
  if (staged) {
    for (uint32_t q = threadIdx.x; q < T; q += blockDim.x) {
      const uint32_t c = tile_start[q + 1] - tile_start[q];
      cnt[plan_slot(q)] = c;
      m = max(m, c);
    }
    __syncthreads();
    order_from_counts(T, m, [&](uint32_t q) { return cnt[plan_slot(q)]; }, o, order);
  } else {
    for (uint32_t q = threadIdx.x; q < T; q += blockDim.x)
      m = max(m, tile_start[q + 1] - tile_start[q]);
    order_from_counts(T, m, [&](uint32_t q) { return tile_start[q + 1] - tile_start[q]; }, o,
                      order);
  }
}

// Views of at most kPlanTiles tiles, after the counting pass: one workgroup writes each
// counter set's offset within its tile (tile_sub), tile_start (the exclusive scan of the
// tile totals, tile_start[T] and *total_out the entry count) and the dispatch order.
__global__ void __launch_bounds__(kPlanThreads) tile_plan_kernel(
    uint32_t T, const uint32_t* __restrict__ tile_count, uint32_t* __restrict__ tile_sub,
    uint32_t* __restrict__ tile_start, uint32_t* __restrict__ total_out,
    uint32_t* __restrict__ order) {
  __shared__ uint32_t cnt[kPlanTiles + kPlanTiles / 32];
  __shared__ uint32_t sh[kPlanThreads / 64 + 1];
  __shared__ OrderShared o;
  const uint32_t t = threadIdx.x;
  // 1. per tile (coalesced over tiles): the counter sets' offsets and the total, four
  // tiles per thread at a time so that their 32 loads are in flight together
  constexpr uint32_t kPlanBatch = 4;
  uint32_t m = 0;
  for (uint32_t q0 = t; q0 < T; q0 += kPlanBatch * kPlanThreads) {
    uint32_t c[kPlanBatch][kTileSub];
#pragma unroll
    for (uint32_t j = 0; j < kPlanBatch; j++) {
      const uint32_t q = q0 + j * kPlanThreads;
#pragma unroll
      for (uint32_t k = 0; k < kTileSub; k++) c[j][k] = q < T ? tile_count[k * T + q] : 0u;
    }
#pragma unroll
    for (uint32_t j = 0; j < kPlanBatch; j++) {
      const uint32_t q = q0 + j * kPlanThreads;
      if (q >= T) break;
      uint32_t acc = 0;
#pragma unroll
      for (uint32_t k = 0; k < kTileSub; k++) {
        tile_sub[k * T + q] = acc;
        acc += c[j][k];
      }
      cnt[plan_slot(q)] = acc;
      m = max(m, acc);
    }
  }
  __syncthreads();
  // 2. tile_start: thread t scans tiles [16 t, 16 t + 16)
  uint32_t vals[kPlanItems];
  uint32_t sum = 0;
  const uint32_t base = t * kPlanItems;
#pragma unroll
  for (uint32_t k = 0; k < kPlanItems; k++) {
    vals[k] = base + k < T ? cnt[plan_slot(base + k)] : 0u;
    sum += vals[k];
  }
  uint32_t total;
  const uint32_t incl = block_inclusive_scan<kPlanThreads>(sum, sh, &total);
  uint32_t run = incl - sum;
#pragma unroll
  for (uint32_t k = 0; k < kPlanItems; k++) {
    if (base + k < T) tile_start[base + k] = run;
    run += vals[k];
  }
  if (t == 0) {
    tile_start[T] = total;
    if (total_out) *total_out = total;
  }
  // 3. the dispatch order
  if (order) order_from_counts(T, m, [&](uint32_t q) { return cnt[plan_slot(q)]; }, o, order);
}

void tile_order(splat_ctx* ctx, const BinBuffers& bb, size_t T) {
  if (!bb.tile_order || T == 0) return;
  launch(ctx, "bin_tile_order", tile_order_kernel, dim3(1), dim3(1024), 0,
         (const uint32_t*)bb.tile_start, (uint32_t)T, bb.tile_order);
}

// After the counting pass: tile_sub, tile_start (total to *total_dev) and the dispatch order.
void tile_layout(splat_ctx* ctx, const BinBuffers& bb, size_t T, uint32_t* total_dev,
                 uint32_t* scan_sums) {
  if (T <= kPlanTiles) {
    launch(ctx, "bin_tile_plan", tile_plan_kernel, dim3(1), dim3(kPlanThreads), 0, (uint32_t)T,
           (const uint32_t*)bb.tile_count, bb.tile_sub, bb.tile_start, total_dev, bb.tile_order);
    return;
  }
  tile_starts(ctx, bb, T, total_dev, scan_sums);
  tile_order(ctx, bb, T);
}

// counts[0] (big tiles) and counts[2] (mid tiles) start at zero; `queue` holds T words.
void sort_tiles(splat_ctx* ctx, const BinBuffers& bb, size_t T, uint64_t* keys_out,
                uint32_t* queue, uint32_t* counts) {
  launch(ctx, "bin_sort_tiles", sort_tiles_kernel, dim3(blocks_for(T, 4)), dim3(256), 0,
         (uint32_t)T, (const uint32_t*)bb.tile_start, (const uint4*)bb.bucket,
         bb.sorted_entry, bb.sorted_gauss, keys_out, queue,
         counts, counts + 2);
  launch(ctx, "bin_sort_mid", sort_mid_kernel, dim3(512), dim3(256), 0, (uint32_t)T,
         (const uint32_t*)bb.tile_start, (const uint4*)bb.bucket,
         bb.sorted_entry, bb.sorted_gauss, keys_out,
         (const uint32_t*)queue, (const uint32_t*)(counts + 2));
  launch(ctx, "bin_sort_big", sort_big_kernel, dim3(256), dim3(1024), 0,
         (const uint32_t*)bb.tile_start, bb.bucket, (const uint32_t*)bb.entry_gaussian,
         bb.sorted_entry, bb.sorted_gauss, keys_out, (const uint32_t*)queue,
         (const uint32_t*)counts);
}

}  // namespace

size_t bin_build(splat_ctx* ctx, ProjBuffers pb, size_t n, uint32_t tiles_x, uint32_t tiles_y,
                 BinBuffers& bb, size_t* capacity_io, std::map<std::string, DevBuf>& store,
                 uint64_t* keys_out) {
  const size_t T = (size_t)tiles_x * tiles_y;
  bb.tiles_x = tiles_x;
  bb.tiles_y = tiles_y;
  bb.entry_offset = (uint32_t*)store["entry_offset"].reserve((n + 1) * 4);
  bb.tile_count = (uint32_t*)store["tile_count"].reserve(kTileSub * T * 4);
  bb.tile_sub = (uint32_t*)store["tile_sub"].reserve((kTileSub + 1) * T * 4);
  bb.tile_start = (uint32_t*)store["tile_start"].reserve((T + 1) * 4);
  bb.tile_order = (uint32_t*)store["tile_order"].reserve(T * 4 + 4);
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
    store["bucket"].reserve(cap * 16);
    store["sorted_entry"].reserve(cap * 4);
    store["sorted_gauss"].reserve(cap * 4);
  }
  bb.entry_gaussian = store["entry_gaussian"].as<uint32_t>();
  bb.entry_tile = store["entry_tile"].as<uint32_t>();
  bb.entry_slot = store["entry_slot"].as<uint32_t>();
  bb.bucket = store["bucket"].as<uint4>();
  bb.sorted_entry = store["sorted_entry"].as<uint32_t>();
  bb.sorted_gauss = store["sorted_gauss"].as<uint32_t>();
  if (ctx->bin_method == 1) {  // radix sorts (same result)
    bb.block_marks = (uint32_t*)store["block_marks"].reserve((E / kKeygenEntries + 2) * 4);
    const RadixScratch rs = radix_scratch(store, n, E);
    uint32_t* sums = (uint32_t*)store["radix_scan_sums"].reserve(
        scan_temp_bytes(radix_count_words(std::max<size_t>(n + 1, E))));
    bin_build_radix_device(ctx, pb, n, bb, E, rs, small, total64, sums, nullptr, nullptr,
                           keys_out, (uint32_t)E, true);
    return E;
  }

  bb.block_marks = (uint32_t*)store["block_marks"].reserve((E / kKeygenEntries + 2) * 4);
  // 2. keygen + per-tile counts
  fill_u32(ctx, bb.tile_count, kTileSub * T, 0);
  fill_u32(ctx, small + 1, 3, 0);  // the queued-tile counts small[1] and small[3]
  launch(ctx, "bin_marks", block_marks_kernel<>, dim3(blocks_for(n, 256)), dim3(256), 0,
         (const uint32_t*)bb.entry_offset, (uint32_t)n, bb.block_marks,
         (const StepStatus*)nullptr);
  launch(ctx, "bin_keygen", keygen_kernel, dim3(blocks_for(E, kKeygenEntries)),
         dim3(kKeygenThreads), 0, (const uint32_t*)bb.entry_offset, (uint32_t)n, (uint32_t)E,
         (const uint2*)pb.rect, tiles_x, bb.entry_gaussian, bb.entry_tile, bb.entry_slot,
         bb.tile_count, (uint32_t)T, (const uint32_t*)bb.block_marks,
         (const StepStatus*)nullptr);
  // 3. tile starts and the dispatch order
  tile_layout(ctx, bb, T, small + 2, nullptr);
  // 4. scatter into buckets
  launch(ctx, "bin_scatter", scatter_kernel, dim3(blocks_for(E, 256)), dim3(256), 0,
         (uint32_t)E, (const uint32_t*)bb.entry_gaussian, (const uint32_t*)bb.entry_tile,
         (const uint32_t*)bb.entry_slot, (const uint32_t*)bb.tile_start,
         (const uint32_t*)bb.tile_sub, (uint32_t)T, (const float*)pb.depth, bb.bucket,
         (const StepStatus*)nullptr);
  // 5. per-tile sorts
  sort_tiles(ctx, bb, T, keys_out, big_list, small + 1);
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
  // the capacity guard and the keygen marks ride on the scan's last pass when it has one
  const ScanTail tail{bb.block_marks, (uint32_t)(capacity / kKeygenEntries + 2), kKeygenEntries,
                      (unsigned long long)capacity, total64, status};
  if (!exclusive_scan_u32(ctx, pb.ntiles, bb.entry_offset, n, small, scan_sums, total64,
                          bd ? nullptr : &tail)) {
    bin_guard(ctx, total64, capacity, status);
    // The entry kernels cover the capacity and read the step's count from `status`.
    launch(ctx, "bin_marks", block_marks_kernel<>, dim3(blocks_for(n, 256)), dim3(256), 0,
           (const uint32_t*)bb.entry_offset, (uint32_t)n, bb.block_marks,
           (const StepStatus*)status);
  }
  if (bd)
    launch(ctx, "bin_keygen", keygen_batch_kernel, dim3(blocks_for(capacity, kKeygenEntries)),
           dim3(kKeygenThreads), 0, (const uint32_t*)bb.entry_offset, (uint32_t)n, *bd,
           (const uint2*)pb.rect, bb.entry_gaussian, bb.entry_tile, bb.entry_slot,
           bb.tile_count, (uint32_t)T, (const uint32_t*)bb.block_marks,
           (const StepStatus*)status);
  else
    launch(ctx, "bin_keygen", keygen_kernel, dim3(blocks_for(capacity, kKeygenEntries)),
           dim3(kKeygenThreads), 0, (const uint32_t*)bb.entry_offset, (uint32_t)n, 0u,
           (const uint2*)pb.rect, bb.tiles_x, bb.entry_gaussian, bb.entry_tile, bb.entry_slot,
           bb.tile_count, (uint32_t)T, (const uint32_t*)bb.block_marks,
           (const StepStatus*)status);
  tile_layout(ctx, bb, T, small + 2, scan_sums);
  launch(ctx, "bin_scatter", scatter_kernel, dim3(blocks_for(capacity, 256)), dim3(256), 0, 0u,
         (const uint32_t*)bb.entry_gaussian, (const uint32_t*)bb.entry_tile,
         (const uint32_t*)bb.entry_slot, (const uint32_t*)bb.tile_start,
         (const uint32_t*)bb.tile_sub, (uint32_t)T, (const float*)pb.depth, bb.bucket,
         (const StepStatus*)status);
  sort_tiles(ctx, bb, T, nullptr, big_list, small + 1);
}


RadixScratch radix_scratch(std::map<std::string, DevBuf>& store, size_t n, size_t capacity) {
  RadixScratch rs;
  for (int k = 0; k < 2; k++) {
    const std::string sk = std::to_string(k);
    rs.rk[k] = (uint32_t*)store["radix_rk" + sk].reserve((n + 1) * 4);
    rs.rv[k] = (uint32_t*)store["radix_rv" + sk].reserve((n + 1) * 4);
    rs.ek[k] = (uint32_t*)store["radix_ek" + sk].reserve((capacity + 1) * 4);
    rs.ev[k] = (uint32_t*)store["radix_ev" + sk].reserve((capacity + 1) * 4);
    rs.eg[k] = (uint32_t*)store["radix_eg" + sk].reserve((capacity + 1) * 4);
  }
  rs.soff = (uint32_t*)store["radix_soff"].reserve((n + 1) * 4);
  rs.rows = (uint4*)store["radix_rows"].reserve((n + 1) * 16);
  rs.recs = (uint4*)store["radix_recs"].reserve((n + 1) * 16);
  rs.counts = (uint32_t*)store["radix_counts"].reserve(radix_count_words(std::max(n, capacity)) * 4);
  return rs;
}

void bin_build_radix_device(splat_ctx* ctx, ProjBuffers pb, size_t n, const BinBuffers& bb,
                            size_t capacity, const RadixScratch& rs, uint32_t* small,
                            unsigned long long* total64, uint32_t* scan_sums, StepStatus* status,
                            const BatchDev* bd, uint64_t* keys_out, uint32_t e_host,
                            bool offsets_done) {
  const size_t T = (size_t)bb.tiles_x * bb.tiles_y;
  SPLAT_REQUIRE(bd || T < kMaxViewTiles, "image too large for radix binning");
  if (!offsets_done) {
    fill_u32(ctx, reinterpret_cast<uint32_t*>(total64), 2, 0u);
    exclusive_scan_u32(ctx, pb.ntiles, bb.entry_offset, n, small, scan_sums, total64);
    if (status) bin_guard(ctx, total64, capacity, status);
  }
  const uint32_t* n_entries = status ? &status->entries : nullptr;
  // 1. rows by depth (and the rows' records, in row order)
  if (bd)
    launch(ctx, "bin_depth_keys", depth_keys_kernel<true>, dim3(blocks_for(n, 256)), dim3(256),
           0, (const float*)pb.depth, (const uint32_t*)pb.ntiles, (const uint2*)pb.rect,
           (const uint32_t*)bb.entry_offset, (uint32_t)n, bb.tiles_x, *bd, rs.rk[0], rs.rv[0],
           rs.recs, (const StepStatus*)status);
  else
    launch(ctx, "bin_depth_keys", depth_keys_kernel<false>, dim3(blocks_for(n, 256)), dim3(256),
           0, (const float*)pb.depth, (const uint32_t*)pb.ntiles, (const uint2*)pb.rect,
           (const uint32_t*)bb.entry_offset, (uint32_t)n, bb.tiles_x, BatchDev{}, rs.rk[0],
           rs.rv[0], rs.recs, (const StepStatus*)status);
  const int c = radix_sort_pairs(ctx, rs.rk[0], rs.rv[0], rs.rk[1], rs.rv[1], n, (uint32_t)n,
                                 nullptr, 32, rs.counts, scan_sums, status);
  const uint32_t* order = rs.rv[c];
  // 2. the rows' entry offsets and records in depth order
  launch(ctx, "bin_ordered_rows", ordered_rows_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
         order, (const uint32_t*)pb.ntiles, (const uint4*)rs.recs, (uint32_t)n, rs.rk[1 - c],
         rs.rows, (const StepStatus*)status);
  exclusive_scan_u32(ctx, rs.rk[1 - c], rs.soff, n, small + 4, scan_sums);
  launch(ctx, "bin_marks", block_marks_kernel<kOrderedEntries>, dim3(blocks_for(n, 256)),
         dim3(256), 0, (const uint32_t*)rs.soff, (uint32_t)n, bb.block_marks,
         (const StepStatus*)status);
  // 3. entries in depth order, keyed by tile
  launch(ctx, "bin_keygen", keygen_ordered_kernel,
         dim3(blocks_for(status ? capacity : e_host, kOrderedEntries)), dim3(kKeygenThreads), 0,
         (const uint32_t*)rs.soff, (const uint4*)rs.rows, (uint32_t)n, e_host,
         (const uint32_t*)bb.block_marks, rs.ek[0], rs.ev[0], rs.eg[0], bb.entry_gaussian,
         (const StepStatus*)status);
  // 4. stable sort by tile
  uint32_t tile_bits = 0;
  while (tile_bits < 32 && (T >> tile_bits) > 1) tile_bits++;  // T - 1 < 2^tile_bits
  if (((size_t)1 << tile_bits) < T) tile_bits++;
  // (the last pass writes the lists in place; a single tile needs no pass)
  const bool direct = tile_bits > 0;
  const int c2 = radix_sort_pairs(ctx, rs.ek[0], rs.ev[0], rs.ek[1], rs.ev[1],
                                  status ? capacity : e_host, e_host, n_entries, tile_bits,
                                  rs.counts, scan_sums, status, rs.eg[0], rs.eg[1],
                                  direct ? bb.sorted_entry : nullptr,
                                  direct ? bb.sorted_gauss : nullptr);
  // 5. lists and tile ranges
  fill_u32(ctx, bb.tile_start, T + 1, 0u);
  launch(ctx, "bin_finalize", radix_finalize_kernel,
         dim3(blocks_for(blocks_for(status ? capacity : e_host, kFinalizePerLane), 256)),
         dim3(256), 0,
         (const uint32_t*)rs.ek[c2], direct ? nullptr : (const uint32_t*)rs.ev[c2],
         direct ? nullptr : (const uint32_t*)rs.eg[c2],
         e_host, (uint32_t)T, (const float*)pb.depth, bb.sorted_entry,
         bb.sorted_gauss, bb.tile_start, keys_out, (const StepStatus*)status);
  tile_order(ctx, bb, T);
}

}  // namespace splat
