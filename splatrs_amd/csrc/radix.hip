// Stable LSD radix sort of (key, value[, value2]) tuples of 32-bit words, up to 9 key bits
// per pass (the fewest passes for the key width, digits as even as possible).
//
// Each pass: an upsweep writes every block's digit histogram, digit-major
// (counts[digit * nblocks + block]); an exclusive scan of that array gives every
// (digit, block) its first output position; a downsweep loads the block's 4096 items
// (16 per lane, prefetched), ranks them in index order (16 rounds of 256): each item's
// position in a block-local digit order stays in a register; then keys and values pass, one
// array at a time, through one LDS buffer in that order and are written out run by run, so
// consecutive lanes store consecutive addresses (one 16 KB buffer instead of one per
// array: five blocks per CU instead of two).  Within a round, a lane's rank among the wave's lanes with
// the same digit comes from ballots over the digit bits (multi-split), and the waves'
// counts are combined in LDS in wave order, so equal digits keep their input order: the
// pass is stable and the sort is exact.
//
// Item counts may live on the device (graph-captured steps): kernels then read the count
// through a pointer and cover a grid sized for the capacity.
#include <utility>

#include "internal.hpp"

namespace splat {
namespace {

constexpr uint32_t kRadixThreads = 256;
// Items per block: 32 rounds of 256 for large sorts (a 9-bit digit's output run in a block
// then averages 16 items, a 64-byte line, and the digit histograms are half as many), 16
// rounds below kLargeSort items (more blocks to fill the device).
constexpr uint32_t kRadixRounds = 16;
constexpr uint32_t kRadixRoundsLarge = 32;
constexpr uint32_t kRadixItems = kRadixThreads * kRadixRounds;  // the smaller block
constexpr size_t kLargeSort = (size_t)1 << 26;
constexpr uint32_t kRadixBins = 512;  // digits of the widest pass (9 bits)
constexpr uint32_t kRadixMaxBits = 9;
constexpr uint32_t kRadixWaves = kRadixThreads / 64;

__device__ __forceinline__ uint32_t item_count(uint32_t n, const uint32_t* n_dev) {
  return n_dev ? *n_dev : n;
}


template <class F, int... I>
__device__ __forceinline__ void rounds_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
// f(integral_constant<0>) ... f(integral_constant<ROUNDS - 1>): register arrays indexed by
// the round stay in registers.
template <uint32_t ROUNDS, class F>
__device__ __forceinline__ void for_rounds(F&& f) {
  rounds_impl(f, std::make_integer_sequence<int, ROUNDS>{});
}

template <uint32_t ROUNDS>
__global__ void __launch_bounds__(kRadixThreads) radix_upsweep_kernel(
    const uint32_t* __restrict__ keys, uint32_t n, const uint32_t* __restrict__ n_dev,
    uint32_t shift, uint32_t mask, uint32_t nblocks, uint32_t* __restrict__ counts,
    const StepStatus* __restrict__ status) {
  __shared__ uint32_t hist[kRadixBins];
  if (status && status->abort) return;  // uniform
  n = item_count(n, n_dev);
  for (uint32_t d = threadIdx.x; d < kRadixBins; d += kRadixThreads) hist[d] = 0;
  __syncthreads();
  // neighbouring blocks' digit counts are adjacent words: their lines fill in one L2
  const uint32_t blk = die_block(blockIdx.x, nblocks);
  const uint32_t base = blk * (kRadixThreads * ROUNDS);
  for (uint32_t r = 0; r < ROUNDS; r++) {
    const uint32_t i = base + r * kRadixThreads + threadIdx.x;
    if (i < n) atomicAdd(&hist[(keys[i] >> shift) & mask], 1u);
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d <= mask; d += kRadixThreads)
    counts[d * nblocks + blk] = hist[d];
}

template <bool TWO, uint32_t ROUNDS>
__global__ void __launch_bounds__(kRadixThreads) radix_downsweep_kernel(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    const uint32_t* __restrict__ vals2_in, uint32_t n, const uint32_t* __restrict__ n_dev,
    uint32_t shift, uint32_t nbits, uint32_t nblocks, const uint32_t* __restrict__ offsets,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
    uint32_t* __restrict__ vals2_out, const StepStatus* __restrict__ status) {
  constexpr uint32_t kItems = kRadixThreads * ROUNDS;
  __shared__ uint32_t s_buf[kItems];  // one array at a time, in block-local order
  __shared__ uint32_t gbase[kRadixBins];  // first output position of each digit
  __shared__ uint32_t lbase[kRadixBins];  // first block-local position of each digit
  __shared__ uint32_t run[kRadixBins];    // next block-local position of each digit
  __shared__ uint32_t wcnt[kRadixWaves][kRadixBins];  // this round's counts per wave
  __shared__ uint32_t wsum[kRadixWaves];
  if (status && status->abort) return;  // uniform
  n = item_count(n, n_dev);
  // neighbouring blocks' output runs of each digit are adjacent: their lines fill in one L2
  const uint32_t blk = die_block(blockIdx.x, nblocks);
  const uint32_t base = blk * kItems;
  if (base >= n) return;  // uniform: the block holds no item
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const uint32_t mask = (1u << nbits) - 1u;
  const uint32_t items = min(kItems, n - base);
  // items of the block, prefetched
  uint32_t k[ROUNDS], v[ROUNDS], v2[ROUNDS], lp[ROUNDS];
  for_rounds<ROUNDS>([&](auto step) {
    constexpr int r = decltype(step)::value;
    const uint32_t i = base + r * kRadixThreads + t;
    const bool valid = i < n;
    k[r] = valid ? keys_in[i] : 0u;
    v[r] = valid ? vals_in[i] : 0u;
    if (TWO) v2[r] = valid ? vals2_in[i] : 0u;
  });
  // this block's digit counts (consecutive words of the scanned histogram) and their
  // block-local exclusive scan; thread t holds digits 2t and 2t + 1
  const uint32_t bins = mask + 1;
  uint32_t g0[2], c[2];
#pragma unroll
  for (uint32_t h = 0; h < 2; h++) {
    const uint32_t d = 2 * t + h;
    g0[h] = d < bins ? offsets[d * nblocks + blk] : 0u;
    c[h] = d < bins ? offsets[d * nblocks + blk + 1] - g0[h] : 0u;
  }
  const uint32_t pair = c[0] + c[1];
  uint32_t x = pair;
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  for (uint32_t d = t; d < kRadixBins; d += kRadixThreads) {
#pragma unroll
    for (uint32_t w = 0; w < kRadixWaves; w++) wcnt[w][d] = 0;
  }
  __syncthreads();
  uint32_t lb = x - pair;
  for (uint32_t w = 0; w < wave; w++) lb += wsum[w];
#pragma unroll
  for (uint32_t h = 0; h < 2; h++) {
    const uint32_t d = 2 * t + h;
    if (d < bins) {
      gbase[d] = g0[h];
      lbase[d] = lb;
      run[d] = lb;
    }
    lb += c[h];
  }
  __syncthreads();
  for_rounds<ROUNDS>([&](auto step) {
    constexpr int r = decltype(step)::value;
    lp[r] = kItems;  // (no item)
    if (base + r * kRadixThreads >= n) return;  // uniform
    const bool valid = base + r * kRadixThreads + t < n;
    const uint32_t d = (k[r] >> shift) & mask;
    // lanes of this wave holding the same digit
    uint64_t same = __ballot(valid);
    for (uint32_t b = 0; b < nbits; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t ones = __ballot(valid && bit);
      same &= bit ? ones : ~ones;
    }
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(same >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)same, 0u));
    if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__builtin_popcountll(same);
    __syncthreads();
    if (valid) {
      uint32_t pos = run[d] + rank;
      for (uint32_t w = 0; w < wave; w++) pos += wcnt[w][d];
      lp[r] = pos;
    }
    __syncthreads();
    for (uint32_t d = t; d < bins; d += kRadixThreads) {
      uint32_t add = 0;
#pragma unroll
      for (uint32_t w = 0; w < kRadixWaves; w++) {
        add += wcnt[w][d];
        wcnt[w][d] = 0;
      }
      run[d] += add;
    }
    __syncthreads();
  });
  // block-local order -> output, run by run, one array at a time: first the keys (kept in
  // registers by the lanes that write them, for the digit of the run), then the values
  uint32_t okey[ROUNDS];
  for_rounds<ROUNDS>([&](auto step) {
    constexpr int r = decltype(step)::value;
    if (lp[r] < kItems) s_buf[lp[r]] = k[r];
  });
  __syncthreads();
  for_rounds<ROUNDS>([&](auto step) {
    constexpr int q = decltype(step)::value;
    const uint32_t at = q * kRadixThreads + t;
    okey[q] = at < items ? s_buf[at] : 0u;
    if (at < items) {
      const uint32_t d = (okey[q] >> shift) & mask;
      keys_out[gbase[d] + (at - lbase[d])] = okey[q];
    }
  });
  const auto pass = [&](const uint32_t(&src)[ROUNDS], uint32_t* __restrict__ out) {
    __syncthreads();
    for_rounds<ROUNDS>([&](auto step) {
      constexpr int r = decltype(step)::value;
      if (lp[r] < kItems) s_buf[lp[r]] = src[r];
    });
    __syncthreads();
    for_rounds<ROUNDS>([&](auto step) {
      constexpr int q = decltype(step)::value;
      const uint32_t at = q * kRadixThreads + t;
      if (at < items) {
        const uint32_t d = (okey[q] >> shift) & mask;
        out[gbase[d] + (at - lbase[d])] = s_buf[at];
      }
    });
  };
  pass(v, vals_out);
  if (TWO) pass(v2, vals2_out);
}

}  // namespace

size_t radix_count_words(size_t capacity) {
  return (size_t)kRadixBins * blocks_for(capacity == 0 ? 1 : capacity, kRadixItems) + 1;
}

int radix_sort_pairs(splat_ctx* ctx, uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1,
                     size_t capacity, uint32_t n_host, const uint32_t* n_dev, uint32_t key_bits,
                     uint32_t* counts, uint32_t* scan_sums, const StepStatus* status,
                     uint32_t* w0, uint32_t* w1, uint32_t* v_final, uint32_t* w_final) {
  const bool large = capacity >= kLargeSort;
  const uint32_t items = kRadixThreads * (large ? kRadixRoundsLarge : kRadixRounds);
  const uint32_t nblocks = blocks_for(capacity == 0 ? 1 : capacity, items);
  int cur = 0;
  uint32_t* keys[2] = {k0, k1};
  uint32_t* vals[2] = {v0, v1};
  uint32_t* vals2[2] = {w0, w1};
  const bool two = w0 != nullptr;
  // 9-bit digits while their (digit, block) counts fit one scan, else 8
  const uint32_t max_bits =
      ((size_t)kRadixBins * nblocks + 1 <= scan_max_items()) ? kRadixMaxBits : 8u;
  const uint32_t passes = (key_bits + max_bits - 1) / max_bits;
  const uint32_t width = passes ? (key_bits + passes - 1) / passes : 0u;
  for (uint32_t shift = 0; shift < key_bits; shift += width) {
    const uint32_t nbits = std::min(width, key_bits - shift);
    const bool last = shift + nbits >= key_bits;
    uint32_t* vo = last && v_final ? v_final : vals[1 - cur];
    uint32_t* wo = last && w_final ? w_final : vals2[1 - cur];
    launch(ctx, "radix_upsweep",
           large ? radix_upsweep_kernel<kRadixRoundsLarge> : radix_upsweep_kernel<kRadixRounds>,
           dim3(nblocks), dim3(kRadixThreads), 0, (const uint32_t*)keys[cur], n_host, n_dev,
           shift, (1u << nbits) - 1u, nblocks, counts, status);
    // in place; the total lands in the array's last word
    const size_t len = ((size_t)1 << nbits) * nblocks;
    exclusive_scan_u32(ctx, counts, counts, len, counts + len, scan_sums);
    const auto down = two ? (large ? radix_downsweep_kernel<true, kRadixRoundsLarge>
                                   : radix_downsweep_kernel<true, kRadixRounds>)
                          : (large ? radix_downsweep_kernel<false, kRadixRoundsLarge>
                                   : radix_downsweep_kernel<false, kRadixRounds>);
    launch(ctx, "radix_downsweep", down, dim3(nblocks), dim3(kRadixThreads), 0,
           (const uint32_t*)keys[cur], (const uint32_t*)vals[cur],
           (const uint32_t*)(two ? vals2[cur] : nullptr), n_host, n_dev, shift, nbits, nblocks,
           (const uint32_t*)counts, keys[1 - cur], vo, two ? wo : (uint32_t*)nullptr, status);
    cur = 1 - cur;
  }
  return cur;
}

}  // namespace splat
