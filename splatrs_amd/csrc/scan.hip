// Device-wide exclusive prefix sum of uint32 (three passes: block sums, scan of the block
// sums in one workgroup, block-local rescan plus offset).  out[n] receives the total.
#include "internal.hpp"

namespace splat {
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096 items per block
constexpr int kTopThreads = 1024;
constexpr int kTopItems = 16;  // one workgroup scans up to 16384 block sums

// Sum of one 64-bit value per thread over the workgroup, added to *total (one atomic per
// workgroup; *total starts at zero).
template <int THREADS>
__device__ __forceinline__ void block_add_u64(unsigned long long v, unsigned long long* sh,
                                              unsigned long long* total) {
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < THREADS / 64; w++) t += sh[w];
    atomicAdd(total, t);
  }
}

// total64 (optional): the exact 64-bit sum of the input, for callers that must detect a
// total beyond the 32-bit offsets.
__global__ void __launch_bounds__(kScanThreads) scan_reduce_kernel(const uint32_t* __restrict__ in,
                                                                   size_t n,
                                                                   uint32_t* __restrict__ block_sums,
                                                                   unsigned long long* total64) {
  __shared__ uint32_t sh[kScanThreads / 64 + 1];
  __shared__ unsigned long long sh64[kScanThreads / 64];
  const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
  uint32_t s = 0;
  unsigned long long s64 = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++)
    if (base + k < n) {
      s += in[base + k];
      s64 += in[base + k];
    }
  uint32_t total;
  block_inclusive_scan<kScanThreads>(s, sh, &total);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
  if (total64) block_add_u64<kScanThreads>(s64, sh64, total64);
}

// The whole scan in one workgroup (n <= kTopThreads * kTopItems): out[n] and *total_out
// receive the total.
__global__ void __launch_bounds__(kTopThreads) scan_single_kernel(const uint32_t* __restrict__ in,
                                                                  uint32_t n,
                                                                  uint32_t* __restrict__ out,
                                                                  uint32_t* __restrict__ total_out,
                                                                  unsigned long long* total64) {
  __shared__ uint32_t sh[kTopThreads / 64 + 1];
  __shared__ unsigned long long sh64[kTopThreads / 64];
  uint32_t vals[kTopItems];
  uint32_t s = 0;
  unsigned long long s64 = 0;
  const uint32_t base = threadIdx.x * kTopItems;
#pragma unroll
  for (int k = 0; k < kTopItems; k++) {
    vals[k] = (base + k < n) ? in[base + k] : 0u;
    s += vals[k];
    s64 += vals[k];
  }
  uint32_t total;
  const uint32_t incl = block_inclusive_scan<kTopThreads>(s, sh, &total);
  uint32_t run = incl - s;
#pragma unroll
  for (int k = 0; k < kTopItems; k++) {
    if (base + k < n) out[base + k] = run;
    run += vals[k];
  }
  if (threadIdx.x == 0) {
    out[n] = total;
    if (total_out) *total_out = total;
  }
  if (total64) block_add_u64<kTopThreads>(s64, sh64, total64);
}

__global__ void __launch_bounds__(kTopThreads) scan_top_kernel(uint32_t* __restrict__ sums, uint32_t nb,
                                                               uint32_t* __restrict__ total_out) {
  __shared__ uint32_t sh[kTopThreads / 64 + 1];
  uint32_t vals[kTopItems];
  uint32_t s = 0;
  const uint32_t base = threadIdx.x * kTopItems;
#pragma unroll
  for (int k = 0; k < kTopItems; k++) {
    vals[k] = (base + k < nb) ? sums[base + k] : 0u;
    s += vals[k];
  }
  uint32_t total;
  const uint32_t incl = block_inclusive_scan<kTopThreads>(s, sh, &total);
  uint32_t run = incl - s;
#pragma unroll
  for (int k = 0; k < kTopItems; k++) {
    if (base + k < nb) sums[base + k] = run;
    run += vals[k];
  }
  if (threadIdx.x == 0 && total_out) *total_out = total;
}

// TAIL (the entry offsets of a binning, `tail` set): also the capacity guard of a
// graph-captured step (block 0: the 64-bit total against the entry capacity, the sticky
// abort flag) and the keygen marks: marks[b] = the item (Gaussian) whose range holds entry
// b * epb, for b < marks_cap (the marks of an aborted step are never read).
template <bool TAIL>
__global__ void __launch_bounds__(kScanThreads) scan_apply_kernel(const uint32_t* __restrict__ in,
                                                                  size_t n,
                                                                  const uint32_t* __restrict__ block_offsets,
                                                                  uint32_t* __restrict__ out,
                                                                  const uint32_t* __restrict__ total,
                                                                  ScanTail tail) {
  __shared__ uint32_t sh[kScanThreads / 64 + 1];
  const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
  uint32_t vals[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    vals[k] = (base + k < n) ? in[base + k] : 0u;
    s += vals[k];
  }
  uint32_t tot;
  const uint32_t incl = block_inclusive_scan<kScanThreads>(s, sh, &tot);
  uint32_t run = block_offsets[blockIdx.x] + incl - s;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    if (base + k < n) {
      out[base + k] = run;
      if (TAIL)
        for (uint32_t b = (uint32_t)(((uint64_t)run + tail.epb - 1) / tail.epb);
             (uint64_t)b * tail.epb < (uint64_t)run + vals[k] && b < tail.marks_cap; b++)
          tail.marks[b] = (uint32_t)(base + k);
    }
    run += vals[k];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    out[n] = *total;
    if (TAIL) {
      const unsigned long long t = *tail.total64;
      tail.status->total = t;
      tail.status->entries = t > 0xffffffffull ? 0xffffffffu : (uint32_t)t;
      if (t > tail.capacity) tail.status->abort = 1u;
    }
  }
}

__global__ void fill_u32_kernel(uint32_t* p, size_t n, uint32_t v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

void fill_u32(splat_ctx* ctx, uint32_t* p, size_t n, uint32_t v) {
  launch(ctx, "fill", fill_u32_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, p, n, v);
}

size_t scan_max_items() { return (size_t)kTopThreads * kTopItems * kScanTile; }

size_t scan_temp_bytes(size_t n) {
  return (blocks_for(n == 0 ? 1 : n, kScanTile) + 1) * sizeof(uint32_t);
}

bool exclusive_scan_u32(splat_ctx* ctx, const uint32_t* in, uint32_t* out, size_t n,
                        uint32_t* total_dev, uint32_t* sums, unsigned long long* total64,
                        const ScanTail* tail) {
  if (n <= (size_t)kTopThreads * kTopItems) {
    launch(ctx, "scan_single", scan_single_kernel, dim3(1), dim3(kTopThreads), 0, in,
           (uint32_t)n, out, total_dev, total64);
    return false;
  }
  const unsigned nb = blocks_for(n, kScanTile);
  SPLAT_REQUIRE(nb <= (unsigned)(kTopThreads * kTopItems), "scan: input too large");
  if (!sums) sums = ctx->buf<uint32_t>("scan_block_sums", nb + 1);
  launch(ctx, "scan_reduce", scan_reduce_kernel, dim3(nb), dim3(kScanThreads), 0, in, n, sums,
         total64);
  launch(ctx, "scan_top", scan_top_kernel, dim3(1), dim3(kTopThreads), 0, sums, nb, total_dev);
  if (tail) {
    launch(ctx, "scan_apply", scan_apply_kernel<true>, dim3(nb), dim3(kScanThreads), 0, in, n,
           (const uint32_t*)sums, out, (const uint32_t*)total_dev, *tail);
    return true;
  }
  launch(ctx, "scan_apply", scan_apply_kernel<false>, dim3(nb), dim3(kScanThreads), 0, in, n,
         (const uint32_t*)sums, out, (const uint32_t*)total_dev, ScanTail{});
  return false;
}

}  // namespace splat
