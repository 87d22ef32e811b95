// Photometric loss (1 - lambda) L1 + lambda (1 - SSIM) and its image gradient
// (src/loss.rs:14-190), in two fused passes.
//
// Pass 1 blurs x, y, x^2, y^2, xy (11-tap separable zero-padded, sigma 1.5) and writes
// the SSIM partials d_mu, d_exx, d_exy plus per-block SSIM and L1 sums.  Pass 2 blurs the
// three partials (the blur is its own adjoint) and combines them with the L1 sign term.
//
// Each workgroup owns a strip of 256 interleaved row elements (x * 3 + c) and a chunk of
// rows.  Rows are staged in LDS; each lane keeps the last 11 horizontally blurred rows of
// every map in registers and emits one vertically blurred row per step.  Taps are summed
// in the reference's order (horizontal: k = 0..10; vertical: out += kv * line, k =
// 0..10); out-of-image taps add +0.0, which is exact, so every intermediate equals the
// reference's value bit for bit.
#include <utility>

#include "internal.hpp"

namespace splat {
namespace {

constexpr int kStrip = 256;
constexpr int kHalo = 15;        // 5 pixels x 3 channels
constexpr int kRowsPerBlock = 32;

struct Taps {
  float k[11];
};

// View::to_real of an RGB8 target byte (src/dataset.rs:23-25): v / 255.0, the correctly
// rounded binary32 quotient.  A byte has 256 values, so each 256-lane workgroup divides
// once per value into an LDS table (every thread calls fill_u8_table, which synchronises)
// and the loads look the quotients up: the same bits for a fraction of the instructions.
template <bool U8>
__device__ __forceinline__ void fill_u8_table(float* lut) {
  if (U8) {
    lut[threadIdx.x] = (float)threadIdx.x / 255.0f;
    __syncthreads();
  }
}

template <bool U8>
__device__ __forceinline__ float load_target(const void* t, size_t idx, const float* lut) {
  if (U8) return lut[static_cast<const uint8_t*>(t)[idx]];
  return static_cast<const float*>(t)[idx];
}

// Horizontal 11-tap blur at flat row position q (channel stride 3) from an LDS row
// segment whose element 0 is row position q0 - kHalo.  The segment holds +0.0 outside the
// image, which is what the reference's zero padding adds.
__device__ __forceinline__ float hblur(const float* seg, int q, int q0, const Taps& k) {
  float acc = 0.0f;
#pragma unroll
  for (int t = 0; t < 11; t++) acc += k.k[t] * seg[q + (t - 5) * 3 - q0 + kHalo];
  return acc;
}

// f(integral_constant<int, 0>) ... f(integral_constant<int, N - 1>), in order.
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// The vertical window is a ring of 11 horizontally blurred rows per map: the row loop is
// unrolled 11 times (static_for) so that row step p writes slot p and the taps read the slots oldest
// first, (p + 1 + t) % 11 for t = 0..10, as compile-time register names (no shifting).
template <int M>
__device__ __forceinline__ void vblur_ring(const float (&h)[M][11], int p, const Taps& k,
                                           float (&out)[M]) {
#pragma unroll
  for (int m = 0; m < M; m++) {
    float acc = 0.0f;
#pragma unroll
    for (int t = 0; t < 11; t++) acc += k.k[t] * h[m][(p + 1 + t) % 11];
    out[m] = acc;
  }
}

__device__ __forceinline__ float block_sum(float v, float* sh) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  return ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// Pass 1.  partial[block] = {sum SSIM, sum |x - y|} over the block's outputs.
// Block coordinates of a loss kernel: (strip, row chunk) of one image.  BATCH: the 1D grid
// covers every pair's blocks (BatchDev::loss_block_base); the image pointers are advanced
// to the pair's pixels, whose maps lie at the same pixel offsets.
struct LossBlock {
  int bx, by;
  uint32_t flat;  // partial-sum index
};

template <bool BATCH>
__device__ __forceinline__ LossBlock loss_block(const BatchDev& bd, uint32_t* W, uint32_t* H,
                                                size_t* pix0, uint32_t* pair) {
  LossBlock b;
  if (BATCH) {
    const uint32_t v = find_segment(bd.loss_block_base, bd.npairs, blockIdx.x);
    const uint32_t local = blockIdx.x - bd.loss_block_base[v];
    *W = bd.width[v];
    *H = bd.height[v];
    const uint32_t gx = (*W * 3 + kStrip - 1) / kStrip;
    b.bx = (int)(local % gx);
    b.by = (int)(local / gx);
    b.flat = blockIdx.x;
    *pix0 = bd.pix_base[v];
    *pair = v;
  } else {
    b.bx = (int)blockIdx.x;
    b.by = (int)blockIdx.y;
    b.flat = blockIdx.y * gridDim.x + blockIdx.x;
    *pix0 = 0;
    *pair = 0;
  }
  return b;
}

template <bool U8, bool BATCH>
__global__ void __launch_bounds__(256) ssim_stats_kernel(const float* __restrict__ xin,
                                                         const void* __restrict__ yin,
                                                         uint32_t W, uint32_t H, Taps taps,
                                                         float c1, float c2,
                                                         float* __restrict__ d_mu,
                                                         float* __restrict__ d_exx,
                                                         float* __restrict__ d_exy,
                                                         float2* __restrict__ partial,
                                                         const StepDyn* __restrict__ dyn,
                                                         BatchDev bd) {
  if (dyn) yin = dyn->target;
  size_t pix0;
  uint32_t pair;
  const LossBlock lb = loss_block<BATCH>(bd, &W, &H, &pix0, &pair);
  if (BATCH) {
    xin += pix0 * 3;
    yin = reinterpret_cast<const void*>(bd.targets[pair]);
    d_mu += pix0 * 3;
    d_exx += pix0 * 3;
    d_exy += pix0 * 3;
  }
  __shared__ float sx[kStrip + 2 * kHalo], sy[kStrip + 2 * kHalo];
  __shared__ float red[4];
  __shared__ float s_u8[U8 ? 256 : 1];
  fill_u8_table<U8>(s_u8);
  const int width3 = (int)W * 3;
  const int q0 = lb.bx * kStrip;
  const int q = q0 + threadIdx.x;
  const bool col_ok = q < width3;
  const int ybeg = lb.by * kRowsPerBlock;
  const int yend = min(ybeg + kRowsPerBlock, (int)H);
  float h[5][11];
#pragma unroll
  for (int m = 0; m < 5; m++)
#pragma unroll
    for (int t = 0; t < 11; t++) h[m][t] = 0.0f;
  float ssum = 0.0f, l1sum = 0.0f;
  // Software pipeline: the global loads of row rr + 1 (segment elements threadIdx.x and
  // threadIdx.x + 256) and of output row rr - 4's x, y are issued before row rr is
  // blurred, and land in LDS at the next step.
  float fx[2], fy[2], ox = 0.0f, oy = 0.0f;
  const auto fetch = [&](int r) {
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const int t = threadIdx.x + k * 256;
      const int qq = q0 - kHalo + t;
      const bool in = r >= 0 && r < (int)H && t < kStrip + 2 * kHalo && qq >= 0 && qq < width3;
      const size_t idx = (size_t)(in ? r : 0) * width3 + (in ? qq : 0);
      fx[k] = in ? xin[idx] : 0.0f;
      fy[k] = in ? load_target<U8>(yin, idx, s_u8) : 0.0f;
    }
    const int orow = r - 5;
    if (orow >= ybeg && orow < yend && col_ok) {
      const size_t idx = (size_t)orow * width3 + q;
      ox = xin[idx];
      oy = load_target<U8>(yin, idx, s_u8);
    }
  };
  fetch(ybeg - 5);
  // Row rr enters the window; output row rr - 5 is then complete.
  for (int rr0 = ybeg - 5; rr0 < yend + 5; rr0 += 11) {
  static_for<11>([&](auto step) {
    constexpr int p = decltype(step)::value;
    const int rr = rr0 + p;
    if (rr >= yend + 5) return;  // uniform over the workgroup
    float nv[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const int t = threadIdx.x + k * 256;
      if (t < kStrip + 2 * kHalo) {
        sx[t] = fx[k];
        sy[t] = fy[k];
      }
    }
    const float xo = ox, yo = oy;  // output row rr - 5
    __syncthreads();
    fetch(rr + 1);
    if (rr >= 0 && rr < (int)H && col_ok) {
      // x, y, x^2, y^2, xy blurred horizontally (products formed per element first; the
      // segment's +0.0 outside the image stands for the zero padding)
      float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f, a4 = 0.0f;
#pragma unroll
      for (int t = 0; t < 11; t++) {
        const int o = q + (t - 5) * 3 - q0 + kHalo;
        const float xv = sx[o];
        const float yv = sy[o];
        const float kk = taps.k[t];
        a0 += kk * xv;
        a1 += kk * yv;
        a2 += kk * (xv * xv);
        a3 += kk * (yv * yv);
        a4 += kk * (xv * yv);
      }
      nv[0] = a0;
      nv[1] = a1;
      nv[2] = a2;
      nv[3] = a3;
      nv[4] = a4;
    }
#pragma unroll
    for (int m = 0; m < 5; m++) h[m][p] = nv[m];
    const int out_row = rr - 5;
    if (out_row >= ybeg && out_row < yend && col_ok) {
      float mu[5];
      vblur_ring<5>(h, p, taps, mu);
      const float mx = mu[0], my = mu[1];
      const float sx2 = mu[2] - mx * mx;
      const float sy2 = mu[3] - my * my;
      const float sxy = mu[4] - mx * my;
      const float n1 = 2.0f * mx * my + c1;
      const float n2 = 2.0f * sxy + c2;
      const float d1 = mx * mx + my * my + c1;
      const float d2 = sx2 + sy2 + c2;
      const float num = n1 * n2;
      const float den = d1 * d2;
      const float s = num / den;
      const float dn_dmx = 2.0f * my * n2 - 2.0f * my * n1;
      const float dd_dmx = 2.0f * mx * d2 - 2.0f * mx * d1;
      const size_t idx = (size_t)out_row * width3 + q;
      d_mu[idx] = (dn_dmx * den - num * dd_dmx) / (den * den);
      d_exx[idx] = -num * d1 / (den * den);
      d_exy[idx] = 2.0f * n1 / den;
      ssum += s;
      l1sum += fabsf(xo - yo);
    }
  });
  }
  const float bs = block_sum(ssum, red);
  const float bl = block_sum(l1sum, red);
  if (threadIdx.x == 0) partial[lb.flat] = make_float2(bs, bl);
}

// Sum of the per-block partials in binary64 by one 256-lane workgroup (lane-strided sums,
// then a fixed tree), so the result does not depend on timing.
__device__ __forceinline__ void sum_partials(const float2* __restrict__ partial, uint32_t nb,
                                             double* a_out, double* b_out) {
  __shared__ double sa[256], sb[256];
  double a = 0, b = 0;
  for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) {
    a += partial[i].x;
    b += partial[i].y;
  }
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sa[threadIdx.x] += sa[threadIdx.x + s];
      sb[threadIdx.x] += sb[threadIdx.x + s];
    }
    __syncthreads();
  }
  *a_out = sa[0];
  *b_out = sb[0];
}

// A graph-captured train step's loss value (to dyn->loss_out when set) and its
// host-visible report, from the loss pass's partial sums: every thread of one workgroup
// calls it.  The report's status words are final by then (only the binning's capacity
// guard writes them), and the host reads a report only after the step's completion event.
__device__ __forceinline__ void write_step_report(const StepLossReport& rep, float lambda,
                                                  const StepDyn* __restrict__ dyn) {
  float* out = dyn->loss_out;
  if (out) {
    double a, b;
    sum_partials(rep.partial, rep.nb, &a, &b);
    if (threadIdx.x == 0) {
      const float inv_n = 1.0f / rep.n;
      const float l1 = (float)b * inv_n;
      const float ssim = (float)a / rep.n;
      *out = (1.0f - lambda) * l1 + lambda * (1.0f - ssim);
    }
  }
  if (threadIdx.x == 0) {
    StepReport r;
    r.seq = dyn->seq;
    r.abort = rep.status->abort;
    r.total = rep.status->total;
    rep.reports[dyn->report_slot] = r;
  }
}

// Pass 2.  MODE 0: photometric-loss gradient; MODE 1: d(mean SSIM)/dx only.
template <bool U8, int MODE, bool BATCH>
__global__ void __launch_bounds__(256) ssim_grad_kernel(const float* __restrict__ xin,
                                                        const void* __restrict__ yin,
                                                        uint32_t W, uint32_t H, Taps taps,
                                                        const float* __restrict__ d_mu,
                                                        const float* __restrict__ d_exx,
                                                        const float* __restrict__ d_exy,
                                                        float inv_n, float w1, float lambda,
                                                        float* __restrict__ grad,
                                                        const StepDyn* __restrict__ dyn,
                                                        BatchDev bd, StepLossReport rep) {
  if (dyn) yin = dyn->target;
  // the partial sums of pass 1 are complete: workgroup 0 also writes the step's report
  if (rep.reports && blockIdx.x == 0 && blockIdx.y == 0) write_step_report(rep, lambda, dyn);
  size_t pix0;
  uint32_t pair;
  const LossBlock lb = loss_block<BATCH>(bd, &W, &H, &pix0, &pair);
  if (BATCH) {
    xin += pix0 * 3;
    yin = reinterpret_cast<const void*>(bd.targets[pair]);
    d_mu += pix0 * 3;
    d_exx += pix0 * 3;
    d_exy += pix0 * 3;
    grad += pix0 * 3;
    // per image: n = W * H * 3 (the same operations as the host computes)
    inv_n = 1.0f / (float)((size_t)W * H * 3);
    w1 = (1.0f - lambda) * inv_n;
  }
  __shared__ float s0[kStrip + 2 * kHalo], s1[kStrip + 2 * kHalo], s2[kStrip + 2 * kHalo];
  __shared__ float s_u8[U8 ? 256 : 1];
  fill_u8_table<U8>(s_u8);
  const int width3 = (int)W * 3;
  const int q0 = lb.bx * kStrip;
  const int q = q0 + threadIdx.x;
  const bool col_ok = q < width3;
  const int ybeg = lb.by * kRowsPerBlock;
  const int yend = min(ybeg + kRowsPerBlock, (int)H);
  float h[3][11];
#pragma unroll
  for (int m = 0; m < 3; m++)
#pragma unroll
    for (int t = 0; t < 11; t++) h[m][t] = 0.0f;
  // Software pipeline as in ssim_stats_kernel: row rr + 1 and output row rr - 4's x, y
  // are fetched while row rr is blurred.
  float f0[2], f1[2], f2[2], ox = 0.0f, oy = 0.0f;
  const auto fetch = [&](int r) {
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const int t = threadIdx.x + k * 256;
      const int qq = q0 - kHalo + t;
      const bool in = r >= 0 && r < (int)H && t < kStrip + 2 * kHalo && qq >= 0 && qq < width3;
      const size_t idx = (size_t)(in ? r : 0) * width3 + (in ? qq : 0);
      f0[k] = in ? d_mu[idx] : 0.0f;
      f1[k] = in ? d_exx[idx] : 0.0f;
      f2[k] = in ? d_exy[idx] : 0.0f;
    }
    const int orow = r - 5;
    if (orow >= ybeg && orow < yend && col_ok) {
      const size_t idx = (size_t)orow * width3 + q;
      ox = xin[idx];
      oy = load_target<U8>(yin, idx, s_u8);
    }
  };
  fetch(ybeg - 5);
  for (int rr0 = ybeg - 5; rr0 < yend + 5; rr0 += 11) {
  static_for<11>([&](auto step) {
    constexpr int p = decltype(step)::value;
    const int rr = rr0 + p;
    if (rr >= yend + 5) return;  // uniform over the workgroup
    float nv[3] = {0.0f, 0.0f, 0.0f};
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const int t = threadIdx.x + k * 256;
      if (t < kStrip + 2 * kHalo) {
        s0[t] = f0[k];
        s1[t] = f1[k];
        s2[t] = f2[k];
      }
    }
    const float xv = ox, yv = oy;  // output row rr - 5
    __syncthreads();
    fetch(rr + 1);
    if (rr >= 0 && rr < (int)H && col_ok) {
      nv[0] = hblur(s0, q, q0, taps);
      nv[1] = hblur(s1, q, q0, taps);
      nv[2] = hblur(s2, q, q0, taps);
    }
#pragma unroll
    for (int m = 0; m < 3; m++) h[m][p] = nv[m];
    const int out_row = rr - 5;
    if (out_row >= ybeg && out_row < yend && col_ok) {
      float b[3];
      vblur_ring<3>(h, p, taps, b);
      const size_t idx = (size_t)out_row * width3 + q;
      const float g = (b[0] + 2.0f * xv * b[1] + yv * b[2]) * inv_n;
      if (MODE == 1) {
        grad[idx] = g;
      } else {
        const float d = xv - yv;
        const float sign = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
        grad[idx] = w1 * sign - lambda * g;
      }
    }
  });
  }
}

// Separable blur for the host-tier blur() (arbitrary channel count).
__global__ void blur_h_kernel(const float* __restrict__ src, uint32_t W, uint32_t H, uint32_t C,
                              Taps taps, float* __restrict__ tmp) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t row = (size_t)W * C;
  if (idx >= row * H) return;
  const size_t y = idx / row, within = idx % row;
  const int x = (int)(within / C), c = (int)(within % C);
  float acc = 0.0f;
  for (int k = 0; k < 11; k++) {
    const int xs = x + k - 5;
    if (xs >= 0 && xs < (int)W) acc += taps.k[k] * src[y * row + (size_t)xs * C + c];
  }
  tmp[idx] = acc;
}

__global__ void blur_v_kernel(const float* __restrict__ tmp, uint32_t W, uint32_t H, uint32_t C,
                              Taps taps, float* __restrict__ dst) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t row = (size_t)W * C;
  if (idx >= row * H) return;
  const int y = (int)(idx / row);
  const size_t within = idx % row;
  float acc = 0.0f;
  for (int k = 0; k < 11; k++) {
    const int ys = y + k - 5;
    if (ys >= 0 && ys < (int)H) acc += taps.k[k] * tmp[(size_t)ys * row + within];
  }
  dst[idx] = acc;
}

// loss = (1 - lambda) * (sum |x - y| * inv_n) + lambda * (1 - sum SSIM / n), from the
// per-block partial sums (summed in binary64), written to *out without a host round trip.
__global__ void __launch_bounds__(256) loss_value_kernel(const float2* __restrict__ partial,
                                                         uint32_t nb, float lambda, float n,
                                                         float* __restrict__ out,
                                                         const StepDyn* __restrict__ dyn) {
  if (dyn) out = dyn->loss_out;
  if (!out) return;
  double a, b;
  sum_partials(partial, nb, &a, &b);
  if (threadIdx.x != 0) return;
  const float inv_n = 1.0f / n;
  const float l1 = (float)b * inv_n;
  const float ssim = (float)a / n;
  *out = (1.0f - lambda) * l1 + lambda * (1.0f - ssim);
}

// Per-pair loss values of a batched step: one workgroup per pair.
__global__ void __launch_bounds__(256) loss_values_batch_kernel(const float2* __restrict__ partial,
                                                                BatchDev bd, float lambda,
                                                                float* __restrict__ out) {
  const uint32_t v = blockIdx.x;
  double a, b;
  sum_partials(partial + bd.loss_block_base[v], bd.loss_block_base[v + 1] - bd.loss_block_base[v],
               &a, &b);
  if (threadIdx.x != 0) return;
  const float n = (float)((size_t)bd.width[v] * bd.height[v] * 3);
  const float inv_n = 1.0f / n;
  const float l1 = (float)b * inv_n;
  const float ssim = (float)a / n;
  out[v] = (1.0f - lambda) * l1 + lambda * (1.0f - ssim);
}

}  // namespace

unsigned loss_blocks(uint32_t width, uint32_t height) {
  return blocks_for((size_t)width * 3, kStrip) * blocks_for(height, kRowsPerBlock);
}

void loss_forward_backward_batch(splat_ctx* ctx, const float* pred, const BatchDev& bd,
                                 unsigned total_blocks, size_t total_values, float lambda,
                                 float* dimage, float* scratch, float* losses) {
  Taps taps;
  gaussian_kernel_host(taps.k);
  const float c1 = 0.01f * 0.01f, c2 = 0.03f * 0.03f;
  float* d_mu = scratch;
  float* d_exx = scratch + total_values;
  float* d_exy = scratch + 2 * total_values;
  float2* partial = reinterpret_cast<float2*>(scratch + 3 * total_values);
  launch(ctx, "loss_ssim_stats", ssim_stats_kernel<true, true>, dim3(total_blocks), dim3(256), 0,
         pred, (const void*)nullptr, 0u, 0u, taps, c1, c2, d_mu, d_exx, d_exy, partial,
         (const StepDyn*)nullptr, bd);
  launch(ctx, "loss_grad", ssim_grad_kernel<true, 0, true>, dim3(total_blocks), dim3(256), 0,
         pred, (const void*)nullptr, 0u, 0u, taps, (const float*)d_mu, (const float*)d_exx,
         (const float*)d_exy, 0.0f, 0.0f, lambda, dimage, (const StepDyn*)nullptr, bd,
         StepLossReport{});
  if (losses)
    launch(ctx, "loss_value", loss_values_batch_kernel, dim3(bd.npairs), dim3(256), 0,
           (const float2*)partial, bd, lambda, losses);
}

void loss_value_to(splat_ctx* ctx, const float* scratch, uint32_t width, uint32_t height,
                   float lambda, float* out_dev, const StepDyn* dyn) {
  const size_t n = (size_t)width * height * 3;
  const unsigned nb = blocks_for((size_t)width * 3, kStrip) * blocks_for(height, kRowsPerBlock);
  launch(ctx, "loss_value", loss_value_kernel, dim3(1), dim3(256), 0,
         reinterpret_cast<const float2*>(scratch + 3 * n), nb, lambda, (float)n, out_dev, dyn);
}

// Both passes and the loss value on caller-provided device buffers, with `scratch` of
// loss_scratch_floats(width, height) floats: plain launches on `stream` (no timing events,
// no host-side state), so nothing is allocated on the device or the host.
hipError_t loss_launch_bare(hipStream_t stream, const float* pred, const float* target,
                            uint32_t width, uint32_t height, float lambda, float* dimage,
                            float* loss_out, float* scratch) {
  const size_t n = (size_t)width * height * 3;
  Taps taps;
  gaussian_kernel_host(taps.k);
  const float c1 = 0.01f * 0.01f, c2 = 0.03f * 0.03f;
  float* d_mu = scratch;
  float* d_exx = scratch + n;
  float* d_exy = scratch + 2 * n;
  float2* partial = reinterpret_cast<float2*>(scratch + 3 * n);
  const dim3 grid(blocks_for((size_t)width * 3, kStrip), blocks_for(height, kRowsPerBlock));
  const float inv_n = 1.0f / (float)n;
  const float w1 = (1.0f - lambda) * inv_n;
  hipLaunchKernelGGL((ssim_stats_kernel<false, false>), grid, dim3(256), 0, stream, pred,
                     (const void*)target, width, height, taps, c1, c2, d_mu, d_exx, d_exy,
                     partial, (const StepDyn*)nullptr, BatchDev{});
  hipLaunchKernelGGL((ssim_grad_kernel<false, 0, false>), grid, dim3(256), 0, stream, pred,
                     (const void*)target, width, height, taps, (const float*)d_mu,
                     (const float*)d_exx, (const float*)d_exy, inv_n, w1, lambda, dimage,
                     (const StepDyn*)nullptr, BatchDev{}, StepLossReport{});
  if (loss_out)
    hipLaunchKernelGGL(loss_value_kernel, dim3(1), dim3(256), 0, stream,
                       (const float2*)partial, grid.x * grid.y, lambda, (float)n, loss_out,
                       (const StepDyn*)nullptr);
  return hipGetLastError();
}

// gaussian_kernel, src/loss.rs:14-27 (host, glibc-compatible expf).
void gaussian_kernel_host(float k[11]) {
  const float sigma = 1.5f;
  float sum = 0.0f;
  for (int i = 0; i < 11; i++) {
    const float x = (float)i - (float)(11 / 2);
    k[i] = expf(-(x * x) / (2.0f * sigma * sigma));
    sum += k[i];
  }
  for (int i = 0; i < 11; i++) k[i] /= sum;
}

size_t loss_scratch_floats(uint32_t width, uint32_t height) {
  const size_t n = (size_t)width * height * 3;
  const size_t blocks = (size_t)blocks_for(n / height, kStrip) * blocks_for(height, kRowsPerBlock);
  return 3 * n + 2 * blocks + 64;
}

// Runs both passes.  target_u8 (trainer views) or target_f32 (host tier) must be given.
// loss_parts_host, if non-null, receives {sum SSIM, sum |x - y|} (synchronises).
// want_l1 = false computes d(mean SSIM)/dx into dimage (ssim_with_grad).
void loss_forward_backward(splat_ctx* ctx, const float* pred, const float* target_f32,
                           const uint8_t* target_u8, uint32_t width, uint32_t height,
                           float lambda, bool want_l1, float* dimage, float* scratch,
                           double* loss_parts_host, const StepDyn* dyn, StepReport* reports,
                           const StepStatus* status) {
  const size_t n = (size_t)width * height * 3;
  Taps taps;
  gaussian_kernel_host(taps.k);
  const float c1 = 0.01f * 0.01f, c2 = 0.03f * 0.03f;
  float* d_mu = scratch;
  float* d_exx = scratch + n;
  float* d_exy = scratch + 2 * n;
  float2* partial = reinterpret_cast<float2*>(scratch + 3 * n);
  const dim3 grid(blocks_for((size_t)width * 3, kStrip), blocks_for(height, kRowsPerBlock));
  const float inv_n = 1.0f / (float)n;
  const float w1 = (1.0f - lambda) * inv_n;
  StepLossReport rep{};
  if (reports) {
    rep.partial = partial;
    rep.nb = grid.x * grid.y;
    rep.n = (float)n;
    rep.status = status;
    rep.reports = reports;
  }
  if (target_u8) {
    launch(ctx, "loss_ssim_stats", ssim_stats_kernel<true, false>, grid, dim3(256), 0, pred,
           (const void*)target_u8, width, height, taps, c1, c2, d_mu, d_exx, d_exy, partial, dyn, BatchDev{});
    if (want_l1)
      launch(ctx, "loss_grad", ssim_grad_kernel<true, 0, false>, grid, dim3(256), 0, pred,
             (const void*)target_u8, width, height, taps, (const float*)d_mu,
             (const float*)d_exx, (const float*)d_exy, inv_n, w1, lambda, dimage, dyn,
             BatchDev{}, rep);
    else
      launch(ctx, "loss_grad", ssim_grad_kernel<true, 1, false>, grid, dim3(256), 0, pred,
             (const void*)target_u8, width, height, taps, (const float*)d_mu,
             (const float*)d_exx, (const float*)d_exy, inv_n, w1, lambda, dimage, dyn,
             BatchDev{}, rep);
  } else {
    launch(ctx, "loss_ssim_stats", ssim_stats_kernel<false, false>, grid, dim3(256), 0, pred,
           (const void*)target_f32, width, height, taps, c1, c2, d_mu, d_exx, d_exy, partial, dyn, BatchDev{});
    if (want_l1)
      launch(ctx, "loss_grad", ssim_grad_kernel<false, 0, false>, grid, dim3(256), 0, pred,
             (const void*)target_f32, width, height, taps, (const float*)d_mu,
             (const float*)d_exx, (const float*)d_exy, inv_n, w1, lambda, dimage, dyn,
             BatchDev{}, rep);
    else
      launch(ctx, "loss_grad", ssim_grad_kernel<false, 1, false>, grid, dim3(256), 0, pred,
             (const void*)target_f32, width, height, taps, (const float*)d_mu,
             (const float*)d_exx, (const float*)d_exy, inv_n, w1, lambda, dimage, dyn,
             BatchDev{}, rep);
  }
  if (loss_parts_host) {
    const size_t nb = (size_t)grid.x * grid.y;
    std::vector<float2> parts(nb);
    SPLAT_HIP(hipMemcpyAsync(parts.data(), partial, nb * sizeof(float2), hipMemcpyDeviceToHost,
                             ctx->stream));
    SPLAT_HIP(hipStreamSynchronize(ctx->stream));
    double a = 0, b = 0;
    for (auto& p : parts) {
      a += p.x;
      b += p.y;
    }
    loss_parts_host[0] = a;
    loss_parts_host[1] = b;
  }
}

void blur_image(splat_ctx* ctx, const float* src, uint32_t width, uint32_t height,
                uint32_t channels, float* dst) {
  Taps taps;
  gaussian_kernel_host(taps.k);
  const size_t n = (size_t)width * height * channels;
  float* tmp = ctx->buf<float>("blur_tmp", n);
  launch(ctx, "blur_h", blur_h_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, src, width, height,
         channels, taps, tmp);
  launch(ctx, "blur_v", blur_v_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
         (const float*)tmp, width, height, channels, taps, dst);
}

}  // namespace splat
