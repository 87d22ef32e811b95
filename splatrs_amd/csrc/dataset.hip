// Dataset loading on the device side (src/dataset.rs:131-205, 239-250): the integer
// downscale of training images, their order and the evaluation split.
//
// dataset::load resizes each decoded RGB8 image with image::imageops::resize(..,
// FilterType::Lanczos3) from the `image` crate (0.25.10, Cargo.lock), a separable
// resampler: a vertical pass into a binary32 RGBA image, then a horizontal pass that clamps
// to [0, 255] and rounds to nearest (ties away from zero).  For each output row (column)
// the filter is centred at (o + 0.5) * ratio in input coordinates, its support widened by
// the ratio when shrinking, and its taps lanczos3((i - centre + 0.5) / max(ratio, 1)) are
// normalised to sum to one; every weight and sum is binary32.  The taps depend only on the
// two sizes, so they are computed here on the host (glibc sinf, as the crate's f32::sin)
// and the device does the per-pixel weighted sums in tap order, one lane per output value.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "internal.hpp"

namespace splat {
namespace {

// sinc(t) = sin(pi t) / (pi t), lanczos(x, 3) = sinc(x) sinc(x / 3) for |x| < 3.
float lanczos3(float x) {
  const auto sinc = [](float t) {
    if (t == 0.0f) return 1.0f;
    const float a = t * 3.14159265358979323846f;
    return std::sin(a) / a;
  };
  return std::fabs(x) < 3.0f ? sinc(x) * sinc(x / 3.0f) : 0.0f;
}

// One axis of the resampler: output o reads inputs [first[o], first[o] + count[o]) with
// weights w[offset[o] + i].
struct AxisTaps {
  std::vector<uint32_t> first, count, offset;
  std::vector<float> w;
};

AxisTaps axis_taps(uint32_t in, uint32_t out) {
  AxisTaps a;
  const float ratio = (float)in / (float)out;
  const float sratio = ratio < 1.0f ? 1.0f : ratio;
  const float support = 3.0f * sratio;
  a.first.resize(out);
  a.count.resize(out);
  a.offset.resize(out);
  for (uint32_t o = 0; o < out; o++) {
    const float centre = ((float)o + 0.5f) * ratio;
    const int64_t left = std::clamp<int64_t>((int64_t)std::floor(centre - support), 0,
                                             (int64_t)in - 1);
    const int64_t right = std::clamp<int64_t>((int64_t)std::ceil(centre + support), left + 1,
                                              (int64_t)in);
    const float c = centre - 0.5f;
    a.first[o] = (uint32_t)left;
    a.count[o] = (uint32_t)(right - left);
    a.offset[o] = (uint32_t)a.w.size();
    float sum = 0.0f;
    for (int64_t i = left; i < right; i++) {
      const float w = lanczos3(((float)i - c) / sratio);
      a.w.push_back(w);
      sum += w;
    }
    for (size_t i = a.offset[o]; i < a.w.size(); i++) a.w[i] /= sum;
  }
  return a;
}

// Vertical pass: tmp[oy][x][c] = sum_i src[first + i][x][c] * w[i] (binary32, tap order).
__global__ void __launch_bounds__(256) resize_rows_kernel(
    const uint8_t* __restrict__ src, uint32_t row_values, uint32_t out_rows,
    const uint32_t* __restrict__ first, const uint32_t* __restrict__ count,
    const uint32_t* __restrict__ offset, const float* __restrict__ w, float* __restrict__ tmp) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;  // x * 3 + c
  const uint32_t oy = blockIdx.y;
  if (q >= row_values || oy >= out_rows) return;
  const uint32_t f = first[oy], n = count[oy], o = offset[oy];
  float t = 0.0f;
  for (uint32_t i = 0; i < n; i++) t += (float)src[(size_t)(f + i) * row_values + q] * w[o + i];
  tmp[(size_t)oy * row_values + q] = t;
}

// Horizontal pass: dst[y][ox][c] = round(clamp(sum_i tmp[y][first + i][c] * w[i], 0, 255)).
__global__ void __launch_bounds__(256) resize_cols_kernel(
    const float* __restrict__ tmp, uint32_t in_width, uint32_t out_width, uint32_t rows,
    const uint32_t* __restrict__ first, const uint32_t* __restrict__ count,
    const uint32_t* __restrict__ offset, const float* __restrict__ w, uint8_t* __restrict__ dst) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;  // ox * 3 + c
  const uint32_t y = blockIdx.y;
  if (q >= out_width * 3 || y >= rows) return;
  const uint32_t ox = q / 3, c = q % 3;
  const uint32_t f = first[ox], n = count[ox], o = offset[ox];
  const float* row = tmp + (size_t)y * in_width * 3 + c;
  float t = 0.0f;
  for (uint32_t i = 0; i < n; i++) t += row[(size_t)(f + i) * 3] * w[o + i];
  t = fminf(fmaxf(t, 0.0f), 255.0f);  // clamp (t is finite: weights and inputs are)
  dst[(size_t)y * out_width * 3 + q] = (uint8_t)roundf(t);
}

}  // namespace

void resize_rgb8_device(splat_ctx* ctx, const uint8_t* src, uint32_t w, uint32_t h,
                        uint32_t out_w, uint32_t out_h, uint8_t* dst) {
  SPLAT_REQUIRE(out_w > 0 && out_h > 0, "empty output image");
  if (w == 0 || h == 0) {  // nothing to sample from: a black image, as the crate returns
    SPLAT_HIP(hipMemsetAsync(dst, 0, (size_t)out_w * out_h * 3, ctx->stream));
    return;
  }
  if (w == out_w && h == out_h) {  // the crate copies
    SPLAT_HIP(hipMemcpyAsync(dst, src, (size_t)w * h * 3, hipMemcpyDeviceToDevice, ctx->stream));
    return;
  }
  const AxisTaps ty = axis_taps(h, out_h), tx = axis_taps(w, out_w);
  // one upload for both axes: first, count, offset of rows then columns, then the weights
  std::vector<uint32_t> idx;
  for (const AxisTaps* a : {&ty, &tx})
    for (const auto* v : {&a->first, &a->count, &a->offset}) idx.insert(idx.end(), v->begin(), v->end());
  std::vector<float> wts(ty.w);
  wts.insert(wts.end(), tx.w.begin(), tx.w.end());
  uint32_t* d_idx = ctx->buf<uint32_t>("resize_idx", idx.size());
  float* d_w = ctx->buf<float>("resize_w", wts.size());
  float* d_tmp = ctx->buf<float>("resize_tmp", (size_t)w * out_h * 3);
  SPLAT_HIP(hipMemcpyAsync(d_idx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  SPLAT_HIP(hipMemcpyAsync(d_w, wts.data(), wts.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  const uint32_t* yf = d_idx;
  const uint32_t* yc = yf + out_h;
  const uint32_t* yo = yc + out_h;
  const uint32_t* xf = yo + out_h;
  const uint32_t* xc = xf + out_w;
  const uint32_t* xo = xc + out_w;
  launch(ctx, "resize_rows", resize_rows_kernel, dim3(blocks_for((size_t)w * 3, 256), out_h),
         dim3(256), 0, src, w * 3, out_h, yf, yc, yo, (const float*)d_w, d_tmp);
  launch(ctx, "resize_cols", resize_cols_kernel, dim3(blocks_for((size_t)out_w * 3, 256), out_h),
         dim3(256), 0, (const float*)d_tmp, w, out_w, out_h, xf, xc, xo,
         (const float*)(d_w + ty.w.size()), dst);
  // the host vectors above are copied before they go out of scope
  SPLAT_HIP(hipStreamSynchronize(ctx->stream));
}

}  // namespace splat

using namespace splat;

extern "C" {

int splat_resize_rgb8(splat_ctx* ctx, const uint8_t* src, uint32_t width, uint32_t height,
                      uint32_t out_width, uint32_t out_height, uint8_t* dst) {
  return ctx_guard(ctx, [&] {
    SPLAT_REQUIRE(dst && (src || (size_t)width * height == 0), "null argument");
    SPLAT_REQUIRE(out_width > 0 && out_height > 0, "empty output image");
    const size_t in_bytes = (size_t)width * height * 3, out_bytes = (size_t)out_width * out_height * 3;
    uint8_t* d_src = ctx->buf<uint8_t>("resize_src", in_bytes + 4);
    uint8_t* d_dst = ctx->buf<uint8_t>("resize_dst", out_bytes + 4);
    if (in_bytes)
      SPLAT_HIP(hipMemcpyAsync(d_src, src, in_bytes, hipMemcpyHostToDevice, ctx->stream));
    resize_rgb8_device(ctx, d_src, width, height, out_width, out_height, d_dst);
    SPLAT_HIP(hipMemcpyAsync(dst, d_dst, out_bytes, hipMemcpyDeviceToHost, ctx->stream));
    SPLAT_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int splat_downscaled_size(uint32_t width, uint32_t height, uint32_t downscale,
                          uint32_t* out_width, uint32_t* out_height) {
  if (!out_width || !out_height) return SPLAT_ERR_INVALID;
  if (downscale == 0) return SPLAT_ERR_INVALID;  // "downscale must be at least 1"
  // ((size as f64 / downscale as f64).round() as u32).max(1), src/dataset.rs:172-175
  const auto scaled = [&](uint32_t v) {
    const double r = std::round((double)v / (double)downscale);
    return std::max<uint32_t>((uint32_t)r, 1u);
  };
  *out_width = scaled(width);
  *out_height = scaled(height);
  return SPLAT_OK;
}

int splat_split_eval(uint64_t n, uint32_t hold, uint8_t* is_test) {
  if (n && !is_test) return SPLAT_ERR_INVALID;
  for (uint64_t i = 0; i < n; i++) is_test[i] = hold > 0 && i % hold == 0;  // dataset.rs:239-250
  return SPLAT_OK;
}

int splat_colmap_sort_images(splat_colmap_image* images, uint64_t n) {
  if (n && !images) return SPLAT_ERR_INVALID;
  // model.images.sort_by(|a, b| a.name.cmp(&b.name)), src/dataset.rs:141: stable, by bytes
  std::stable_sort(images, images + n, [](const splat_colmap_image& a, const splat_colmap_image& b) {
    return std::strcmp(a.name, b.name) < 0;
  });
  return SPLAT_OK;
}

}  // extern "C"
