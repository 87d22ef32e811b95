// Initialisation from a point cloud (Model::from_points, src/model.rs:77-104) with the
// mean squared distance to the three nearest other points (mean_nn_dist2,
// src/model.rs:128-186) computed on the GPU.
//
// Brute force in binary64: each lane owns one query point and streams every point through
// LDS, keeping the three nearest other points.  The reference queries a kd-tree over the
// unique positions and walks them by distance adding `distance * take`; here candidates
// at the same position are grouped the same way, so the result is the reference's value.
#include <cmath>
#include <cstring>

#include "internal.hpp"

namespace splat {
namespace {

constexpr int kKnnBlock = 256;

__global__ void __launch_bounds__(kKnnBlock) knn3_kernel(const float* __restrict__ xyz, uint32_t n,
                                                         double* __restrict__ out) {
  __shared__ double sp[kKnnBlock * 3];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  const double qx = valid ? (double)xyz[3 * (size_t)i] : 0.0;
  const double qy = valid ? (double)xyz[3 * (size_t)i + 1] : 0.0;
  const double qz = valid ? (double)xyz[3 * (size_t)i + 2] : 0.0;
  double d0 = INFINITY, d1 = INFINITY, d2 = INFINITY;
  uint32_t j0 = 0xffffffffu, j1 = 0xffffffffu, j2 = 0xffffffffu;
  for (uint32_t base = 0; base < n; base += kKnnBlock) {
    __syncthreads();
    const uint32_t jl = base + threadIdx.x;
    if (jl < n) {
      sp[3 * threadIdx.x] = (double)xyz[3 * (size_t)jl];
      sp[3 * threadIdx.x + 1] = (double)xyz[3 * (size_t)jl + 1];
      sp[3 * threadIdx.x + 2] = (double)xyz[3 * (size_t)jl + 2];
    }
    __syncthreads();
    const uint32_t cnt = min((uint32_t)kKnnBlock, n - base);
    for (uint32_t k = 0; k < cnt; k++) {
      const uint32_t j = base + k;
      const double dx = qx - sp[3 * k], dy = qy - sp[3 * k + 1], dz = qz - sp[3 * k + 2];
      const double d = dx * dx + dy * dy + dz * dz;  // SquaredEuclidean, x, y, z in order
      if (j == i || !(d < d2)) continue;
      if (d < d1) {
        d2 = d1;
        j2 = j1;
        if (d < d0) {
          d1 = d0;
          j1 = j0;
          d0 = d;
          j0 = j;
        } else {
          d1 = d;
          j1 = j;
        }
      } else {
        d2 = d;
        j2 = j;
      }
    }
  }
  if (!valid) return;
  // Walk the candidates by distance, grouping points at the same position.
  const double ds[3] = {d0, d1, d2};
  const uint32_t js[3] = {j0, j1, j2};
  double sum = 0.0;
  uint32_t counted = 0;
  for (int c = 0; c < 3;) {
    if (js[c] == 0xffffffffu) break;
    const uint32_t a = js[c];
    int g = c + 1;
    // same position = the same bits in all three coordinates (the reference's key)
    while (g < 3 && js[g] != 0xffffffffu &&
           f2bits(xyz[3 * (size_t)js[g]]) == f2bits(xyz[3 * (size_t)a]) &&
           f2bits(xyz[3 * (size_t)js[g] + 1]) == f2bits(xyz[3 * (size_t)a + 1]) &&
           f2bits(xyz[3 * (size_t)js[g] + 2]) == f2bits(xyz[3 * (size_t)a + 2]))
      g++;
    const uint32_t take = (uint32_t)(g - c);
    sum += ds[c] * (double)take;
    counted += take;
    c = g;
  }
  out[i] = counted > 0 ? sum / (double)counted : 0.0;
}

}  // namespace

void mean_nn_dist2_dev(splat_ctx* ctx, const float* xyz, size_t n, double* out) {
  launch(ctx, "knn3", knn3_kernel, dim3(blocks_for(n, kKnnBlock)), dim3(kKnnBlock), 0, xyz,
         (uint32_t)n, out);
}

}  // namespace splat

using namespace splat;

extern "C" {

int splat_mean_nn_dist2(splat_ctx* ctx, const float* xyz, uint64_t n, double* out) {
  if (!ctx) return SPLAT_ERR_INVALID;
  try {
    SPLAT_HIP(hipSetDevice(ctx->device));
    SPLAT_REQUIRE(n == 0 || (xyz && out), "null argument");
    if (n < 2) {
      for (uint64_t i = 0; i < n; i++) out[i] = 0.0;
      return SPLAT_OK;
    }
    float* dx = ctx->buf<float>("knn_xyz", 3 * n);
    double* dd = ctx->buf<double>("knn_out", n);
    SPLAT_HIP(hipMemcpyAsync(dx, xyz, 3 * n * 4, hipMemcpyHostToDevice, ctx->stream));
    mean_nn_dist2_dev(ctx, dx, n, dd);
    SPLAT_HIP(hipMemcpyAsync(out, dd, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    SPLAT_HIP(hipStreamSynchronize(ctx->stream));
    ctx->err.clear();
    return SPLAT_OK;
  } catch (const Error& e) {
    ctx->err = e.msg;
    return e.code;
  }
}

int splat_from_points(splat_ctx* ctx, const float* xyz, const uint8_t* rgb, uint64_t n,
                      float* gaussians) {
  if (!ctx) return SPLAT_ERR_INVALID;
  if (n && (!xyz || !rgb || !gaussians)) {
    ctx->err = "null argument";
    return SPLAT_ERR_INVALID;
  }
  std::vector<double> d2(n);
  const int st = splat_mean_nn_dist2(ctx, xyz, n, d2.data());
  if (st != SPLAT_OK) return st;
  const float opacity = logf(0.1f / (1.0f - 0.1f));  // inverse_sigmoid(0.1), math.rs:395-397
  const float sh_c0 = (float)0.28209479177387814;
  for (uint64_t i = 0; i < n; i++) {
    float* g = gaussians + i * kParams;
    std::memset(g, 0, kParams * sizeof(float));
    const float dd = rmax((float)d2[i], 1e-7f);
    const float log_s = logf(sqrtf(dd));
    for (int c = 0; c < 3; c++) {
      g[c] = xyz[3 * i + c];
      g[3 + c] = log_s;
      g[11 + c] = ((float)rgb[3 * i + c] / 255.0f - 0.5f) / sh_c0;  // rgb_to_sh, sh.rs:41-43
    }
    g[6] = 1.0f;
    g[10] = opacity;
  }
  return SPLAT_OK;
}

}  // extern "C"
