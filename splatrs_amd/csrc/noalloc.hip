// The photometric loss for callers that cannot allocate device memory (code under HIP graph
// capture, tight evaluation loops, harnesses that forbid hipMalloc).  Inputs and outputs
// are the caller's device buffers; the loss kernels' scratch (three blurred SSIM partial
// maps and the per-workgroup sums) lives in one static device array that each GPU holds for
// the life of the process.
#include <mutex>

#include "internal.hpp"

namespace {

// Images of up to 2^21 pixels (1920x1080 fits).  The loss needs 3 floats per value for the
// partial maps (9 per pixel) and 2 per workgroup for the sums: at most about 66k
// workgroups for any shape of that many pixels, within the 2^18 floats set aside.
constexpr size_t kMaxPixels = size_t(1) << 21;
constexpr size_t kScratchFloats = 9 * kMaxPixels + (size_t(1) << 18);  // 73 MiB

constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) float scratch[kScratchFloats];

namespace {

std::mutex g_mu;
float* g_scratch_dev[kMaxDevices] = {};  // this device's address of `scratch`

}  // namespace

namespace splat {

// With `device` current: resolves its copy of `scratch` once.
int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return SPLAT_ERR_INVALID;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch_dev[device]) return SPLAT_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SPLAT_ERR_HIP;
  g_scratch_dev[device] = static_cast<float*>(p);
  return SPLAT_OK;
}

}  // namespace splat

extern "C" {

int splat_photometric_loss_device_noalloc(splat_ctx* ctx, const float* pred,
                                          const float* target, uint32_t width, uint32_t height,
                                          float lambda_dssim, float* loss, float* grad) {
  if (!ctx) return SPLAT_ERR_INVALID;
  if (!pred || !target || !grad) {
    ctx->err = "null argument";
    return SPLAT_ERR_PRECONDITION;
  }
  const size_t pixels = (size_t)width * height;
  if (pixels == 0) {
    ctx->err = "empty image";
    return SPLAT_ERR_PRECONDITION;
  }
  if (pixels > kMaxPixels || splat::loss_scratch_floats(width, height) > kScratchFloats) {
    ctx->err = "image of " + std::to_string(width) + "x" + std::to_string(height) +
               " pixels exceeds the no-allocation maximum of " + std::to_string(kMaxPixels) +
               " pixels";
    return SPLAT_ERR_INVALID;
  }
  if (hipSetDevice(ctx->device) != hipSuccess) {
    ctx->err = "hipSetDevice failed";
    return SPLAT_ERR_HIP;
  }
  const int st = splat::noalloc_setup(ctx->device);
  if (st != SPLAT_OK) {
    ctx->err = "cannot resolve the no-allocation scratch array";
    return st;
  }
  const hipError_t e = splat::loss_launch_bare(ctx->stream, pred, target, width, height,
                                               lambda_dssim, grad, loss,
                                               g_scratch_dev[ctx->device]);
  if (e != hipSuccess) {
    ctx->err = std::string("kernel launch failed: ") + hipGetErrorString(e);
    return SPLAT_ERR_HIP;
  }
  ctx->err.clear();
  return SPLAT_OK;
}

}  // extern "C"
