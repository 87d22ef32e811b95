// C ABI of libsplat: context, timing, and the host tier of include/splat.h.
//
// Host-tier calls upload their inputs into context-owned device buffers, run the HIP
// kernels and download the outputs.  Caller-supplied indices (binning arrays) are
// validated on the host before any kernel reads through them; a malformed binning is a
// precondition failure (the reference panics on the equivalent out-of-bounds index).
#include <cmath>
#include <cstring>
#include <new>

#include "internal.hpp"

using namespace splat;


namespace {

template <class F>
int guard(splat_ctx* ctx, F&& f) {
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

void h2d(splat_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes) SPLAT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
}
void d2h(splat_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes) SPLAT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
}
void sync(splat_ctx* ctx) { SPLAT_HIP(hipStreamSynchronize(ctx->stream)); }

// Upload n AoS records of 59 floats into planes (returned pointer, stride set).
float* upload_planes(splat_ctx* ctx, const char* name, const float* aos, size_t n,
                     size_t* stride) {
  *stride = plane_stride(n);
  float* staging = ctx->buf<float>("aos_staging", n * kParams + 1);
  float* planes = ctx->buf<float>(name, *stride * kParams);
  h2d(ctx, staging, aos, n * kParams * 4);
  aos_to_planes(ctx, staging, n, planes, *stride);
  return planes;
}

void download_planes(splat_ctx* ctx, const float* planes, size_t n, size_t stride, float* aos) {
  float* staging = ctx->buf<float>("aos_staging_out", n * kParams + 1);
  planes_to_aos(ctx, planes, n, stride, staging);
  d2h(ctx, aos, staging, n * kParams * 4);
}

ProjBuffers alloc_proj(splat_ctx* ctx, size_t n) {
  ProjBuffers pb;
  const size_t c = n + 1;
  pb.rec = ctx->buf<float4>("proj_rec", 4 * c);
  pb.depth = ctx->buf<float>("proj_depth", c);
  pb.radius = ctx->buf<uint32_t>("proj_radius", c);
  pb.rect = ctx->buf<uint2>("proj_rect", c);
  pb.ntiles = ctx->buf<uint32_t>("proj_ntiles", c);
  return pb;
}

void tile_grid(uint32_t w, uint32_t h, uint32_t* tx, uint32_t* ty) {  // src/project.rs:272-274
  *tx = (w + kTile - 1) / kTile;
  *ty = (h + kTile - 1) / kTile;
}

// Upload host Proj records into rasterizer records, checking the tile rectangles
// against the grid (src/rasterize.rs:65-69).
ProjBuffers upload_projs(splat_ctx* ctx, const splat_proj* projs, size_t n, uint32_t w,
                         uint32_t h) {
  uint32_t tx, ty;
  tile_grid(w, h, &tx, &ty);
  ProjBuffers pb = alloc_proj(ctx, n);
  splat_proj* dp = ctx->buf<splat_proj>("proj_upload", n + 1);
  uint32_t* bad = ctx->buf<uint32_t>("proj_bad", 1);
  h2d(ctx, dp, projs, n * sizeof(splat_proj));
  SPLAT_HIP(hipMemsetAsync(bad, 0, 4, ctx->stream));
  proj_unpack(ctx, dp, n, tx, ty, pb, bad);
  uint32_t b = 0;
  d2h(ctx, &b, bad, 4);
  sync(ctx);
  SPLAT_REQUIRE(b == 0,
                "projection tile rectangle exceeds the tile grid; the projections were made for "
                "a different image size");
  return pb;
}

// Device copy of a caller's binning (validated).
BinBuffers upload_binning(splat_ctx* ctx, const splat_binning* b, size_t n, uint32_t w,
                          uint32_t h) {
  uint32_t tx, ty;
  tile_grid(w, h, &tx, &ty);
  SPLAT_REQUIRE(b && b->tiles_x == tx && b->tiles_y == ty, "binning tile grid mismatch");
  const size_t E = b->num_entries, T = (size_t)tx * ty;
  SPLAT_REQUIRE(E < 0xffffffffull, "too many tile entries");
  SPLAT_REQUIRE(E == 0 || (b->entry_gaussian && b->sorted_entries), "binning arrays missing");
  SPLAT_REQUIRE(b->entry_offset && b->tile_range, "binning arrays missing");
  for (size_t i = 0; i < n; i++)
    SPLAT_REQUIRE(b->entry_offset[i] <= b->entry_offset[i + 1], "entry_offset not monotone");
  SPLAT_REQUIRE(b->entry_offset[n] <= E, "entry_offset exceeds the entry count");
  std::vector<uint32_t> start(T + 1), gauss(E);
  uint32_t run = 0;
  for (size_t t = 0; t < T; t++) {
    const uint32_t s = b->tile_range[2 * t], e = b->tile_range[2 * t + 1];
    if (e > s) {
      SPLAT_REQUIRE(s == run && e <= E, "tile ranges are not a partition of the sorted list");
      start[t] = s;
      run = e;
    } else {
      start[t] = run;
    }
  }
  SPLAT_REQUIRE(run == E, "tile ranges do not cover the sorted list");
  start[T] = (uint32_t)E;
  for (size_t k = 0; k < E; k++) {
    const uint32_t e = b->sorted_entries[k];
    SPLAT_REQUIRE(e < E, "sorted entry index out of range");
    const uint32_t g = b->entry_gaussian[e];
    SPLAT_REQUIRE(g < n, "entry Gaussian index out of range");
    gauss[k] = g;
  }
  BinBuffers bb{};
  bb.tiles_x = tx;
  bb.tiles_y = ty;
  bb.entry_offset = ctx->buf<uint32_t>("hb_entry_offset", n + 1);
  bb.tile_start = ctx->buf<uint32_t>("hb_tile_start", T + 1);
  bb.sorted_entry = ctx->buf<uint32_t>("hb_sorted_entry", E + 1);
  bb.sorted_gauss = ctx->buf<uint32_t>("hb_sorted_gauss", E + 1);
  h2d(ctx, bb.entry_offset, b->entry_offset, (n + 1) * 4);
  h2d(ctx, bb.tile_start, start.data(), (T + 1) * 4);
  h2d(ctx, bb.sorted_entry, b->sorted_entries, E * 4);
  h2d(ctx, bb.sorted_gauss, gauss.data(), E * 4);
  sync(ctx);  // the host vectors go out of scope
  return bb;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- context ----------

int splat_ctx_create(int device, splat_ctx** out) {
  if (!out) return SPLAT_ERR_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    return SPLAT_ERR_NO_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return SPLAT_ERR_NO_DEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return SPLAT_ERR_NO_DEVICE;
  splat_ctx* ctx = new (std::nothrow) splat_ctx();
  if (!ctx) return SPLAT_ERR_OOM;
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
    delete ctx;
    return SPLAT_ERR_HIP;
  }
  // this device's address of the no-allocation entry point's scratch array
  if (splat::noalloc_setup(device) != SPLAT_OK) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return SPLAT_ERR_HIP;
  }
  *out = ctx;
  return SPLAT_OK;
}

int splat_ctx_destroy(splat_ctx* ctx) {
  if (!ctx) return SPLAT_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& p : ctx->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  ctx->ws.clear();
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return SPLAT_OK;
}

const char* splat_last_error(const splat_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int splat_sync(splat_ctx* ctx) {
  return guard(ctx, [&] { sync(ctx); });
}

void* splat_ctx_stream(splat_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int splat_set_timing(splat_ctx* ctx, int enabled) {
  return guard(ctx, [&] {
    ctx->drain_timing();
    ctx->timing = enabled != 0;
  });
}

int splat_set_bin_method(splat_ctx* ctx, int method) {
  return guard(ctx, [&] {
    SPLAT_REQUIRE(method == 0 || method == 1, "bin method must be 0 or 1");
    ctx->bin_method = method;
  });
}

int splat_reset_timing(splat_ctx* ctx) {
  return guard(ctx, [&] {
    ctx->drain_timing();
    ctx->kstats.clear();
  });
}

int splat_kernel_times(splat_ctx* ctx, char* names, size_t names_len, uint64_t* launches,
                       double* total_ms, uint32_t cap, uint32_t* count) {
  return guard(ctx, [&] {
    ctx->drain_timing();
    uint32_t k = 0;
    size_t off = 0;
    for (auto& kv : ctx->kstats) {
      if (k >= cap) break;
      const size_t len = kv.first.size() + 1;
      if (names && off + len <= names_len) {
        std::memcpy(names + off, kv.first.c_str(), len);
        off += len;
      }
      if (launches) launches[k] = kv.second.launches;
      if (total_ms) total_ms[k] = kv.second.ms;
      k++;
    }
    if (count) *count = k;
  });
}

// ---------------------------------------------------------------- host tier --------

int splat_project_all(splat_ctx* ctx, const float* gaussians, uint64_t n, const splat_camera* cam,
                      const splat_proj_opts* opts, splat_proj* out) {
  return guard(ctx, [&] {
    SPLAT_REQUIRE(cam && opts && (n == 0 || (gaussians && out)), "null argument");
    SPLAT_REQUIRE(opts->sh_degree >= 0 && opts->sh_degree <= 3, "sh_degree must be 0..3");
    if (n == 0) return;
    size_t stride;
    float* planes = upload_planes(ctx, "params", gaussians, n, &stride);
    ProjBuffers pb = alloc_proj(ctx, n);
    project_forward(ctx, planes, n, stride, *cam, *opts, pb);
    splat_proj* dp = ctx->buf<splat_proj>("proj_out", n);
    proj_pack(ctx, pb, n, dp);
    d2h(ctx, out, dp, n * sizeof(splat_proj));
    sync(ctx);
  });
}

int splat_project_backward_all(splat_ctx* ctx, const float* gaussians, uint64_t n,
                               const splat_camera* cam, const splat_proj_opts* opts,
                               const splat_proj_grad* grads, float* out) {
  return guard(ctx, [&] {
    SPLAT_REQUIRE(cam && opts && (n == 0 || (gaussians && grads && out)), "null argument");
    SPLAT_REQUIRE(opts->sh_degree >= 0 && opts->sh_degree <= 3, "sh_degree must be 0..3");
    if (n == 0) return;
    size_t stride;
    float* planes = upload_planes(ctx, "params", gaussians, n, &stride);
    splat_proj_grad* dg = ctx->buf<splat_proj_grad>("pgrad_in", n);
    h2d(ctx, dg, grads, n * sizeof(splat_proj_grad));
    float* gp = ctx->buf<float>("grad_planes", stride * kParams);
    project_backward(ctx, planes, n, stride, *cam, *opts, dg, gp);
    download_planes(ctx, gp, n, stride, out);
    sync(ctx);
  });
}

int splat_bin(splat_ctx* ctx, const splat_proj* projs, uint64_t n, uint32_t width, uint32_t height,
              splat_binning* out) {
  return guard(ctx, [&] {
    SPLAT_REQUIRE(out && (n == 0 || projs), "null argument");
    uint32_t tx, ty;
    tile_grid(width, height, &tx, &ty);
    out->tiles_x = tx;
    out->tiles_y = ty;
    const size_t T = (size_t)tx * ty;
    ProjBuffers pb = upload_projs(ctx, projs, n, width, height);
    BinBuffers bb{};
    size_t cap = 0;
    // The first pass counts the entries; the second writes the sorted keys too.
    size_t E = bin_build(ctx, pb, n, tx, ty, bb, &cap, ctx->ws, nullptr);
    out->num_entries = E;
    if (out->capacity < E) fail(SPLAT_ERR_CAPACITY, "binning capacity too small");
    uint64_t* keys = ctx->buf<uint64_t>("bin_keys", E + 1);
    E = bin_build(ctx, pb, n, tx, ty, bb, &cap, ctx->ws, keys);
    std::vector<uint32_t> tstart(T + 1);
    d2h(ctx, out->entry_gaussian, bb.entry_gaussian, E * 4);
    d2h(ctx, out->entry_offset, bb.entry_offset, (n + 1) * 4);
    d2h(ctx, out->sorted_keys, keys, E * 8);
    d2h(ctx, out->sorted_entries, bb.sorted_entry, E * 4);
    d2h(ctx, tstart.data(), bb.tile_start, (T + 1) * 4);
    sync(ctx);
    for (size_t t = 0; t < T; t++) {
      const bool nonempty = tstart[t + 1] > tstart[t];
      out->tile_range[2 * t] = nonempty ? tstart[t] : 0;
      out->tile_range[2 * t + 1] = nonempty ? tstart[t + 1] : 0;
    }
  });
}

int splat_rasterize_forward(splat_ctx* ctx, const splat_proj* projs, uint64_t n,
                            const splat_binning* binning, const splat_raster_settings* settings,
                            float* image, float* final_t, uint32_t* n_contrib) {
  return guard(ctx, [&] {
    SPLAT_REQUIRE(settings && binning && image && final_t && n_contrib, "null argument");
    const uint32_t w = settings->width, h = settings->height;
    const size_t npx = (size_t)w * h;
    ProjBuffers pb = upload_projs(ctx, projs, n, w, h);
    BinBuffers bb = upload_binning(ctx, binning, n, w, h);
    float* dimg = ctx->buf<float>("image", npx * 3 + 1);
    float* dft = ctx->buf<float>("final_t", npx + 1);
    uint32_t* dnc = ctx->buf<uint32_t>("n_contrib", npx + 1);
    raster_forward(ctx, pb, bb, w, h, settings->background, dimg, dft, dnc);
    d2h(ctx, image, dimg, npx * 12);
    d2h(ctx, final_t, dft, npx * 4);
    d2h(ctx, n_contrib, dnc, npx * 4);
    sync(ctx);
  });
}

int splat_rasterize_backward(splat_ctx* ctx, const splat_proj* projs, uint64_t n,
                             const splat_binning* binning, const splat_raster_settings* settings,
                             const float* final_t, const uint32_t* n_contrib, const float* dimage,
                             int ordered, splat_proj_grad* grads, float* abs_grads) {
  return guard(ctx, [&] {
    SPLAT_REQUIRE(settings && binning && final_t && n_contrib && dimage && grads && abs_grads,
                  "null argument");
    const uint32_t w = settings->width, h = settings->height;
    const size_t npx = (size_t)w * h;
    ProjBuffers pb = upload_projs(ctx, projs, n, w, h);
    BinBuffers bb = upload_binning(ctx, binning, n, w, h);
    const size_t E = binning->num_entries;
    float* dft = ctx->buf<float>("final_t", npx + 1);
    uint32_t* dnc = ctx->buf<uint32_t>("n_contrib", npx + 1);
    float* ddi = ctx->buf<float>("dimage", npx * 3 + 1);
    float* slots = ctx->buf<float>("slots", E * kSlots + 1);
    h2d(ctx, dft, final_t, npx * 4);
    h2d(ctx, dnc, n_contrib, npx * 4);
    h2d(ctx, ddi, dimage, npx * 12);
    raster_backward(ctx, pb, bb, w, h, settings->background, dft, dnc, ddi, ordered != 0, slots,
                    true);
    splat_proj_grad* dg = ctx->buf<splat_proj_grad>("pgrad_out", n + 1);
    float* dab = ctx->buf<float>("abs_out", 2 * n + 1);
    reduce_slots(ctx, bb.entry_offset, slots, n, dg, dab);
    d2h(ctx, grads, dg, n * sizeof(splat_proj_grad));
    d2h(ctx, abs_grads, dab, n * 8);
    sync(ctx);
  });
}

int splat_render(splat_ctx* ctx, const float* gaussians, uint64_t n, const splat_camera* cam,
                 const splat_proj_opts* opts, const float background[3], float* image) {
  return guard(ctx, [&] {
    SPLAT_REQUIRE(cam && opts && background && image && (n == 0 || gaussians), "null argument");
    SPLAT_REQUIRE(opts->sh_degree >= 0 && opts->sh_degree <= 3, "sh_degree must be 0..3");
    const uint32_t w = cam->width, h = cam->height;
    const size_t npx = (size_t)w * h;
    uint32_t tx, ty;
    tile_grid(w, h, &tx, &ty);
    size_t stride = plane_stride(n);
    float* planes = n ? upload_planes(ctx, "params", gaussians, n, &stride) : nullptr;
    ProjBuffers pb = alloc_proj(ctx, n);
    if (n) project_forward(ctx, planes, n, stride, *cam, *opts, pb);
    BinBuffers bb{};
    size_t cap = 0;
    bin_build(ctx, pb, n, tx, ty, bb, &cap, ctx->ws, nullptr);
    float* dimg = ctx->buf<float>("image", npx * 3 + 1);
    float* dft = ctx->buf<float>("final_t", npx + 1);
    uint32_t* dnc = ctx->buf<uint32_t>("n_contrib", npx + 1);
    raster_forward(ctx, pb, bb, w, h, background, dimg, dft, dnc);
    d2h(ctx, image, dimg, npx * 12);
    sync(ctx);
  });
}

int splat_to_rgb8(const float* image, uint64_t len, uint8_t* out) {
  if (len && (!image || !out)) return SPLAT_ERR_INVALID;
  for (uint64_t i = 0; i < len; i++) {
    float v = rclamp(image[i], 0.0f, 1.0f) * 255.0f;  // src/render.rs:26-31
    out[i] = (uint8_t)std::min(as_u32(roundf(v)), 255u);
  }
  return SPLAT_OK;
}

int splat_photometric_loss(splat_ctx* ctx, const float* pred, const float* target, uint32_t width,
                           uint32_t height, float lambda_dssim, float* loss, float* grad) {
  return guard(ctx, [&] {
    SPLAT_REQUIRE(pred && target && loss && grad, "null argument");
    const size_t n = (size_t)width * height * 3;
    SPLAT_REQUIRE(n > 0, "empty image");
    float* dp = ctx->buf<float>("loss_pred", n);
    float* dt = ctx->buf<float>("loss_target", n);
    float* dg = ctx->buf<float>("loss_grad", n);
    float* scratch = ctx->buf<float>("loss_scratch", loss_scratch_floats(width, height));
    h2d(ctx, dp, pred, n * 4);
    h2d(ctx, dt, target, n * 4);
    double parts[2];
    loss_forward_backward(ctx, dp, dt, nullptr, width, height, lambda_dssim, true, dg, scratch,
                          parts);
    d2h(ctx, grad, dg, n * 4);
    sync(ctx);
    const float inv_n = 1.0f / (float)n;
    const float l1 = (float)parts[1] * inv_n;
    const float ssim = (float)parts[0] / (float)n;
    *loss = (1.0f - lambda_dssim) * l1 + lambda_dssim * (1.0f - ssim);
  });
}

int splat_ssim_with_grad(splat_ctx* ctx, const float* x, const float* y, uint32_t width,
                         uint32_t height, float* ssim, float* grad) {
  return guard(ctx, [&] {
    SPLAT_REQUIRE(x && y && ssim && grad, "null argument");
    const size_t n = (size_t)width * height * 3;
    SPLAT_REQUIRE(n > 0, "empty image");
    float* dp = ctx->buf<float>("loss_pred", n);
    float* dt = ctx->buf<float>("loss_target", n);
    float* dg = ctx->buf<float>("loss_grad", n);
    float* scratch = ctx->buf<float>("loss_scratch", loss_scratch_floats(width, height));
    h2d(ctx, dp, x, n * 4);
    h2d(ctx, dt, y, n * 4);
    double parts[2];
    loss_forward_backward(ctx, dp, dt, nullptr, width, height, 0.0f, false, dg, scratch, parts);
    d2h(ctx, grad, dg, n * 4);
    sync(ctx);
    *ssim = (float)parts[0] / (float)n;
  });
}

int splat_blur(splat_ctx* ctx, const float* src, uint32_t width, uint32_t height, uint32_t channels,
               float* dst) {
  return guard(ctx, [&] {
    SPLAT_REQUIRE(src && dst, "null argument");
    const size_t n = (size_t)width * height * channels;
    if (n == 0) return;
    float* ds = ctx->buf<float>("blur_src", n);
    float* dd = ctx->buf<float>("blur_dst", n);
    h2d(ctx, ds, src, n * 4);
    blur_image(ctx, ds, width, height, channels, dd);
    d2h(ctx, dst, dd, n * 4);
    sync(ctx);
  });
}

int splat_psnr(const float* pred, const float* target, uint64_t len, float* out) {
  if (!out || (len && (!pred || !target))) return SPLAT_ERR_INVALID;
  double mse = 0;  // src/loss.rs:175-190
  for (uint64_t i = 0; i < len; i++) {
    const double d = (double)rclamp(pred[i], 0.0f, 1.0f) - (double)rclamp(target[i], 0.0f, 1.0f);
    mse += d * d;
  }
  mse /= (double)(len > 0 ? len : 1);
  *out = mse <= 0 ? INFINITY : (float)(-10.0 * std::log10(mse));
  return SPLAT_OK;
}

int splat_adam_step(splat_ctx* ctx, float* params, const float* grads, float* m, float* v,
                    uint64_t n, splat_adam_state* state, const splat_learning_rates* lr) {
  return guard(ctx, [&] {
    SPLAT_REQUIRE(state && lr && (n == 0 || (params && grads && m && v)), "null argument");
    state->step += 1;
    const float c1 = 1.0f - powi(state->beta1, (int32_t)state->step);  // src/adam.rs:82-83
    const float c2 = 1.0f - powi(state->beta2, (int32_t)state->step);
    if (n == 0) return;
    size_t stride;
    float* dp = upload_planes(ctx, "adam_p", params, n, &stride);
    float* dg = upload_planes(ctx, "adam_g", grads, n, &stride);
    float* dm = upload_planes(ctx, "adam_m", m, n, &stride);
    float* dv = upload_planes(ctx, "adam_v", v, n, &stride);
    adam_planes(ctx, dp, dg, dm, dv, n, stride, c1, c2, *state, *lr);
    download_planes(ctx, dp, n, stride, params);
    sync(ctx);
    download_planes(ctx, dm, n, stride, m);
    sync(ctx);
    download_planes(ctx, dv, n, stride, v);
    sync(ctx);
  });
}

int splat_stats_accumulate(splat_ctx* ctx, const splat_proj* projs, const splat_proj_grad* grads,
                           const float* abs_grads, uint64_t n, uint32_t width, uint32_t height,
                           float* grad_accum, uint32_t* denom, uint32_t* max_radii) {
  return guard(ctx, [&] {
    SPLAT_REQUIRE(n == 0 || (projs && grads && grad_accum && denom && max_radii), "null argument");
    if (n == 0) return;
    std::vector<uint32_t> radius(n);
    for (size_t i = 0; i < n; i++) radius[i] = projs[i].radius;
    uint32_t* dr = ctx->buf<uint32_t>("st_radius", n);
    splat_proj_grad* dg = ctx->buf<splat_proj_grad>("st_grads", n);
    float* dab = abs_grads ? ctx->buf<float>("st_abs", 2 * n) : nullptr;
    float* dacc = ctx->buf<float>("st_acc", n);
    uint32_t* dden = ctx->buf<uint32_t>("st_den", n);
    uint32_t* dmr = ctx->buf<uint32_t>("st_mr", n);
    h2d(ctx, dr, radius.data(), n * 4);
    h2d(ctx, dg, grads, n * sizeof(splat_proj_grad));
    if (dab) h2d(ctx, dab, abs_grads, n * 8);
    h2d(ctx, dacc, grad_accum, n * 4);
    h2d(ctx, dden, denom, n * 4);
    h2d(ctx, dmr, max_radii, n * 4);
    stats_accumulate_dev(ctx, dr, dg, dab, n, width, height, dacc, dden, dmr);
    d2h(ctx, grad_accum, dacc, n * 4);
    d2h(ctx, denom, dden, n * 4);
    d2h(ctx, max_radii, dmr, n * 4);
    sync(ctx);
  });
}

int splat_densify_and_prune(splat_ctx* ctx, const float* params, const float* m, const float* v,
                            const float* grad_accum, const uint32_t* denom,
                            const uint32_t* max_radii, uint64_t n, const splat_densify_config* cfg,
                            float extent, const float* normals, uint64_t normals_len,
                            float* params_out, float* m_out, float* v_out, uint64_t out_capacity,
                            uint64_t* n_out, splat_densify_report* report) {
  return guard(ctx, [&] {
    SPLAT_REQUIRE(cfg && n_out && report, "null argument");
    SPLAT_REQUIRE(n == 0 || (params && m && v && grad_accum && denom && max_radii),
                  "null argument");
    *report = splat_densify_report{0, 0, 0};
    *n_out = 0;
    if (n == 0) return;
    size_t stride;
    float* dp = upload_planes(ctx, "dn_p", params, n, &stride);
    float* dm = upload_planes(ctx, "dn_m", m, n, &stride);
    float* dv = upload_planes(ctx, "dn_v", v, n, &stride);
    float* dacc = ctx->buf<float>("dn_acc", n);
    uint32_t* dden = ctx->buf<uint32_t>("dn_den", n);
    uint32_t* dmr = ctx->buf<uint32_t>("dn_mr", n);
    float* dnorm = ctx->buf<float>("dn_normals", normals_len + 1);
    h2d(ctx, dacc, grad_accum, n * 4);
    h2d(ctx, dden, denom, n * 4);
    h2d(ctx, dmr, max_radii, n * 4);
    if (normals_len) h2d(ctx, dnorm, normals, normals_len * 4);
    const size_t cap_rows = n * (1 + (size_t)std::max<uint32_t>(cfg->split_children, 1));
    const size_t ostride = plane_stride(cap_rows);
    float* po = ctx->buf<float>("dn_po", ostride * kParams);
    float* mo = ctx->buf<float>("dn_mo", ostride * kParams);
    float* vo = ctx->buf<float>("dn_vo", ostride * kParams);
    const size_t k = densify_planes(ctx, dp, dm, dv, n, stride, dacc, dden, dmr, *cfg, extent,
                                    dnorm, normals_len, 0, po, mo, vo, cap_rows, ostride, report);
    if (k == (size_t)-1) fail(SPLAT_ERR_CAPACITY, "not enough normal draws for the split children");
    *n_out = k;
    if (k > out_capacity) fail(SPLAT_ERR_CAPACITY, "output capacity too small");
    if (k == 0) return;
    download_planes(ctx, po, k, ostride, params_out);
    sync(ctx);
    download_planes(ctx, mo, k, ostride, m_out);
    sync(ctx);
    download_planes(ctx, vo, k, ostride, v_out);
    sync(ctx);
  });
}

int splat_reset_opacity(splat_ctx* ctx, float* params, float* m, float* v, uint64_t n,
                        float max_alpha) {
  return guard(ctx, [&] {
    SPLAT_REQUIRE(n == 0 || (params && m && v), "null argument");
    if (n == 0) return;
    const float limit = logf(max_alpha / (1.0f - max_alpha));  // inverse_sigmoid, math.rs:395
    size_t stride;
    float* dp = upload_planes(ctx, "ro_p", params, n, &stride);
    float* dm = upload_planes(ctx, "ro_m", m, n, &stride);
    float* dv = upload_planes(ctx, "ro_v", v, n, &stride);
    reset_opacity_planes(ctx, dp, dm, dv, n, stride, limit);
    download_planes(ctx, dp, n, stride, params);
    sync(ctx);
    download_planes(ctx, dm, n, stride, m);
    sync(ctx);
    download_planes(ctx, dv, n, stride, v);
    sync(ctx);
  });
}

}  // extern "C"
