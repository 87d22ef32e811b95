// C ABI of the learned densification scorer: create / weights / forward / gradient / fit.
// The kernels are in mlp.hip; the trainer drives the same device functions for its online
// training and learned densification.
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "mlp.hpp"
#include "scorer.hpp"

using namespace splat;

namespace splat {

namespace {
uint64_t sm64(uint64_t& s) {  // SplitMix64
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
double uniform(uint64_t& s) { return (double)(sm64(s) >> 11) * 0x1.0p-53; }
}  // namespace

void scorer_upload(splat_scorer* s, const float* host_params) {
  splat_ctx* ctx = s->ctx;
  SPLAT_HIP(hipMemcpyAsync(s->params.ptr, host_params, kMlpParams * 4, hipMemcpyHostToDevice,
                           ctx->stream));
  fill_u32(ctx, s->m.as<uint32_t>(), kMlpParams, 0u);
  fill_u32(ctx, s->v.as<uint32_t>(), kMlpParams, 0u);
  s->adam_step = 0;
  mlp_pack(ctx, s->params.as<float>(), s->packed.ptr);
  SPLAT_HIP(hipStreamSynchronize(ctx->stream));
}

float scorer_fit_device(splat_scorer* s, const float* feats, const uint8_t* labels,
                        const float* weights, size_t n, const splat_scorer_fit_config& cfg) {
  splat_ctx* ctx = s->ctx;
  SPLAT_REQUIRE(cfg.batch > 0 && cfg.epochs > 0, "batch and epochs must be positive");
  if (n == 0) return 0.0f;
  const size_t batch = std::min<size_t>(cfg.batch, n);
  const unsigned rows = mlp_grad_rows(batch);
  float* partial = (float*)s->scratch["partial"].reserve((size_t)rows * kMlpParams * 4);
  float* ploss = (float*)s->scratch["ploss"].reserve((size_t)rows * 4 + 4);
  float* losses = (float*)s->scratch["losses"].reserve(4 * (size_t)(n / batch + 2));
  uint32_t* order = (uint32_t*)s->scratch["order"].reserve(n * 4);
  std::vector<uint32_t> perm(n);
  uint64_t rs = cfg.seed;
  float last = 0.0f;
  for (uint32_t ep = 0; ep < cfg.epochs; ep++) {
    // per-epoch Fisher-Yates permutation of the samples
    for (size_t i = 0; i < n; i++) perm[i] = (uint32_t)i;
    for (size_t i = n; i-- > 1;) std::swap(perm[i], perm[(size_t)(sm64(rs) % (i + 1))]);
    SPLAT_HIP(hipMemcpyAsync(order, perm.data(), n * 4, hipMemcpyHostToDevice, ctx->stream));
    size_t nb = 0;
    for (size_t b0 = 0; b0 + batch <= n; b0 += batch, nb++) {
      mlp_grad(ctx, s->packed.ptr, s->params.as<float>(), feats, labels, weights, order + b0,
               batch, 1.0f / (float)batch, partial, ploss);
      s->adam_step++;
      const float c1 = 1.0f - powi(0.9f, (int32_t)s->adam_step);
      const float c2 = 1.0f - powi(0.999f, (int32_t)s->adam_step);
      mlp_reduce_update(ctx, s->params.as<float>(), s->m.as<float>(), s->v.as<float>(), partial,
                        rows, nullptr, true, cfg.lr, c1, c2);
      mlp_pack(ctx, s->params.as<float>(), s->packed.ptr);
      mlp_loss_sum(ctx, ploss, rows, losses + nb);
    }
    std::vector<float> lv(nb);
    SPLAT_HIP(hipMemcpyAsync(lv.data(), losses, nb * 4, hipMemcpyDeviceToHost, ctx->stream));
    SPLAT_HIP(hipStreamSynchronize(ctx->stream));
    double sum = 0;
    for (float x : lv) sum += x;
    last = nb ? (float)(sum / (double)nb) : 0.0f;
  }
  return last;
}

}  // namespace splat

extern "C" {

int splat_scorer_create(splat_ctx* ctx, const float* params, uint64_t seed, splat_scorer** out) {
  if (!ctx || !out) return SPLAT_ERR_INVALID;
  *out = nullptr;
  splat_scorer* s = new (std::nothrow) splat_scorer();
  if (!s) return SPLAT_ERR_OOM;
  s->ctx = ctx;
  s->device = ctx->device;
  const int st = ctx_guard(ctx, [&] {
    s->params.reserve(kMlpParams * 4);
    s->m.reserve(kMlpParams * 4);
    s->v.reserve(kMlpParams * 4);
    s->packed.reserve(mlp_packed_bytes());
    std::vector<float> init(kMlpParams, 0.0f);
    if (params) {
      std::memcpy(init.data(), params, kMlpParams * 4);
    } else {
      // weights ~ N(0, 1/fan_in) (Box-Muller on SplitMix64), biases 0
      uint64_t rs = seed;
      auto fill = [&](int off, int nout, int nin) {
        const double sd = 1.0 / std::sqrt((double)nin);
        for (int k = 0; k < nout * nin; k++) {
          const double u1 = std::max(uniform(rs), 1e-12), u2 = uniform(rs);
          init[off + k] =
              (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
        }
      };
      fill(0, 64, 12);
      fill(832, 64, 64);
      fill(4992, 64, 64);
      fill(9152, 3, 64);
    }
    scorer_upload(s, init.data());
  });
  if (st != SPLAT_OK) {
    delete s;
    return st;
  }
  *out = s;
  return SPLAT_OK;
}

int splat_scorer_destroy(splat_scorer* s) {
  if (!s) return SPLAT_OK;
  (void)hipSetDevice(s->device);
  (void)hipDeviceSynchronize();
  delete s;
  return SPLAT_OK;
}

int splat_scorer_get_params(splat_scorer* s, float* params) {
  if (!s) return SPLAT_ERR_INVALID;
  return ctx_guard(s->ctx, [&] {
    SPLAT_REQUIRE(params, "null argument");
    SPLAT_HIP(hipMemcpyAsync(params, s->params.ptr, kMlpParams * 4, hipMemcpyDeviceToHost,
                             s->ctx->stream));
    SPLAT_HIP(hipStreamSynchronize(s->ctx->stream));
  });
}

int splat_scorer_set_params(splat_scorer* s, const float* params) {
  if (!s) return SPLAT_ERR_INVALID;
  return ctx_guard(s->ctx, [&] {
    SPLAT_REQUIRE(params, "null argument");
    scorer_upload(s, params);
  });
}

int splat_scorer_forward(splat_scorer* s, const float* features, uint64_t n, float* logits) {
  if (!s) return SPLAT_ERR_INVALID;
  return ctx_guard(s->ctx, [&] {
    SPLAT_REQUIRE(n == 0 || (features && logits), "null argument");
    if (n == 0) return;
    splat_ctx* ctx = s->ctx;
    float* f = (float*)s->scratch["feats"].reserve(n * kMlpFeatures * 4);
    float* z = (float*)s->scratch["logits"].reserve(n * 3 * 4);
    SPLAT_HIP(hipMemcpyAsync(f, features, n * kMlpFeatures * 4, hipMemcpyHostToDevice,
                             ctx->stream));
    mlp_forward(ctx, s->packed.ptr, s->params.as<float>(), f, n, z);
    SPLAT_HIP(hipMemcpyAsync(logits, z, n * 3 * 4, hipMemcpyDeviceToHost, ctx->stream));
    SPLAT_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int splat_scorer_loss_grad(splat_scorer* s, const float* features, const uint8_t* labels,
                           const float* weights, uint64_t n, float* grad, float* loss) {
  if (!s) return SPLAT_ERR_INVALID;
  return ctx_guard(s->ctx, [&] {
    SPLAT_REQUIRE(n > 0 && features && labels && weights && grad && loss, "bad argument");
    splat_ctx* ctx = s->ctx;
    float* f = (float*)s->scratch["feats"].reserve(n * kMlpFeatures * 4);
    uint8_t* y = (uint8_t*)s->scratch["labels"].reserve(n);
    float* w = (float*)s->scratch["weights"].reserve(n * 4);
    const unsigned rows = mlp_grad_rows(n);
    float* partial = (float*)s->scratch["partial"].reserve((size_t)rows * kMlpParams * 4);
    float* ploss = (float*)s->scratch["ploss"].reserve((size_t)rows * 4 + 4);
    float* g = (float*)s->scratch["grad"].reserve(kMlpParams * 4 + 4);
    SPLAT_HIP(hipMemcpyAsync(f, features, n * kMlpFeatures * 4, hipMemcpyHostToDevice,
                             ctx->stream));
    SPLAT_HIP(hipMemcpyAsync(y, labels, n, hipMemcpyHostToDevice, ctx->stream));
    SPLAT_HIP(hipMemcpyAsync(w, weights, n * 4, hipMemcpyHostToDevice, ctx->stream));
    mlp_grad(ctx, s->packed.ptr, s->params.as<float>(), f, y, w, nullptr, n, 1.0f / (float)n,
             partial, ploss);
    mlp_reduce_update(ctx, s->params.as<float>(), s->m.as<float>(), s->v.as<float>(), partial,
                      rows, g, false, 0.0f, 1.0f, 1.0f);
    mlp_loss_sum(ctx, ploss, rows, g + kMlpParams);
    SPLAT_HIP(hipMemcpyAsync(grad, g, kMlpParams * 4, hipMemcpyDeviceToHost, ctx->stream));
    SPLAT_HIP(hipMemcpyAsync(loss, g + kMlpParams, 4, hipMemcpyDeviceToHost, ctx->stream));
    SPLAT_HIP(hipStreamSynchronize(ctx->stream));
  });
}

void splat_scorer_fit_config_default(splat_scorer_fit_config* c) {
  if (!c) return;
  c->epochs = 4;
  c->batch = 4096;
  c->lr = 1e-3f;
  c->pad_ = 0;
  c->seed = 0;
}

int splat_scorer_fit(splat_scorer* s, const float* features, const uint8_t* labels,
                     const float* weights, uint64_t n, const splat_scorer_fit_config* cfg,
                     float* loss) {
  if (!s) return SPLAT_ERR_INVALID;
  return ctx_guard(s->ctx, [&] {
    SPLAT_REQUIRE(cfg && features && labels && weights, "null argument");
    splat_ctx* ctx = s->ctx;
    float* f = (float*)s->scratch["fit_feats"].reserve(n * kMlpFeatures * 4 + 4);
    uint8_t* y = (uint8_t*)s->scratch["fit_labels"].reserve(n + 4);
    float* w = (float*)s->scratch["fit_weights"].reserve(n * 4 + 4);
    SPLAT_HIP(hipMemcpyAsync(f, features, n * kMlpFeatures * 4, hipMemcpyHostToDevice,
                             ctx->stream));
    SPLAT_HIP(hipMemcpyAsync(y, labels, n, hipMemcpyHostToDevice, ctx->stream));
    SPLAT_HIP(hipMemcpyAsync(w, weights, n * 4, hipMemcpyHostToDevice, ctx->stream));
    const float l = scorer_fit_device(s, f, y, w, n, *cfg);
    if (loss) *loss = l;
  });
}

}  // extern "C"
