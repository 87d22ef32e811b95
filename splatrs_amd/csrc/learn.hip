// Learned densification inside the trainer (no counterpart in the reference, whose
// densification is the threshold heuristic of src/densify.rs:125-221, kept here as the
// fallback and as the teacher).
//
// Per Gaussian the trainer can track, next to the reference's statistics, the largest
// per-step gradient norm, the number of steps whose norm reached the threshold, the birth
// iteration and a lineage tag.  At each densification event:
//   * 12 features are computed from the model and these statistics;
//   * mode 1 (learn): the reference's decisions are applied unchanged, and a sample of them
//     is logged (features, decision bits, parent tag).  Once `outcome_delay` iterations have
//     passed, each logged clone/split is labelled by whether any descendant of it is still
//     alive with an opacity of at least useful_ratio times its parent's at the decision (and
//     useful_opacity): a densification whose offspring were pruned or faded is labelled
//     negative; or (credit_views > 0) by its credit, the share of the images of
//     credit_views evenly spaced views (the sum of alpha T over their pixels) that the
//     Gaussians it added hold at the outcome, against the median Gaussian's share;
//     samples are weighted by the PSNR change over the same window, and the scorer is fit
//     on every labelled sample so far;
//   * mode 2 (learned): the scorer's logits decide clone / split / prune; the reference's
//     prune rules still apply on top (fallback).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "mlp.hpp"
#include "scorer.hpp"
#include "trainer.hpp"

using namespace splat;

struct LearnEvent {
  uint32_t iteration, base, count;
  float psnr;
};

struct LearnState {
  splat_learn_config cfg;
  size_t capacity = 0, used = 0, resolved = 0, next_event = 0;
  DevBuf feats, dec, parent, label, weight, alive, counter;
  std::vector<LearnEvent> events;
  uint64_t updates = 0, positives = 0;
  float last_loss = 0.0f;
};

splat_trainer::splat_trainer() = default;
splat_trainer::~splat_trainer() = default;

namespace splat {
namespace {

constexpr float kShC0 = 0.28209479177387814f;  // src/sh.rs:16

struct FeatParams {
  float thr;          // densify gradient threshold
  float log_size;     // ln(percent_dense * extent), the clone/split size boundary
  uint32_t iteration;
};

// The 12 scorer features of Gaussian i.
__global__ void __launch_bounds__(256) features_kernel(
    const float* __restrict__ params, uint32_t n, uint32_t stride,
    const float* __restrict__ grad_accum, const uint32_t* __restrict__ denom,
    const uint32_t* __restrict__ max_radii, const float* __restrict__ max_grad,
    const uint32_t* __restrict__ strong, const uint32_t* __restrict__ birth, FeatParams fp,
    float* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t d = denom[i];
  const float mean = d > 0 ? grad_accum[i] / (float)d : 0.0f;
  const float l0 = params[3 * (size_t)stride + i], l1 = params[4 * (size_t)stride + i],
              l2 = params[5 * (size_t)stride + i];
  float f[kMlpFeatures];
  f[0] = log2f(mean / fp.thr + 1e-6f) * 0.125f;
  f[1] = log2f(max_grad[i] / fp.thr + 1e-6f) * 0.125f;
  f[2] = (float)strong[i] / (float)max(d, 1u);
  f[3] = (l0 - fp.log_size) * 0.25f;
  f[4] = (l1 - fp.log_size) * 0.25f;
  f[5] = (l2 - fp.log_size) * 0.25f;
  f[6] = 1.0f / (1.0f + splat_expf(-params[10 * (size_t)stride + i]));
  f[7] = log2f(1.0f + (float)max_radii[i]) * 0.125f;
  f[8] = log2f(1.0f + (float)d) * 0.125f;
  const float c0 = kShC0 * params[11 * (size_t)stride + i] + 0.5f;
  const float c1 = kShC0 * params[12 * (size_t)stride + i] + 0.5f;
  const float c2 = kShC0 * params[13 * (size_t)stride + i] + 0.5f;
  f[9] = 0.299f * c0 + 0.587f * c1 + 0.114f * c2;
  f[10] = log2f(1.0f + (float)(fp.iteration - birth[i])) / 12.0f;
  f[11] = (rmax(rmax(l0, l1), l2) - fp.log_size) * 0.25f;
#pragma unroll
  for (int k = 0; k < kMlpFeatures; k++) out[(size_t)i * kMlpFeatures + k] = f[k];
}

__device__ __forceinline__ uint32_t hash32(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return (uint32_t)((x ^ (x >> 31)) >> 32);
}

// Sample selection: a densified row is kept with probability pos_keep / 2^32, any other row
// with neg_keep / 2^32 (counter-based, deterministic).
__global__ void __launch_bounds__(256) select_kernel(const uint8_t* __restrict__ dec, uint32_t n,
                                                     uint64_t key, uint64_t pos_keep,
                                                     uint64_t neg_keep,
                                                     uint32_t* __restrict__ flag) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool densified = (dec[i] & (kDecClone | kDecSplit)) != 0;
  const uint64_t h = hash32(key ^ ((uint64_t)i * 0xd1b54a32d192ed03ull));
  flag[i] = h < (densified ? pos_keep : neg_keep) ? 1u : 0u;
}

__global__ void __launch_bounds__(256) log_kernel(
    const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos, uint32_t n,
    uint32_t budget, uint32_t base, const float* __restrict__ feats,
    const uint8_t* __restrict__ dec, const uint32_t* __restrict__ origin,
    float* __restrict__ s_feats, uint8_t* __restrict__ s_dec, uint32_t* __restrict__ s_parent,
    uint32_t* __restrict__ sample_of) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t id = kNoSample;
  if (flag[i] && pos[i] < budget) {
    id = base + pos[i];
#pragma unroll
    for (int k = 0; k < kMlpFeatures; k++)
      s_feats[(size_t)id * kMlpFeatures + k] = feats[(size_t)i * kMlpFeatures + k];
    s_dec[id] = dec[i];
    s_parent[id] = origin[i];
  }
  sample_of[i] = id;
}

// Useful descendants of the samples [b, b + c): every live row walks up its lineage
// (sample ids decrease towards older decisions) and credits the samples of this event it
// meets when its opacity is at least the larger of min_opacity and `ratio` times the
// opacity the logged Gaussian had when it was densified (feature 6) -- a clone or child the
// optimiser faded well below its parent did not earn its place.  any_survivor credits every live descendant
// (used when an opacity reset fell inside the window).
__global__ void __launch_bounds__(256) alive_kernel(const uint32_t* __restrict__ origin,
                                                    uint32_t n,
                                                    const float* __restrict__ opacity_logit,
                                                    float min_opacity, float ratio,
                                                    int any_survivor,
                                                    const float* __restrict__ s_feats,
                                                    const uint32_t* __restrict__ s_parent,
                                                    uint32_t b, uint32_t c,
                                                    uint32_t* __restrict__ alive) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float op = 1.0f / (1.0f + splat_expf(-opacity_logit[i]));
  uint32_t s = origin[i];
  for (int depth = 0; depth < 64 && s != kNoSample && s >= b; depth++) {
    if (s < b + c &&
        (any_survivor || op >= rmax(min_opacity, ratio * s_feats[(size_t)s * kMlpFeatures + 6])))
      atomicAdd(&alive[s - b], 1u);
    s = s_parent[s];
  }
}

// Credits of the samples [b, b + c): every live row walks up its lineage and adds its image
// share to each sample of this event it meets, in fixed point (2^-16 pixel units), so the
// sums do not depend on the order of the atomics.
__global__ void __launch_bounds__(256) credit_kernel(const uint32_t* __restrict__ origin,
                                                     uint32_t n, const float* __restrict__ share,
                                                     const uint32_t* __restrict__ s_parent,
                                                     uint32_t b, uint32_t c,
                                                     unsigned long long* __restrict__ credit) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float sh = share[i];
  if (!(sh > 0.0f)) return;
  const unsigned long long q = (unsigned long long)(rmin(sh, 1e12f) * 65536.0f);
  uint32_t s = origin[i];
  for (int depth = 0; depth < 64 && s != kNoSample && s >= b; depth++) {
    if (s < b + c) atomicAdd(&credit[s - b], q);
    s = s_parent[s];
  }
}

// alive[k] = 1 where the credit reaches `need` (fixed point), the input of label_kernel.
__global__ void __launch_bounds__(256) credit_alive_kernel(
    const unsigned long long* __restrict__ credit, uint32_t c, unsigned long long need,
    uint32_t* __restrict__ alive) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= c) return;
  alive[k] = credit[k] >= need ? 1u : 0u;
}

// Labels: clone / split only if the reference densified the row and a descendant is
// alive; prune where the reference pruned or the row's image share was low when logged.
// Densified samples carry the event's weight.
__global__ void __launch_bounds__(256) label_kernel(const uint8_t* __restrict__ s_dec,
                                                    const uint32_t* __restrict__ alive,
                                                    uint32_t b, uint32_t c, float wpos,
                                                    uint8_t* __restrict__ s_label,
                                                    float* __restrict__ s_weight,
                                                    uint32_t* __restrict__ positives) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= c) return;
  const uint32_t d = s_dec[b + k];
  const bool ok = alive[k] > 0;
  uint32_t y = 0;
  if ((d & kDecClone) && ok) y |= 1u;
  if ((d & kDecSplit) && ok) y |= 2u;
  if (d & (kDecPrune | kDecLowShare)) y |= 4u;
  s_label[b + k] = (uint8_t)y;
  s_weight[b + k] = (d & (kDecClone | kDecSplit)) ? wpos : 1.0f;
  if (y & 3u) atomicAdd(positives, 1u);
}

// Rows whose image share is below `bar` are pruned (and neither cloned nor split), or
// (mark_only) tagged kDecLowShare.
__global__ void __launch_bounds__(256) share_prune_kernel(const float* __restrict__ share,
                                                          uint32_t n, float bar, int mark_only,
                                                          uint8_t* __restrict__ dec) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !(share[i] < bar)) return;
  dec[i] = mark_only ? (uint8_t)(dec[i] | kDecLowShare)
                     : (uint8_t)((dec[i] & ~(kDecClone | kDecSplit)) | kDecPrune);
}

// Learned decisions from the logits; the reference's prune rules stay in force.
// (The prune head is applied only where prune_head is set.)
__global__ void __launch_bounds__(256) learned_decide_kernel(const float* __restrict__ logits,
                                                             const uint8_t* __restrict__ heur,
                                                             uint32_t n, float t_dens,
                                                             float t_prune, int prune_head,
                                                             uint8_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float zc = logits[3 * (size_t)i], zs = logits[3 * (size_t)i + 1],
              zp = logits[3 * (size_t)i + 2];
  uint32_t d = heur[i] & (kDecPrune | kDecPruneCopy);
  if (zc > t_dens || zs > t_dens) d |= zc >= zs ? kDecClone : kDecSplit;
  if (prune_head && zp > t_prune) d |= kDecPrune | kDecPruneCopy;
  out[i] = (uint8_t)d;
}

float logit(float p) { return logf(p / (1.0f - p)); }

}  // namespace

void learn_track_alloc(splat_trainer* t) {
  splat_ctx* ctx = t->ctx;
  const size_t c = t->cap + 1;
  t->max_grad.reserve(c * 4);
  t->strong.reserve(c * 4);
  t->birth.reserve(c * 4);
  t->origin.reserve(c * 4);
  fill_u32(ctx, t->birth.as<uint32_t>(), t->n, 0u);
  fill_u32(ctx, t->origin.as<uint32_t>(), t->n, kNoSample);
  t->track = true;
  learn_zero_stats(t);
  SPLAT_HIP(hipStreamSynchronize(ctx->stream));
}

void learn_zero_stats(splat_trainer* t) {
  fill_u32(t->ctx, t->max_grad.as<uint32_t>(), t->n, 0u);
  fill_u32(t->ctx, t->strong.as<uint32_t>(), t->n, 0u);
}

namespace {
void compute_features(splat_trainer* t, float extent, float thr, float percent_dense,
                      uint32_t iteration, float* out) {
  FeatParams fp;
  fp.thr = thr;
  fp.log_size = logf(percent_dense * extent);
  fp.iteration = iteration;
  launch(t->ctx, "learn_features", features_kernel, dim3(blocks_for(t->n, 256)), dim3(256), 0,
         (const float*)t->params.as<float>(), (uint32_t)t->n, (uint32_t)t->stride,
         (const float*)t->grad_accum.as<float>(), (const uint32_t*)t->denom.as<uint32_t>(),
         (const uint32_t*)t->max_radii.as<uint32_t>(), (const float*)t->max_grad.as<float>(),
         (const uint32_t*)t->strong.as<uint32_t>(), (const uint32_t*)t->birth.as<uint32_t>(), fp,
         out);
}

uint32_t scan_total(splat_trainer* t, const uint32_t* flag, uint32_t* pos, size_t n) {
  uint32_t* total = (uint32_t*)t->store["learn_total"].reserve(16);
  exclusive_scan_u32(t->ctx, flag, pos, n, total);
  uint32_t h = 0;
  SPLAT_HIP(hipMemcpyAsync(&h, total, 4, hipMemcpyDeviceToHost, t->ctx->stream));
  SPLAT_HIP(hipStreamSynchronize(t->ctx->stream));
  return h;
}
}  // namespace

namespace {
// Image share of every row over all views, and `fraction` times its median.
float all_view_share(splat_trainer* t, int32_t sh_degree, float fraction, float* share) {
  const size_t n = t->n;
  const uint32_t V = (uint32_t)t->views.size();
  std::vector<uint32_t> list(V);
  for (uint32_t k = 0; k < V; k++) list[k] = k;
  trainer_image_share(t, list.data(), V, sh_degree, share);
  std::vector<float> h(n);
  SPLAT_HIP(hipMemcpyAsync(h.data(), share, n * 4, hipMemcpyDeviceToHost, t->ctx->stream));
  SPLAT_HIP(hipStreamSynchronize(t->ctx->stream));
  std::nth_element(h.begin(), h.begin() + n / 2, h.end());
  return fraction * h[n / 2];
}
}  // namespace

void learn_densify(splat_trainer* t, const splat_densify_config& cfg, float extent,
                   uint64_t seed, uint32_t iteration, int32_t sh_degree, const float bg[3],
                   uint32_t last_reset, splat_densify_report* report) {
  trainer_drain(t);
  splat_ctx* ctx = t->ctx;
  LearnState& L = *t->learn;
  const size_t n = t->n;
  if (n == 0) {
    *report = splat_densify_report{0, 0, 0};
    return;
  }
  uint8_t* heur = (uint8_t*)t->store["learn_heur"].reserve(n + 4);
  densify_decide(ctx, t->params.as<float>(), n, t->stride, t->grad_accum.as<float>(),
                 t->denom.as<uint32_t>(), t->max_radii.as<uint32_t>(), cfg, extent, heur);
  float* feats = (float*)t->store["learn_feats"].reserve(n * kMlpFeatures * 4);
  compute_features(t, extent, cfg.grad_threshold, cfg.percent_dense, iteration, feats);
  const uint8_t* decide = heur;
  uint32_t* sample_of = nullptr;
  const bool learn_mode = L.cfg.mode == 1;
  const bool due = learn_mode && L.next_event < L.events.size() &&
                   iteration >= L.events[L.next_event].iteration + L.cfg.outcome_delay;
  float psnr = 0.0f;
  if (learn_mode && (L.used < L.capacity || due))
    psnr = trainer_eval_psnr(t, L.cfg.eval_views, sh_degree, bg);
  if (learn_mode && L.used < L.capacity && L.cfg.share_prune > 0.0f &&
      iteration >= L.cfg.share_prune_from) {  // prune-head labels: low image share now
    float* share = (float*)t->store["learn_share"].reserve(n * 4 + 4);
    const float bar = all_view_share(t, sh_degree, L.cfg.share_prune, share);
    launch(ctx, "learn_low_share", share_prune_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
           (const float*)share, (uint32_t)n, bar, 1, heur);
  }
  if (learn_mode && L.used < L.capacity) {
    uint32_t* flag = (uint32_t*)t->store["learn_flag"].reserve((n + 1) * 4);
    uint32_t* pos = (uint32_t*)t->store["learn_pos"].reserve((n + 1) * 4);
    const uint64_t key = seed ^ ((uint64_t)iteration << 32);
    launch(ctx, "learn_select", select_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
           (const uint8_t*)heur, (uint32_t)n, key, 1ull << 32, 0ull, flag);
    const uint32_t npos = scan_total(t, flag, pos, n);
    const size_t budget = std::min<size_t>(L.cfg.samples_per_event, L.capacity - L.used);
    const double half = 0.5 * (double)budget;
    const double pk = npos ? std::min(1.0, half / npos) : 1.0;
    const double rest = (double)budget - std::min<double>(npos, half);
    const double nk = n > npos ? std::min(1.0, rest / (double)(n - npos)) : 0.0;
    launch(ctx, "learn_select", select_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
           (const uint8_t*)heur, (uint32_t)n, key, (uint64_t)(pk * 4294967296.0),
           (uint64_t)(nk * 4294967296.0), flag);
    const uint32_t sel = scan_total(t, flag, pos, n);
    const uint32_t count = (uint32_t)std::min<size_t>(sel, budget);
    sample_of = (uint32_t*)t->store["learn_sample_of"].reserve((n + 1) * 4);
    launch(ctx, "learn_log", log_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
           (const uint32_t*)flag, (const uint32_t*)pos, (uint32_t)n, count, (uint32_t)L.used,
           (const float*)feats, (const uint8_t*)heur, (const uint32_t*)t->origin.as<uint32_t>(),
           L.feats.as<float>(), L.dec.as<uint8_t>(), L.parent.as<uint32_t>(), sample_of);
    L.events.push_back(LearnEvent{iteration, (uint32_t)L.used, count, psnr});
    L.used += count;
  } else if (L.cfg.mode == 2) {
    float* logits = (float*)t->store["learn_logits"].reserve(n * 3 * 4);
    mlp_forward(ctx, t->scorer->packed.ptr, t->scorer->params.as<float>(), feats, n, logits);
    uint8_t* dec = (uint8_t*)t->store["learn_dec"].reserve(n + 4);
    launch(ctx, "learn_decide", learned_decide_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
           (const float*)logits, (const uint8_t*)heur, (uint32_t)n, logit(L.cfg.threshold),
           logit(L.cfg.prune_threshold), iteration >= L.cfg.share_prune_from ? 1 : 0, dec);
    if (L.cfg.share_prune > 0.0f && iteration >= L.cfg.share_prune_from) {  // measured rule
      float* share = (float*)t->store["learn_share"].reserve(n * 4 + 4);
      const float bar = all_view_share(t, sh_degree, L.cfg.share_prune, share);
      launch(ctx, "learn_share_prune", share_prune_kernel, dim3(blocks_for(n, 256)), dim3(256),
             0, (const float*)share, (uint32_t)n, bar, 0, dec);
    }
    decide = dec;
  }
  // label the events whose outcome window has passed, then refit on all labels so far
  bool fresh = false;
  const auto resolve_due = [&] {
    std::vector<unsigned long long> need_q;  // credit labels: the usefulness bar
    float* share = nullptr;
    while (L.next_event < L.events.size() &&
           iteration >= L.events[L.next_event].iteration + L.cfg.outcome_delay) {
      const LearnEvent& e = L.events[L.next_event];
      if (e.count) {
        uint32_t* alive = (uint32_t*)L.alive.reserve((size_t)e.count * 4 + 4);
        if (L.cfg.credit_views > 0) {
          if (!share) {  // the model's image shares, once per call
            const uint32_t V = (uint32_t)t->views.size();
            const uint32_t cv = std::min<uint32_t>(L.cfg.credit_views, V);
            std::vector<uint32_t> list(cv);
            for (uint32_t k = 0; k < cv; k++) list[k] = (uint32_t)(((uint64_t)k * V) / cv);
            share = (float*)t->store["learn_share"].reserve(t->n * 4 + 4);
            trainer_image_share(t, list.data(), cv, sh_degree, share);
            std::vector<float> h(t->n);
            SPLAT_HIP(hipMemcpyAsync(h.data(), share, t->n * 4, hipMemcpyDeviceToHost,
                                     ctx->stream));
            SPLAT_HIP(hipStreamSynchronize(ctx->stream));
            std::nth_element(h.begin(), h.begin() + h.size() / 2, h.end());
            const double med = h.empty() ? 0.0 : (double)h[h.size() / 2];
            need_q.push_back((unsigned long long)(std::max(L.cfg.credit_ratio * med, 1e-6) *
                                                  65536.0) + 1ull);
          }
          unsigned long long* credit =
              (unsigned long long*)t->store["learn_credit"].reserve((size_t)e.count * 8 + 8);
          fill_u32(ctx, reinterpret_cast<uint32_t*>(credit), (size_t)e.count * 2, 0u);
          launch(ctx, "learn_credit", credit_kernel, dim3(blocks_for(t->n, 256)), dim3(256), 0,
                 (const uint32_t*)t->origin.as<uint32_t>(), (uint32_t)t->n, (const float*)share,
                 (const uint32_t*)L.parent.as<uint32_t>(), e.base, e.count, credit);
          launch(ctx, "learn_credit_alive", credit_alive_kernel, dim3(blocks_for(e.count, 256)),
                 dim3(256), 0, (const unsigned long long*)credit, e.count, need_q[0], alive);
        } else {
          fill_u32(ctx, alive, e.count, 0u);
          // a descendant counts while its opacity reached its parent's at the decision (and
          // useful_opacity); after an opacity reset inside the window every survivor counts
          const int reset_inside = last_reset > e.iteration ? 1 : 0;
          launch(ctx, "learn_alive", alive_kernel, dim3(blocks_for(t->n, 256)), dim3(256), 0,
                 (const uint32_t*)t->origin.as<uint32_t>(), (uint32_t)t->n,
                 (const float*)(t->params.as<float>() + 10 * t->stride), L.cfg.useful_opacity,
                 L.cfg.useful_ratio, reset_inside, (const float*)L.feats.as<float>(),
                 (const uint32_t*)L.parent.as<uint32_t>(), e.base, e.count, alive);
        }
        const float wpos = rclamp(1.0f + (psnr - e.psnr), 0.25f, 4.0f);
        uint32_t* cnt = (uint32_t*)L.counter.reserve(16);
        fill_u32(ctx, cnt, 1, 0u);
        launch(ctx, "learn_label", label_kernel, dim3(blocks_for(e.count, 256)), dim3(256), 0,
               (const uint8_t*)L.dec.as<uint8_t>(), (const uint32_t*)alive, e.base, e.count,
               wpos, L.label.as<uint8_t>(), L.weight.as<float>(), cnt);
        uint32_t h = 0;
        SPLAT_HIP(hipMemcpyAsync(&h, cnt, 4, hipMemcpyDeviceToHost, ctx->stream));
        SPLAT_HIP(hipStreamSynchronize(ctx->stream));
        L.positives += h;
        fresh = true;
      }
      L.resolved = e.base + e.count;
      L.next_event++;
    }
  };
  // credits are taken on the model as trained through the window, before this event's
  // densification; survival on the model after it
  if (learn_mode && L.cfg.credit_views > 0) resolve_due();
  trainer_densify(t, cfg, extent, seed, decide, sample_of, iteration, report);
  if (!learn_mode) return;
  if (L.cfg.credit_views == 0) resolve_due();
  if (fresh && L.resolved >= L.cfg.fit.batch) {
    L.last_loss = scorer_fit_device(t->scorer, L.feats.as<float>(), L.label.as<uint8_t>(),
                                    L.weight.as<float>(), L.resolved, L.cfg.fit);
    L.updates++;
  }
}

}  // namespace splat

extern "C" {

void splat_learn_config_default(splat_learn_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->mode = 1;
  c->samples_per_event = 16384;
  c->max_events = 64;
  c->outcome_delay = 500;
  c->threshold = 0.5f;
  c->prune_threshold = 0.5f;
  c->eval_views = 8;
  c->useful_opacity = 0.05f;
  c->useful_ratio = 0.5f;
  c->credit_views = 0;
  c->credit_ratio = 1.0f;
  c->share_prune = 0.0f;
  c->share_prune_from = 0;
  splat_scorer_fit_config_default(&c->fit);
}

int splat_trainer_attach_scorer(splat_trainer* t, splat_scorer* s, const splat_learn_config* cfg) {
  if (!t || !t->ctx) return SPLAT_ERR_INVALID;
  return ctx_guard(t->ctx, [&] {
    trainer_drain(t);
    if (!s) {
      t->scorer = nullptr;
      t->learn.reset();
      return;
    }
    SPLAT_REQUIRE(cfg && (cfg->mode == 1 || cfg->mode == 2), "mode must be 1 (learn) or 2 (learned)");
    SPLAT_REQUIRE(cfg->threshold > 0.0f && cfg->threshold < 1.0f && cfg->prune_threshold > 0.0f &&
                      cfg->prune_threshold < 1.0f && cfg->useful_opacity > 0.0f &&
                      cfg->useful_opacity < 1.0f && cfg->useful_ratio >= 0.0f,
                  "thresholds must lie in (0, 1)");
    SPLAT_REQUIRE(cfg->credit_views == 0 || cfg->credit_ratio > 0.0f,
                  "credit_ratio must be positive");
    SPLAT_REQUIRE(s->ctx == t->ctx, "scorer and trainer must share a context");
    auto L = std::make_unique<LearnState>();
    L->cfg = *cfg;
    if (cfg->mode == 1) {
      L->capacity = (size_t)cfg->samples_per_event * cfg->max_events;
      const size_t c = L->capacity + 1;
      L->feats.reserve(c * kMlpFeatures * 4);
      L->dec.reserve(c);
      L->parent.reserve(c * 4);
      L->label.reserve(c);
      L->weight.reserve(c * 4);
    }
    if (!t->track) learn_track_alloc(t);
    t->scorer = s;
    t->learn = std::move(L);
  });
}

int splat_trainer_learn_stats(splat_trainer* t, splat_learn_stats* out) {
  if (!t || !out) return SPLAT_ERR_INVALID;
  std::memset(out, 0, sizeof(*out));
  if (!t->learn) return SPLAT_OK;
  out->logged = t->learn->used;
  out->resolved = t->learn->resolved;
  out->updates = t->learn->updates;
  out->positives = t->learn->positives;
  out->events = t->learn->events.size();
  out->last_loss = t->learn->last_loss;
  return SPLAT_OK;
}

int splat_trainer_features(splat_trainer* t, float extent, float grad_threshold,
                           float percent_dense, uint32_t iteration, float* out) {
  if (!t || !t->ctx) return SPLAT_ERR_INVALID;
  return ctx_guard(t->ctx, [&] {
    SPLAT_REQUIRE(out, "null argument");
    SPLAT_REQUIRE(t->track, "feature statistics are tracked only with a scorer attached");
    trainer_drain(t);
    if (t->n == 0) return;
    float* f = (float*)t->store["learn_feats"].reserve(t->n * kMlpFeatures * 4);
    compute_features(t, extent, grad_threshold, percent_dense, iteration, f);
    SPLAT_HIP(hipMemcpyAsync(out, f, t->n * kMlpFeatures * 4, hipMemcpyDeviceToHost,
                             t->ctx->stream));
    SPLAT_HIP(hipStreamSynchronize(t->ctx->stream));
  });
}

int splat_trainer_download_lineage(splat_trainer* t, float* max_grad, uint32_t* strong,
                                   uint32_t* birth, uint32_t* origin, uint64_t capacity) {
  if (!t || !t->ctx) return SPLAT_ERR_INVALID;
  return ctx_guard(t->ctx, [&] {
    SPLAT_REQUIRE(t->track, "feature statistics are tracked only with a scorer attached");
    SPLAT_REQUIRE(capacity >= t->n, "output capacity too small");
    trainer_drain(t);
    const size_t b = t->n * 4;
    if (max_grad) SPLAT_HIP(hipMemcpyAsync(max_grad, t->max_grad.ptr, b, hipMemcpyDeviceToHost, t->ctx->stream));
    if (strong) SPLAT_HIP(hipMemcpyAsync(strong, t->strong.ptr, b, hipMemcpyDeviceToHost, t->ctx->stream));
    if (birth) SPLAT_HIP(hipMemcpyAsync(birth, t->birth.ptr, b, hipMemcpyDeviceToHost, t->ctx->stream));
    if (origin) SPLAT_HIP(hipMemcpyAsync(origin, t->origin.ptr, b, hipMemcpyDeviceToHost, t->ctx->stream));
    SPLAT_HIP(hipStreamSynchronize(t->ctx->stream));
  });
}

int splat_trainer_learn_samples(splat_trainer* t, float* features, uint8_t* labels,
                                float* weights, uint64_t capacity, uint64_t* count) {
  if (!t || !t->ctx || !count) return SPLAT_ERR_INVALID;
  return ctx_guard(t->ctx, [&] {
    *count = 0;
    if (!t->learn) return;
    const size_t k = std::min<size_t>(t->learn->resolved, capacity);
    LearnState& L = *t->learn;
    if (k && features)
      SPLAT_HIP(hipMemcpyAsync(features, L.feats.ptr, k * kMlpFeatures * 4,
                               hipMemcpyDeviceToHost, t->ctx->stream));
    if (k && labels)
      SPLAT_HIP(hipMemcpyAsync(labels, L.label.ptr, k, hipMemcpyDeviceToHost, t->ctx->stream));
    if (k && weights)
      SPLAT_HIP(hipMemcpyAsync(weights, L.weight.ptr, k * 4, hipMemcpyDeviceToHost,
                               t->ctx->stream));
    SPLAT_HIP(hipStreamSynchronize(t->ctx->stream));
    *count = k;
  });
}

}  // extern "C"
