// Adam (src/adam.rs:44-161), densification statistics (src/densify.rs:36-60), densify and
// prune (src/densify.rs:125-221) and opacity reset (src/densify.rs:224-230) on the
// parameter planes.
#include "internal.hpp"

namespace splat {
namespace {

// blockIdx.y = plane; four consecutive Gaussians per lane.
__global__ void __launch_bounds__(256) adam_kernel(float* __restrict__ params,
                                                   const float* __restrict__ grads,
                                                   float* __restrict__ m, float* __restrict__ v,
                                                   uint32_t n, uint32_t stride, float c1,
                                                   float c2, float b1, float b2, float eps,
                                                   splat_learning_rates lr,
                                                   const StepDyn* __restrict__ dyn,
                                                   const StepStatus* __restrict__ status) {
  const int j = blockIdx.y;
  const uint32_t i0 = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (i0 >= n) return;
  if (status && status->abort) return;
  if (dyn) {
    c1 = dyn->c1;
    c2 = dyn->c2;
    lr = dyn->lr;
  }
  const float rate = plane_lr(j, lr);
  const size_t base = (size_t)j * stride + i0;
  if (i0 + 4 <= n) {
    float4 P = load_stream(params + base);
    const float4 G = load_stream(grads + base);
    float4 M = load_stream(m + base);
    float4 V = load_stream(v + base);
    adam_update(P.x, G.x, M.x, V.x, rate, b1, b2, c1, c2, eps);
    adam_update(P.y, G.y, M.y, V.y, rate, b1, b2, c1, c2, eps);
    adam_update(P.z, G.z, M.z, V.z, rate, b1, b2, c1, c2, eps);
    adam_update(P.w, G.w, M.w, V.w, rate, b1, b2, c1, c2, eps);
    store_stream(params + base, P);
    store_stream(m + base, M);
    store_stream(v + base, V);
  } else {
    for (uint32_t k = 0; i0 + k < n; k++)
      adam_update(params[base + k], grads[base + k], m[base + k], v[base + k], rate, b1, b2, c1,
                  c2, eps);
  }
}

__global__ void __launch_bounds__(256) stats_kernel(const uint32_t* __restrict__ radius,
                                                    const splat_proj_grad* __restrict__ grads,
                                                    const float* __restrict__ abs_grads,
                                                    uint32_t n, float sx, float sy,
                                                    float* __restrict__ grad_accum,
                                                    uint32_t* __restrict__ denom,
                                                    uint32_t* __restrict__ max_radii,
                                                    const StepStatus* __restrict__ status,
                                                    float* __restrict__ max_grad,
                                                    uint32_t* __restrict__ strong,
                                                    float strong_threshold) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (status && status->abort) return;
  const uint32_t r = radius[i];
  if (r == 0) return;
  float g0, g1;
  if (abs_grads) {
    g0 = abs_grads[2 * (size_t)i];
    g1 = abs_grads[2 * (size_t)i + 1];
  } else {
    g0 = grads[i].mean2d[0];
    g1 = grads[i].mean2d[1];
  }
  const float gx = g0 * sx;
  const float gy = g1 * sy;
  const float norm = sqrtf(gx * gx + gy * gy);
  grad_accum[i] += norm;
  denom[i] += 1u;
  max_radii[i] = max(max_radii[i], r);
  if (max_grad) {  // extra statistics of the learned scorer's features
    max_grad[i] = rmax(max_grad[i], norm);
    strong[i] += norm >= strong_threshold ? 1u : 0u;
  }
}


__device__ __forceinline__ float max_scale(float l0, float l1, float l2) {  // scales().max_elem()
  return rmax(rmax(splat_expf(l0), splat_expf(l1)), splat_expf(l2));
}

__device__ __forceinline__ float sigmoid_dev(float x) { return 1.0f / (1.0f + splat_expf(-x)); }

// The reference's decisions as DecisionBits (src/densify.rs:125-214).
__global__ void __launch_bounds__(256) densify_decide_kernel(const float* __restrict__ params,
                                                             uint32_t n, uint32_t stride,
                                                             const float* __restrict__ grad_accum,
                                                             const uint32_t* __restrict__ denom,
                                                             const uint32_t* __restrict__ max_radii,
                                                             DensifyParams dp,
                                                             uint8_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float grad = denom[i] > 0 ? grad_accum[i] / (float)denom[i] : 0.0f;
  const float smax = max_scale(params[3 * (size_t)stride + i], params[4 * (size_t)stride + i],
                               params[5 * (size_t)stride + i]);
  const float alpha = sigmoid_dev(params[10 * (size_t)stride + i]);
  bool prune_same = alpha < dp.min_opacity;
  if (dp.prune_big_world) prune_same = prune_same || smax > dp.world_limit;
  bool prune_orig = prune_same;
  if (dp.has_max_screen) prune_orig = prune_orig || max_radii[i] > dp.max_screen;
  uint32_t d = 0;
  if (grad >= dp.grad_threshold) d |= smax <= dp.size_limit ? kDecClone : kDecSplit;
  if (prune_orig) d |= kDecPrune;
  if (prune_same) d |= kDecPruneCopy;
  out[i] = (uint8_t)d;
}

// Per original Gaussian: keep-original, keep-clone, split, keep-children counts.
__global__ void __launch_bounds__(256) densify_flags_kernel(const float* __restrict__ params,
                                                            uint32_t n, uint32_t stride,
                                                            const float* __restrict__ grad_accum,
                                                            const uint32_t* __restrict__ denom,
                                                            const uint32_t* __restrict__ max_radii,
                                                            DensifyParams dp,
                                                            uint32_t* __restrict__ keep_orig,
                                                            uint32_t* __restrict__ keep_clone,
                                                            uint32_t* __restrict__ split,
                                                            uint32_t* __restrict__ keep_child,
                                                            uint32_t* __restrict__ clone,
                                                            const uint8_t* __restrict__ decide) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float grad = denom[i] > 0 ? grad_accum[i] / (float)denom[i] : 0.0f;
  const float l0 = params[3 * (size_t)stride + i], l1 = params[4 * (size_t)stride + i],
              l2 = params[5 * (size_t)stride + i];
  const float smax = max_scale(l0, l1, l2);
  bool is_clone = grad >= dp.grad_threshold && smax <= dp.size_limit;
  bool is_split = grad >= dp.grad_threshold && smax > dp.size_limit;
  const float alpha = sigmoid_dev(params[10 * (size_t)stride + i]);
  // prune of a row with these parameters (src/densify.rs:205-214)
  bool prune_same = alpha < dp.min_opacity;
  if (dp.prune_big_world) prune_same = prune_same || smax > dp.world_limit;
  bool prune_orig = prune_same;
  if (dp.has_max_screen) prune_orig = prune_orig || max_radii[i] > dp.max_screen;
  if (decide) {  // decisions made elsewhere (DecisionBits)
    const uint32_t d = decide[i];
    is_clone = (d & kDecClone) != 0;
    is_split = (d & kDecSplit) != 0 && !is_clone;
    prune_orig = (d & kDecPrune) != 0;
    prune_same = (d & kDecPruneCopy) != 0;
  }
  bool prune_child = alpha < dp.min_opacity;
  if (dp.prune_big_world)
    prune_child = prune_child ||
                  max_scale(l0 - dp.shrink, l1 - dp.shrink, l2 - dp.shrink) > dp.world_limit;
  keep_orig[i] = (!is_split && !prune_orig) ? 1u : 0u;
  keep_clone[i] = (is_clone && !prune_same) ? 1u : 0u;
  split[i] = is_split ? 1u : 0u;
  keep_child[i] = (is_split && !prune_child) ? 1u : 0u;
  clone[i] = is_clone ? 1u : 0u;
}

// Kept originals (parameters and both moments) and kept clones (parameters, zero moments:
// src/densify.rs:148-158, the retains of :192-197 and :203-218) to their output rows,
// kDensifyPlanes planes per workgroup row (blockIdx.y): each lane
// has all of its loads for a row in flight before it stores, and a wave's accesses stay
// within a few planes at a time.  Group 0 also writes the lineage of those rows.  A row's
// keep flags are the steps of the exclusive scans of the flags (pos[i + 1] - pos[i]; the
// scans hold n + 1 words), so each group reads two arrays per row instead of four.
constexpr int kDensifyPlanes = 3;
constexpr int kDensifyGroups = (kParams + kDensifyPlanes - 1) / kDensifyPlanes;

__global__ void __launch_bounds__(256) densify_copy_kernel(
    const float* __restrict__ params, const float* __restrict__ m, const float* __restrict__ v,
    uint32_t n, uint32_t stride, const uint32_t* __restrict__ pos_orig,
    const uint32_t* __restrict__ pos_clone, uint32_t n_orig_kept, float* __restrict__ po,
    float* __restrict__ mo, float* __restrict__ vo, uint32_t out_stride, DensifyLineage ln) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t oo = pos_orig[i], oc = n_orig_kept + pos_clone[i];
  const bool ko = pos_orig[i + 1] != oo, kc = n_orig_kept + pos_clone[i + 1] != oc;
  if (!ko && !kc) return;
  if (blockIdx.y == 0 && ln.birth_out) {
    // lineage: kept rows keep theirs; a clone is born now and descends from the logged
    // decision on its parent (or from the parent's own origin when it was not logged)
    if (ko) {
      ln.birth_out[oo] = ln.birth_in[i];
      ln.origin_out[oo] = ln.origin_in[i];
    }
    if (kc) {
      ln.birth_out[oc] = ln.iteration;
      ln.origin_out[oc] =
          ln.sample_of && ln.sample_of[i] != kNoSample ? ln.sample_of[i] : ln.origin_in[i];
    }
  }
  const int j0 = (int)blockIdx.y * kDensifyPlanes;
  float tp[kDensifyPlanes], tm[kDensifyPlanes], tv[kDensifyPlanes];
#pragma unroll
  for (int c = 0; c < kDensifyPlanes; c++) {
    const size_t q = (size_t)(j0 + c) * stride + i;
    if (j0 + c < kParams) {
      tp[c] = params[q];
      if (ko) {
        tm[c] = m[q];
        tv[c] = v[q];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < kDensifyPlanes; c++) {
    if (j0 + c >= kParams) break;
    const size_t plane = (size_t)(j0 + c) * out_stride;
    if (ko) {
      po[plane + oo] = tp[c];
      mo[plane + oo] = tm[c];
      vo[plane + oo] = tv[c];
    }
    if (kc) {
      po[plane + oc] = tp[c];
      mo[plane + oc] = 0.0f;
      vo[plane + oc] = 0.0f;
    }
  }
}

// The kept children of split rows (src/densify.rs:160-186) and their lineage.
__global__ void __launch_bounds__(256) densify_children_kernel(
    const float* __restrict__ params, uint32_t n, uint32_t stride,
    const uint32_t* __restrict__ split, const uint32_t* __restrict__ keep_child,
    const uint32_t* __restrict__ pos_split, const uint32_t* __restrict__ pos_child,
    uint32_t n_orig_kept, uint32_t n_clone_kept, DensifyParams dp,
    const float* __restrict__ normals, uint64_t normals_seed, float* __restrict__ po,
    float* __restrict__ mo, float* __restrict__ vo, uint32_t out_stride, DensifyLineage ln) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !split[i] || !keep_child[i]) return;
  if (ln.birth_out) {  // children are born now, descended as clones are
    const uint32_t new_origin =
        ln.sample_of && ln.sample_of[i] != kNoSample ? ln.sample_of[i] : ln.origin_in[i];
    for (uint32_t c = 0; c < dp.children; c++) {
      const uint32_t o = n_orig_kept + n_clone_kept + pos_child[i] * dp.children + c;
      ln.birth_out[o] = ln.iteration;
      ln.origin_out[o] = new_origin;
    }
  }
  {
    float g[kParams];
    for (int j = 0; j < kParams; j++) g[j] = params[(size_t)j * stride + i];
    const float s[3] = {splat_expf(g[3]), splat_expf(g[4]), splat_expf(g[5])};
    // g.rotation().to_mat(), src/math.rs:294-330
    float q[4];
    {
      const float w = g[6], x = g[7], y = g[8], z = g[9];
      const float nn = sqrtf(w * w + x * x + y * y + z * z);
      if (nn > 0.0f) {
        q[0] = w / nn;
        q[1] = x / nn;
        q[2] = y / nn;
        q[3] = z / nn;
      } else {
        q[0] = 1.0f;
        q[1] = q[2] = q[3] = 0.0f;
      }
    }
    const float w = q[0], x = q[1], y = q[2], z = q[3];
    const float R[3][3] = {
        {1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y - w * z), 2.0f * (x * z + w * y)},
        {2.0f * (x * y + w * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z - w * x)},
        {2.0f * (x * z - w * y), 2.0f * (y * z + w * x), 1.0f - 2.0f * (x * x + y * y)}};
    for (uint32_t c = 0; c < dp.children; c++) {
      const size_t draw = ((size_t)pos_split[i] * dp.children + c) * 3;
      float nd[3];
      for (int k = 0; k < 3; k++)
        nd[k] = normals ? normals[draw + k] : normal_draw(normals_seed, draw + k);
      const float loc[3] = {nd[0] * s[0], nd[1] * s[1], nd[2] * s[2]};
      const uint32_t o = n_orig_kept + n_clone_kept + pos_child[i] * dp.children + c;
      for (int j = 0; j < kParams; j++) {
        float val = g[j];
        if (j < 3) val = g[j] + (R[j][0] * loc[0] + R[j][1] * loc[1] + R[j][2] * loc[2]);
        else if (j < 6) val = g[j] - dp.shrink;
        po[(size_t)j * out_stride + o] = val;
        mo[(size_t)j * out_stride + o] = 0.0f;
        vo[(size_t)j * out_stride + o] = 0.0f;
      }
    }
  }
}

__global__ void reset_opacity_kernel(float* __restrict__ params, float* __restrict__ m,
                                     float* __restrict__ v, uint32_t n, uint32_t stride,
                                     float limit) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t o = 10 * (size_t)stride + i;
  params[o] = rmin(params[o], limit);
  m[o] = 0.0f;
  v[o] = 0.0f;
}

}  // namespace

void adam_planes(splat_ctx* ctx, float* params, const float* grads, float* m, float* v, size_t n,
                 size_t stride, float c1, float c2, const splat_adam_state& st,
                 const splat_learning_rates& lr, const StepDyn* dyn, const StepStatus* status) {
  launch(ctx, "adam", adam_kernel, dim3(blocks_for((n + 3) / 4, 256), kParams), dim3(256), 0,
         params, grads, m, v, (uint32_t)n, (uint32_t)stride, c1, c2, st.beta1, st.beta2, st.eps,
         lr, dyn, status);
}

void stats_accumulate_dev(splat_ctx* ctx, const uint32_t* radius, const splat_proj_grad* grads,
                          const float* abs_grads, size_t n, uint32_t width, uint32_t height,
                          float* grad_accum, uint32_t* denom, uint32_t* max_radii,
                          const StepStatus* status, float* max_grad, uint32_t* strong,
                          float strong_threshold) {
  const float sx = (float)width * 0.5f, sy = (float)height * 0.5f;
  launch(ctx, "stats", stats_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, radius, grads,
         abs_grads, (uint32_t)n, sx, sy, grad_accum, denom, max_radii, status, max_grad, strong,
         strong_threshold);
}

DensifyParams densify_params(const splat_densify_config& cfg, float extent) {
  DensifyParams dp;
  dp.grad_threshold = cfg.grad_threshold;
  dp.size_limit = cfg.percent_dense * extent;
  dp.min_opacity = cfg.min_opacity;
  dp.world_limit = cfg.max_world_size_frac * extent;
  dp.shrink = logf(0.8f * (float)cfg.split_children);  // src/densify.rs:169
  dp.prune_big_world = cfg.prune_big_world;
  dp.has_max_screen = cfg.has_max_screen_size;
  dp.max_screen = cfg.max_screen_size;
  dp.children = cfg.split_children;
  return dp;
}

void densify_decide(splat_ctx* ctx, const float* params, size_t n, size_t stride,
                    const float* grad_accum, const uint32_t* denom, const uint32_t* max_radii,
                    const splat_densify_config& cfg, float extent, uint8_t* out) {
  launch(ctx, "densify_decide", densify_decide_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
         params, (uint32_t)n, (uint32_t)stride, grad_accum, denom, max_radii,
         densify_params(cfg, extent), out);
}

// Densify on planes.  Returns the new count; fills report.  Output planes have stride
// plane_stride(out_capacity) and must hold the result (checked).
size_t densify_planes(splat_ctx* ctx, const float* params, const float* m, const float* v,
                      size_t n, size_t stride, const float* grad_accum, const uint32_t* denom,
                      const uint32_t* max_radii, const splat_densify_config& cfg, float extent,
                      const float* normals_dev, size_t normals_len, uint64_t normals_seed,
                      float* po, float* mo, float* vo, size_t out_capacity, size_t out_stride,
                      splat_densify_report* report, const uint8_t* decide,
                      const DensifyLineage* lineage) {
  const DensifyParams dp = densify_params(cfg, extent);
  DensifyLineage ln{};
  if (lineage) ln = *lineage;
  uint32_t* flags = ctx->buf<uint32_t>("densify_flags", 5 * (n + 1));
  uint32_t* pos = ctx->buf<uint32_t>("densify_pos", 5 * (n + 1));
  uint32_t* totals = ctx->buf<uint32_t>("densify_totals", 5);
  uint32_t *ko = flags, *kc = flags + (n + 1), *sp = flags + 2 * (n + 1),
           *kch = flags + 3 * (n + 1), *cl = flags + 4 * (n + 1);
  uint32_t *po_o = pos, *po_c = pos + (n + 1), *po_s = pos + 2 * (n + 1),
           *po_ch = pos + 3 * (n + 1), *po_cl = pos + 4 * (n + 1);
  launch(ctx, "densify_flags", densify_flags_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
         params, (uint32_t)n, (uint32_t)stride, grad_accum, denom, max_radii, dp, ko, kc, sp, kch,
         cl, decide);
  exclusive_scan_u32(ctx, ko, po_o, n, totals + 0);
  exclusive_scan_u32(ctx, kc, po_c, n, totals + 1);
  exclusive_scan_u32(ctx, sp, po_s, n, totals + 2);
  exclusive_scan_u32(ctx, kch, po_ch, n, totals + 3);
  exclusive_scan_u32(ctx, cl, po_cl, n, totals + 4);
  uint32_t t[5];
  SPLAT_HIP(hipMemcpyAsync(t, totals, 20, hipMemcpyDeviceToHost, ctx->stream));
  SPLAT_HIP(hipStreamSynchronize(ctx->stream));
  const uint32_t n_orig_kept = t[0], n_clone_kept = t[1], n_split = t[2], n_child_par = t[3],
                 n_clone = t[4];
  const size_t children = cfg.split_children;
  const size_t n_out = (size_t)n_orig_kept + n_clone_kept + (size_t)n_child_par * children;
  report->cloned = n_clone;
  report->split = n_split;
  report->pruned = ((size_t)n - n_split - n_orig_kept) + ((size_t)n_clone - n_clone_kept) +
                   ((size_t)n_split - n_child_par) * children;
  if (normals_dev && normals_len < 3ull * children * n_split) return (size_t)-1;
  if (n_out > out_capacity) return n_out;
  launch(ctx, "densify_copy", densify_copy_kernel, dim3(blocks_for(n, 256), kDensifyGroups),
         dim3(256), 0, params, m, v, (uint32_t)n, (uint32_t)stride, (const uint32_t*)po_o,
         (const uint32_t*)po_c, n_orig_kept, po, mo, vo, (uint32_t)out_stride, ln);
  if (n_split > 0)
    launch(ctx, "densify_children", densify_children_kernel, dim3(blocks_for(n, 256)), dim3(256),
           0, params, (uint32_t)n, (uint32_t)stride, (const uint32_t*)sp, (const uint32_t*)kch,
           (const uint32_t*)po_s, (const uint32_t*)po_ch, n_orig_kept, n_clone_kept, dp,
           normals_dev, normals_seed, po, mo, vo, (uint32_t)out_stride, ln);
  return n_out;
}

void reset_opacity_planes(splat_ctx* ctx, float* params, float* m, float* v, size_t n,
                          size_t stride, float limit) {
  launch(ctx, "reset_opacity", reset_opacity_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
         params, m, v, (uint32_t)n, (uint32_t)stride, limit);
}

}  // namespace splat
