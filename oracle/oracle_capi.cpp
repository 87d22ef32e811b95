// C entry points of the CPU oracle -- TEST INFRASTRUCTURE ONLY (see splat_oracle.hpp).
//
// The f32 entry points take the same plain records as libsplat's host tier
// (include/splat.h), so a test feeds both the same bytes.  The f64 entry points use flat
// double arrays (layouts documented per function) for the finite-difference checks.
#include <cstdio>
#include <cstring>
#include <vector>

#include "../include/splat.h"
#include "splat_oracle.hpp"
#include "mlp_oracle.hpp"

using namespace oracle;

namespace {

template <class R>
Camera<R> camera_from(const splat_camera* c) {
  Camera<R> cam;
  cam.width = c->width;
  cam.height = c->height;
  cam.fx = c->fx;
  cam.fy = c->fy;
  cam.cx = c->cx;
  cam.cy = c->cy;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) cam.rot.m[i][j] = c->rot[3 * i + j];
  cam.trans = V3<R>(c->trans[0], c->trans[1], c->trans[2]);
  return cam;
}

template <class R>
ProjOpts<R> opts_from(const splat_proj_opts* o) {
  ProjOpts<R> p;
  p.near_plane = o->near_plane;
  p.eps2d = o->eps2d;
  p.antialias = o->antialias != 0;
  p.sh_degree = o->sh_degree;
  return p;
}

static_assert(sizeof(Proj<float>) == sizeof(splat_proj), "Proj layout");
static_assert(sizeof(ProjGrad<float>) == sizeof(splat_proj_grad), "ProjGrad layout");

// f64 Proj: 15 doubles [mean2d 2, conic 3, depth, radius, opacity, color 3, tile_min 2,
// tile_max 2].
void proj_to64(const Proj<double>& p, double* o) {
  o[0] = p.mean2d[0];
  o[1] = p.mean2d[1];
  for (int i = 0; i < 3; i++) o[2 + i] = p.conic[i];
  o[5] = p.depth;
  o[6] = p.radius;
  o[7] = p.opacity;
  for (int i = 0; i < 3; i++) o[8 + i] = p.color[i];
  o[11] = p.tile_min[0];
  o[12] = p.tile_min[1];
  o[13] = p.tile_max[0];
  o[14] = p.tile_max[1];
}
Proj<double> proj_from64(const double* o) {
  Proj<double> p;
  p.mean2d[0] = o[0];
  p.mean2d[1] = o[1];
  for (int i = 0; i < 3; i++) p.conic[i] = o[2 + i];
  p.depth = o[5];
  p.radius = (uint32_t)o[6];
  p.opacity = o[7];
  for (int i = 0; i < 3; i++) p.color[i] = o[8 + i];
  p.tile_min[0] = (uint32_t)o[11];
  p.tile_min[1] = (uint32_t)o[12];
  p.tile_max[0] = (uint32_t)o[13];
  p.tile_max[1] = (uint32_t)o[14];
  return p;
}

Binning binning_from(const splat_binning* b, uint64_t n) {
  Binning out;
  out.tiles_x = b->tiles_x;
  out.tiles_y = b->tiles_y;
  size_t e = b->num_entries;
  out.entry_gaussian.assign(b->entry_gaussian, b->entry_gaussian + e);
  out.entry_offset.assign(b->entry_offset, b->entry_offset + n + 1);
  out.sorted.resize(e);
  for (size_t i = 0; i < e; i++) out.sorted[i] = std::make_pair(b->sorted_keys[i], b->sorted_entries[i]);
  size_t t = (size_t)b->tiles_x * b->tiles_y;
  out.tile_range.resize(t);
  for (size_t i = 0; i < t; i++) out.tile_range[i] = std::make_pair(b->tile_range[2 * i], b->tile_range[2 * i + 1]);
  return out;
}

int binning_to(const Binning& b, splat_binning* out) {
  out->tiles_x = b.tiles_x;
  out->tiles_y = b.tiles_y;
  out->num_entries = b.sorted.size();
  if (out->capacity < b.sorted.size()) return SPLAT_ERR_CAPACITY;
  size_t e = b.sorted.size();
  std::memcpy(out->entry_gaussian, b.entry_gaussian.data(), e * 4);
  std::memcpy(out->entry_offset, b.entry_offset.data(), b.entry_offset.size() * 4);
  for (size_t i = 0; i < e; i++) {
    out->sorted_keys[i] = b.sorted[i].first;
    out->sorted_entries[i] = b.sorted[i].second;
  }
  for (size_t i = 0; i < b.tile_range.size(); i++) {
    out->tile_range[2 * i] = b.tile_range[i].first;
    out->tile_range[2 * i + 1] = b.tile_range[i].second;
  }
  return SPLAT_OK;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- f32 --------------

void oracle_sh_basis_f32(int degree, const float dir[3], float out[16]) {
  sh_basis<float>(degree, V3<float>(dir[0], dir[1], dir[2]), out);
}
void oracle_sh_basis_grad_f32(int degree, const float dir[3], float out[48]) {
  V3<float> g[16];
  sh_basis_grad<float>(degree, V3<float>(dir[0], dir[1], dir[2]), g);
  for (int k = 0; k < 16; k++) {
    out[3 * k] = g[k].x;
    out[3 * k + 1] = g[k].y;
    out[3 * k + 2] = g[k].z;
  }
}

int oracle_project_all_f32(const float* g, uint64_t n, const splat_camera* cam,
                           const splat_proj_opts* opts, splat_proj* out) {
  Camera<float> c = camera_from<float>(cam);
  ProjOpts<float> o = opts_from<float>(opts);
  for (uint64_t i = 0; i < n; i++) {
    Proj<float> p = project_one<float>(g + i * 59, c, o);
    std::memcpy(&out[i], &p, sizeof(p));
  }
  return SPLAT_OK;
}

int oracle_project_backward_all_f32(const float* g, uint64_t n, const splat_camera* cam,
                                    const splat_proj_opts* opts, const splat_proj_grad* grads,
                                    float* out) {
  Camera<float> c = camera_from<float>(cam);
  ProjOpts<float> o = opts_from<float>(opts);
  for (uint64_t i = 0; i < n; i++) {
    ProjGrad<float> dp;
    std::memcpy(&dp, &grads[i], sizeof(dp));
    project_backward_one<float>(g + i * 59, c, o, dp, out + i * 59);
  }
  return SPLAT_OK;
}

int oracle_bin_f32(const splat_proj* projs, uint64_t n, uint32_t width, uint32_t height,
                   splat_binning* out) {
  Binning b;
  if (!bin<float>(reinterpret_cast<const Proj<float>*>(projs), n, width, height, b))
    return SPLAT_ERR_PRECONDITION;
  return binning_to(b, out);
}

int oracle_rasterize_forward_f32(const splat_proj* projs, uint64_t n, const splat_binning* binning,
                                 const splat_raster_settings* s, float* image, float* final_t,
                                 uint32_t* n_contrib) {
  Binning b = binning_from(binning, n);
  rasterize_forward<float>(reinterpret_cast<const Proj<float>*>(projs), b, s->width, s->height,
                           s->background, image, final_t, n_contrib);
  return SPLAT_OK;
}

int oracle_rasterize_backward_f32(const splat_proj* projs, uint64_t n, const splat_binning* binning,
                                  const splat_raster_settings* s, const float* final_t,
                                  const uint32_t* n_contrib, const float* dimage,
                                  splat_proj_grad* grads, float* abs_grads) {
  Binning b = binning_from(binning, n);
  rasterize_backward<float>(reinterpret_cast<const Proj<float>*>(projs), n, b, s->width, s->height,
                            s->background, final_t, n_contrib, dimage,
                            reinterpret_cast<ProjGrad<float>*>(grads), abs_grads);
  return SPLAT_OK;
}

int oracle_raster_pair_counts(const splat_proj* projs, uint64_t n, const splat_binning* binning,
                              const splat_raster_settings* s, const uint32_t* n_contrib,
                              uint64_t* counts) {
  Binning b = binning_from(binning, n);
  raster_pair_counts(reinterpret_cast<const Proj<float>*>(projs), b, s->width, s->height,
                     n_contrib, counts);
  return SPLAT_OK;
}

int oracle_resize_rgb8(const uint8_t* src, uint32_t w, uint32_t h, uint32_t ow, uint32_t oh,
                       uint8_t* dst) {
  resize_rgb8(src, w, h, ow, oh, dst);
  return SPLAT_OK;
}

int oracle_blur_f32(const float* src, uint32_t w, uint32_t h, uint32_t ch, float* dst) {
  float k[SSIM_WINDOW];
  gaussian_kernel<float>(k);
  blur<float>(src, w, h, ch, k, dst);
  return SPLAT_OK;
}
void oracle_gaussian_kernel_f32(float out[11]) { gaussian_kernel<float>(out); }

int oracle_ssim_with_grad_f32(const float* x, const float* y, uint32_t w, uint32_t h, float* ssim,
                              float* grad) {
  *ssim = ssim_with_grad<float>(x, y, w, h, grad);
  return SPLAT_OK;
}

int oracle_photometric_loss_f32(const float* pred, const float* target, uint32_t w, uint32_t h,
                                float lambda, float* loss, float* grad) {
  *loss = photometric_loss<float>(pred, target, w, h, lambda, grad);
  return SPLAT_OK;
}

// The binary64 sums of the binary32 SSIM map and L1 terms of photometric_loss.
int oracle_loss_sums_f32(const float* pred, const float* target, uint32_t w, uint32_t h,
                         double* sums) {
  std::vector<float> grad((size_t)w * h * 3);
  photometric_loss<float>(pred, target, w, h, 0.2f, grad.data(), sums);
  return SPLAT_OK;
}

float oracle_psnr_f32(const float* pred, const float* target, uint64_t len) {
  return psnr<float>(pred, target, len);
}

int oracle_adam_step_f32(float* params, const float* grads, float* m, float* v, uint64_t n,
                         splat_adam_state* st, const splat_learning_rates* lr) {
  float l[6] = {lr->mean, lr->log_scale, lr->quat, lr->opacity, lr->sh_dc, lr->sh_rest};
  adam_step<float>(params, grads, m, v, n, &st->step, st->beta1, st->beta2, st->eps, l);
  return SPLAT_OK;
}

int oracle_stats_accumulate_f32(const splat_proj* projs, const splat_proj_grad* grads,
                                const float* abs_grads, uint64_t n, uint32_t width,
                                uint32_t height, float* grad_accum, uint32_t* denom,
                                uint32_t* max_radii) {
  stats_accumulate<float>(reinterpret_cast<const Proj<float>*>(projs),
                          reinterpret_cast<const ProjGrad<float>*>(grads), abs_grads, n, width,
                          height, grad_accum, denom, max_radii);
  return SPLAT_OK;
}

int oracle_densify_and_prune_f32(const float* params, const float* m, const float* v,
                                 const float* grad_accum, const uint32_t* denom,
                                 const uint32_t* max_radii, uint64_t n,
                                 const splat_densify_config* cfg, float extent,
                                 const float* normals, uint64_t normals_len, float* params_out,
                                 float* m_out, float* v_out, uint64_t out_capacity,
                                 uint64_t* n_out, splat_densify_report* report) {
  std::vector<float> p(params, params + n * 59), mm(m, m + n * 59), vv(v, v + n * 59);
  DensifyConfig<float> c;
  c.grad_threshold = cfg->grad_threshold;
  c.percent_dense = cfg->percent_dense;
  c.min_opacity = cfg->min_opacity;
  c.prune_big_world = cfg->prune_big_world != 0;
  c.max_world_size_frac = cfg->max_world_size_frac;
  c.has_max_screen_size = cfg->has_max_screen_size != 0;
  c.max_screen_size = cfg->max_screen_size;
  c.split_children = cfg->split_children;
  DensifyReport r;
  bool ok = densify_and_prune<float>(p, mm, vv, grad_accum, denom, max_radii, c, extent, normals,
                                     normals_len, r);
  report->cloned = r.cloned;
  report->split = r.split;
  report->pruned = r.pruned;
  if (!ok) return SPLAT_ERR_CAPACITY;
  uint64_t no = p.size() / 59;
  *n_out = no;
  if (no > out_capacity) return SPLAT_ERR_CAPACITY;
  std::memcpy(params_out, p.data(), p.size() * 4);
  std::memcpy(m_out, mm.data(), mm.size() * 4);
  std::memcpy(v_out, vv.data(), vv.size() * 4);
  return SPLAT_OK;
}

int oracle_reset_opacity_f32(float* params, float* m, float* v, uint64_t n, float max_alpha) {
  reset_opacity<float>(params, m, v, n, max_alpha);
  return SPLAT_OK;
}

float oracle_sigmoid_f32(float x) { return sigmoid<float>(x); }
float oracle_inverse_sigmoid_f32(float x) { return inverse_sigmoid<float>(x); }
float oracle_powi_f32(float a, int b) { return powi<float>(a, b); }

void oracle_mean_nn_dist2(const double* pts, uint64_t n, double* out) { mean_nn_dist2(pts, n, out); }

// One training iteration (src/train.rs:211-247): project, bin, composite, loss against an
// RGB8 target (View::to_real), backward, optional statistics, Adam.  In place on params,
// m, v, the statistics and *step.  Returns the loss; *num_entries receives E.
float oracle_train_step_f32(float* params, float* m, float* v, uint64_t n, uint64_t* step,
                            const splat_camera* cam, const uint8_t* rgb, int sh_degree,
                            int antialias, float lambda, const float bg[3],
                            const splat_learning_rates* lr, int accumulate_stats, int absgrad,
                            float* grad_accum, uint32_t* denom, uint32_t* max_radii,
                            uint64_t* num_entries) {
  Camera<float> c = camera_from<float>(cam);
  ProjOpts<float> o;
  o.near_plane = 0.2f;
  o.eps2d = 0.3f;
  o.antialias = antialias != 0;
  o.sh_degree = sh_degree;
  const uint32_t w = cam->width, h = cam->height;
  std::vector<Proj<float>> p(n);
  ORACLE_PAR
  for (uint64_t i = 0; i < n; i++) p[i] = project_one<float>(params + i * 59, c, o);
  Binning b;
  if (!bin<float>(p.data(), n, w, h, b)) return NAN;
  *num_entries = b.sorted.size();
  const size_t npx = (size_t)w * h;
  std::vector<float> image(npx * 3), final_t(npx), target(npx * 3), dimage(npx * 3);
  std::vector<uint32_t> nc(npx);
  rasterize_forward<float>(p.data(), b, w, h, bg, image.data(), final_t.data(), nc.data());
  for (size_t i = 0; i < npx * 3; i++) target[i] = (float)rgb[i] / 255.0f;
  const float loss = photometric_loss<float>(image.data(), target.data(), w, h, lambda,
                                             dimage.data());
  std::vector<ProjGrad<float>> pg(n);
  std::vector<float> ab(2 * n), grads(n * 59);
  rasterize_backward<float>(p.data(), n, b, w, h, bg, final_t.data(), nc.data(), dimage.data(),
                            pg.data(), ab.data());
  ORACLE_PAR
  for (uint64_t i = 0; i < n; i++) project_backward_one<float>(params + i * 59, c, o, pg[i], &grads[i * 59]);
  if (accumulate_stats)
    stats_accumulate<float>(p.data(), pg.data(), absgrad ? ab.data() : nullptr, n, w, h,
                            grad_accum, denom, max_radii);
  float l[6] = {lr->mean, lr->log_scale, lr->quat, lr->opacity, lr->sh_dc, lr->sh_rest};
  adam_step<float>(params, grads.data(), m, v, n, step, 0.9f, 0.999f, 1e-15f, l);
  return loss;
}

// Binary64 accuracy reference of rasterize_backward (rasterize_backward_truth): the binary32
// forward's decisions, binary64 values.  grads: n x 9 doubles, abs_grads: n x 2.
int oracle_rasterize_backward_truth(const splat_proj* projs, uint64_t n,
                                    const splat_binning* binning, const splat_raster_settings* s,
                                    const uint32_t* n_contrib, const float* dimage,
                                    double* grads, double* abs_grads) {
  Binning b = binning_from(binning, n);
  rasterize_backward_truth(reinterpret_cast<const Proj<float>*>(projs), n, b, s->width,
                           s->height, s->background, n_contrib, dimage, grads, abs_grads);
  return SPLAT_OK;
}

// Parameter gradients of one training step's loss (src/train.rs:211-240, no Adam) at
// `params`: grads_ref (n x 59 binary32, optional) in the reference's arithmetic, and
// grads_true (n x 59 binary64, optional) = the binary64 projection backward of the
// parameters applied to rasterize_backward_truth's screen-space gradients, given the
// binary32 forward's image and loss gradient.  pg_ref / pg_true (n ProjGrad, n x 9 doubles)
// optionally receive the screen-space gradients.  Returns 0 or a precondition code.
int oracle_step_gradients(const float* params, uint64_t n, const splat_camera* cam,
                          const uint8_t* rgb, int sh_degree, float lambda, const float bg[3],
                          float* grads_ref, double* grads_true, splat_proj_grad* pg_ref,
                          double* pg_true) {
  Camera<float> c = camera_from<float>(cam);
  ProjOpts<float> o;
  o.near_plane = 0.2f;
  o.eps2d = 0.3f;
  o.antialias = false;
  o.sh_degree = sh_degree;
  const uint32_t w = cam->width, h = cam->height;
  std::vector<Proj<float>> p(n);
  ORACLE_PAR
  for (uint64_t i = 0; i < n; i++) p[i] = project_one<float>(params + i * 59, c, o);
  Binning b;
  if (!bin<float>(p.data(), n, w, h, b)) return SPLAT_ERR_PRECONDITION;
  const size_t npx = (size_t)w * h;
  std::vector<float> image(npx * 3), final_t(npx), target(npx * 3), dimage(npx * 3);
  std::vector<uint32_t> nc(npx);
  rasterize_forward<float>(p.data(), b, w, h, bg, image.data(), final_t.data(), nc.data());
  for (size_t i = 0; i < npx * 3; i++) target[i] = (float)rgb[i] / 255.0f;
  photometric_loss<float>(image.data(), target.data(), w, h, lambda, dimage.data());
  if (grads_ref || pg_ref) {
    std::vector<ProjGrad<float>> pg(n);
    std::vector<float> ab(2 * n);
    rasterize_backward<float>(p.data(), n, b, w, h, bg, final_t.data(), nc.data(),
                              dimage.data(), pg.data(), ab.data());
    if (pg_ref) std::memcpy(pg_ref, pg.data(), n * sizeof(splat_proj_grad));
    if (grads_ref) {
      ORACLE_PAR
      for (uint64_t i = 0; i < n; i++)
        project_backward_one<float>(params + i * 59, c, o, pg[i], grads_ref + i * 59);
    }
  }
  if (grads_true || pg_true) {
    std::vector<double> pg9(n * 9), ab(2 * n);
    rasterize_backward_truth(p.data(), n, b, w, h, bg, nc.data(), dimage.data(), pg9.data(),
                             ab.data());
    if (pg_true) std::memcpy(pg_true, pg9.data(), n * 9 * sizeof(double));
    if (grads_true) {
      Camera<double> c64;
      c64.width = c.width;
      c64.height = c.height;
      c64.fx = c.fx;
      c64.fy = c.fy;
      c64.cx = c.cx;
      c64.cy = c.cy;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) c64.rot.m[i][j] = c.rot.m[i][j];
      c64.trans = V3<double>(c.trans.x, c.trans.y, c.trans.z);
      ProjOpts<double> o64;
      o64.near_plane = 0.2;
      o64.eps2d = 0.3;
      o64.antialias = false;
      o64.sh_degree = sh_degree;
      ORACLE_PAR
      for (uint64_t i = 0; i < n; i++) {
        double g64[59];
        for (int k = 0; k < 59; k++) g64[k] = params[i * 59 + k];
        ProjGrad<double> dp;
        std::memcpy(&dp, &pg9[i * 9], sizeof(dp));
        project_backward_one<double>(g64, c64, o64, dp, grads_true + i * 59);
      }
    }
  }
  return SPLAT_OK;
}

// ---------------------------------------------------------------- f64 --------------
// Layouts: gaussians 59 doubles each; camera 18 doubles [w, h, fx, fy, cx, cy, rot 9,
// trans 3]; opts [near, eps2d, antialias, sh_degree]; Proj 15 doubles (proj_to64);
// ProjGrad 9 doubles.

static Camera<double> cam64(const double* c) {
  Camera<double> cam;
  cam.width = (uint32_t)c[0];
  cam.height = (uint32_t)c[1];
  cam.fx = c[2];
  cam.fy = c[3];
  cam.cx = c[4];
  cam.cy = c[5];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) cam.rot.m[i][j] = c[6 + 3 * i + j];
  cam.trans = V3<double>(c[15], c[16], c[17]);
  return cam;
}
static ProjOpts<double> opts64(const double* o) {
  ProjOpts<double> p;
  p.near_plane = o[0];
  p.eps2d = o[1];
  p.antialias = o[2] != 0;
  p.sh_degree = (int)o[3];
  return p;
}

void oracle_sh_basis_f64(int degree, const double dir[3], double out[16]) {
  sh_basis<double>(degree, V3<double>(dir[0], dir[1], dir[2]), out);
}
void oracle_sh_basis_grad_f64(int degree, const double dir[3], double out[48]) {
  V3<double> g[16];
  sh_basis_grad<double>(degree, V3<double>(dir[0], dir[1], dir[2]), g);
  for (int k = 0; k < 16; k++) {
    out[3 * k] = g[k].x;
    out[3 * k + 1] = g[k].y;
    out[3 * k + 2] = g[k].z;
  }
}

void oracle_project_all_f64(const double* g, uint64_t n, const double* cam, const double* opts,
                            double* out) {
  Camera<double> c = cam64(cam);
  ProjOpts<double> o = opts64(opts);
  for (uint64_t i = 0; i < n; i++) proj_to64(project_one<double>(g + i * 59, c, o), out + i * 15);
}

void oracle_project_backward_all_f64(const double* g, uint64_t n, const double* cam,
                                     const double* opts, const double* grads, double* out) {
  Camera<double> c = cam64(cam);
  ProjOpts<double> o = opts64(opts);
  for (uint64_t i = 0; i < n; i++) {
    ProjGrad<double> dp;
    std::memcpy(&dp, grads + i * 9, sizeof(dp));
    project_backward_one<double>(g + i * 59, c, o, dp, out + i * 59);
  }
}

// Forward render of f64 projections (bins internally).  Returns 0 or a precondition code.
int oracle_render_projs_f64(const double* projs, uint64_t n, uint32_t w, uint32_t h,
                            const double bg[3], double* image, double* final_t,
                            uint32_t* n_contrib) {
  std::vector<Proj<double>> p(n);
  for (uint64_t i = 0; i < n; i++) p[i] = proj_from64(projs + i * 15);
  Binning b;
  if (!bin<double>(p.data(), n, w, h, b)) return SPLAT_ERR_PRECONDITION;
  rasterize_forward<double>(p.data(), b, w, h, bg, image, final_t, n_contrib);
  return SPLAT_OK;
}

// Backward of the f64 render: grads 9 doubles per Gaussian, abs 2.
int oracle_raster_backward_projs_f64(const double* projs, uint64_t n, uint32_t w, uint32_t h,
                                     const double bg[3], const double* dimage, double* grads,
                                     double* abs_grads) {
  std::vector<Proj<double>> p(n);
  for (uint64_t i = 0; i < n; i++) p[i] = proj_from64(projs + i * 15);
  Binning b;
  if (!bin<double>(p.data(), n, w, h, b)) return SPLAT_ERR_PRECONDITION;
  std::vector<double> image((size_t)w * h * 3), final_t((size_t)w * h);
  std::vector<uint32_t> nc((size_t)w * h);
  rasterize_forward<double>(p.data(), b, w, h, bg, image.data(), final_t.data(), nc.data());
  rasterize_backward<double>(p.data(), n, b, w, h, bg, final_t.data(), nc.data(), dimage,
                             reinterpret_cast<ProjGrad<double>*>(grads), abs_grads);
  return SPLAT_OK;
}

double oracle_ssim_with_grad_f64(const double* x, const double* y, uint32_t w, uint32_t h,
                                 double* grad) {
  return ssim_with_grad<double>(x, y, w, h, grad);
}

double oracle_photometric_loss_f64(const double* pred, const double* target, uint32_t w,
                                   uint32_t h, double lambda, double* grad) {
  return photometric_loss<double>(pred, target, w, h, lambda, grad);
}

// Whole chain parameters -> projection -> raster -> photometric loss (f64), returning the
// loss and, if `grads` is non-null, d loss / d params (59 per Gaussian).
double oracle_chain_loss_f64(const double* g, uint64_t n, const double* cam, const double* opts,
                             const double bg[3], const double* target, double lambda,
                             double* grads) {
  Camera<double> c = cam64(cam);
  ProjOpts<double> o = opts64(opts);
  uint32_t w = c.width, h = c.height;
  std::vector<Proj<double>> p(n);
  for (uint64_t i = 0; i < n; i++) p[i] = project_one<double>(g + i * 59, c, o);
  Binning b;
  if (!bin<double>(p.data(), n, w, h, b)) return NAN;
  size_t npx = (size_t)w * h;
  std::vector<double> image(npx * 3), final_t(npx), dimage(npx * 3);
  std::vector<uint32_t> nc(npx);
  rasterize_forward<double>(p.data(), b, w, h, bg, image.data(), final_t.data(), nc.data());
  double loss = photometric_loss<double>(image.data(), target, w, h, lambda, dimage.data());
  if (grads) {
    std::vector<ProjGrad<double>> pg(n);
    std::vector<double> ab(2 * n);
    rasterize_backward<double>(p.data(), n, b, w, h, bg, final_t.data(), nc.data(), dimage.data(),
                               pg.data(), ab.data());
    for (uint64_t i = 0; i < n; i++) project_backward_one<double>(g + i * 59, c, o, pg[i], grads + i * 59);
  }
  return loss;
}


// ---------------------------------------------------------------- scorer MLP -------

void oracle_mlp_forward_f32(const float* params, const float* feats, int64_t n, float* logits,
                            int bf16) {
  oracle::mlp::Trace<float> t;
  for (int64_t s = 0; s < n; s++) {
    if (bf16)
      oracle::mlp::forward<float, true>(params, feats + s * 12, t);
    else
      oracle::mlp::forward<float, false>(params, feats + s * 12, t);
    for (int o = 0; o < 3; o++) logits[s * 3 + o] = t.z4[o];
  }
}

float oracle_mlp_loss_grad_f32(const float* params, const float* feats, const uint8_t* labels,
                               const float* weights, int64_t n, float* grad, int bf16) {
  if (bf16)
    return oracle::mlp::loss_and_grad<float, true>(params, feats, labels, weights, n, grad);
  return oracle::mlp::loss_and_grad<float, false>(params, feats, labels, weights, n, grad);
}

void oracle_mlp_forward_f64(const double* params, const double* feats, int64_t n,
                            double* logits) {
  oracle::mlp::Trace<double> t;
  for (int64_t s = 0; s < n; s++) {
    oracle::mlp::forward<double, false>(params, feats + s * 12, t);
    for (int o = 0; o < 3; o++) logits[s * 3 + o] = t.z4[o];
  }
}

double oracle_mlp_loss_grad_f64(const double* params, const double* feats, const uint8_t* labels,
                                const double* weights, int64_t n, double* grad) {
  return oracle::mlp::loss_and_grad<double, false>(params, feats, labels, weights, n, grad);
}

// The learned scorer's 12 features of each Gaussian (AoS parameters, flat order of
// tests/common/mod.rs:111-139), as specified next to splat_learn_config.
void oracle_scorer_features_f32(const float* g, int64_t n, const float* grad_accum,
                                const uint32_t* denom, const uint32_t* max_radii,
                                const float* max_grad, const uint32_t* strong,
                                const uint32_t* birth, float thr, float log_size,
                                uint32_t iteration, float* out) {
  const float c0s = 0.28209479177387814f;
  for (int64_t i = 0; i < n; i++) {
    const float* p = g + i * 59;
    float* f = out + i * 12;
    const uint32_t d = denom[i];
    const float mean = d > 0 ? grad_accum[i] / (float)d : 0.0f;
    f[0] = log2f(mean / thr + 1e-6f) * 0.125f;
    f[1] = log2f(max_grad[i] / thr + 1e-6f) * 0.125f;
    f[2] = (float)strong[i] / (float)(d > 1u ? d : 1u);
    for (int k = 0; k < 3; k++) f[3 + k] = (p[3 + k] - log_size) * 0.25f;
    f[6] = 1.0f / (1.0f + expf(-p[10]));
    f[7] = log2f(1.0f + (float)max_radii[i]) * 0.125f;
    f[8] = log2f(1.0f + (float)d) * 0.125f;
    const float r = c0s * p[11] + 0.5f, gg = c0s * p[12] + 0.5f, b = c0s * p[13] + 0.5f;
    f[9] = 0.299f * r + 0.587f * gg + 0.114f * b;
    f[10] = log2f(1.0f + (float)(iteration - birth[i])) / 12.0f;
    f[11] = (fmaxf(fmaxf(p[3], p[4]), p[5]) - log_size) * 0.25f;
  }
}

// Threads of the parallel loops (1 = the scalar restatement); returns the count in use.
int oracle_set_threads(int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
  return omp_get_max_threads();
#else
  (void)threads;
  return 1;
#endif
}

void oracle_bf16_round(const float* in, float* out, int64_t n) {
  for (int64_t i = 0; i < n; i++) out[i] = oracle::mlp::bf16_round(in[i]);
}

void oracle_mlp_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1,
                     float b2, float eps, float c1, float c2) {
  oracle::mlp::adam(p, g, m, v, n, lr, b1, b2, eps, c1, c2);
}

}  // extern "C"
