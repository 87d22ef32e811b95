// Projection of Gaussians to screen space and its backward pass (src/project.rs).
//
// One lane per Gaussian; the 59 parameter planes are read with unit stride across a wave.
// The arithmetic restates src/project.rs:137-474 in the reference's evaluation order
// (compiled without FMA contraction) so the screen records match it bit for bit.
#include <cstring>

#include "internal.hpp"

namespace splat {
namespace {

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

struct M3 {
  float m[3][3];
};

// SH constants, src/sh.rs:16-33 (f64 literals rounded to f32).
#define SH_C0 ((float)0.28209479177387814)
#define SH_C1 ((float)0.4886025119029199)
#define SH_C20 ((float)1.0925484305920792)
#define SH_C21 ((float)-1.0925484305920792)
#define SH_C22 ((float)0.31539156525252005)
#define SH_C23 ((float)-1.0925484305920792)
#define SH_C24 ((float)0.5462742152960396)
#define SH_C30 ((float)-0.5900435899266435)
#define SH_C31 ((float)2.890611442640554)
#define SH_C32 ((float)-0.4570457994644658)
#define SH_C33 ((float)0.3731763325901154)
#define SH_C34 ((float)-0.4570457994644658)
#define SH_C35 ((float)1.445305721320277)
#define SH_C36 ((float)-0.5900435899266435)

// basis, src/sh.rs:47-75.
__device__ __forceinline__ void sh_basis(int degree, V3 d, float b[16]) {
#pragma unroll
  for (int i = 0; i < 16; i++) b[i] = 0.0f;
  const float x = d.x, y = d.y, z = d.z;
  b[0] = SH_C0;
  if (degree >= 1) {
    b[1] = -SH_C1 * y;
    b[2] = SH_C1 * z;
    b[3] = -SH_C1 * x;
  }
  if (degree >= 2) {
    const float xx = x * x, yy = y * y, zz = z * z;
    const float xy = x * y, yz = y * z, xz = x * z;
    b[4] = SH_C20 * xy;
    b[5] = SH_C21 * yz;
    b[6] = SH_C22 * (2.0f * zz - xx - yy);
    b[7] = SH_C23 * xz;
    b[8] = SH_C24 * (xx - yy);
    if (degree >= 3) {
      b[9] = SH_C30 * y * (3.0f * xx - yy);
      b[10] = SH_C31 * xy * z;
      b[11] = SH_C32 * y * (4.0f * zz - xx - yy);
      b[12] = SH_C33 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
      b[13] = SH_C34 * x * (4.0f * zz - xx - yy);
      b[14] = SH_C35 * z * (xx - yy);
      b[15] = SH_C36 * x * (xx - 3.0f * yy);
    }
  }
}

// basis_grad, src/sh.rs:80-130.
__device__ __forceinline__ void sh_basis_grad(int degree, V3 d, V3 g[16]) {
#pragma unroll
  for (int i = 0; i < 16; i++) g[i] = V3{0.0f, 0.0f, 0.0f};
  const float x = d.x, y = d.y, z = d.z;
  if (degree >= 1) {
    g[1] = V3{0.0f, -SH_C1, 0.0f};
    g[2] = V3{0.0f, 0.0f, SH_C1};
    g[3] = V3{-SH_C1, 0.0f, 0.0f};
  }
  if (degree >= 2) {
    const float xx = x * x, yy = y * y, zz = z * z;
    g[4] = V3{SH_C20 * y, SH_C20 * x, 0.0f};
    g[5] = V3{0.0f, SH_C21 * z, SH_C21 * y};
    g[6] = V3{-2.0f * SH_C22 * x, -2.0f * SH_C22 * y, 4.0f * SH_C22 * z};
    g[7] = V3{SH_C23 * z, 0.0f, SH_C23 * x};
    g[8] = V3{2.0f * SH_C24 * x, -2.0f * SH_C24 * y, 0.0f};
    if (degree >= 3) {
      g[9] = V3{SH_C30 * 6.0f * x * y, SH_C30 * (3.0f * xx - 3.0f * yy), 0.0f};
      g[10] = V3{SH_C31 * y * z, SH_C31 * x * z, SH_C31 * x * y};
      g[11] = V3{SH_C32 * (-2.0f * x * y), SH_C32 * (4.0f * zz - xx - 3.0f * yy),
                 SH_C32 * 8.0f * y * z};
      g[12] = V3{SH_C33 * (-6.0f * x * z), SH_C33 * (-6.0f * y * z),
                 SH_C33 * (6.0f * zz - 3.0f * xx - 3.0f * yy)};
      g[13] = V3{SH_C34 * (4.0f * zz - 3.0f * xx - yy), SH_C34 * (-2.0f * x * y),
                 SH_C34 * 8.0f * x * z};
      g[14] = V3{SH_C35 * 2.0f * x * z, -SH_C35 * 2.0f * y * z, SH_C35 * (xx - yy)};
      g[15] = V3{SH_C36 * (3.0f * xx - 3.0f * yy), SH_C36 * (-6.0f * x * y), 0.0f};
    }
  }
}

// Everything core() computes (src/project.rs:96-125).
struct Core {
  V3 t;
  float txc, tyc;
  bool clamped_x, clamped_y;
  V3 s;
  float q[4];
  M3 rq, m, sigma;
  float tmat[2][3];
  float a0, b0, c0, det0, a, b, c, det;
  float conic[3];
  float comp;
  float mean2d[2];
  float radius;
  float o_raw;
  V3 dir;
  float dist;
  float basis[16];
  float color_raw[3];
};

__device__ __forceinline__ float clamp_flag(float v, float lo, float hi, bool* f) {
  if (v < lo) {
    *f = true;
    return lo;
  }
  if (v > hi) {
    *f = true;
    return hi;
  }
  *f = false;
  return v;
}

__device__ __forceinline__ float sigmoidf_(float x) { return 1.0f / (1.0f + splat_expf(-x)); }

// core, src/project.rs:137-269.  g: the 59 parameters.
__device__ bool core(const float* g, const CameraDev& cam, float near_plane, float eps2d,
                     bool antialias, int degree, Core& k) {
  const V3 p{g[0], g[1], g[2]};
  const V3 r0{cam.rot[0], cam.rot[1], cam.rot[2]};
  const V3 r1{cam.rot[3], cam.rot[4], cam.rot[5]};
  const V3 r2{cam.rot[6], cam.rot[7], cam.rot[8]};
  k.t = V3{dot3(r0, p) + cam.trans[0], dot3(r1, p) + cam.trans[1], dot3(r2, p) + cam.trans[2]};
  if (__builtin_isnan(k.t.z) || k.t.z <= near_plane) return false;
  k.s = V3{splat_expf(g[3]), splat_expf(g[4]), splat_expf(g[5])};
  {  // Quat::normalized, src/math.rs:294-306
    const float w = g[6], x = g[7], y = g[8], z = g[9];
    const float n = sqrtf(w * w + x * x + y * y + z * z);
    if (n > 0.0f) {
      k.q[0] = w / n;
      k.q[1] = x / n;
      k.q[2] = y / n;
      k.q[3] = z / n;
    } else {
      k.q[0] = 1.0f;
      k.q[1] = k.q[2] = k.q[3] = 0.0f;
    }
  }
  {  // Quat::to_mat, src/math.rs:309-330
    const float w = k.q[0], x = k.q[1], y = k.q[2], z = k.q[3];
    k.rq.m[0][0] = 1.0f - 2.0f * (y * y + z * z);
    k.rq.m[0][1] = 2.0f * (x * y - w * z);
    k.rq.m[0][2] = 2.0f * (x * z + w * y);
    k.rq.m[1][0] = 2.0f * (x * y + w * z);
    k.rq.m[1][1] = 1.0f - 2.0f * (x * x + z * z);
    k.rq.m[1][2] = 2.0f * (y * z - w * x);
    k.rq.m[2][0] = 2.0f * (x * z - w * y);
    k.rq.m[2][1] = 2.0f * (y * z + w * x);
    k.rq.m[2][2] = 1.0f - 2.0f * (x * x + y * y);
  }
  for (int i = 0; i < 3; i++) {
    k.m.m[i][0] = k.rq.m[i][0] * k.s.x;
    k.m.m[i][1] = k.rq.m[i][1] * k.s.y;
    k.m.m[i][2] = k.rq.m[i][2] * k.s.z;
  }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      float acc = 0.0f;
      for (int kk = 0; kk < 3; kk++) acc += k.m.m[i][kk] * k.m.m[j][kk];
      k.sigma.m[i][j] = acc;
    }

  const float W = (float)cam.width, H = (float)cam.height;
  const float tan_x = 0.5f * W / cam.fx;
  const float tan_y = 0.5f * H / cam.fy;
  const float lim_x_pos = (W - cam.cx) / cam.fx + 0.3f * tan_x;
  const float lim_x_neg = cam.cx / cam.fx + 0.3f * tan_x;
  const float lim_y_pos = (H - cam.cy) / cam.fy + 0.3f * tan_y;
  const float lim_y_neg = cam.cy / cam.fy + 0.3f * tan_y;
  const float rz = 1.0f / k.t.z;
  const float rx = k.t.x * rz;
  const float ry = k.t.y * rz;
  const float rxc = clamp_flag(rx, -lim_x_neg, lim_x_pos, &k.clamped_x);
  const float ryc = clamp_flag(ry, -lim_y_neg, lim_y_pos, &k.clamped_y);
  k.txc = rxc * k.t.z;
  k.tyc = ryc * k.t.z;
  const float rz2 = rz * rz;
  const float j[2][3] = {{cam.fx * rz, 0.0f, -cam.fx * k.txc * rz2},
                         {0.0f, cam.fy * rz, -cam.fy * k.tyc * rz2}};
  for (int r = 0; r < 2; r++)
    for (int c = 0; c < 3; c++) {
      float acc = 0.0f;
      for (int kk = 0; kk < 3; kk++) acc += j[r][kk] * cam.rot[3 * kk + c];
      k.tmat[r][c] = acc;
    }
  float ts[2][3];
  for (int r = 0; r < 2; r++)
    for (int c = 0; c < 3; c++) {
      float acc = 0.0f;
      for (int kk = 0; kk < 3; kk++) acc += k.tmat[r][kk] * k.sigma.m[kk][c];
      ts[r][c] = acc;
    }
  float a0 = 0.0f, b0 = 0.0f, c0 = 0.0f;
  for (int c = 0; c < 3; c++) {
    a0 += ts[0][c] * k.tmat[0][c];
    b0 += ts[0][c] * k.tmat[1][c];
    c0 += ts[1][c] * k.tmat[1][c];
  }
  k.a0 = a0;
  k.b0 = b0;
  k.c0 = c0;
  k.det0 = a0 * c0 - b0 * b0;
  k.a = a0 + eps2d;
  k.b = b0;
  k.c = c0 + eps2d;
  k.det = k.a * k.c - k.b * k.b;
  if (__builtin_isnan(k.det) || k.det <= 0.0f) return false;
  const float inv_det = 1.0f / k.det;
  k.conic[0] = k.c * inv_det;
  k.conic[1] = -k.b * inv_det;
  k.conic[2] = k.a * inv_det;
  if (antialias)
    k.comp = k.det0 > 0.0f ? sqrtf(k.det0 * inv_det) : 0.0f;
  else
    k.comp = 1.0f;
  const float mid = 0.5f * (k.a + k.c);
  const float lambda1 = mid + sqrtf(rmax(mid * mid - k.det, 0.1f));
  k.radius = ceilf(3.0f * sqrtf(lambda1));
  k.mean2d[0] = cam.fx * rx + cam.cx;
  k.mean2d[1] = cam.fy * ry + cam.cy;
  k.o_raw = sigmoidf_(g[10]);

  // Camera::center, src/camera.rs:41-43: -(rot^T trans).
  const V3 tr{cam.trans[0], cam.trans[1], cam.trans[2]};
  const V3 center{-dot3(V3{cam.rot[0], cam.rot[3], cam.rot[6]}, tr),
                  -dot3(V3{cam.rot[1], cam.rot[4], cam.rot[7]}, tr),
                  -dot3(V3{cam.rot[2], cam.rot[5], cam.rot[8]}, tr)};
  const V3 d{p.x - center.x, p.y - center.y, p.z - center.z};
  k.dist = sqrtf(dot3(d, d));
  const float inv = 1.0f / k.dist;
  k.dir = V3{d.x * inv, d.y * inv, d.z * inv};
  sh_basis(degree, k.dir, k.basis);
  const int n = (degree + 1) * (degree + 1);
  k.color_raw[0] = k.color_raw[1] = k.color_raw[2] = 0.5f;
#pragma unroll
  for (int kk = 0; kk < 16; kk++) {
    if (kk < n) {
#pragma unroll
      for (int ch = 0; ch < 3; ch++) k.color_raw[ch] += k.basis[kk] * g[11 + 3 * kk + ch];
    }
  }
  return true;
}

__device__ __forceinline__ uint32_t clamp_tile(float v, uint32_t n) {  // src/project.rs:287-295
  if (v <= 0.0f) return 0u;
  if (v >= (float)n) return n;
  return as_u32(v);
}

// Rasterizer record of one screen-space Gaussian (TileGauss::new's per-Gaussian part,
// src/rasterize.rs:149-157): sigma_max = ln(opacity / (1/255)) and the half extents of
// the ellipse's bounding box; half_x = -1 marks an empty footprint.
__device__ __forceinline__ void raster_record(float mx, float my, float A, float B, float C,
                                              float o, float r, float gcol, float bcol,
                                              float4* rec0, float4* rec1, float4* rec2) {
  const float smax = splat_logf(o / kAlphaMin);
  const float det = A * C - B * B;
  float hx = -1.0f, hy = -1.0f;
  if (smax > 0.0f && det > 0.0f) {
    hx = sqrtf(2.0f * smax * C / det) + 1e-3f;
    hy = sqrtf(2.0f * smax * A / det) + 1e-3f;
  }
  *rec0 = make_float4(mx, my, A, B);
  *rec1 = make_float4(C, o, smax, hx);
  *rec2 = make_float4(hy, r, gcol, bcol);
}

// project_one (src/project.rs:277-330) of parameter row pi into projection row i.
__device__ __forceinline__ void load_row(const float* __restrict__ params, size_t stride,
                                         size_t pi, int degree, float g[kParams]) {
  const int ncoef = (degree + 1) * (degree + 1);
#pragma unroll
  for (int j = 0; j < kParams; j++) {
    // SH coefficients beyond the active degree are never read by the forward pass.  The
    // parameter planes are streamed once per pass: non-temporal loads.
    g[j] = (j < 11 + 3 * ncoef) ? __builtin_nontemporal_load(params + (size_t)j * stride + pi)
                                : 0.0f;
  }
}

// project_one of the parameters g into projection row i.
__device__ __forceinline__ void project_g(const float g[kParams], const CameraDev& cam,
                                          float near_plane, float eps2d, int antialias,
                                          int degree, const ProjBuffers& pb, size_t i) {
  Core k;
  bool ok = core(g, cam, near_plane, eps2d, antialias != 0, degree, k);
  uint32_t tiles_x = (cam.width + kTile - 1) / kTile, tiles_y = (cam.height + kTile - 1) / kTile;
  uint32_t tx0 = 0, tx1 = 0, ty0 = 0, ty1 = 0;
  const float r = k.radius;
  if (ok) {
    if (__builtin_isnan(r) || r <= 0.0f) ok = false;
  }
  if (ok) {
    const float tile = (float)kTile;
    tx0 = clamp_tile(floorf((k.mean2d[0] - r) / tile), tiles_x);
    tx1 = clamp_tile(ceilf((k.mean2d[0] + r) / tile), tiles_x);
    ty0 = clamp_tile(floorf((k.mean2d[1] - r) / tile), tiles_y);
    ty1 = clamp_tile(ceilf((k.mean2d[1] + r) / tile), tiles_y);
    if (tx0 >= tx1 || ty0 >= ty1) ok = false;
  }
  if (!ok) {
    pb.rec[4 * (size_t)i] = make_float4(0.f, 0.f, 0.f, 0.f);
    pb.rec[4 * (size_t)i + 1] = make_float4(0.f, 0.f, 0.f, -1.f);
    pb.rec[4 * (size_t)i + 2] = make_float4(-1.f, 0.f, 0.f, 0.f);
    pb.depth[i] = 0.0f;
    pb.radius[i] = 0u;
    pb.rect[i] = make_uint2(0u, 0u);
    pb.ntiles[i] = 0u;
    return;
  }
  float4 a, b, c;
  raster_record(k.mean2d[0], k.mean2d[1], k.conic[0], k.conic[1], k.conic[2],
                k.o_raw * k.comp, rmax(k.color_raw[0], 0.0f), rmax(k.color_raw[1], 0.0f),
                rmax(k.color_raw[2], 0.0f), &a, &b, &c);
  pb.rec[4 * (size_t)i] = a;
  pb.rec[4 * (size_t)i + 1] = b;
  pb.rec[4 * (size_t)i + 2] = c;
  pb.depth[i] = k.t.z;
  pb.radius[i] = as_u32(r);
  pb.rect[i] = make_uint2(tx0 | (ty0 << 16), tx1 | (ty1 << 16));
  pb.ntiles[i] = (tx1 - tx0) * (ty1 - ty0);
}

__device__ __forceinline__ void project_row(const float* __restrict__ params, size_t stride,
                                            size_t pi, const CameraDev& cam, float near_plane,
                                            float eps2d, int antialias, int degree,
                                            const ProjBuffers& pb, size_t i) {
  float g[kParams];
  load_row(params, stride, pi, degree, g);
  project_g(g, cam, near_plane, eps2d, antialias, degree, pb, i);
}

__global__ void __launch_bounds__(256) project_fwd_kernel(const float* __restrict__ params,
                                                          uint32_t n, uint32_t stride,
                                                          CameraDev cam, float near_plane,
                                                          float eps2d, int antialias, int degree,
                                                          ProjBuffers pb,
                                                          const StepDyn* __restrict__ dyn) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (dyn) cam = dyn->cam;
  project_row(params, stride, i, cam, near_plane, eps2d, antialias, degree, pb, i);
}

// All (scene, view) pairs of a batched step: row w of the projection buffers is Gaussian
// w - work_base[v] of pair v's scene seen through pair v's camera.
// Batched: one lane per scene row (scene rows start at multiples of 64, so a wave's rows
// share a scene); the row's parameters are read once and projected for each of its scene's
// views into that view's work row.
__global__ void __launch_bounds__(256) project_fwd_batch_kernel(const float* __restrict__ params,
                                                                uint32_t stride, BatchDev bd,
                                                                uint32_t rows, float near_plane,
                                                                float eps2d, int antialias,
                                                                int degree, ProjBuffers pb) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= rows) return;
  const uint32_t s = find_segment(bd.scene_off, bd.nscenes, p);
  const uint32_t i = p - bd.scene_off[s];
  const uint32_t k0 = bd.scene_pair_begin[s], k1 = bd.scene_pair_begin[s + 1];
  if (i >= bd.scene_n[s] || k0 == k1) return;  // padding, or a scene without views
  float g[kParams];
  load_row(params, stride, p, degree, g);
  for (uint32_t k = k0; k < k1; k++) {
    const uint32_t v = bd.scene_pairs[k];
    project_g(g, bd.cams[v], near_plane, eps2d, antialias, degree, pb, bd.work_base[v] + i);
  }
}

// ProjBuffers -> Proj records (src/project.rs:47-64).
__global__ void proj_pack_kernel(ProjBuffers pb, uint32_t n, splat_proj* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  splat_proj p;
  const float4 a = pb.rec[4 * (size_t)i], b = pb.rec[4 * (size_t)i + 1],
               c = pb.rec[4 * (size_t)i + 2];
  const uint32_t rad = pb.radius[i];
  if (rad == 0) {
    memset(&p, 0, sizeof(p));
  } else {
    const uint2 rc = pb.rect[i];
    p.mean2d[0] = a.x;
    p.mean2d[1] = a.y;
    p.conic[0] = a.z;
    p.conic[1] = a.w;
    p.conic[2] = b.x;
    p.depth = pb.depth[i];
    p.radius = rad;
    p.opacity = b.y;
    p.color[0] = c.y;
    p.color[1] = c.z;
    p.color[2] = c.w;
    p.tile_min[0] = rc.x & 0xffffu;
    p.tile_min[1] = rc.x >> 16;
    p.tile_max[0] = rc.y & 0xffffu;
    p.tile_max[1] = rc.y >> 16;
  }
  out[i] = p;
}

// Proj records -> ProjBuffers, checking the rectangle against the tile grid (the
// precondition of src/rasterize.rs:65-69).
__global__ void proj_unpack_kernel(const splat_proj* __restrict__ in, uint32_t n,
                                   uint32_t tiles_x, uint32_t tiles_y, ProjBuffers pb,
                                   uint32_t* bad) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const splat_proj p = in[i];
  if (p.radius == 0) {
    pb.rec[4 * (size_t)i] = make_float4(0.f, 0.f, 0.f, 0.f);
    pb.rec[4 * (size_t)i + 1] = make_float4(0.f, 0.f, 0.f, -1.f);
    pb.rec[4 * (size_t)i + 2] = make_float4(-1.f, 0.f, 0.f, 0.f);
    pb.depth[i] = p.depth;
    pb.radius[i] = 0u;
    pb.rect[i] = make_uint2(0u, 0u);
    pb.ntiles[i] = 0u;
    return;
  }
  // Rectangles outside the grid (the reference asserts) or inverted (its u32 tile count
  // would wrap) are rejected before any entry is generated from them.
  if (p.tile_max[0] > tiles_x || p.tile_max[1] > tiles_y || p.tile_min[0] > p.tile_max[0] ||
      p.tile_min[1] > p.tile_max[1] || p.tile_max[0] > 0xffffu || p.tile_max[1] > 0xffffu)
    atomicOr(bad, 1u);
  float4 a, b, c;
  raster_record(p.mean2d[0], p.mean2d[1], p.conic[0], p.conic[1], p.conic[2], p.opacity,
                p.color[0], p.color[1], p.color[2], &a, &b, &c);
  pb.rec[4 * (size_t)i] = a;
  pb.rec[4 * (size_t)i + 1] = b;
  pb.rec[4 * (size_t)i + 2] = c;
  pb.depth[i] = p.depth;
  pb.radius[i] = p.radius;
  const uint32_t x0 = p.tile_min[0], y0 = p.tile_min[1], x1 = p.tile_max[0], y1 = p.tile_max[1];
  pb.rect[i] = make_uint2((x0 & 0xffffu) | (y0 << 16), (x1 & 0xffffu) | (y1 << 16));
  // Proj::num_tiles (src/project.rs:73-79), wrapping like u32 arithmetic.
  pb.ntiles[i] = (x1 - x0) * (y1 - y0);
}

// project_backward_one, src/project.rs:335-474.
// Gradient of one Gaussian (parameters g) for one view: put(j, x) receives the gradient x
// of each plane j that has a nonzero term, at most once per plane; the other planes' are
// zero.
// sh_put(ch, dc, basis, nco) receives, for each colour channel with a nonzero term, the
// factors of that channel's SH gradients: plane 11 + 3 kk + ch gets dc * basis[kk] for
// kk < nco (src/project.rs:346-360).
template <class Put, class ShPut>
__device__ __forceinline__ void project_bwd_put_sh(const float* g, const CameraDev& cam,
                                                   float near_plane, float eps2d, int antialias,
                                                   int degree,
                                                   const splat_proj_grad* __restrict__ dproj_row,
                                                   Put&& put, ShPut&& sh_put) {
  Core k;
  if (core(g, cam, near_plane, eps2d, antialias != 0, degree, k)) {
    const splat_proj_grad dp = *dproj_row;
    const int nco = (degree + 1) * (degree + 1);
    V3 bgrad[16];
    sh_basis_grad(degree, k.dir, bgrad);
    V3 ddir{0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ch = 0; ch < 3; ch++) {
      if (k.color_raw[ch] > 0.0f) {
        const float dc = dp.color[ch];
        sh_put(ch, dc, k.basis, nco);
#pragma unroll
        for (int kk = 0; kk < 16; kk++) {
          if (kk < nco) {
            const float w = dc * g[11 + 3 * kk + ch];
            ddir.x += bgrad[kk].x * w;
            ddir.y += bgrad[kk].y * w;
            ddir.z += bgrad[kk].z * w;
          }
        }
      }
    }
    const float dd = dot3(k.dir, ddir);
    const float idist = 1.0f / k.dist;
    const V3 dmean_color{(ddir.x - k.dir.x * dd) * idist, (ddir.y - k.dir.y * dd) * idist,
                         (ddir.z - k.dir.z * dd) * idist};

    const float do_raw = dp.opacity * k.comp;
    const float dcomp = dp.opacity * k.o_raw;
    put(10, do_raw * k.o_raw * (1.0f - k.o_raw));

    const float a = k.a, b = k.b, c = k.c, det = k.det;
    const float da_ = dp.conic[0], db_ = dp.conic[1], dc_ = dp.conic[2];
    const float inv_det2 = 1.0f / (det * det);
    float da = (-c * c * da_ + b * c * db_ - b * b * dc_) * inv_det2;
    float db = (2.0f * b * c * da_ - (a * c + b * b) * db_ + 2.0f * a * b * dc_) * inv_det2;
    float dcc = (-b * b * da_ + a * b * db_ - a * a * dc_) * inv_det2;
    if (antialias && k.det0 > 0.0f) {
      const float ddet0 = dcomp * 0.5f * k.comp / k.det0;
      const float ddet = -dcomp * 0.5f * k.comp / det;
      da += ddet0 * k.c0 + ddet * c;
      dcc += ddet0 * k.a0 + ddet * a;
      db += ddet0 * (-2.0f * k.b0) + ddet * (-2.0f * b);
    }
    const float gm[2][2] = {{da, 0.5f * db}, {0.5f * db, dcc}};
    float gt[2][3];
    for (int r = 0; r < 2; r++)
      for (int col = 0; col < 3; col++)
        gt[r][col] = gm[r][0] * k.tmat[0][col] + gm[r][1] * k.tmat[1][col];
    M3 dsigma;
    for (int ii = 0; ii < 3; ii++)
      for (int jj = 0; jj < 3; jj++)
        dsigma.m[ii][jj] = k.tmat[0][ii] * gt[0][jj] + k.tmat[1][ii] * gt[1][jj];
    float dt[2][3];
    for (int r = 0; r < 2; r++)
      for (int col = 0; col < 3; col++) {
        float acc = 0.0f;
        for (int kk = 0; kk < 3; kk++) acc += gt[r][kk] * k.sigma.m[kk][col];
        dt[r][col] = 2.0f * acc;
      }
    float dj[2][3];
    for (int r = 0; r < 2; r++)
      for (int col = 0; col < 3; col++) {
        float acc = 0.0f;
        for (int kk = 0; kk < 3; kk++) acc += dt[r][kk] * cam.rot[3 * col + kk];
        dj[r][col] = acc;
      }
    const float fx = cam.fx, fy = cam.fy;
    const float rz = 1.0f / k.t.z;
    const float rz2 = rz * rz;
    const float rz3 = rz2 * rz;
    const float dtxc = dj[0][2] * (-fx * rz2);
    const float dtyc = dj[1][2] * (-fy * rz2);
    float dtx = 0.0f, dty = 0.0f;
    float dtz = dj[0][0] * (-fx * rz2) + dj[1][1] * (-fy * rz2) +
                dj[0][2] * (2.0f * fx * k.txc * rz3) + dj[1][2] * (2.0f * fy * k.tyc * rz3);
    if (k.clamped_x)
      dtz += dtxc * k.txc * rz;
    else
      dtx += dtxc;
    if (k.clamped_y)
      dtz += dtyc * k.tyc * rz;
    else
      dty += dtyc;
    dtx += dp.mean2d[0] * fx * rz;
    dtz += dp.mean2d[0] * (-fx * k.t.x * rz2);
    dty += dp.mean2d[1] * fy * rz;
    dtz += dp.mean2d[1] * (-fy * k.t.y * rz2);
    const V3 dtv{dtx, dty, dtz};
    // rot^T * dt (rows of the transpose are the columns of rot)
    put(0, dot3(V3{cam.rot[0], cam.rot[3], cam.rot[6]}, dtv) + dmean_color.x);
    put(1, dot3(V3{cam.rot[1], cam.rot[4], cam.rot[7]}, dtv) + dmean_color.y);
    put(2, dot3(V3{cam.rot[2], cam.rot[5], cam.rot[8]}, dtv) + dmean_color.z);

    M3 dm;
    for (int ii = 0; ii < 3; ii++)
      for (int jj = 0; jj < 3; jj++) {
        float acc = 0.0f;
        for (int kk = 0; kk < 3; kk++) acc += dsigma.m[ii][kk] * k.m.m[kk][jj];
        dm.m[ii][jj] = acc * 2.0f;
      }
    float ds[3] = {0.0f, 0.0f, 0.0f};
    M3 dr;
    const float sv[3] = {k.s.x, k.s.y, k.s.z};
    for (int r = 0; r < 3; r++)
      for (int col = 0; col < 3; col++) {
        ds[col] += dm.m[r][col] * k.rq.m[r][col];
        dr.m[r][col] = dm.m[r][col] * sv[col];
      }
    for (int ii = 0; ii < 3; ii++) put(3 + ii, ds[ii] * sv[ii]);
    // quat_to_mat_backward, src/math.rs:336-372.
    const float w = k.q[0], x = k.q[1], y = k.q[2], z = k.q[3];
    const float Dw[3][3] = {{0.0f, -2.0f * z, 2.0f * y},
                            {2.0f * z, 0.0f, -2.0f * x},
                            {-2.0f * y, 2.0f * x, 0.0f}};
    const float Dx[3][3] = {{0.0f, 2.0f * y, 2.0f * z},
                            {2.0f * y, -4.0f * x, -2.0f * w},
                            {2.0f * z, 2.0f * w, -4.0f * x}};
    const float Dy[3][3] = {{-4.0f * y, 2.0f * x, 2.0f * w},
                            {2.0f * x, 0.0f, 2.0f * z},
                            {-2.0f * w, 2.0f * z, -4.0f * y}};
    const float Dz[3][3] = {{-4.0f * z, -2.0f * w, 2.0f * x},
                            {2.0f * w, -4.0f * z, 2.0f * y},
                            {2.0f * x, 2.0f * y, 0.0f}};
    float dq[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int ii = 0; ii < 3; ii++)
      for (int jj = 0; jj < 3; jj++) {
        dq[0] += dr.m[ii][jj] * Dw[ii][jj];
        dq[1] += dr.m[ii][jj] * Dx[ii][jj];
        dq[2] += dr.m[ii][jj] * Dy[ii][jj];
        dq[3] += dr.m[ii][jj] * Dz[ii][jj];
      }
    // normalize_backward, src/math.rs:375-387.
    const float* raw = g + 6;
    const float nn = sqrtf(raw[0] * raw[0] + raw[1] * raw[1] + raw[2] * raw[2] + raw[3] * raw[3]);
    if (nn > 0.0f) {
      const float u[4] = {raw[0] / nn, raw[1] / nn, raw[2] / nn, raw[3] / nn};
      const float dot = u[0] * dq[0] + u[1] * dq[1] + u[2] * dq[2] + u[3] * dq[3];
      for (int t = 0; t < 4; t++) put(6 + t, (dq[t] - u[t] * dot) / nn);
    }
  }
}

template <class Put>
__device__ __forceinline__ void project_bwd_put(const float* g, const CameraDev& cam,
                                                float near_plane, float eps2d, int antialias,
                                                int degree,
                                                const splat_proj_grad* __restrict__ dproj_row,
                                                Put&& put) {
  project_bwd_put_sh(g, cam, near_plane, eps2d, antialias, degree, dproj_row, put,
                     [&](int ch, float dc, const float* basis, int nco) {
#pragma unroll
                       for (int kk = 0; kk < 16; kk++)
                         if (kk < nco) put(11 + 3 * kk + ch, dc * basis[kk]);
                     });
}

// The same gradient into out[j * ostride]: zeroed and written, or with ACC added to what
// out holds.
template <bool ACC = false>
__device__ __forceinline__ void project_bwd_row(const float* g, const CameraDev& cam,
                                                float near_plane, float eps2d, int antialias,
                                                int degree,
                                                const splat_proj_grad* __restrict__ dproj_row,
                                                float* out, int ostride = 1) {
  if (!ACC) {
#pragma unroll
    for (int j = 0; j < kParams; j++) out[j * ostride] = 0.0f;
  }
  project_bwd_put(g, cam, near_plane, eps2d, antialias, degree, dproj_row,
                  [&](int j, float x) {
                    if (ACC)
                      out[j * ostride] += x;
                    else
                      out[j * ostride] = x;
                  });
}

__global__ void __launch_bounds__(256) project_bwd_kernel(const float* __restrict__ params,
                                                          uint32_t n, uint32_t stride,
                                                          CameraDev cam, float near_plane,
                                                          float eps2d, int antialias, int degree,
                                                          const splat_proj_grad* __restrict__ dproj,
                                                          float* __restrict__ grads,
                                                          const StepDyn* __restrict__ dyn) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (dyn) cam = dyn->cam;
  float g[kParams];
#pragma unroll
  for (int j = 0; j < kParams; j++) g[j] = params[(size_t)j * stride + i];
  float out[kParams];
  project_bwd_row(g, cam, near_plane, eps2d, antialias, degree, dproj + i, out);
#pragma unroll
  for (int j = 0; j < kParams; j++) grads[(size_t)j * stride + i] = out[j];
}

// Adam (src/adam.rs:44-61) over a workgroup's 59 x R parameter and moment values, as
// float4 groups of 4 consecutive rows (plane-major) starting at row `ib`; grad(j, r) gives
// the float4 gradient of plane j, rows r..r+3.  kU groups per lane are in flight: their
// parameter and moment loads are issued together, before any of their updates (one group's
// 48 bytes per lane between waits leaves HBM latency exposed at a few waves per SIMD).
// Groups that run past `rows` take the scalar path.  `ib` and `stride` are multiples of 64.
template <uint32_t R, typename GradFn>
__device__ __forceinline__ void adam_groups(float* __restrict__ params, float* __restrict__ m,
                                            float* __restrict__ v, uint32_t stride, uint32_t ib,
                                            uint32_t rows, const splat_learning_rates& lr,
                                            float b1, float b2, float c1, float c2, float eps,
                                            GradFn grad) {
  constexpr uint32_t kPerPlane = R / 4;  // float4 groups per plane
  constexpr uint32_t kGroups = kParams * kPerPlane;
  constexpr uint32_t kU = 4;
  for (uint32_t q0 = threadIdx.x; q0 < kGroups; q0 += kU * R) {
    float4 P[kU], M[kU], V[kU];
#pragma unroll
    for (uint32_t u = 0; u < kU; u++) {
      // unconditional, so the loads issue back to back: a group past the rows (used by the
      // scalar path or not at all) reads rows ib..ib+3 of plane 0 instead, which exist since
      // ib < n <= stride and both ib and stride are multiples of 64
      const uint32_t q = q0 + u * R;
      const uint32_t j = q / kPerPlane, r = (q % kPerPlane) * 4u;
      const size_t at = q < kGroups && r + 4 <= rows ? (size_t)j * stride + ib + r : (size_t)ib;
      // the parameter planes are read and written with ordinary (cached) accesses: the
      // projection backward above has just read them, and the next step's projection
      // reads what is stored here; the moments stream with the non-temporal hint
      P[u] = *reinterpret_cast<const float4*>(params + at);
      M[u] = load_stream(m + at);
      V[u] = load_stream(v + at);
    }
#pragma unroll
    for (uint32_t u = 0; u < kU; u++) {
      const uint32_t q = q0 + u * R;
      const uint32_t j = q / kPerPlane, r = (q % kPerPlane) * 4u;
      if (q >= kGroups || r >= rows) continue;
      const float rate = plane_lr((int)j, lr);
      const size_t at = (size_t)j * stride + ib + r;
      const float4 G = grad(j, r);
      if (r + 4 <= rows) {
        adam_update(P[u].x, G.x, M[u].x, V[u].x, rate, b1, b2, c1, c2, eps);
        adam_update(P[u].y, G.y, M[u].y, V[u].y, rate, b1, b2, c1, c2, eps);
        adam_update(P[u].z, G.z, M[u].z, V[u].z, rate, b1, b2, c1, c2, eps);
        adam_update(P[u].w, G.w, M[u].w, V[u].w, rate, b1, b2, c1, c2, eps);
        *reinterpret_cast<float4*>(params + at) = P[u];
        store_stream(m + at, M[u]);
        store_stream(v + at, V[u]);
      } else {
        const float gq[4] = {G.x, G.y, G.z, G.w};
        for (uint32_t k = 0; r + k < rows; k++)
          adam_update(params[at + k], gq[k], m[at + k], v[at + k], rate, b1, b2, c1, c2, eps);
      }
    }
  }
}

// The per-Gaussian part of a train step's backward and the optimizer in one pass, one
// lane per Gaussian row: the row's slots summed over its entries in entry order
// (src/rasterize.rs:480-503; the workgroup's rows own one contiguous span of entries,
// staged through LDS with coalesced loads), the densification statistics
// (src/densify.rs:36-60, when accumulating) and project_backward_one (src/project.rs:
// 335-474).  Its gradients go to LDS (the staging buffer, reused, [plane][rows]): the 11
// planes before the SH coefficients as values, the 48 SH planes as their factors (the 16
// basis values and the 3 colour gradients of the row, zero where a plane has no term), so
// 30 words per row instead of 59.  After a barrier the workgroup runs Adam
// (src/adam.rs:44-61) over its 59 x kGradAdamRows parameter and moment values as float4
// groups of 4 consecutive rows (plane-major), an SH plane's gradient formed as
// dc * basis, the product project_backward_one takes.  The arithmetic is that of the
// separate host-tier kernels (a plane without a term gets a zero gradient of either sign,
// which Adam treats alike).
constexpr uint32_t kGradAdamRows = 128;  // Gaussian rows per grad_adam workgroup
constexpr uint32_t kGradWords = 11 + 16 + 3;  // LDS words per row: planes 0-10, basis, dc

__global__ void __launch_bounds__(kGradAdamRows) grad_adam_kernel(
    float* __restrict__ params, float* __restrict__ m, float* __restrict__ v, uint32_t n,
    uint32_t stride, CameraDev cam, float near_plane, float eps2d, int antialias, int degree,
    const uint32_t* __restrict__ entry_offset, const float* __restrict__ slots, StatsArgs sa,
    float c1, float c2, float b1, float b2, float eps, splat_learning_rates lr,
    const StepDyn* __restrict__ dyn, const StepStatus* __restrict__ status) {
  constexpr uint32_t R = kGradAdamRows;
  constexpr uint32_t kChunk = 2 * R;  // entries staged per pass
  constexpr uint32_t kShared =
      kChunk * kSlots > kGradWords * R ? kChunk * kSlots : kGradWords * R;
  __shared__ __align__(16) float sh[kShared];
  if (status && status->abort) return;  // uniform: every lane reads the same flag
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (dyn) {
    cam = dyn->cam;
    c1 = dyn->c1;
    c2 = dyn->c2;
    lr = dyn->lr;
  }
  float sum[kSlots];
#pragma unroll
  for (int k = 0; k < kSlots; k++) sum[k] = 0.0f;
  const uint32_t ib = blockIdx.x * blockDim.x, ie = min(ib + blockDim.x, n);
  const uint32_t span0 = entry_offset[ib], span1 = entry_offset[ie];
  const uint32_t e0 = i < n ? entry_offset[i] : 0u, e1 = i < n ? entry_offset[i + 1] : 0u;
  // words per entry: the backward writes the absolute sums (slots 9, 10) exactly when the
  // step accumulates AbsGS statistics
  const bool abs = sa.grad_accum && sa.absgrad;
  const uint32_t ns = abs ? kSlots : 9;
  for (uint32_t c0 = span0; c0 < span1; c0 += kChunk) {
    const uint32_t c1e = min(c0 + kChunk, span1);
    __syncthreads();
    {  // 16 coalesced loads per lane in flight before their LDS stores
      constexpr uint32_t kLoads = 16;
      const uint32_t words = (c1e - c0) * ns;
      const float* src = slots + (size_t)c0 * ns;
      for (uint32_t q0 = threadIdx.x; q0 < words; q0 += kLoads * R) {
        float x[kLoads];
#pragma unroll
        for (uint32_t u = 0; u < kLoads; u++) {
          const uint32_t q = q0 + u * R;
          x[u] = src[min(q, words - 1)];  // in the span: no branch between the loads
        }
#pragma unroll
        for (uint32_t u = 0; u < kLoads; u++) {
          const uint32_t q = q0 + u * R;
          if (q < words) sh[q] = x[u];
        }
      }
    }
    __syncthreads();
    for (uint32_t e = max(e0, c0); e < min(e1, c1e); e++) {
      const float* q = sh + (e - c0) * ns;
#pragma unroll
      for (int k = 0; k < 9; k++) sum[k] += q[k];
      if (abs) {
        sum[9] += q[9];
        sum[10] += q[10];
      }
    }
  }
  __syncthreads();  // the staging buffer now holds gradients
  float* const gw = sh + threadIdx.x;  // this row's column: word w at gw[w * R]
#pragma unroll
  for (uint32_t w = 0; w < kGradWords; w++) gw[w * R] = 0.0f;
  if (i < n) {
  splat_proj_grad pg;
  pg.mean2d[0] = sum[0];
  pg.mean2d[1] = sum[1];
  pg.conic[0] = sum[2];
  pg.conic[1] = sum[3];
  pg.conic[2] = sum[4];
  pg.opacity = sum[5];
  pg.color[0] = sum[6];
  pg.color[1] = sum[7];
  pg.color[2] = sum[8];
  if (sa.grad_accum) {
    const uint32_t r = sa.radius[i];
    if (r != 0) {
      const float g0 = sa.absgrad ? sum[9] : sum[0];
      const float g1 = sa.absgrad ? sum[10] : sum[1];
      const float gx = g0 * sa.sx;
      const float gy = g1 * sa.sy;
      const float norm = sqrtf(gx * gx + gy * gy);
      sa.grad_accum[i] += norm;
      sa.denom[i] += 1u;
      sa.max_radii[i] = max(sa.max_radii[i], r);
      if (sa.max_grad) {
        sa.max_grad[i] = rmax(sa.max_grad[i], norm);
        sa.strong[i] += norm >= sa.strong_threshold ? 1u : 0u;
      }
    }
  }
  float g[kParams];
#pragma unroll
  for (int j = 0; j < kParams; j++) g[j] = params[(size_t)j * stride + i];
  project_bwd_put_sh(g, cam, near_plane, eps2d, antialias, degree, &pg,
                     [&](int j, float x) { gw[j * R] = x; },  // planes 0-10 only
                     [&](int ch, float dc, const float* basis, int nco) {
#pragma unroll
                       for (int kk = 0; kk < 16; kk++)
                         if (kk < nco) gw[(11 + kk) * R] = basis[kk];
                       gw[(27 + ch) * R] = dc;
                     });
  }
  __syncthreads();  // every row's gradient words are in LDS
  adam_groups<R>(params, m, v, stride, ib, ie - ib, lr, b1, b2, c1, c2, eps,
                 [&](uint32_t j, uint32_t r) {
                   if (j < 11) return *reinterpret_cast<const float4*>(sh + j * R + r);
                   const uint32_t kk = (j - 11) / 3, ch = (j - 11) % 3;
                   const float4 dc = *reinterpret_cast<const float4*>(sh + (27 + ch) * R + r);
                   const float4 bs = *reinterpret_cast<const float4*>(sh + (11 + kk) * R + r);
                   return make_float4(dc.x * bs.x, dc.y * bs.y, dc.z * bs.z, dc.w * bs.w);
                 });
}

// Batched backward tail and optimizer, one wave of 64 scene rows per workgroup (scene rows
// start at multiples of 64, so the workgroup's rows share a scene).  For each of the
// scene's views in step order, project_backward_one (src/project.rs:335-474) adds the
// view's parameter gradients (from the row's summed slots, dproj) to accumulators in LDS
// ([plane][row]).  Adam (src/adam.rs:44-61, the scene's own position learning rate) then
// runs over the workgroup's 59 x 64 values as float4 groups: no gradient planes in HBM.
// A view none of whose 64 rows has an entry adds nothing and is skipped: an all-zero slot
// sum gives zero gradients, and Adam's moments and step do not depend on the sign of a
// zero gradient.
constexpr uint32_t kBatchRows = 64;  // scene rows per grad_adam_batch workgroup

__global__ void __launch_bounds__(kBatchRows) grad_adam_batch_kernel(
    float* __restrict__ params, float* __restrict__ m, float* __restrict__ v, uint32_t stride,
    BatchDev bd, float near_plane, float eps2d, int antialias, int degree,
    const uint32_t* __restrict__ entry_offset, const splat_proj_grad* __restrict__ dproj,
    float c1, float c2, float b1, float b2, float eps, splat_learning_rates lr,
    const StepDyn* __restrict__ dyn, const StepStatus* __restrict__ status) {
  constexpr uint32_t R = kBatchRows;
  __shared__ __align__(16) float acc[kParams * R];  // gradient sums, [plane][row]
  if (status && status->abort) return;  // uniform
  if (dyn) {
    c1 = dyn->c1;
    c2 = dyn->c2;
    lr = dyn->lr;
  }
  const uint32_t ib = blockIdx.x * R, t = threadIdx.x;
  const uint32_t s = find_segment(bd.scene_off, bd.nscenes, ib);
  const uint32_t i0 = ib - bd.scene_off[s], ns = bd.scene_n[s];
  const uint32_t rows = i0 < ns ? min(R, ns - i0) : 0u;  // padding rows are left alone
  if (rows == 0) return;  // uniform
  const bool live = t < rows;
  float g[kParams];
#pragma unroll
  for (int j = 0; j < kParams; j++) {
    g[j] = live ? params[(size_t)j * stride + ib + t] : 0.0f;
    acc[j * R + t] = 0.0f;
  }
  const uint32_t k0 = bd.scene_pair_begin[s], k1 = bd.scene_pair_begin[s + 1];
  for (uint32_t k = k0; k < k1; k++) {
    const uint32_t view = bd.scene_pairs[k];
    const uint32_t w0 = bd.work_base[view] + i0;
    if (entry_offset[w0] == entry_offset[w0 + rows]) continue;  // uniform: view sees no row
    splat_proj_grad pg{};
    if (live) pg = dproj[w0 + t];
    project_bwd_row<true>(g, bd.cams[view], near_plane, eps2d, antialias, degree, &pg, acc + t, R);
  }
  __syncthreads();
  lr.mean = bd.lr_mean[s];
  adam_groups<R>(params, m, v, stride, ib, rows, lr, b1, b2, c1, c2, eps,
                 [&](uint32_t j, uint32_t r) {
                   return *reinterpret_cast<const float4*>(acc + j * R + r);
                 });
}

// AoS (n x 59) <-> planes, staged through LDS so both sides are read and written with
// unit stride.  One block handles 64 Gaussians.
__global__ void __launch_bounds__(256) aos_to_planes_kernel(const float* __restrict__ aos,
                                                            uint32_t n, float* __restrict__ planes,
                                                            uint32_t stride) {
  __shared__ float tile[64 * kParams + 64];
  const uint32_t g0 = blockIdx.x * 64;
  const uint32_t cnt = min(64u, n - g0);
  for (uint32_t t = threadIdx.x; t < cnt * kParams; t += blockDim.x)
    tile[(t / kParams) * (kParams + 1) + t % kParams] = aos[(size_t)g0 * kParams + t];
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < 64 * kParams; t += blockDim.x) {
    const uint32_t j = t / 64, gi = t % 64;
    if (gi < cnt) planes[(size_t)j * stride + g0 + gi] = tile[gi * (kParams + 1) + j];
  }
}

__global__ void __launch_bounds__(256) planes_to_aos_kernel(const float* __restrict__ planes,
                                                            uint32_t n, uint32_t stride,
                                                            float* __restrict__ aos) {
  __shared__ float tile[64 * kParams + 64];
  const uint32_t g0 = blockIdx.x * 64;
  const uint32_t cnt = min(64u, n - g0);
  for (uint32_t t = threadIdx.x; t < 64 * kParams; t += blockDim.x) {
    const uint32_t j = t / 64, gi = t % 64;
    if (gi < cnt) tile[gi * (kParams + 1) + j] = planes[(size_t)j * stride + g0 + gi];
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < cnt * kParams; t += blockDim.x)
    aos[(size_t)g0 * kParams + t] = tile[(t / kParams) * (kParams + 1) + t % kParams];
}

}  // namespace

void project_forward(splat_ctx* ctx, const float* params, size_t n, size_t stride,
                     const splat_camera& cam, const splat_proj_opts& opts, ProjBuffers pb,
                     const StepDyn* dyn) {
  launch(ctx, "project_fwd", project_fwd_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, params,
         (uint32_t)n, (uint32_t)stride, camera_dev(cam), opts.near_plane, opts.eps2d,
         opts.antialias, opts.sh_degree, pb, dyn);
}

void proj_pack(splat_ctx* ctx, ProjBuffers pb, size_t n, splat_proj* out) {
  launch(ctx, "proj_pack", proj_pack_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, pb,
         (uint32_t)n, out);
}

void proj_unpack(splat_ctx* ctx, const splat_proj* in, size_t n, uint32_t tiles_x,
                 uint32_t tiles_y, ProjBuffers pb, uint32_t* bad_flag) {
  launch(ctx, "proj_unpack", proj_unpack_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, in,
         (uint32_t)n, tiles_x, tiles_y, pb, bad_flag);
}

void project_backward(splat_ctx* ctx, const float* params, size_t n, size_t stride,
                      const splat_camera& cam, const splat_proj_opts& opts,
                      const splat_proj_grad* dproj, float* grads, const StepDyn* dyn) {
  launch(ctx, "project_bwd", project_bwd_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, params,
         (uint32_t)n, (uint32_t)stride, camera_dev(cam), opts.near_plane, opts.eps2d,
         opts.antialias, opts.sh_degree, dproj, grads, dyn);
}

void grad_adam(splat_ctx* ctx, float* params, float* m, float* v, size_t n, size_t stride,
               const splat_camera& cam, const splat_proj_opts& opts, const uint32_t* entry_offset,
               const float* slots, const StatsArgs& sa, float c1, float c2,
               const splat_adam_state& st, const splat_learning_rates& lr, const StepDyn* dyn,
               const StepStatus* status) {
  launch(ctx, "grad_adam", grad_adam_kernel, dim3(blocks_for(n, kGradAdamRows)),
         dim3(kGradAdamRows), 0, params, m,
         v, (uint32_t)n, (uint32_t)stride, camera_dev(cam), opts.near_plane, opts.eps2d,
         opts.antialias, opts.sh_degree, entry_offset, slots, sa, c1, c2, st.beta1, st.beta2,
         st.eps, lr, dyn, status);
}

void project_forward_batch(splat_ctx* ctx, const float* params, size_t stride,
                           const BatchDev& bd, size_t rows, float near_plane, int antialias,
                           int degree, ProjBuffers pb) {
  launch(ctx, "project_fwd", project_fwd_batch_kernel, dim3(blocks_for(rows, 256)), dim3(256), 0,
         params, (uint32_t)stride, bd, (uint32_t)rows, near_plane, 0.3f, antialias, degree, pb);
}

void grad_adam_batch(splat_ctx* ctx, float* params, float* m, float* v, size_t stride,
                     const BatchDev& bd, size_t rows, float near_plane, int antialias, int degree,
                     const uint32_t* entry_offset, const splat_proj_grad* dproj,
                     const splat_adam_state& st, const StepDyn* dyn, const StepStatus* status) {
  SPLAT_REQUIRE(rows % kBatchRows == 0, "batched rows must be a multiple of 64");
  splat_learning_rates lr0;
  std::memset(&lr0, 0, sizeof(lr0));
  launch(ctx, "grad_adam", grad_adam_batch_kernel, dim3((unsigned)(rows / kBatchRows)),
         dim3(kBatchRows), 0, params, m, v, (uint32_t)stride, bd, near_plane, 0.3f, antialias,
         degree, entry_offset, dproj, 1.0f, 1.0f, st.beta1, st.beta2, st.eps, lr0, dyn, status);
}

void aos_to_planes(splat_ctx* ctx, const float* aos, size_t n, float* planes, size_t stride) {
  launch(ctx, "aos_to_planes", aos_to_planes_kernel, dim3(blocks_for(n, 64)), dim3(256), 0, aos,
         (uint32_t)n, planes, (uint32_t)stride);
}

void planes_to_aos(splat_ctx* ctx, const float* planes, size_t n, size_t stride, float* aos) {
  launch(ctx, "planes_to_aos", planes_to_aos_kernel, dim3(blocks_for(n, 64)), dim3(256), 0,
         planes, (uint32_t)n, (uint32_t)stride, aos);
}

}  // namespace splat
