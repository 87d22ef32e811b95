// CPU oracle for the splat training hot path -- TEST INFRASTRUCTURE ONLY.
//
// An op-for-op restatement of the reference (harryprayiv/splatrs, Rust) used to check
// the HIP library.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
// may load it; the product never links or calls it.
//
// Every function cites the reference function it restates.  Arithmetic follows the
// reference's evaluation order literally (left-to-right sums, the same association of
// products, explicit 0.0 accumulator starts) and is compiled with -ffp-contract=off
// -fno-fast-math, which is what rustc emits for baseline x86-64 (no FMA contraction).
// Transcendentals call glibc expf/logf/exp/log, as Rust's f32::exp / f32::ln do on
// Linux.  Templated on the scalar type R (float = the reference default, double = its
// `f64` cargo feature, src/lib.rs:18-23) so the finite-difference checks can run in f64.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#include <parallel/algorithm>
// Parallel loops along the reference's rayon axes (Gaussians, tiles, image rows, elements).
// Every parallel loop writes disjoint outputs and keeps each output's arithmetic order, so
// results are identical for any thread count.
#define ORACLE_PAR _Pragma("omp parallel for schedule(dynamic, 64)")
#define ORACLE_SORT __gnu_parallel::sort
#else
#define ORACLE_PAR
#define ORACLE_SORT std::sort
#endif

namespace oracle {

// Literal of type R spelled as the reference spells it (`0.3` is an f32 literal in the
// default build, an f64 literal in the f64 build).
#define RL(x) (std::is_same<R, float>::value ? (R)(x##f) : (R)(x))

inline float rexp(float x) { return expf(x); }
inline double rexp(double x) { return exp(x); }
inline float rln(float x) { return logf(x); }
inline double rln(double x) { return log(x); }
inline float rsqrt_(float x) { return sqrtf(x); }
inline double rsqrt_(double x) { return sqrt(x); }
inline float rceil(float x) { return ceilf(x); }
inline double rceil(double x) { return ceil(x); }
inline float rfloor(float x) { return floorf(x); }
inline double rfloor(double x) { return floor(x); }
inline float rabs(float x) { return fabsf(x); }
inline double rabs(double x) { return fabs(x); }
// Rust f32::max / f32::min: IEEE maxNum/minNum (a NaN operand yields the other one).
inline float rmax(float a, float b) { return fmaxf(a, b); }
inline double rmax(double a, double b) { return fmax(a, b); }
inline float rmin(float a, float b) { return fminf(a, b); }
inline double rmin(double a, double b) { return fmin(a, b); }

// Rust `f as u32`: saturating, NaN -> 0.
template <class R>
inline uint32_t as_u32(R v) {
  if (!(v > (R)0)) return 0;
  if (v >= (R)4294967296.0) return 0xffffffffu;
  return (uint32_t)v;
}
// Rust `f.clamp(lo, hi)` (NaN passes through).
template <class R>
inline R rclamp(R v, R lo, R hi) {
  if (v < lo) v = lo;
  if (v > hi) v = hi;
  return v;
}
// Bits of `depth as f32`, src/rasterize.rs:82.
template <class R>
inline uint32_t f32_bits(R v) {
  float f = (float)v;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

// ---------------------------------------------------------------- math.rs ----------

template <class R>
struct V3 {
  R x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(R a, R b, R c) : x(a), y(b), z(c) {}
  R dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }  // math.rs:46-48
  R norm() const { return rsqrt_(dot(*this)); }                      // math.rs:60-67
  V3 cross(const V3& o) const {                                      // math.rs:51-57
    return V3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x);
  }
  V3 normalized() const {  // math.rs:70-77
    R n = norm();
    if (n > (R)0) return (*this) * ((R)1 / n);
    return *this;
  }
  R get(int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
  V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
  V3 operator*(R s) const { return V3(x * s, y * s, z * s); }
  V3 operator-() const { return V3(-x, -y, -z); }
  V3& operator+=(const V3& o) {
    x += o.x;
    y += o.y;
    z += o.z;
    return *this;
  }
  R max_elem() const { return rmax(rmax(x, y), z); }  // math.rs:90-92
};

// Row-major 3x3, math.rs:142-236.
template <class R>
struct M3 {
  R m[3][3];
  M3() {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) m[i][j] = 0;
  }
  V3<R> row(int i) const { return V3<R>(m[i][0], m[i][1], m[i][2]); }
  V3<R> mul_v(const V3<R>& v) const {  // math.rs:174-176
    return V3<R>(row(0).dot(v), row(1).dot(v), row(2).dot(v));
  }
  M3 mul_m(const M3& o) const {  // math.rs:179-191
    M3 r;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        R s = 0;
        for (int k = 0; k < 3; k++) s += m[i][k] * o.m[k][j];
        r.m[i][j] = s;
      }
    return r;
  }
  M3 transpose() const {
    M3 r;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) r.m[i][j] = m[j][i];
    return r;
  }
  M3 scale_cols(const V3<R>& s) const {  // math.rs:205-213
    M3 r = *this;
    for (int i = 0; i < 3; i++) {
      r.m[i][0] *= s.x;
      r.m[i][1] *= s.y;
      r.m[i][2] *= s.z;
    }
    return r;
  }
  M3 scaled(R s) const {  // math.rs:216-224
    M3 r = *this;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) r.m[i][j] *= s;
    return r;
  }
  R frob_dot(const M3& o) const {  // math.rs:227-235
    R s = 0;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) s += m[i][j] * o.m[i][j];
    return s;
  }
};

template <class R>
struct Quat {
  R w, x, y, z;
  R norm() const { return rsqrt_(w * w + x * x + y * y + z * z); }  // math.rs:289-291
  Quat normalized() const {                                          // math.rs:294-306
    R n = norm();
    if (n > (R)0) return Quat{w / n, x / n, y / n, z / n};
    return Quat{1, 0, 0, 0};
  }
  M3<R> to_mat() const {  // math.rs:309-330
    M3<R> r;
    r.m[0][0] = RL(1.0) - RL(2.0) * (y * y + z * z);
    r.m[0][1] = RL(2.0) * (x * y - w * z);
    r.m[0][2] = RL(2.0) * (x * z + w * y);
    r.m[1][0] = RL(2.0) * (x * y + w * z);
    r.m[1][1] = RL(1.0) - RL(2.0) * (x * x + z * z);
    r.m[1][2] = RL(2.0) * (y * z - w * x);
    r.m[2][0] = RL(2.0) * (x * z - w * y);
    r.m[2][1] = RL(2.0) * (y * z + w * x);
    r.m[2][2] = RL(1.0) - RL(2.0) * (x * x + y * y);
    return r;
  }
};

// quat_to_mat_backward, math.rs:336-372.
template <class R>
inline void quat_to_mat_backward(const Quat<R>& q, const M3<R>& dr, R out[4]) {
  const R w = q.w, x = q.x, y = q.y, z = q.z;
  M3<R> dw, dx, dy, dz;
  R a[3][3] = {{0, RL(-2.0) * z, RL(2.0) * y},
               {RL(2.0) * z, 0, RL(-2.0) * x},
               {RL(-2.0) * y, RL(2.0) * x, 0}};
  R b[3][3] = {{0, RL(2.0) * y, RL(2.0) * z},
               {RL(2.0) * y, RL(-4.0) * x, RL(-2.0) * w},
               {RL(2.0) * z, RL(2.0) * w, RL(-4.0) * x}};
  R c[3][3] = {{RL(-4.0) * y, RL(2.0) * x, RL(2.0) * w},
               {RL(2.0) * x, 0, RL(2.0) * z},
               {RL(-2.0) * w, RL(2.0) * z, RL(-4.0) * y}};
  R d[3][3] = {{RL(-4.0) * z, RL(-2.0) * w, RL(2.0) * x},
               {RL(2.0) * w, RL(-4.0) * z, RL(2.0) * y},
               {RL(2.0) * x, RL(2.0) * y, 0}};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      dw.m[i][j] = a[i][j];
      dx.m[i][j] = b[i][j];
      dy.m[i][j] = c[i][j];
      dz.m[i][j] = d[i][j];
    }
  out[0] = dr.frob_dot(dw);
  out[1] = dr.frob_dot(dx);
  out[2] = dr.frob_dot(dy);
  out[3] = dr.frob_dot(dz);
}

// normalize_backward, math.rs:375-387.
template <class R>
inline void normalize_backward(const R raw[4], const R du[4], R out[4]) {
  R n = rsqrt_(raw[0] * raw[0] + raw[1] * raw[1] + raw[2] * raw[2] + raw[3] * raw[3]);
  if (n <= (R)0) {
    out[0] = out[1] = out[2] = out[3] = 0;
    return;
  }
  R u[4] = {raw[0] / n, raw[1] / n, raw[2] / n, raw[3] / n};
  R dot = u[0] * du[0] + u[1] * du[1] + u[2] * du[2] + u[3] * du[3];
  for (int i = 0; i < 4; i++) out[i] = (du[i] - u[i] * dot) / n;
}

template <class R>
inline R sigmoid(R x) {  // math.rs:390-392
  return RL(1.0) / (RL(1.0) + rexp(-x));
}
template <class R>
inline R inverse_sigmoid(R x) {  // math.rs:395-397
  return rln(x / (RL(1.0) - x));
}

// ---------------------------------------------------------------- sh.rs ------------

template <class R>
struct SH {
  // sh.rs:16-33: f64 literals cast to Real.
  static R c0() { return (R)0.28209479177387814; }
  static R c1() { return (R)0.4886025119029199; }
  static R c2(int i) {
    static const double v[5] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005,
                                -1.0925484305920792, 0.5462742152960396};
    return (R)v[i];
  }
  static R c3(int i) {
    static const double v[7] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658,
                                0.3731763325901154,  -0.4570457994644658, 1.445305721320277,
                                -0.5900435899266435};
    return (R)v[i];
  }
};

inline int num_coeffs(int degree) { return (degree + 1) * (degree + 1); }  // sh.rs:36

// basis, sh.rs:47-75.
template <class R>
inline void sh_basis(int degree, const V3<R>& dir, R b[16]) {
  for (int i = 0; i < 16; i++) b[i] = 0;
  const R x = dir.x, y = dir.y, z = dir.z;
  b[0] = SH<R>::c0();
  if (degree >= 1) {
    b[1] = -SH<R>::c1() * y;
    b[2] = SH<R>::c1() * z;
    b[3] = -SH<R>::c1() * x;
  }
  if (degree >= 2) {
    const R xx = x * x, yy = y * y, zz = z * z;
    const R xy = x * y, yz = y * z, xz = x * z;
    b[4] = SH<R>::c2(0) * xy;
    b[5] = SH<R>::c2(1) * yz;
    b[6] = SH<R>::c2(2) * (RL(2.0) * zz - xx - yy);
    b[7] = SH<R>::c2(3) * xz;
    b[8] = SH<R>::c2(4) * (xx - yy);
    if (degree >= 3) {
      b[9] = SH<R>::c3(0) * y * (RL(3.0) * xx - yy);
      b[10] = SH<R>::c3(1) * xy * z;
      b[11] = SH<R>::c3(2) * y * (RL(4.0) * zz - xx - yy);
      b[12] = SH<R>::c3(3) * z * (RL(2.0) * zz - RL(3.0) * xx - RL(3.0) * yy);
      b[13] = SH<R>::c3(4) * x * (RL(4.0) * zz - xx - yy);
      b[14] = SH<R>::c3(5) * z * (xx - yy);
      b[15] = SH<R>::c3(6) * x * (xx - RL(3.0) * yy);
    }
  }
}

// basis_grad, sh.rs:80-130.
template <class R>
inline void sh_basis_grad(int degree, const V3<R>& dir, V3<R> g[16]) {
  for (int i = 0; i < 16; i++) g[i] = V3<R>();
  const R x = dir.x, y = dir.y, z = dir.z;
  if (degree >= 1) {
    g[1] = V3<R>(0, -SH<R>::c1(), 0);
    g[2] = V3<R>(0, 0, SH<R>::c1());
    g[3] = V3<R>(-SH<R>::c1(), 0, 0);
  }
  if (degree >= 2) {
    const R xx = x * x, yy = y * y, zz = z * z;
    const R c20 = SH<R>::c2(0), c21 = SH<R>::c2(1), c22 = SH<R>::c2(2), c23 = SH<R>::c2(3),
            c24 = SH<R>::c2(4);
    g[4] = V3<R>(c20 * y, c20 * x, 0);
    g[5] = V3<R>(0, c21 * z, c21 * y);
    g[6] = V3<R>(RL(-2.0) * c22 * x, RL(-2.0) * c22 * y, RL(4.0) * c22 * z);
    g[7] = V3<R>(c23 * z, 0, c23 * x);
    g[8] = V3<R>(RL(2.0) * c24 * x, RL(-2.0) * c24 * y, 0);
    if (degree >= 3) {
      const R c30 = SH<R>::c3(0), c31 = SH<R>::c3(1), c32 = SH<R>::c3(2), c33 = SH<R>::c3(3),
              c34 = SH<R>::c3(4), c35 = SH<R>::c3(5), c36 = SH<R>::c3(6);
      g[9] = V3<R>(c30 * RL(6.0) * x * y, c30 * (RL(3.0) * xx - RL(3.0) * yy), 0);
      g[10] = V3<R>(c31 * y * z, c31 * x * z, c31 * x * y);
      g[11] = V3<R>(c32 * (RL(-2.0) * x * y), c32 * (RL(4.0) * zz - xx - RL(3.0) * yy),
                    c32 * RL(8.0) * y * z);
      g[12] = V3<R>(c33 * (RL(-6.0) * x * z), c33 * (RL(-6.0) * y * z),
                    c33 * (RL(6.0) * zz - RL(3.0) * xx - RL(3.0) * yy));
      g[13] = V3<R>(c34 * (RL(4.0) * zz - RL(3.0) * xx - yy), c34 * (RL(-2.0) * x * y),
                    c34 * RL(8.0) * x * z);
      g[14] = V3<R>(c35 * RL(2.0) * x * z, -c35 * RL(2.0) * y * z, c35 * (xx - yy));
      g[15] = V3<R>(c36 * (RL(3.0) * xx - RL(3.0) * yy), c36 * (RL(-6.0) * x * y), 0);
    }
  }
}

// ---------------------------------------------------------------- model / camera ---

// Gaussian record (model.rs:17-28) as 59 scalars, flat order tests/common/mod.rs:111-139.
template <class R>
struct GaussianView {
  const R* p;
  V3<R> mean() const { return V3<R>(p[0], p[1], p[2]); }
  V3<R> scales() const { return V3<R>(rexp(p[3]), rexp(p[4]), rexp(p[5])); }  // model.rs:37-39
  Quat<R> rotation() const { return Quat<R>{p[6], p[7], p[8], p[9]}.normalized(); }
  R opacity() const { return p[10]; }
  R sh(int k, int c) const { return p[11 + 3 * k + c]; }
  R alpha() const { return sigmoid(p[10]); }  // model.rs:47-49
};

template <class R>
struct Camera {  // camera.rs:13-32
  uint32_t width, height;
  R fx, fy, cx, cy;
  M3<R> rot;
  V3<R> trans;
  V3<R> world_to_cam(const V3<R>& p) const { return rot.mul_v(p) + trans; }  // camera.rs:36-38
  V3<R> center() const { return -(rot.transpose().mul_v(trans)); }          // camera.rs:41-43
};

template <class R>
struct ProjOpts {
  R near_plane;
  R eps2d;
  bool antialias;
  int sh_degree;
};

template <class R>
struct Proj {  // project.rs:47-64
  R mean2d[2];
  R conic[3];
  R depth;
  uint32_t radius;
  R opacity;
  R color[3];
  uint32_t tile_min[2];
  uint32_t tile_max[2];
  bool visible() const { return radius > 0; }
  uint32_t num_tiles() const {
    return radius == 0 ? 0 : (tile_max[0] - tile_min[0]) * (tile_max[1] - tile_min[1]);
  }
};

template <class R>
struct ProjGrad {  // project.rs:84-93
  R mean2d[2];
  R conic[3];
  R opacity;
  R color[3];
};

static const uint32_t TILE = 16;  // project.rs:18

inline void tile_grid(uint32_t w, uint32_t h, uint32_t* tx, uint32_t* ty) {  // project.rs:272-274
  *tx = (w + TILE - 1) / TILE;
  *ty = (h + TILE - 1) / TILE;
}

// ---------------------------------------------------------------- project.rs -------

template <class R>
struct Core {  // project.rs:96-125
  V3<R> t;
  R txc, tyc;
  bool clamped_x, clamped_y;
  V3<R> s;
  Quat<R> q;
  M3<R> rq, m, sigma;
  R tmat[2][3];
  R a0, b0, c0, det0, a, b, c, det;
  R conic[3];
  R comp;
  R mean2d[2];
  R radius;
  R o_raw;
  V3<R> dir;
  R dist;
  R basis[16];
  R color_raw[3];
};

template <class R>
inline R clamp_flag(R v, R lo, R hi, bool* flag) {  // project.rs:127-135
  if (v < lo) {
    *flag = true;
    return lo;
  }
  if (v > hi) {
    *flag = true;
    return hi;
  }
  *flag = false;
  return v;
}

// core, project.rs:137-269.  Returns false when culled.
template <class R>
bool core(const R* gp, const Camera<R>& cam, const ProjOpts<R>& opts, Core<R>& k) {
  GaussianView<R> g{gp};
  V3<R> p = g.mean();
  k.t = cam.world_to_cam(p);
  if (std::isnan(k.t.z) || k.t.z <= opts.near_plane) return false;
  k.s = g.scales();
  k.q = g.rotation();
  k.rq = k.q.to_mat();
  k.m = k.rq.scale_cols(k.s);
  k.sigma = k.m.mul_m(k.m.transpose());

  const R W = (R)cam.width, H = (R)cam.height;
  R tan_x = RL(0.5) * W / cam.fx;
  R tan_y = RL(0.5) * H / cam.fy;
  R lim_x_pos = (W - cam.cx) / cam.fx + RL(0.3) * tan_x;
  R lim_x_neg = cam.cx / cam.fx + RL(0.3) * tan_x;
  R lim_y_pos = (H - cam.cy) / cam.fy + RL(0.3) * tan_y;
  R lim_y_neg = cam.cy / cam.fy + RL(0.3) * tan_y;
  R rz = RL(1.0) / k.t.z;
  R rx = k.t.x * rz;
  R ry = k.t.y * rz;
  R rxc = clamp_flag(rx, -lim_x_neg, lim_x_pos, &k.clamped_x);
  R ryc = clamp_flag(ry, -lim_y_neg, lim_y_pos, &k.clamped_y);
  k.txc = rxc * k.t.z;
  k.tyc = ryc * k.t.z;
  R rz2 = rz * rz;
  R j[2][3] = {{cam.fx * rz, 0, -cam.fx * k.txc * rz2}, {0, cam.fy * rz, -cam.fy * k.tyc * rz2}};

  for (int r = 0; r < 2; r++)
    for (int c = 0; c < 3; c++) {
      R acc = 0;
      for (int kk = 0; kk < 3; kk++) acc += j[r][kk] * cam.rot.m[kk][c];
      k.tmat[r][c] = acc;
    }
  R ts[2][3];
  for (int r = 0; r < 2; r++)
    for (int c = 0; c < 3; c++) {
      R acc = 0;
      for (int kk = 0; kk < 3; kk++) acc += k.tmat[r][kk] * k.sigma.m[kk][c];
      ts[r][c] = acc;
    }
  R a0 = 0, b0 = 0, c0 = 0;
  for (int c = 0; c < 3; c++) {
    a0 += ts[0][c] * k.tmat[0][c];
    b0 += ts[0][c] * k.tmat[1][c];
    c0 += ts[1][c] * k.tmat[1][c];
  }
  k.a0 = a0;
  k.b0 = b0;
  k.c0 = c0;
  k.det0 = a0 * c0 - b0 * b0;
  k.a = a0 + opts.eps2d;
  k.b = b0;
  k.c = c0 + opts.eps2d;
  k.det = k.a * k.c - k.b * k.b;
  if (std::isnan(k.det) || k.det <= (R)0) return false;
  R inv_det = RL(1.0) / k.det;
  k.conic[0] = k.c * inv_det;
  k.conic[1] = -k.b * inv_det;
  k.conic[2] = k.a * inv_det;
  if (opts.antialias)
    k.comp = k.det0 > (R)0 ? rsqrt_(k.det0 * inv_det) : (R)0;
  else
    k.comp = RL(1.0);

  R mid = RL(0.5) * (k.a + k.c);
  R lambda1 = mid + rsqrt_(rmax(mid * mid - k.det, RL(0.1)));
  k.radius = rceil(RL(3.0) * rsqrt_(lambda1));

  k.mean2d[0] = cam.fx * rx + cam.cx;
  k.mean2d[1] = cam.fy * ry + cam.cy;
  k.o_raw = sigmoid(g.opacity());

  V3<R> d = p - cam.center();
  k.dist = d.norm();
  k.dir = d * (RL(1.0) / k.dist);
  sh_basis(opts.sh_degree, k.dir, k.basis);
  int n = num_coeffs(opts.sh_degree);
  k.color_raw[0] = k.color_raw[1] = k.color_raw[2] = RL(0.5);
  for (int kk = 0; kk < n; kk++)
    for (int ch = 0; ch < 3; ch++) k.color_raw[ch] += k.basis[kk] * g.sh(kk, ch);
  return true;
}

// clamp_tile closure, project.rs:287-295.
template <class R>
inline uint32_t clamp_tile(R v, uint32_t n) {
  if (v <= (R)0) return 0;
  if (v >= (R)n) return n;
  return as_u32(v);
}

// project_one, project.rs:277-317.
template <class R>
Proj<R> project_one(const R* gp, const Camera<R>& cam, const ProjOpts<R>& opts) {
  Proj<R> out;
  std::memset(&out, 0, sizeof(out));
  Core<R> k;
  if (!core(gp, cam, opts, k)) return out;
  uint32_t tiles_x, tiles_y;
  tile_grid(cam.width, cam.height, &tiles_x, &tiles_y);
  R r = k.radius;
  if (std::isnan(r) || r <= (R)0) return out;
  const R tile = (R)TILE;
  uint32_t tx0 = clamp_tile(rfloor((k.mean2d[0] - r) / tile), tiles_x);
  uint32_t tx1 = clamp_tile(rceil((k.mean2d[0] + r) / tile), tiles_x);
  uint32_t ty0 = clamp_tile(rfloor((k.mean2d[1] - r) / tile), tiles_y);
  uint32_t ty1 = clamp_tile(rceil((k.mean2d[1] + r) / tile), tiles_y);
  if (tx0 >= tx1 || ty0 >= ty1) return out;
  out.mean2d[0] = k.mean2d[0];
  out.mean2d[1] = k.mean2d[1];
  for (int i = 0; i < 3; i++) out.conic[i] = k.conic[i];
  out.depth = k.t.z;
  out.radius = as_u32(r);
  out.opacity = k.o_raw * k.comp;
  for (int ch = 0; ch < 3; ch++) out.color[ch] = rmax(k.color_raw[ch], (R)0);
  out.tile_min[0] = tx0;
  out.tile_min[1] = ty0;
  out.tile_max[0] = tx1;
  out.tile_max[1] = ty1;
  return out;
}

// project_backward_one, project.rs:335-474.  out: 59 scalars.
template <class R>
void project_backward_one(const R* gp, const Camera<R>& cam, const ProjOpts<R>& opts,
                          const ProjGrad<R>& dp, R* out) {
  for (int i = 0; i < 59; i++) out[i] = 0;
  Core<R> k;
  if (!core(gp, cam, opts, k)) return;
  GaussianView<R> g{gp};

  int n = num_coeffs(opts.sh_degree);
  V3<R> bgrad[16];
  sh_basis_grad(opts.sh_degree, k.dir, bgrad);
  V3<R> ddir;
  for (int ch = 0; ch < 3; ch++) {
    if (k.color_raw[ch] > (R)0) {
      R dc = dp.color[ch];
      for (int kk = 0; kk < n; kk++) {
        out[11 + 3 * kk + ch] = dc * k.basis[kk];
        ddir += bgrad[kk] * (dc * g.sh(kk, ch));
      }
    }
  }
  V3<R> dmean_color = (ddir - k.dir * k.dir.dot(ddir)) * (RL(1.0) / k.dist);

  R do_raw = dp.opacity * k.comp;
  R dcomp = dp.opacity * k.o_raw;
  out[10] = do_raw * k.o_raw * (RL(1.0) - k.o_raw);

  const R a = k.a, b = k.b, c = k.c, det = k.det;
  const R da_ = dp.conic[0], db_ = dp.conic[1], dc_ = dp.conic[2];
  R inv_det2 = RL(1.0) / (det * det);
  R da = (-c * c * da_ + b * c * db_ - b * b * dc_) * inv_det2;
  R db = (RL(2.0) * b * c * da_ - (a * c + b * b) * db_ + RL(2.0) * a * b * dc_) * inv_det2;
  R dcc = (-b * b * da_ + a * b * db_ - a * a * dc_) * inv_det2;

  if (opts.antialias && k.det0 > (R)0) {
    R ddet0 = dcomp * RL(0.5) * k.comp / k.det0;
    R ddet = -dcomp * RL(0.5) * k.comp / det;
    da += ddet0 * k.c0 + ddet * c;
    dcc += ddet0 * k.a0 + ddet * a;
    db += ddet0 * (RL(-2.0) * k.b0) + ddet * (RL(-2.0) * b);
  }

  R gmat[2][2] = {{da, RL(0.5) * db}, {RL(0.5) * db, dcc}};
  R gt[2][3];
  for (int r = 0; r < 2; r++)
    for (int col = 0; col < 3; col++)
      gt[r][col] = gmat[r][0] * k.tmat[0][col] + gmat[r][1] * k.tmat[1][col];
  M3<R> dsigma;
  for (int i = 0; i < 3; i++)
    for (int jj = 0; jj < 3; jj++)
      dsigma.m[i][jj] = k.tmat[0][i] * gt[0][jj] + k.tmat[1][i] * gt[1][jj];
  R dt[2][3];
  for (int r = 0; r < 2; r++)
    for (int col = 0; col < 3; col++) {
      R acc = 0;
      for (int kk = 0; kk < 3; kk++) acc += gt[r][kk] * k.sigma.m[kk][col];
      dt[r][col] = RL(2.0) * acc;
    }
  R dj[2][3];
  for (int r = 0; r < 2; r++)
    for (int col = 0; col < 3; col++) {
      R acc = 0;
      for (int kk = 0; kk < 3; kk++) acc += dt[r][kk] * cam.rot.m[col][kk];
      dj[r][col] = acc;
    }

  const R fx = cam.fx, fy = cam.fy;
  R rz = RL(1.0) / k.t.z;
  R rz2 = rz * rz;
  R rz3 = rz2 * rz;
  R dtxc = dj[0][2] * (-fx * rz2);
  R dtyc = dj[1][2] * (-fy * rz2);
  R dtx = 0, dty = 0;
  R dtz = dj[0][0] * (-fx * rz2) + dj[1][1] * (-fy * rz2) +
          dj[0][2] * (RL(2.0) * fx * k.txc * rz3) + dj[1][2] * (RL(2.0) * fy * k.tyc * rz3);
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

  V3<R> dmean = cam.rot.transpose().mul_v(V3<R>(dtx, dty, dtz)) + dmean_color;
  out[0] = dmean.x;
  out[1] = dmean.y;
  out[2] = dmean.z;

  M3<R> dm = dsigma.mul_m(k.m).scaled(RL(2.0));
  R ds[3] = {0, 0, 0};
  M3<R> dr;
  for (int r = 0; r < 3; r++)
    for (int col = 0; col < 3; col++) {
      ds[col] += dm.m[r][col] * k.rq.m[r][col];
      dr.m[r][col] = dm.m[r][col] * k.s.get(col);
    }
  for (int i = 0; i < 3; i++) out[3 + i] = ds[i] * k.s.get(i);
  R dq_unit[4];
  quat_to_mat_backward(k.q, dr, dq_unit);
  normalize_backward(gp + 6, dq_unit, out + 6);
}

// ---------------------------------------------------------------- rasterize.rs -----

struct Binning {  // rasterize.rs:26-39
  uint32_t tiles_x, tiles_y;
  std::vector<uint32_t> entry_gaussian;
  std::vector<uint32_t> entry_offset;
  std::vector<std::pair<uint64_t, uint32_t>> sorted;
  std::vector<std::pair<uint32_t, uint32_t>> tile_range;
};

// bin, rasterize.rs:60-128.  Returns false on the precondition failures the reference
// asserts (tile rectangle outside the grid, entry overflow).
template <class R>
bool bin(const Proj<R>* projs, size_t n, uint32_t width, uint32_t height, Binning& b) {
  tile_grid(width, height, &b.tiles_x, &b.tiles_y);
  b.entry_offset.assign(n + 1, 0);
  uint64_t total = 0;
  for (size_t i = 0; i < n; i++) {
    const Proj<R>& p = projs[i];
    if (p.visible() && !(p.tile_max[0] <= b.tiles_x && p.tile_max[1] <= b.tiles_y)) return false;
    b.entry_offset[i] = (uint32_t)total;
    total += p.num_tiles();
  }
  if (total >= 0xffffffffull) return false;
  b.entry_offset[n] = (uint32_t)total;
  b.entry_gaussian.resize(total);
  b.sorted.resize(total);
  ORACLE_PAR
  for (size_t gi = 0; gi < n; gi++) {
    const Proj<R>& p = projs[gi];
    uint64_t depth_bits = f32_bits(p.depth);
    if (!p.visible()) continue;
    size_t e = b.entry_offset[gi];  // this Gaussian's entries, in the reference's order
    for (uint32_t ty = p.tile_min[1]; ty < p.tile_max[1]; ty++)
      for (uint32_t tx = p.tile_min[0]; tx < p.tile_max[0]; tx++) {
        uint64_t tile = (uint64_t)(ty * b.tiles_x + tx);
        b.entry_gaussian[e] = (uint32_t)gi;
        b.sorted[e] = std::make_pair((tile << 32) | depth_bits, (uint32_t)e);
        e++;
      }
  }
  ORACLE_SORT(b.sorted.begin(), b.sorted.end());  // keys are unique: any sort gives one order
  size_t num_tiles = (size_t)b.tiles_x * b.tiles_y;
  b.tile_range.assign(num_tiles, std::make_pair(0u, 0u));
  size_t s = 0;
  while (s < b.sorted.size()) {
    size_t tile = (size_t)(b.sorted[s].first >> 32);
    size_t e2 = s + 1;
    while (e2 < b.sorted.size() && (size_t)(b.sorted[e2].first >> 32) == tile) e2++;
    b.tile_range[tile] = std::make_pair((uint32_t)s, (uint32_t)e2);
    s = e2;
  }
  return true;
}

template <class R>
struct Raster {
  static R alpha_min() { return RL(1.0) / RL(255.0); }  // rasterize.rs:53
  static R alpha_max() { return RL(0.99); }            // rasterize.rs:55
  static R t_min() { return RL(1e-4); }                // rasterize.rs:57
};

template <class R>
struct TileGauss {  // rasterize.rs:132-145
  R mean[2];
  R conic[3];
  R opacity;
  R color[3];
  R sigma_max;
  uint32_t x0, x1, y0, y1;
  bool is_empty() const { return x0 >= x1 || y0 >= y1; }
};

// TileGauss::new, rasterize.rs:149-182.
template <class R>
TileGauss<R> tile_gauss(const Proj<R>& p, uint32_t tx0, uint32_t ty0, uint32_t w, uint32_t h) {
  TileGauss<R> t;
  R sigma_max = rln(p.opacity / Raster<R>::alpha_min());
  R a = p.conic[0], b = p.conic[1], c = p.conic[2];
  R det = a * c - b * b;
  uint32_t x0 = 0, x1 = 0, y0 = 0, y1 = 0;
  if (sigma_max > (R)0 && det > (R)0) {
    R half_x = rsqrt_(RL(2.0) * sigma_max * c / det) + RL(1e-3);
    R half_y = rsqrt_(RL(2.0) * sigma_max * a / det) + RL(1e-3);
    auto clip = [](R lo, R hi, uint32_t offset, uint32_t nn, uint32_t* o0, uint32_t* o1) {
      R l = rceil(lo - RL(0.5)) - (R)offset;
      R hh = rfloor(hi - RL(0.5)) - (R)offset + RL(1.0);
      *o0 = as_u32(rclamp(l, (R)0, (R)nn));
      *o1 = as_u32(rclamp(hh, (R)0, (R)nn));
    };
    clip(p.mean2d[0] - half_x, p.mean2d[0] + half_x, tx0, w, &x0, &x1);
    clip(p.mean2d[1] - half_y, p.mean2d[1] + half_y, ty0, h, &y0, &y1);
  }
  t.mean[0] = p.mean2d[0];
  t.mean[1] = p.mean2d[1];
  for (int i = 0; i < 3; i++) {
    t.conic[i] = p.conic[i];
    t.color[i] = p.color[i];
  }
  t.opacity = p.opacity;
  t.sigma_max = sigma_max;
  t.x0 = x0;
  t.x1 = x1;
  t.y0 = y0;
  t.y1 = y1;
  return t;
}

struct TileGeom {
  uint32_t x0, y0, w, h;
};

inline TileGeom tile_geom(const Binning& b, uint32_t width, uint32_t height, size_t tile) {
  uint32_t tx = (uint32_t)tile % b.tiles_x;  // rasterize.rs:198-209
  uint32_t ty = (uint32_t)tile / b.tiles_x;
  TileGeom g;
  g.x0 = tx * TILE;
  g.y0 = ty * TILE;
  g.w = std::min(width - g.x0, TILE);
  g.h = std::min(height - g.y0, TILE);
  return g;
}

// evaluate, rasterize.rs:231-244.  Returns false when skipped.
template <class R>
inline bool evaluate(const TileGauss<R>& g, R px, R py, R* gauss, R* alpha) {
  R dx = px - g.mean[0];
  R dy = py - g.mean[1];
  R sigma = RL(0.5) * (g.conic[0] * dx * dx + g.conic[2] * dy * dy) + g.conic[1] * dx * dy;
  if (sigma < (R)0 || sigma > g.sigma_max) return false;
  R G = rexp(-sigma);
  R og = g.opacity * G;
  if (og < Raster<R>::alpha_min()) return false;
  *gauss = G;
  *alpha = rmin(og, Raster<R>::alpha_max());
  return true;
}

// rasterize_forward (rasterize.rs:305-337) with forward_tile (rasterize.rs:249-302).
template <class R>
void rasterize_forward(const Proj<R>* projs, const Binning& b, uint32_t width, uint32_t height,
                       const R bg[3], R* image, R* final_t, uint32_t* n_contrib) {
  size_t num_tiles = (size_t)b.tiles_x * b.tiles_y;
  ORACLE_PAR
  for (size_t tile = 0; tile < num_tiles; tile++) {
    TileGeom geom = tile_geom(b, width, height, tile);
    uint32_t s = b.tile_range[tile].first, e = b.tile_range[tile].second;
    size_t n = (size_t)geom.w * geom.h;
    std::vector<R> color(n * 3, (R)0), trans(n, RL(1.0));
    std::vector<uint32_t> nc(n, 0);
    std::vector<char> done(n, 0);
    for (uint32_t k = 0; k < e - s; k++) {
      uint32_t entry = b.sorted[s + k].second;
      TileGauss<R> g = tile_gauss(projs[b.entry_gaussian[entry]], geom.x0, geom.y0, geom.w, geom.h);
      if (g.is_empty()) continue;
      for (uint32_t j = g.y0; j < g.y1; j++) {
        R py = (R)(geom.y0 + j) + RL(0.5);
        for (uint32_t i = g.x0; i < g.x1; i++) {
          size_t pix = (size_t)j * geom.w + i;
          if (done[pix]) continue;
          R px = (R)(geom.x0 + i) + RL(0.5);
          R G, alpha;
          if (!evaluate(g, px, py, &G, &alpha)) continue;
          R t = trans[pix];
          R test_t = t * (RL(1.0) - alpha);
          if (test_t < Raster<R>::t_min()) {
            done[pix] = 1;
            continue;
          }
          for (int ch = 0; ch < 3; ch++) color[pix * 3 + ch] += alpha * t * g.color[ch];
          trans[pix] = test_t;
          nc[pix] = k + 1;
        }
      }
    }
    for (size_t pix = 0; pix < n; pix++)
      for (int ch = 0; ch < 3; ch++) color[pix * 3 + ch] += trans[pix] * bg[ch];
    for (uint32_t j = 0; j < geom.h; j++)
      for (uint32_t i = 0; i < geom.w; i++) {
        size_t src = (size_t)j * geom.w + i;
        size_t dst = (size_t)(geom.y0 + j) * width + geom.x0 + i;
        for (int ch = 0; ch < 3; ch++) image[dst * 3 + ch] = color[src * 3 + ch];
        final_t[dst] = trans[src];
        n_contrib[dst] = nc[src];
      }
  }
}

// (Gaussian, pixel) pair counts of the compositing walks, for checking a device's counters:
// counts[0] pairs forward_tile (rasterize.rs:262-290) evaluates (sigma in [0, sigma_max]
// on a pixel not yet done, so the exponential is computed), [1] of those passing alpha_min,
// [2] composited (the transmittance test passes); [3] pairs backward_tile (rasterize.rs:
// 402-415) evaluates (k < n_contrib), [4] of those passing alpha_min.
inline void raster_pair_counts(const Proj<float>* projs, const Binning& b, uint32_t width,
                               uint32_t height, const uint32_t* n_contrib_img,
                               uint64_t counts[5]) {
  for (int q = 0; q < 5; q++) counts[q] = 0;
  size_t num_tiles = (size_t)b.tiles_x * b.tiles_y;
  for (size_t tile = 0; tile < num_tiles; tile++) {
    TileGeom geom = tile_geom(b, width, height, tile);
    uint32_t s = b.tile_range[tile].first, e = b.tile_range[tile].second;
    size_t n = (size_t)geom.w * geom.h;
    std::vector<float> trans(n, 1.0f);
    std::vector<char> done(n, 0);
    for (uint32_t k = 0; k < e - s; k++) {
      uint32_t entry = b.sorted[s + k].second;
      TileGauss<float> g =
          tile_gauss(projs[b.entry_gaussian[entry]], geom.x0, geom.y0, geom.w, geom.h);
      if (g.is_empty()) continue;
      for (uint32_t j = g.y0; j < g.y1; j++) {
        float py = (float)(geom.y0 + j) + 0.5f;
        for (uint32_t i = g.x0; i < g.x1; i++) {
          size_t pix = (size_t)j * geom.w + i;
          float px = (float)(geom.x0 + i) + 0.5f;
          float dx = px - g.mean[0], dy = py - g.mean[1];
          float sigma = 0.5f * (g.conic[0] * dx * dx + g.conic[2] * dy * dy) + g.conic[1] * dx * dy;
          bool evaluated = !(sigma < 0.0f || sigma > g.sigma_max);
          float G = 0, alpha = 0;
          bool passing = evaluate(g, px, py, &G, &alpha);
          size_t gpix = (size_t)(geom.y0 + j) * width + geom.x0 + i;
          if (k < n_contrib_img[gpix]) {
            counts[3] += evaluated;
            counts[4] += passing;
          }
          if (done[pix] || !evaluated) continue;
          counts[0]++;
          if (!passing) continue;
          counts[1]++;
          float test_t = trans[pix] * (1.0f - alpha);
          if (test_t < Raster<float>::t_min()) {
            done[pix] = 1;
            continue;
          }
          counts[2]++;
          trans[pix] = test_t;
        }
      }
    }
  }
}

// image::imageops::resize(.., FilterType::Lanczos3) of an RGB8 image, as dataset::load
// calls it (src/dataset.rs:176-180).  The `image` crate (0.25.10, Cargo.lock) is not
// vendored; this restates its published resampler (imageops/sample.rs): a copy for equal
// sizes, else a vertical pass into binary32 RGBA, then a horizontal pass that clamps to
// [0, 255] and rounds half away from zero; per output index o the filter centre is
// (o + 0.5) * ratio, the support 3 * max(ratio, 1), taps lanczos3((i - centre + 0.5) /
// max(ratio, 1)) on the clamped range [floor(c - s), ceil(c + s)), normalised by their sum.
inline float resize_lanczos3(float x) {
  auto sinc = [](float t) {
    if (t == 0.0f) return 1.0f;
    float a = t * 3.14159265358979323846f;
    return std::sin(a) / a;
  };
  return std::fabs(x) < 3.0f ? sinc(x) * sinc(x / 3.0f) : 0.0f;
}

inline void resize_taps(uint32_t in, uint32_t out, uint32_t o, int64_t* left_out,
                        std::vector<float>* ws) {
  float ratio = (float)in / (float)out;
  float sratio = ratio < 1.0f ? 1.0f : ratio;
  float support = 3.0f * sratio;
  float centre = ((float)o + 0.5f) * ratio;
  int64_t left = std::min<int64_t>(std::max<int64_t>((int64_t)std::floor(centre - support), 0),
                                   (int64_t)in - 1);
  int64_t right = std::min<int64_t>(
      std::max<int64_t>((int64_t)std::ceil(centre + support), left + 1), (int64_t)in);
  float c = centre - 0.5f;
  ws->clear();
  float sum = 0.0f;
  for (int64_t i = left; i < right; i++) {
    float w = resize_lanczos3(((float)i - c) / sratio);
    ws->push_back(w);
    sum += w;
  }
  for (auto& w : *ws) w /= sum;
  *left_out = left;
}

inline void resize_rgb8(const uint8_t* src, uint32_t w, uint32_t h, uint32_t ow, uint32_t oh,
                        uint8_t* dst) {
  if (w == 0 || h == 0) {
    std::fill(dst, dst + (size_t)ow * oh * 3, (uint8_t)0);
    return;
  }
  if (w == ow && h == oh) {
    std::copy(src, src + (size_t)w * h * 3, dst);
    return;
  }
  std::vector<float> tmp((size_t)w * oh * 3);
  std::vector<float> ws;
  for (uint32_t oy = 0; oy < oh; oy++) {  // vertical_sample
    int64_t left;
    resize_taps(h, oh, oy, &left, &ws);
    for (uint32_t x = 0; x < w; x++)
      for (int c = 0; c < 3; c++) {
        float t = 0.0f;
        for (size_t i = 0; i < ws.size(); i++)
          t += (float)src[((size_t)(left + i) * w + x) * 3 + c] * ws[i];
        tmp[((size_t)oy * w + x) * 3 + c] = t;
      }
  }
  for (uint32_t ox = 0; ox < ow; ox++) {  // horizontal_sample
    int64_t left;
    resize_taps(w, ow, ox, &left, &ws);
    for (uint32_t y = 0; y < oh; y++)
      for (int c = 0; c < 3; c++) {
        float t = 0.0f;
        for (size_t i = 0; i < ws.size(); i++)
          t += tmp[((size_t)y * w + left + i) * 3 + c] * ws[i];
        t = t < 0.0f ? 0.0f : (t > 255.0f ? 255.0f : t);
        dst[((size_t)y * ow + ox) * 3 + c] = (uint8_t)std::round(t);
      }
  }
}

static const int SLOTS = 11;  // rasterize.rs:341

// rasterize_backward (rasterize.rs:459-505) with backward_tile (rasterize.rs:354-455).
// grads: n ProjGrad; abs_grads: 2n.
template <class R>
void rasterize_backward(const Proj<R>* projs, size_t np, const Binning& b, uint32_t width,
                        uint32_t height, const R bgc[3], const R* final_t_img,
                        const uint32_t* n_contrib_img, const R* dimage, ProjGrad<R>* grads,
                        R* abs_grads) {
  size_t num_entries = b.sorted.size();
  std::vector<R> slots(num_entries * SLOTS, (R)0);
  size_t num_tiles = (size_t)b.tiles_x * b.tiles_y;
  ORACLE_PAR
  for (size_t tile = 0; tile < num_tiles; tile++) {
    TileGeom geom = tile_geom(b, width, height, tile);
    uint32_t start = b.tile_range[tile].first, end = b.tile_range[tile].second;
    if (end == start) continue;
    std::vector<TileGauss<R>> list;
    for (uint32_t i = start; i < end; i++)
      list.push_back(tile_gauss(projs[b.entry_gaussian[b.sorted[i].second]], geom.x0, geom.y0,
                                geom.w, geom.h));
    size_t n = (size_t)geom.w * geom.h;
    struct PixelBack {
      R t;
      R accum[3];
      R last_alpha;
      R last_color[3];
    };
    std::vector<PixelBack> state(n);
    std::vector<R> dl(n * 3), bg_dot(n), t_final(n);
    std::vector<size_t> last(n);
    size_t q = 0;
    for (uint32_t j = 0; j < geom.h; j++)
      for (uint32_t i = 0; i < geom.w; i++, q++) {
        size_t pix = (size_t)(geom.y0 + j) * width + geom.x0 + i;
        R d[3] = {dimage[pix * 3], dimage[pix * 3 + 1], dimage[pix * 3 + 2]};
        for (int c = 0; c < 3; c++) dl[q * 3 + c] = d[c];
        bg_dot[q] = bgc[0] * d[0] + bgc[1] * d[1] + bgc[2] * d[2];
        last[q] = n_contrib_img[pix];
        t_final[q] = final_t_img[pix];
        state[q].t = final_t_img[pix];
        for (int c = 0; c < 3; c++) {
          state[q].accum[c] = 0;
          state[q].last_color[c] = 0;
        }
        state[q].last_alpha = 0;
      }
    std::vector<R> acc(list.size() * SLOTS, (R)0);
    for (size_t kk = list.size(); kk-- > 0;) {
      const TileGauss<R>& g = list[kk];
      if (g.is_empty()) continue;
      R* a = &acc[kk * SLOTS];
      for (uint32_t j = g.y0; j < g.y1; j++) {
        R py = (R)(geom.y0 + j) + RL(0.5);
        for (uint32_t i = g.x0; i < g.x1; i++) {
          size_t pix = (size_t)j * geom.w + i;
          if (kk >= last[pix]) continue;
          R px = (R)(geom.x0 + i) + RL(0.5);
          R gauss, alpha;
          if (!evaluate(g, px, py, &gauss, &alpha)) continue;
          PixelBack& s = state[pix];
          const R* d = &dl[pix * 3];
          s.t /= RL(1.0) - alpha;
          R t = s.t;
          R dalpha = 0;
          for (int ch = 0; ch < 3; ch++) {
            s.accum[ch] = s.last_alpha * s.last_color[ch] + (RL(1.0) - s.last_alpha) * s.accum[ch];
            dalpha += (g.color[ch] - s.accum[ch]) * d[ch];
            a[6 + ch] += alpha * t * d[ch];
          }
          dalpha *= t;
          for (int ch = 0; ch < 3; ch++) s.last_color[ch] = g.color[ch];
          s.last_alpha = alpha;
          dalpha += -t_final[pix] / (RL(1.0) - alpha) * bg_dot[pix];
          if (g.opacity * gauss < Raster<R>::alpha_max()) {
            a[5] += gauss * dalpha;
            R dgauss = g.opacity * dalpha;
            R dsigma = -dgauss * gauss;
            R dx = px - g.mean[0];
            R dy = py - g.mean[1];
            R ddx = dsigma * (g.conic[0] * dx + g.conic[1] * dy);
            R ddy = dsigma * (g.conic[1] * dx + g.conic[2] * dy);
            a[0] -= ddx;
            a[1] -= ddy;
            a[2] += dsigma * RL(0.5) * dx * dx;
            a[3] += dsigma * dx * dy;
            a[4] += dsigma * RL(0.5) * dy * dy;
            a[9] += rabs(ddx);
            a[10] += rabs(ddy);
          }
        }
      }
    }
    for (size_t kk = 0; kk < list.size(); kk++) {
      uint32_t entry = b.sorted[start + kk].second;
      for (int s = 0; s < SLOTS; s++) slots[(size_t)entry * SLOTS + s] = acc[kk * SLOTS + s];
    }
  }
  ORACLE_PAR
  for (size_t g = 0; g < np; g++) {
    R sum[SLOTS];
    for (int k = 0; k < SLOTS; k++) sum[k] = 0;
    for (size_t entry = b.entry_offset[g]; entry < b.entry_offset[g + 1]; entry++)
      for (int k = 0; k < SLOTS; k++) sum[k] += slots[entry * SLOTS + k];
    grads[g].mean2d[0] = sum[0];
    grads[g].mean2d[1] = sum[1];
    grads[g].conic[0] = sum[2];
    grads[g].conic[1] = sum[3];
    grads[g].conic[2] = sum[4];
    grads[g].opacity = sum[5];
    grads[g].color[0] = sum[6];
    grads[g].color[1] = sum[7];
    grads[g].color[2] = sum[8];
    abs_grads[2 * g] = sum[9];
    abs_grads[2 * g + 1] = sum[10];
  }
}

// Accuracy reference for binary32 backward passes (not a restatement of one reference
// function): the exact gradient of the binary32 forward's piecewise-smooth branch,
// evaluated in binary64.  It walks the pairs backward_tile (rasterize.rs:354-455) walks and
// takes every decision as the binary32 code takes it -- the sigma range and alpha_min
// tests of evaluate (rasterize.rs:231-244), the alpha_max clamp, the last contributor
// n_contrib of the binary32 forward -- but computes every value (G, alpha, the
// transmittance products, the accumulated colour, all slot sums) in binary64 from the
// binary32 inputs.  The final transmittance is recomputed in binary64 by the forward walk
// over the same pairs.  grads9: n x 9 doubles (ProjGrad order), abs2: n x 2.
inline void rasterize_backward_truth(const Proj<float>* projs, size_t np, const Binning& b,
                                     uint32_t width, uint32_t height, const float bgc[3],
                                     const uint32_t* n_contrib_img, const float* dimage,
                                     double* grads9, double* abs2) {
  size_t num_entries = b.sorted.size();
  std::vector<double> slots(num_entries * SLOTS, 0.0);
  size_t num_tiles = (size_t)b.tiles_x * b.tiles_y;
  const double amax32 = (double)Raster<float>::alpha_max();
  ORACLE_PAR
  for (size_t tile = 0; tile < num_tiles; tile++) {
    TileGeom geom = tile_geom(b, width, height, tile);
    uint32_t start = b.tile_range[tile].first, end = b.tile_range[tile].second;
    if (end == start) continue;
    std::vector<TileGauss<float>> list;
    for (uint32_t i = start; i < end; i++)
      list.push_back(tile_gauss(projs[b.entry_gaussian[b.sorted[i].second]], geom.x0, geom.y0,
                                geom.w, geom.h));
    // Binary64 value of one pair that the binary32 code keeps; false when it skips it.
    auto pair64 = [&](const TileGauss<float>& g, uint32_t i, uint32_t j, double* G64,
                      double* a64, bool* unclamped) {
      if (g.is_empty() || i < g.x0 || i >= g.x1 || j < g.y0 || j >= g.y1) return false;
      const float px = (float)(geom.x0 + i) + 0.5f, py = (float)(geom.y0 + j) + 0.5f;
      float G32, a32;
      if (!evaluate(g, px, py, &G32, &a32)) return false;
      const double dx = (double)px - g.mean[0], dy = (double)py - g.mean[1];
      const double sigma = 0.5 * ((double)g.conic[0] * dx * dx + (double)g.conic[2] * dy * dy) +
                           (double)g.conic[1] * dx * dy;
      *G64 = exp(-sigma);
      *unclamped = g.opacity * G32 < Raster<float>::alpha_max();
      *a64 = *unclamped ? (double)g.opacity * *G64 : amax32;
      return true;
    };
    std::vector<double> acc(list.size() * SLOTS, 0.0);
    for (uint32_t j = 0; j < geom.h; j++)
      for (uint32_t i = 0; i < geom.w; i++) {
        const size_t pix = (size_t)(geom.y0 + j) * width + geom.x0 + i;
        const uint32_t last = n_contrib_img[pix];
        const double d[3] = {dimage[pix * 3], dimage[pix * 3 + 1], dimage[pix * 3 + 2]};
        const double bg_dot = (double)bgc[0] * d[0] + (double)bgc[1] * d[1] + (double)bgc[2] * d[2];
        double t_final = 1.0;
        for (uint32_t kk = 0; kk < last && kk < list.size(); kk++) {
          double G, a;
          bool un;
          if (pair64(list[kk], i, j, &G, &a, &un)) t_final *= 1.0 - a;
        }
        double t = t_final, accum[3] = {0, 0, 0}, last_alpha = 0, last_color[3] = {0, 0, 0};
        const double px = (double)((float)(geom.x0 + i) + 0.5f);
        const double py = (double)((float)(geom.y0 + j) + 0.5f);
        for (size_t kk = std::min<size_t>(last, list.size()); kk-- > 0;) {
          const TileGauss<float>& g = list[kk];
          double G, alpha;
          bool unclamped;
          if (!pair64(g, i, j, &G, &alpha, &unclamped)) continue;
          double* a = &acc[kk * SLOTS];
          t /= 1.0 - alpha;
          double dalpha = 0;
          for (int ch = 0; ch < 3; ch++) {
            accum[ch] = last_alpha * last_color[ch] + (1.0 - last_alpha) * accum[ch];
            dalpha += ((double)g.color[ch] - accum[ch]) * d[ch];
            a[6 + ch] += alpha * t * d[ch];
          }
          dalpha *= t;
          for (int ch = 0; ch < 3; ch++) last_color[ch] = g.color[ch];
          last_alpha = alpha;
          dalpha += -t_final / (1.0 - alpha) * bg_dot;
          if (unclamped) {
            a[5] += G * dalpha;
            const double dsigma = -(double)g.opacity * dalpha * G;
            const double dx = px - g.mean[0], dy = py - g.mean[1];
            const double ddx = dsigma * ((double)g.conic[0] * dx + (double)g.conic[1] * dy);
            const double ddy = dsigma * ((double)g.conic[1] * dx + (double)g.conic[2] * dy);
            a[0] -= ddx;
            a[1] -= ddy;
            a[2] += dsigma * 0.5 * dx * dx;
            a[3] += dsigma * dx * dy;
            a[4] += dsigma * 0.5 * dy * dy;
            a[9] += fabs(ddx);
            a[10] += fabs(ddy);
          }
        }
      }
    for (size_t kk = 0; kk < list.size(); kk++) {
      uint32_t entry = b.sorted[start + kk].second;
      for (int s = 0; s < SLOTS; s++) slots[(size_t)entry * SLOTS + s] = acc[kk * SLOTS + s];
    }
  }
  ORACLE_PAR
  for (size_t g = 0; g < np; g++) {
    double sum[SLOTS] = {0};
    for (size_t entry = b.entry_offset[g]; entry < b.entry_offset[g + 1]; entry++)
      for (int k = 0; k < SLOTS; k++) sum[k] += slots[entry * SLOTS + k];
    for (int k = 0; k < 9; k++) grads9[g * 9 + k] = sum[k];
    abs2[2 * g] = sum[9];
    abs2[2 * g + 1] = sum[10];
  }
}

// ---------------------------------------------------------------- loss.rs ----------

static const int SSIM_WINDOW = 11;  // loss.rs:12

template <class R>
void gaussian_kernel(R k[SSIM_WINDOW]) {  // loss.rs:14-27
  R sigma = RL(1.5);
  R sum = 0;
  for (int i = 0; i < SSIM_WINDOW; i++) {
    R x = (R)i - (R)(SSIM_WINDOW / 2);
    k[i] = rexp(-(x * x) / (RL(2.0) * sigma * sigma));
    sum += k[i];
  }
  for (int i = 0; i < SSIM_WINDOW; i++) k[i] /= sum;
}

// blur, loss.rs:31-72.
template <class R>
void blur(const R* src, size_t width, size_t height, size_t channels, const R kernel[SSIM_WINDOW],
          R* dst) {
  const long half = SSIM_WINDOW / 2;
  const size_t row = width * channels;
  std::vector<R> tmp(width * height * channels, (R)0);
  ORACLE_PAR
  for (size_t y = 0; y < height; y++) {
    const R* line = src + y * row;
    R* out = &tmp[y * row];
    for (size_t x = 0; x < width; x++)
      for (size_t c = 0; c < channels; c++) {
        R acc = 0;
        for (int k = 0; k < SSIM_WINDOW; k++) {
          long xs = (long)x + k - half;
          if (xs >= 0 && (size_t)xs < width) acc += kernel[k] * line[(size_t)xs * channels + c];
        }
        out[x * channels + c] = acc;
      }
  }
  for (size_t i = 0; i < width * height * channels; i++) dst[i] = 0;
  ORACLE_PAR
  for (size_t y = 0; y < height; y++) {
    R* out = dst + y * row;
    for (int k = 0; k < SSIM_WINDOW; k++) {
      long ys = (long)y + k - half;
      if (ys >= 0 && (size_t)ys < height) {
        const R* line = &tmp[(size_t)ys * row];
        R kv = kernel[k];
        for (size_t i = 0; i < row; i++) out[i] += kv * line[i];
      }
    }
  }
}

// ssim_with_grad, loss.rs:75-134.  The reference sums the SSIM map with rayon's parallel
// `sum` (thread-dependent order); this sums it sequentially.  sum64, if given, receives
// the binary64 sum of the same map values (the accuracy reference for reordered sums).
template <class R>
R ssim_with_grad(const R* x, const R* y, size_t width, size_t height, R* grad,
                 double* sum64 = nullptr) {
  size_t n = width * height * 3;
  R kernel[SSIM_WINDOW];
  gaussian_kernel(kernel);
  R c1 = RL(0.01) * RL(0.01);
  R c2 = RL(0.03) * RL(0.03);
  std::vector<R> mu_x(n), mu_y(n), xx(n), yy(n), xy(n), e_xx(n), e_yy(n), e_xy(n);
  blur(x, width, height, 3, kernel, mu_x.data());
  blur(y, width, height, 3, kernel, mu_y.data());
  for (size_t i = 0; i < n; i++) {
    xx[i] = x[i] * x[i];
    yy[i] = y[i] * y[i];
    xy[i] = x[i] * y[i];
  }
  blur(xx.data(), width, height, 3, kernel, e_xx.data());
  blur(yy.data(), width, height, 3, kernel, e_yy.data());
  blur(xy.data(), width, height, 3, kernel, e_xy.data());
  std::vector<R> d_mu(n), d_exx(n), d_exy(n), smap(n);
  ORACLE_PAR
  for (size_t i = 0; i < n; i++) {
    R mx = mu_x[i], my = mu_y[i];
    R sx2 = e_xx[i] - mx * mx;
    R sy2 = e_yy[i] - my * my;
    R sxy = e_xy[i] - mx * my;
    R n1 = RL(2.0) * mx * my + c1;
    R n2 = RL(2.0) * sxy + c2;
    R d1 = mx * mx + my * my + c1;
    R d2 = sx2 + sy2 + c2;
    R num = n1 * n2;
    R den = d1 * d2;
    R s = num / den;
    R dn_dmx = RL(2.0) * my * n2 - RL(2.0) * my * n1;
    R dd_dmx = RL(2.0) * mx * d2 - RL(2.0) * mx * d1;
    d_mu[i] = (dn_dmx * den - num * dd_dmx) / (den * den);
    d_exx[i] = -num * d1 / (den * den);
    d_exy[i] = RL(2.0) * n1 / den;
    smap[i] = s;
  }
  R ssim_sum = 0;
  for (size_t i = 0; i < n; i++) ssim_sum += smap[i];
  if (sum64) {
    double acc = 0;
    for (size_t i = 0; i < n; i++) acc += (double)smap[i];
    *sum64 = acc;
  }
  R mean = ssim_sum / (R)n;
  std::vector<R> b_mu(n), b_exx(n), b_exy(n);
  blur(d_mu.data(), width, height, 3, kernel, b_mu.data());
  blur(d_exx.data(), width, height, 3, kernel, b_exx.data());
  blur(d_exy.data(), width, height, 3, kernel, b_exy.data());
  R inv_n = RL(1.0) / (R)n;
  ORACLE_PAR
  for (size_t i = 0; i < n; i++)
    grad[i] = (b_mu[i] + RL(2.0) * x[i] * b_exx[i] + y[i] * b_exy[i]) * inv_n;
  return mean;
}

// photometric_loss, loss.rs:137-171 (the L1 sum is sequential here, rayon's `sum` in the
// reference).  sums64, if given, receives the binary64 sums of the SSIM map and of the
// |pred - target| terms, each term as the reference rounds it.
template <class R>
R photometric_loss(const R* pred, const R* target, size_t width, size_t height, R lambda,
                   R* grad, double* sums64 = nullptr) {
  size_t n = width * height * 3;
  R inv_n = RL(1.0) / (R)n;
  R l1 = 0;
  for (size_t i = 0; i < n; i++) l1 += rabs(pred[i] - target[i]);
  if (sums64) {
    double acc = 0;
    for (size_t i = 0; i < n; i++) acc += (double)rabs(pred[i] - target[i]);
    sums64[1] = acc;
  }
  l1 = l1 * inv_n;
  std::vector<R> dssim(n);
  R ssim = ssim_with_grad(pred, target, width, height, dssim.data(), sums64);
  R loss = (RL(1.0) - lambda) * l1 + lambda * (RL(1.0) - ssim);
  R w1 = (RL(1.0) - lambda) * inv_n;
  ORACLE_PAR
  for (size_t i = 0; i < n; i++) {
    R d = pred[i] - target[i];
    R sign = d > (R)0 ? RL(1.0) : (d < (R)0 ? RL(-1.0) : RL(0.0));
    grad[i] = w1 * sign - lambda * dssim[i];
  }
  return loss;
}

// psnr, loss.rs:175-190.
template <class R>
R psnr(const R* pred, const R* target, size_t len) {
  double mse = 0;
  for (size_t i = 0; i < len; i++) {
    double d = (double)rclamp(pred[i], (R)0, (R)1) - (double)rclamp(target[i], (R)0, (R)1);
    mse += d * d;
  }
  mse = mse / (double)(len > 0 ? len : 1);
  if (mse <= 0) return (R)INFINITY;
  return (R)(-10.0 * log10(mse));
}

// ---------------------------------------------------------------- adam.rs ----------

// Real::powi via the compiler runtime's square-and-multiply (__powisf2 / __powidf2).
template <class R>
R powi(R a, int b) {
  bool recip = b < 0;
  unsigned int e = recip ? 0u - (unsigned)b : (unsigned)b;
  R r = 1;
  while (true) {
    if (e & 1u) r *= a;
    e >>= 1;
    if (e == 0) break;
    a *= a;
  }
  return recip ? (R)1 / r : r;
}

// update, adam.rs:44-61.
template <class R>
inline void adam_update(R* p, R g, R* m, R* v, R lr, R b1, R b2, R c1, R c2, R eps) {
  *m = b1 * *m + (RL(1.0) - b1) * g;
  *v = b2 * *v + (RL(1.0) - b2) * g * g;
  R m_hat = *m / c1;
  R v_hat = *v / c2;
  *p -= lr * m_hat / (rsqrt_(v_hat) + eps);
}

// Adam::step, adam.rs:77-161.  lr = {mean, log_scale, quat, opacity, sh_dc, sh_rest}.
template <class R>
void adam_step(R* params, const R* grads, R* m, R* v, size_t n, uint64_t* step, R b1, R b2,
               R eps, const R lr[6]) {
  *step += 1;
  R c1 = RL(1.0) - powi(b1, (int)*step);
  R c2 = RL(1.0) - powi(b2, (int)*step);
  ORACLE_PAR
  for (size_t gi = 0; gi < n; gi++) {
    R* p = params + gi * 59;
    const R* g = grads + gi * 59;
    R* mm = m + gi * 59;
    R* vv = v + gi * 59;
    for (int i = 0; i < 3; i++) {
      adam_update(p + i, g[i], mm + i, vv + i, lr[0], b1, b2, c1, c2, eps);
      adam_update(p + 3 + i, g[3 + i], mm + 3 + i, vv + 3 + i, lr[1], b1, b2, c1, c2, eps);
    }
    for (int i = 0; i < 4; i++)
      adam_update(p + 6 + i, g[6 + i], mm + 6 + i, vv + 6 + i, lr[2], b1, b2, c1, c2, eps);
    adam_update(p + 10, g[10], mm + 10, vv + 10, lr[3], b1, b2, c1, c2, eps);
    for (int k = 0; k < 16; k++) {
      R rate = k == 0 ? lr[4] : lr[5];
      for (int c = 0; c < 3; c++) {
        int o = 11 + 3 * k + c;
        adam_update(p + o, g[o], mm + o, vv + o, rate, b1, b2, c1, c2, eps);
      }
    }
  }
}

// ---------------------------------------------------------------- densify.rs -------

// DensifyStats::accumulate, densify.rs:36-60.
template <class R>
void stats_accumulate(const Proj<R>* projs, const ProjGrad<R>* grads, const R* abs_grads,
                      size_t n, uint32_t width, uint32_t height, R* grad_accum, uint32_t* denom,
                      uint32_t* max_radii) {
  R sx = (R)width * RL(0.5);
  R sy = (R)height * RL(0.5);
  for (size_t i = 0; i < n; i++) {
    if (!projs[i].visible()) continue;
    R g0 = abs_grads ? abs_grads[2 * i] : grads[i].mean2d[0];
    R g1 = abs_grads ? abs_grads[2 * i + 1] : grads[i].mean2d[1];
    R gx = g0 * sx;
    R gy = g1 * sy;
    grad_accum[i] += rsqrt_(gx * gx + gy * gy);
    denom[i] += 1;
    max_radii[i] = std::max(max_radii[i], projs[i].radius);
  }
}

template <class R>
struct DensifyConfig {  // densify.rs:69-103
  R grad_threshold, percent_dense, min_opacity;
  bool prune_big_world;
  R max_world_size_frac;
  bool has_max_screen_size;
  uint32_t max_screen_size;
  uint32_t split_children;
};

struct DensifyReport {
  uint64_t cloned, split, pruned;
};

// densify_and_prune, densify.rs:125-221.  The standard normal draws of the split children
// (densify.rs:116-121, 174-179) are an input in draw order.  Rows are 59 scalars; the
// statistics are reset (to n_out zeros) by the caller.  Returns false if `normals` is too
// short (report.split tells how many parents were split).
template <class R>
bool densify_and_prune(std::vector<R>& params, std::vector<R>& m, std::vector<R>& v,
                       const R* grad_accum, const uint32_t* denom, const uint32_t* max_radii,
                       const DensifyConfig<R>& cfg, R extent, const R* normals,
                       size_t normals_len, DensifyReport& report) {
  const size_t P = 59;
  size_t n0 = params.size() / P;
  std::vector<R> grads(n0);
  for (size_t i = 0; i < n0; i++)
    grads[i] = denom[i] > 0 ? grad_accum[i] / (R)denom[i] : (R)0;
  R size_limit = cfg.percent_dense * extent;
  report = DensifyReport{0, 0, 0};
  // origin: index into stats, or -1 for new rows.
  std::vector<long> origin(n0);
  for (size_t i = 0; i < n0; i++) origin[i] = (long)i;

  // Clone (densify.rs:150-159).
  std::vector<size_t> clone_idx;
  for (size_t i = 0; i < n0; i++) {
    GaussianView<R> g{&params[i * P]};
    if (grads[i] >= cfg.grad_threshold && g.scales().max_elem() <= size_limit) clone_idx.push_back(i);
  }
  report.cloned = clone_idx.size();
  for (size_t i : clone_idx)
    for (size_t k = 0; k < P; k++) params.push_back(params[i * P + k]);
  m.resize(params.size(), (R)0);
  v.resize(params.size(), (R)0);
  origin.resize(params.size() / P, -1);

  // Split (densify.rs:162-198).
  std::vector<size_t> split_idx;
  for (size_t i = 0; i < n0; i++) {
    GaussianView<R> g{&params[i * P]};
    if (grads[i] >= cfg.grad_threshold && g.scales().max_elem() > size_limit) split_idx.push_back(i);
  }
  report.split = split_idx.size();
  if (normals_len < 3 * (size_t)cfg.split_children * split_idx.size()) return false;
  R shrink = rln(RL(0.8) * (R)cfg.split_children);
  std::vector<R> children;
  size_t draw = 0;
  for (size_t i : split_idx) {
    std::vector<R> g(params.begin() + i * P, params.begin() + (i + 1) * P);
    GaussianView<R> gv{g.data()};
    V3<R> s = gv.scales();
    M3<R> rot = gv.rotation().to_mat();
    for (uint32_t c = 0; c < cfg.split_children; c++) {
      R n0_ = normals[draw], n1_ = normals[draw + 1], n2_ = normals[draw + 2];
      draw += 3;
      V3<R> local(n0_ * s.x, n1_ * s.y, n2_ * s.z);
      V3<R> mean = gv.mean() + rot.mul_v(local);
      std::vector<R> child = g;
      child[0] = mean.x;
      child[1] = mean.y;
      child[2] = mean.z;
      for (int k = 0; k < 3; k++) child[3 + k] = g[3 + k] - shrink;
      children.insert(children.end(), child.begin(), child.end());
    }
  }
  params.insert(params.end(), children.begin(), children.end());
  m.resize(params.size(), (R)0);
  v.resize(params.size(), (R)0);
  origin.resize(params.size() / P, -1);
  std::vector<char> keep(params.size() / P, 1);
  for (size_t i : split_idx) keep[i] = 0;
  auto retain = [&](std::vector<char>& kp) {
    std::vector<R> p2, m2, v2;
    std::vector<long> o2;
    for (size_t i = 0; i < kp.size(); i++)
      if (kp[i]) {
        p2.insert(p2.end(), params.begin() + i * P, params.begin() + (i + 1) * P);
        m2.insert(m2.end(), m.begin() + i * P, m.begin() + (i + 1) * P);
        v2.insert(v2.end(), v.begin() + i * P, v.begin() + (i + 1) * P);
        o2.push_back(origin[i]);
      }
    params.swap(p2);
    m.swap(m2);
    v.swap(v2);
    origin.swap(o2);
  };
  retain(keep);

  // Prune (densify.rs:203-218).
  size_t n2 = params.size() / P;
  std::vector<char> keep2(n2, 1);
  for (size_t i = 0; i < n2; i++) {
    GaussianView<R> g{&params[i * P]};
    bool prune = g.alpha() < cfg.min_opacity;
    if (cfg.prune_big_world) prune |= g.scales().max_elem() > cfg.max_world_size_frac * extent;
    if (cfg.has_max_screen_size && origin[i] >= 0)
      prune |= max_radii[origin[i]] > cfg.max_screen_size;
    keep2[i] = !prune;
  }
  for (char k : keep2) report.pruned += k ? 0 : 1;
  retain(keep2);
  return true;
}

// reset_opacity, densify.rs:224-230 (+ Adam::reset_opacity_state, adam.rs:176-183).
template <class R>
void reset_opacity(R* params, R* m, R* v, size_t n, R max_alpha) {
  R limit = inverse_sigmoid(max_alpha);
  for (size_t i = 0; i < n; i++) {
    params[i * 59 + 10] = rmin(params[i * 59 + 10], limit);
    m[i * 59 + 10] = 0;
    v[i * 59 + 10] = 0;
  }
}

// ---------------------------------------------------------------- model.rs init ----

// mean_nn_dist2 (model.rs:128-186) with the kd-tree query replaced by brute force over the
// unique positions.  As the reference: coincident points (same bits) collapse to one
// position with a multiplicity; the query takes the min(unique, 4) nearest unique positions
// (self included, distance 0), walks them by increasing distance adding
// `distance * take` until three other points are counted.  f64 throughout.
inline void mean_nn_dist2(const double* pts, size_t n, double* out) {
  if (n < 2) {
    for (size_t i = 0; i < n; i++) out[i] = 0;
    return;
  }
  std::vector<size_t> uniq_of(n);
  std::vector<size_t> rep;       // representative point of each unique position
  std::vector<uint32_t> count;
  {
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; i++) order[i] = i;
    auto key = [&](size_t i) {
      uint64_t k[3];
      std::memcpy(k, pts + 3 * i, 24);
      return std::make_tuple(k[0], k[1], k[2]);
    };
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return key(a) < key(b); });
    for (size_t t = 0; t < n; t++) {
      size_t i = order[t];
      if (t == 0 || key(i) != key(order[t - 1])) {
        rep.push_back(i);
        count.push_back(0);
      }
      uniq_of[i] = rep.size() - 1;
      count.back()++;
    }
  }
  size_t nu = rep.size();
  size_t kq = std::min<size_t>(nu, 4);
  std::vector<std::pair<double, size_t>> d(nu);
  for (size_t i = 0; i < n; i++) {
    size_t u = uniq_of[i];
    for (size_t w = 0; w < nu; w++) {
      size_t j = rep[w];
      double dx = pts[3 * i] - pts[3 * j], dy = pts[3 * i + 1] - pts[3 * j + 1],
             dz = pts[3 * i + 2] - pts[3 * j + 2];
      d[w] = std::make_pair(w == u ? 0.0 : dx * dx + dy * dy + dz * dz, w == u ? 0 : 1 + w);
    }
    std::partial_sort(d.begin(), d.begin() + kq, d.end());
    double sum = 0;
    uint32_t cnt = 0;
    for (size_t t = 0; t < kq; t++) {
      size_t w = d[t].second == 0 ? u : d[t].second - 1;
      uint32_t mult = w == u ? count[w] - 1 : count[w];
      uint32_t take = std::min<uint32_t>(mult, 3 - cnt);
      sum += d[t].first * (double)take;
      cnt += take;
      if (cnt == 3) break;
    }
    out[i] = cnt > 0 ? sum / (double)cnt : 0.0;
  }
}

}  // namespace oracle
