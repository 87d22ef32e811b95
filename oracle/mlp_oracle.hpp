// CPU oracle for the learned densification scorer -- TEST INFRASTRUCTURE ONLY.
//
// The scorer has no counterpart in the reference (its densification is the threshold
// heuristic of src/densify.rs:125-221).  Its specification: 12 per-Gaussian features ->
// three hidden layers of 64 units with the tanh-form GELU -> 3 logits (clone, split,
// prune), weights and activations rounded to bfloat16 (round to nearest even), sums in
// binary32; trained with a weighted binary cross-entropy on the logits.
//
// This file restates that specification plainly (loops in index order) for the GPU
// kernels to be checked against.  Templated on the scalar type and on whether values are
// rounded to bfloat16, so the backward pass can be checked by finite differences and by
// an autograd library in binary64 without rounding.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace oracle {
namespace mlp {

constexpr int kIn = 12, kHid = 64, kOut = 3;
// Flat parameter layout: W1[64][12] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[3][64] b4[3]
constexpr int kW1 = 0, kB1 = kW1 + kHid * kIn, kW2 = kB1 + kHid, kB2 = kW2 + kHid * kHid,
              kW3 = kB2 + kHid, kB3 = kW3 + kHid * kHid, kW4 = kB3 + kHid,
              kB4 = kW4 + kOut * kHid, kNumParams = kB4 + kOut;

// binary32 -> bfloat16 (round to nearest, ties to even) -> binary32.  Finite inputs only.
inline float bf16_round(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

template <class R, bool BF>
inline R rnd(R x) {
  if constexpr (BF) return (R)bf16_round((float)x);
  return x;
}

inline float xexp(float x) { return expf(x); }
inline double xexp(double x) { return exp(x); }

// gelu(z) = 0.5 z (1 + tanh(u)), u = c (z + 0.044715 z^3), tanh(u) = 1 - 2 / (1 + e^{2u}).
template <class R>
inline R gelu_tanh(R z, R* dgelu) {
  const R c = (R)0.7978845608028654;
  const R k = (R)0.044715;
  const R u = c * (z + k * z * z * z);
  const R t = (R)1 - (R)2 / ((R)1 + xexp((R)2 * u));
  if (dgelu) {
    const R du = c * ((R)1 + (R)3 * k * z * z);
    *dgelu = (R)0.5 * ((R)1 + t) + (R)0.5 * z * ((R)1 - t * t) * du;
  }
  return (R)0.5 * z * ((R)1 + t);
}

template <class R>
struct Trace {
  R x[kIn];
  R z1[kHid], a1[kHid], d1[kHid];
  R z2[kHid], a2[kHid], d2[kHid];
  R z3[kHid], a3[kHid], d3[kHid];
  R z4[kOut];
};

// One dense layer: z[o] = b[o] + sum_i w[o][i] a[i] (i ascending), weights rounded.
template <class R, bool BF>
inline void dense(const R* w, const R* b, const R* a, int nin, int nout, R* z) {
  for (int o = 0; o < nout; o++) {
    R s = (R)0;
    for (int i = 0; i < nin; i++) s += rnd<R, BF>(w[o * nin + i]) * a[i];
    z[o] = s + b[o];
  }
}

template <class R, bool BF>
inline void forward(const R* p, const R* feat, Trace<R>& t) {
  for (int i = 0; i < kIn; i++) t.x[i] = rnd<R, BF>(feat[i]);
  dense<R, BF>(p + kW1, p + kB1, t.x, kIn, kHid, t.z1);
  for (int o = 0; o < kHid; o++) t.a1[o] = rnd<R, BF>(gelu_tanh(t.z1[o], &t.d1[o]));
  dense<R, BF>(p + kW2, p + kB2, t.a1, kHid, kHid, t.z2);
  for (int o = 0; o < kHid; o++) t.a2[o] = rnd<R, BF>(gelu_tanh(t.z2[o], &t.d2[o]));
  dense<R, BF>(p + kW3, p + kB3, t.a2, kHid, kHid, t.z3);
  for (int o = 0; o < kHid; o++) t.a3[o] = rnd<R, BF>(gelu_tanh(t.z3[o], &t.d3[o]));
  dense<R, BF>(p + kW4, p + kB4, t.a3, kHid, kOut, t.z4);
}

template <class R>
inline R softplus(R z) {  // log(1 + e^z), stable
  return z > (R)0 ? z + std::log1p(xexp(-z)) : std::log1p(xexp(z));
}

template <class R>
inline R sigm(R z) { return (R)1 / ((R)1 + xexp(-z)); }

// Weighted binary cross-entropy over the batch:
//   L = (1/B) sum_s w_s sum_o [softplus(z_so) - y_so z_so],  y_so = bit o of labels[s].
// Accumulates dL/dparams into grad (zeroed here) and returns L.  The backward rounds each
// layer's error signal to bfloat16 before it is propagated or multiplied into a weight
// gradient (as the GPU's matrix-core operands are), and uses the rounded weights.
template <class R, bool BF>
inline R loss_and_grad(const R* p, const R* feats, const uint8_t* labels, const R* weights,
                       int64_t n, R* grad) {
  for (int k = 0; k < kNumParams; k++) grad[k] = (R)0;
  R loss = (R)0;
  const R invb = (R)1 / (R)n;
  Trace<R> t;
  for (int64_t s = 0; s < n; s++) {
    forward<R, BF>(p, feats + s * kIn, t);
    R e4[kOut], e3[kHid], e2[kHid], e1[kHid];
    for (int o = 0; o < kOut; o++) {
      const R y = (labels[s] >> o) & 1 ? (R)1 : (R)0;
      loss += weights[s] * (softplus(t.z4[o]) - y * t.z4[o]) * invb;
      e4[o] = rnd<R, BF>(weights[s] * (sigm(t.z4[o]) - y) * invb);
    }
    auto back = [&](const R* w, const R* eout, int nin, int nout, const R* dact, R* ein) {
      for (int i = 0; i < nin; i++) {
        R sacc = (R)0;
        for (int o = 0; o < nout; o++) sacc += rnd<R, BF>(w[o * nin + i]) * eout[o];
        ein[i] = rnd<R, BF>(sacc * dact[i]);
      }
    };
    back(p + kW4, e4, kHid, kOut, t.d3, e3);
    back(p + kW3, e3, kHid, kHid, t.d2, e2);
    back(p + kW2, e2, kHid, kHid, t.d1, e1);
    auto wgrad = [&](R* gw, R* gb, const R* eout, const R* ain, int nin, int nout) {
      for (int o = 0; o < nout; o++) {
        for (int i = 0; i < nin; i++) gw[o * nin + i] += eout[o] * ain[i];
        gb[o] += eout[o];
      }
    };
    wgrad(grad + kW4, grad + kB4, e4, t.a3, kHid, kOut);
    wgrad(grad + kW3, grad + kB3, e3, t.a2, kHid, kHid);
    wgrad(grad + kW2, grad + kB2, e2, t.a1, kHid, kHid);
    wgrad(grad + kW1, grad + kB1, e1, t.x, kIn, kHid);
  }
  return loss;
}

// Adam on the scorer's binary32 master weights (Kingma & Ba, bias-corrected; the same
// update rule as src/adam.rs:44-61), with the bias corrections c1, c2 given.
inline void adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1,
                 float b2, float eps, float c1, float c2) {
  for (int k = 0; k < n; k++) {
    m[k] = b1 * m[k] + (1.0f - b1) * g[k];
    v[k] = b2 * v[k] + (1.0f - b2) * g[k] * g[k];
    p[k] -= lr * (m[k] / c1) / (std::sqrt(v[k] / c2) + eps);
  }
}

}  // namespace mlp
}  // namespace oracle
