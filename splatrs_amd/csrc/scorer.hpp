// The scorer object behind splat_scorer_* (shared by scorer.hip and the trainer).
#pragma once

#include "internal.hpp"

struct splat_scorer {
  splat_ctx* ctx = nullptr;
  int device = 0;
  splat::DevBuf params, m, v;  // binary32 master weights and Adam moments
  splat::DevBuf packed;        // bfloat16 matrix-core operand blocks of the weights
  uint64_t adam_step = 0;
  std::map<std::string, splat::DevBuf> scratch;
};

namespace splat {
// Replace the weights (host array of kMlpParams floats); resets the Adam state.
void scorer_upload(splat_scorer* s, const float* host_params);
// Epochs of minibatch Adam over n device-resident samples; returns the last epoch's mean
// minibatch loss.
float scorer_fit_device(splat_scorer* s, const float* feats, const uint8_t* labels,
                        const float* weights, size_t n, const splat_scorer_fit_config& cfg);
}  // namespace splat
