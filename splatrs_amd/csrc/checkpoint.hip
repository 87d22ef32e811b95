// Resumable training checkpoints, host code.  The reference saves only the model (.ply,
// src/ply.rs:19-69; its train loop keeps no optimiser state on disk); a checkpoint adds
// what a resumed run needs to continue exactly: the Adam moments and step count and the
// densification statistics.
//
// Layout (little endian): "SPLATCKP" | u32 version (1) | i32 max_sh_degree | u64 n |
// u64 adam_step | params n x 59 f32 | m n x 59 f32 | v n x 59 f32 | grad_accum n f32 |
// denom n u32 | max_radii n u32.  Rows use the canonical 59-scalar order
// (tests/common/mod.rs:111-139).
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/splat.h"

namespace {

constexpr int kParams = 59;
constexpr char kMagic[8] = {'S', 'P', 'L', 'A', 'T', 'C', 'K', 'P'};
constexpr uint32_t kVersion = 1;

int set_err(char* err, size_t len, int code, const std::string& msg) {
  if (err && len) std::snprintf(err, len, "%s", msg.c_str());
  return code;
}

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

// Writes `count` items of `size` bytes, or zeros when src is null.
bool put(FILE* f, const void* src, size_t size, uint64_t count) {
  if (src) return std::fwrite(src, size, count, f) == count;
  std::vector<char> zero(1 << 16, 0);
  uint64_t bytes = (uint64_t)size * count;
  while (bytes) {
    const size_t k = (size_t)std::min<uint64_t>(bytes, zero.size());
    if (std::fwrite(zero.data(), 1, k, f) != k) return false;
    bytes -= k;
  }
  return true;
}

// Reads `count` items into dst, or skips them when dst is null.
bool get(FILE* f, void* dst, size_t size, uint64_t count) {
  if (dst) return std::fread(dst, size, count, f) == count;
  return std::fseek(f, (long)(size * count), SEEK_CUR) == 0;
}

}  // namespace

extern "C" {

int splat_checkpoint_write(const char* path, const float* gaussians, const float* m,
                           const float* v, const float* grad_accum, const uint32_t* denom,
                           const uint32_t* max_radii, uint64_t n, uint64_t adam_step,
                           int32_t max_sh_degree, char* err, size_t err_len) {
  if (!path || (n && !gaussians)) return set_err(err, err_len, SPLAT_ERR_INVALID, "null argument");
  if (max_sh_degree < 0 || max_sh_degree > 3)
    return set_err(err, err_len, SPLAT_ERR_INVALID, "sh degree must be 0..3");
  File file;
  file.f = std::fopen(path, "wb");
  if (!file.f) return set_err(err, err_len, SPLAT_ERR_IO, std::string("cannot create ") + path);
  FILE* f = file.f;
  const uint64_t rows = n * kParams;
  bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && put(f, &kVersion, 4, 1) &&
            put(f, &max_sh_degree, 4, 1) && put(f, &n, 8, 1) && put(f, &adam_step, 8, 1) &&
            put(f, gaussians, 4, rows) && put(f, m, 4, rows) && put(f, v, 4, rows) &&
            put(f, grad_accum, 4, n) && put(f, denom, 4, n) && put(f, max_radii, 4, n);
  ok = ok && std::fflush(f) == 0;
  if (!ok) return set_err(err, err_len, SPLAT_ERR_IO, std::string("write failed: ") + path);
  return SPLAT_OK;
}

int splat_checkpoint_read(const char* path, float* gaussians, float* m, float* v,
                          float* grad_accum, uint32_t* denom, uint32_t* max_radii,
                          uint64_t capacity, uint64_t* n, uint64_t* adam_step,
                          int32_t* max_sh_degree, char* err, size_t err_len) {
  if (!path || !n) return set_err(err, err_len, SPLAT_ERR_INVALID, "null argument");
  File file;
  file.f = std::fopen(path, "rb");
  if (!file.f) return set_err(err, err_len, SPLAT_ERR_IO, std::string("cannot open ") + path);
  FILE* f = file.f;
  char magic[8];
  uint32_t version = 0;
  int32_t degree = 0;
  uint64_t count = 0, step = 0;
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kMagic, 8) != 0)
    return set_err(err, err_len, SPLAT_ERR_FORMAT, "not a checkpoint file");
  if (!get(f, &version, 4, 1) || version != kVersion)
    return set_err(err, err_len, SPLAT_ERR_FORMAT, "unsupported checkpoint version");
  if (!get(f, &degree, 4, 1) || !get(f, &count, 8, 1) || !get(f, &step, 8, 1) || degree < 0 ||
      degree > 3)
    return set_err(err, err_len, SPLAT_ERR_FORMAT, "malformed checkpoint header");
  // the file must hold exactly the arrays the header announces (the count is checked
  // against the remaining bytes before it is multiplied, so no size can wrap)
  constexpr uint64_t kRowBytes = 3ull * kParams * 4 + 12;
  const long here = std::ftell(f);
  if (here < 0 || std::fseek(f, 0, SEEK_END) != 0)
    return set_err(err, err_len, SPLAT_ERR_IO, "cannot size checkpoint");
  const long end = std::ftell(f);
  if (end < here || count > (uint64_t)(end - here) / kRowBytes ||
      (uint64_t)(end - here) != count * kRowBytes)
    return set_err(err, err_len, SPLAT_ERR_FORMAT, "truncated or oversized checkpoint");
  std::fseek(f, here, SEEK_SET);
  *n = count;
  if (adam_step) *adam_step = step;
  if (max_sh_degree) *max_sh_degree = degree;
  if (capacity < count) return set_err(err, err_len, SPLAT_ERR_CAPACITY, "output capacity too small");
  const uint64_t rows = count * kParams;
  const bool ok = get(f, gaussians, 4, rows) && get(f, m, 4, rows) && get(f, v, 4, rows) &&
                  get(f, grad_accum, 4, count) && get(f, denom, 4, count) &&
                  get(f, max_radii, 4, count);
  if (!ok) return set_err(err, err_len, SPLAT_ERR_IO, "read failed");
  return SPLAT_OK;
}

}  // extern "C"
