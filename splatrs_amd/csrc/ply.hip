// The standard 3DGS .ply layout (src/ply.rs:19-181), host code.
//
// Per vertex, float32 little endian: x y z nx ny nz f_dc_0..2 f_rest_0..(3(K-1)-1)
// opacity scale_0..2 rot_0..3 with K = (degree + 1)^2 and f_rest channel-major
// (f_rest[c (K-1) + (k-1)] = coefficient k of channel c); normals written as zeros.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/splat.h"

namespace {

constexpr int kParams = 59;

int set_err(char* err, size_t len, int code, const std::string& msg) {
  if (err && len) {
    std::snprintf(err, len, "%s", msg.c_str());
  }
  return code;
}

std::vector<std::string> split_ws(const std::string& line) {
  std::vector<std::string> out;
  std::istringstream is(line);
  std::string t;
  while (is >> t) out.push_back(t);
  return out;
}

}  // namespace

extern "C" {

int splat_ply_write(const char* path, const float* gaussians, uint64_t n, int32_t max_sh_degree,
                    char* err, size_t err_len) {
  if (!path || (n && !gaussians)) return set_err(err, err_len, SPLAT_ERR_INVALID, "null argument");
  const int degree = max_sh_degree < 0 ? 0 : (max_sh_degree > 3 ? 3 : max_sh_degree);
  const int k = (degree + 1) * (degree + 1);
  const int n_rest = 3 * (k - 1);
  std::string header = "ply\nformat binary_little_endian 1.0\n";
  header += "element vertex " + std::to_string(n) + "\n";
  for (const char* name : {"x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"})
    header += std::string("property float ") + name + "\n";
  for (int i = 0; i < n_rest; i++) header += "property float f_rest_" + std::to_string(i) + "\n";
  header += "property float opacity\n";
  for (int i = 0; i < 3; i++) header += "property float scale_" + std::to_string(i) + "\n";
  for (int i = 0; i < 4; i++) header += "property float rot_" + std::to_string(i) + "\n";
  header += "end_header\n";
  FILE* f = std::fopen(path, "wb");
  if (!f) return set_err(err, err_len, SPLAT_ERR_IO, std::string("cannot create ") + path);
  bool ok = std::fwrite(header.data(), 1, header.size(), f) == header.size();
  std::vector<float> row(17 + n_rest);
  for (uint64_t i = 0; ok && i < n; i++) {
    const float* g = gaussians + i * kParams;
    size_t r = 0;
    for (int c = 0; c < 3; c++) row[r++] = g[c];
    for (int c = 0; c < 3; c++) row[r++] = 0.0f;
    for (int c = 0; c < 3; c++) row[r++] = g[11 + c];
    for (int c = 0; c < 3; c++)
      for (int kk = 1; kk < k; kk++) row[r++] = g[11 + 3 * kk + c];
    row[r++] = g[10];
    for (int c = 0; c < 3; c++) row[r++] = g[3 + c];
    for (int c = 0; c < 4; c++) row[r++] = g[6 + c];
    // x86-64 and gfx950 hosts are little endian: the floats are the file bytes.
    ok = std::fwrite(row.data(), 4, r, f) == r;
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return set_err(err, err_len, SPLAT_ERR_IO, std::string("write failed: ") + path);
  return SPLAT_OK;
}

int splat_ply_read(const char* path, float* gaussians, uint64_t capacity, uint64_t* n_out,
                   int32_t* sh_degree, char* err, size_t err_len) {
  if (!path || !n_out || !sh_degree) return set_err(err, err_len, SPLAT_ERR_INVALID, "null argument");
  std::ifstream in(path, std::ios::binary);
  if (!in) return set_err(err, err_len, SPLAT_ERR_IO, std::string("cannot open ") + path);
  std::vector<char> bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  const std::string p(path);
  auto fmt = [&](const std::string& m) {
    return set_err(err, err_len, SPLAT_ERR_FORMAT, p + ": " + m);
  };
  static const char kEnd[] = "end_header\n";
  const size_t kEndLen = sizeof(kEnd) - 1;
  size_t end = std::string::npos;
  for (size_t i = 0; i + kEndLen <= bytes.size(); i++)
    if (std::memcmp(&bytes[i], kEnd, kEndLen) == 0) {
      end = i;
      break;
    }
  if (end == std::string::npos) return fmt("missing end_header");
  const std::string header(bytes.data(), end);
  const size_t body_off = end + kEndLen;
  bool have_count = false, in_vertex = false, binary_le = false;
  uint64_t count = 0;
  std::vector<std::string> props;
  std::istringstream hs(header);
  std::string line;
  while (std::getline(hs, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();  // str::lines() strips \r\n
    const auto t = split_ws(line);
    if (t.size() == 3 && t[0] == "format" && t[1] == "binary_little_endian") {
      binary_le = true;
    } else if (!t.empty() && t[0] == "format") {
      return fmt("only binary_little_endian is supported");
    } else if (t.size() == 3 && t[0] == "element" && t[1] == "vertex") {
      in_vertex = true;
      char* e = nullptr;
      const unsigned long long v = std::strtoull(t[2].c_str(), &e, 10);
      if (t[2].empty() || *e != '\0' || t[2][0] == '-') return fmt("bad vertex count");
      count = v;
      have_count = true;
    } else if (!t.empty() && t[0] == "element") {
      in_vertex = false;
    } else if (t.size() == 3 && t[0] == "property" && in_vertex) {
      if (t[1] != "float" && t[1] != "float32")
        return fmt("property " + t[2] + " has type " + t[1] + "; only float is supported");
      props.push_back(t[2]);
    }
  }
  if (!binary_le) return fmt("missing format line");
  if (!have_count) return fmt("missing vertex element");
  const uint64_t stride = props.size() * 4;
  if (stride && count > (uint64_t)-1 / stride) return fmt("vertex count overflows");
  const uint64_t needed = count * stride;
  if (bytes.size() - body_off < needed) return fmt("file is shorter than the header announces");
  std::string missing;
  auto index = [&](const std::string& name) -> int {
    for (size_t i = 0; i < props.size(); i++)
      if (props[i] == name) return (int)i;
    if (missing.empty()) missing = name;
    return -1;
  };
  int ix[3], idc[3], isc[3], irot[4];
  const char* xyz[3] = {"x", "y", "z"};
  for (int i = 0; i < 3; i++) ix[i] = index(xyz[i]);
  for (int i = 0; i < 3 && missing.empty(); i++) idc[i] = index("f_dc_" + std::to_string(i));
  const int iop = missing.empty() ? index("opacity") : -1;
  for (int i = 0; i < 3 && missing.empty(); i++) isc[i] = index("scale_" + std::to_string(i));
  for (int i = 0; i < 4 && missing.empty(); i++) irot[i] = index("rot_" + std::to_string(i));
  if (!missing.empty()) return fmt("missing property " + missing);
  size_t n_rest = 0;
  for (auto& s : props)
    if (s.rfind("f_rest_", 0) == 0) n_rest++;
  int degree;
  switch (n_rest) {
    case 0: degree = 0; break;
    case 9: degree = 1; break;
    case 24: degree = 2; break;
    case 45: degree = 3; break;
    default: return fmt("unsupported number of f_rest properties: " + std::to_string(n_rest));
  }
  const int k = (degree + 1) * (degree + 1);
  std::vector<int> irest(n_rest);
  for (size_t i = 0; i < n_rest; i++) {
    irest[i] = index("f_rest_" + std::to_string(i));
    if (!missing.empty()) return fmt("missing property " + missing);
  }
  *n_out = count;
  *sh_degree = degree;
  if (capacity < count) return SPLAT_ERR_CAPACITY;
  if (count && !gaussians) return set_err(err, err_len, SPLAT_ERR_INVALID, "null output");
  std::vector<float> row(props.size());
  for (uint64_t v = 0; v < count; v++) {
    std::memcpy(row.data(), bytes.data() + body_off + v * stride, stride);
    float* g = gaussians + v * kParams;
    std::memset(g, 0, kParams * sizeof(float));
    for (int i = 0; i < 3; i++) {
      g[i] = row[ix[i]];
      g[3 + i] = row[isc[i]];
      g[11 + i] = row[idc[i]];
    }
    for (int i = 0; i < 4; i++) g[6 + i] = row[irot[i]];
    g[10] = row[iop];
    for (int c = 0; c < 3; c++)
      for (int kk = 1; kk < k; kk++) g[11 + 3 * kk + c] = row[irest[c * (k - 1) + (kk - 1)]];
  }
  return SPLAT_OK;
}

}  // extern "C"
