// COLMAP sparse reconstructions (src/colmap/*.rs) and their conversion to training
// cameras (src/dataset.rs:83-129, 213-237), host code.
//
// Binary (.bin, little endian, COLMAP's reconstruction_io layout) and text (.txt) readers
// for cameras, images and points3D; 2D observations and point tracks are parsed past and
// not kept.  Writers emit images without observations and points without tracks.
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/splat.h"

namespace {

int set_err(char* err, size_t len, int code, const std::string& msg) {
  if (err && len) std::snprintf(err, len, "%s", msg.c_str());
  return code;
}

struct Fail {
  int code;
  std::string msg;
};

// Camera models: COLMAP id order, names and parameter counts (src/colmap/mod.rs:17-135).
constexpr int kModels = 12;
const char* const kModelName[kModels] = {
    "SIMPLE_PINHOLE", "PINHOLE", "SIMPLE_RADIAL", "RADIAL", "OPENCV", "OPENCV_FISHEYE",
    "FULL_OPENCV", "FOV", "SIMPLE_RADIAL_FISHEYE", "RADIAL_FISHEYE", "THIN_PRISM_FISHEYE",
    "RAD_TAN_THIN_PRISM_FISHEYE"};
constexpr uint32_t kModelParams[kModels] = {3, 4, 4, 5, 8, 8, 12, 5, 4, 5, 12, 16};

bool read_file(const std::string& path, std::string* out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::string s;
  char buf[1 << 16];
  size_t k;
  while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, k);
  const bool ok = !std::ferror(f);
  std::fclose(f);
  *out = std::move(s);
  return ok;
}

bool is_file(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fclose(f);
  return true;
}

// Little-endian cursor over a file's bytes.
struct Reader {
  const std::string& buf;
  const std::string& path;
  size_t pos = 0;
  const char* take(size_t n) {
    if (n > buf.size() - pos) throw Fail{SPLAT_ERR_FORMAT, path + ": unexpected end of file"};
    const char* p = buf.data() + pos;
    pos += n;
    return p;
  }
  template <class T>
  T get() {
    T v;
    std::memcpy(&v, take(sizeof(T)), sizeof(T));
    return v;
  }
  std::string cstring() {
    const size_t end = buf.find('\0', pos);
    if (end == std::string::npos) throw Fail{SPLAT_ERR_FORMAT, path + ": unterminated string"};
    std::string s = buf.substr(pos, end - pos);
    pos = end + 1;
    return s;
  }
  // A record count; each record takes at least one byte, so a count beyond the remaining
  // bytes is corrupt.
  uint64_t count(const char* what) {
    const uint64_t n = get<uint64_t>();
    if (n > buf.size() - pos)
      throw Fail{SPLAT_ERR_FORMAT, path + ": implausible " + what + " count " + std::to_string(n)};
    return n;
  }
};

void copy_name(splat_colmap_image* im, const std::string& name, const std::string& path) {
  if (name.size() >= sizeof(im->name))
    throw Fail{SPLAT_ERR_FORMAT, path + ": image name longer than 255 bytes: " + name};
  std::memset(im->name, 0, sizeof(im->name));
  std::memcpy(im->name, name.data(), name.size());
}

std::vector<splat_colmap_camera> read_cameras_bin(const std::string& path) {
  std::string bytes;
  if (!read_file(path, &bytes)) throw Fail{SPLAT_ERR_IO, "cannot read " + path};
  Reader r{bytes, path};
  const uint64_t n = r.count("camera");
  std::vector<splat_colmap_camera> out(n);
  for (auto& c : out) {
    std::memset(&c, 0, sizeof(c));
    c.id = r.get<uint32_t>();
    c.model = r.get<int32_t>();
    if (c.model < 0 || c.model >= kModels)
      throw Fail{SPLAT_ERR_FORMAT, path + ": unknown camera model id " + std::to_string(c.model)};
    c.width = r.get<uint64_t>();
    c.height = r.get<uint64_t>();
    c.num_params = kModelParams[c.model];
    for (uint32_t k = 0; k < c.num_params; k++) c.params[k] = r.get<double>();
  }
  return out;
}

std::vector<splat_colmap_image> read_images_bin(const std::string& path) {
  std::string bytes;
  if (!read_file(path, &bytes)) throw Fail{SPLAT_ERR_IO, "cannot read " + path};
  Reader r{bytes, path};
  const uint64_t n = r.count("image");
  std::vector<splat_colmap_image> out(n);
  for (auto& im : out) {
    im.id = r.get<uint32_t>();
    for (double& q : im.qvec) q = r.get<double>();
    for (double& t : im.tvec) t = r.get<double>();
    im.camera_id = r.get<uint32_t>();
    copy_name(&im, r.cstring(), path);
    const uint64_t n2d = r.count("2D point");
    if (n2d > (bytes.size() - r.pos) / 24)
      throw Fail{SPLAT_ERR_FORMAT, path + ": unexpected end of file"};
    r.take(n2d * 24);  // x, y (f64) and a point3D id (u64) per observation
  }
  return out;
}

std::vector<splat_colmap_point> read_points_bin(const std::string& path) {
  std::string bytes;
  if (!read_file(path, &bytes)) throw Fail{SPLAT_ERR_IO, "cannot read " + path};
  Reader r{bytes, path};
  const uint64_t n = r.count("3D point");
  std::vector<splat_colmap_point> out(n);
  for (auto& p : out) {
    std::memset(&p, 0, sizeof(p));
    p.id = r.get<uint64_t>();
    for (double& x : p.xyz) x = r.get<double>();
    for (uint8_t& c : p.rgb) c = r.get<uint8_t>();
    p.error = r.get<double>();
    const uint64_t len = r.count("track element");
    if (len > (bytes.size() - r.pos) / 8)
      throw Fail{SPLAT_ERR_FORMAT, path + ": unexpected end of file"};
    r.take(len * 8);  // image id and 2D point index (u32 each) per track element
  }
  return out;
}

// Lines with comments and blank lines removed, split on whitespace.
std::vector<std::vector<std::string>> text_lines(const std::string& path) {
  std::string text;
  if (!read_file(path, &text)) throw Fail{SPLAT_ERR_IO, "cannot read " + path};
  std::vector<std::vector<std::string>> out;
  std::istringstream is(text);
  std::string line;
  while (std::getline(is, line)) {
    std::istringstream ls(line);
    std::vector<std::string> toks;
    std::string t;
    while (ls >> t) toks.push_back(t);
    if (toks.empty() || toks[0][0] == '#') continue;
    out.push_back(std::move(toks));
  }
  return out;
}

// Whole-token numeric parses as Rust's str::parse: an optional '+' before an unsigned
// integer, no trailing characters.
bool parse_u64(const std::string& t, uint64_t* v) {
  const char* b = t.data();
  const char* e = t.data() + t.size();
  if (b != e && *b == '+') b++;
  if (b == e || *b == '-' || *b == '+') return false;
  const auto r = std::from_chars(b, e, *v);
  return r.ec == std::errc() && r.ptr == e;
}
bool parse_u32(const std::string& t, uint32_t* v) {
  uint64_t x;
  if (!parse_u64(t, &x) || x > 0xffffffffull) return false;
  *v = (uint32_t)x;
  return true;
}
bool parse_u8(const std::string& t, uint8_t* v) {
  uint64_t x;
  if (!parse_u64(t, &x) || x > 255) return false;
  *v = (uint8_t)x;
  return true;
}
bool parse_f64(const std::string& t, double* v) {
  if (t.empty()) return false;
  const char* b = t.data();
  if (*b == '+') b++;  // Rust's f64 parse accepts a leading '+'
  if (b == t.data() + t.size()) return false;
  std::string low(b, t.data() + t.size());
  for (char& ch : low) ch = (char)std::tolower((unsigned char)ch);
  const bool neg = low[0] == '-';
  const std::string body = neg ? low.substr(1) : low;
  if (body == "inf" || body == "infinity") {
    *v = neg ? -INFINITY : INFINITY;
    return true;
  }
  if (body == "nan") {
    *v = NAN;
    return true;
  }
  const auto r = std::from_chars(b, t.data() + t.size(), *v, std::chars_format::general);
  return r.ec == std::errc() && r.ptr == t.data() + t.size();
}

[[noreturn]] void bad(const std::string& path, const std::string& what, const std::string* tok) {
  throw Fail{SPLAT_ERR_FORMAT,
             path + (tok ? ": bad " + what + ": \"" + *tok + "\"" : ": missing " + what)};
}

std::vector<splat_colmap_camera> read_cameras_txt(const std::string& path) {
  std::vector<splat_colmap_camera> out;
  for (const auto& t : text_lines(path)) {
    splat_colmap_camera c;
    std::memset(&c, 0, sizeof(c));
    if (!parse_u32(t[0], &c.id)) bad(path, "camera id", &t[0]);
    if (t.size() < 2) bad(path, "camera model", nullptr);
    c.model = -1;
    for (int m = 0; m < kModels; m++)
      if (t[1] == kModelName[m]) c.model = m;
    if (c.model < 0) throw Fail{SPLAT_ERR_FORMAT, path + ": unknown camera model " + t[1]};
    if (t.size() < 3) bad(path, "width", nullptr);
    if (!parse_u64(t[2], &c.width)) bad(path, "width", &t[2]);
    if (t.size() < 4) bad(path, "height", nullptr);
    if (!parse_u64(t[3], &c.height)) bad(path, "height", &t[3]);
    const size_t np = t.size() - 4;
    c.num_params = kModelParams[c.model];
    if (np != c.num_params)
      throw Fail{SPLAT_ERR_FORMAT, path + ": camera " + std::to_string(c.id) + ": model " +
                                       kModelName[c.model] + " expects " +
                                       std::to_string(c.num_params) + " parameters, found " +
                                       std::to_string(np)};
    for (size_t k = 0; k < np; k++)
      if (!parse_f64(t[4 + k], &c.params[k])) bad(path, "camera parameter", &t[4 + k]);
    out.push_back(c);
  }
  return out;
}

// Two lines per image (header and 2D observations, the latter sometimes omitted): a line
// is a header when it has at least ten tokens, the first nine numeric and the tenth not
// (src/colmap/text.rs:62-101).
std::vector<splat_colmap_image> read_images_txt(const std::string& path) {
  std::vector<splat_colmap_image> out;
  for (const auto& t : text_lines(path)) {
    uint32_t u;
    double d;
    bool header = t.size() >= 10 && parse_u32(t[0], &u);
    for (int k = 1; header && k < 9; k++) header = parse_f64(t[k], &d);
    header = header && !parse_f64(t[9], &d);
    if (!header) continue;
    splat_colmap_image im;
    im.id = u;
    for (int k = 0; k < 4; k++) parse_f64(t[1 + k], &im.qvec[k]);
    for (int k = 0; k < 3; k++) parse_f64(t[5 + k], &im.tvec[k]);
    if (!parse_u32(t[8], &im.camera_id)) bad(path, "camera id", &t[8]);
    copy_name(&im, t[9], path);
    out.push_back(im);
  }
  return out;
}

std::vector<splat_colmap_point> read_points_txt(const std::string& path) {
  std::vector<splat_colmap_point> out;
  static const char* const kField[] = {"x", "y", "z", "r", "g", "b", "error"};
  for (const auto& t : text_lines(path)) {
    splat_colmap_point p;
    std::memset(&p, 0, sizeof(p));
    if (!parse_u64(t[0], &p.id)) bad(path, "point id", &t[0]);
    for (int k = 0; k < 7; k++) {
      if (t.size() < (size_t)(2 + k)) bad(path, kField[k], nullptr);
      const std::string& tok = t[1 + k];
      const bool ok = k < 3   ? parse_f64(tok, &p.xyz[k])
                      : k < 6 ? parse_u8(tok, &p.rgb[k - 3])
                              : parse_f64(tok, &p.error);
      if (!ok) bad(path, kField[k], &tok);
    }
    out.push_back(p);
  }
  return out;
}

// Rust's Display for f64: the shortest digits that read back exactly, placed without an
// exponent (zeros pad the integer part or follow the point).
std::string fmt_f64(double v) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
  if (v == 0.0) return std::signbit(v) ? "-0" : "0";
  char buf[64];
  const auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::scientific);
  std::string s(buf, r.ptr);
  const bool neg = s[0] == '-';
  if (neg) s.erase(0, 1);
  const size_t e = s.find('e');
  const int exp = std::stoi(s.substr(e + 1));
  std::string digits;
  for (size_t k = 0; k < e; k++)
    if (s[k] != '.') digits += s[k];
  const int point = exp + 1;  // digits before the decimal point
  std::string out;
  if (point <= 0)
    out = "0." + std::string((size_t)-point, '0') + digits;
  else if (point >= (int)digits.size())
    out = digits + std::string((size_t)point - digits.size(), '0');
  else
    out = digits.substr(0, (size_t)point) + "." + digits.substr((size_t)point);
  return neg ? "-" + out : out;
}

bool write_file(const std::string& path, const std::string& data) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  const bool ok = std::fwrite(data.data(), 1, data.size(), f) == data.size();
  return std::fclose(f) == 0 && ok;
}

template <class T>
void put(std::string& s, T v) {
  s.append(reinterpret_cast<const char*>(&v), sizeof(T));
}

std::vector<const splat_colmap_camera*> sorted_cameras(const splat_colmap_camera* c, uint64_t n) {
  std::vector<const splat_colmap_camera*> out(n);
  for (uint64_t i = 0; i < n; i++) {
    if (c[i].model < 0 || c[i].model >= kModels)
      throw Fail{SPLAT_ERR_INVALID, "unknown camera model id " + std::to_string(c[i].model)};
    if (c[i].num_params != kModelParams[c[i].model])
      throw Fail{SPLAT_ERR_INVALID, "camera " + std::to_string(c[i].id) + ": model " +
                                        kModelName[c[i].model] + " expects " +
                                        std::to_string(kModelParams[c[i].model]) +
                                        " parameters, found " + std::to_string(c[i].num_params)};
    out[i] = &c[i];
  }
  std::sort(out.begin(), out.end(), [](auto* a, auto* b) { return a->id < b->id; });
  return out;
}

// Quat::normalized + Quat::to_mat (src/math.rs:289-330) in binary32.
void quat_to_rot(float w, float x, float y, float z, float* rot) {
  const float n = std::sqrt(w * w + x * x + y * y + z * z);
  if (n > 0.0f) {
    w = w / n;
    x = x / n;
    y = y / n;
    z = z / n;
  } else {
    w = 1.0f;
    x = y = z = 0.0f;
  }
  rot[0] = 1.0f - 2.0f * (y * y + z * z);
  rot[1] = 2.0f * (x * y - w * z);
  rot[2] = 2.0f * (x * z + w * y);
  rot[3] = 2.0f * (x * y + w * z);
  rot[4] = 1.0f - 2.0f * (x * x + z * z);
  rot[5] = 2.0f * (y * z - w * x);
  rot[6] = 2.0f * (x * z - w * y);
  rot[7] = 2.0f * (y * z + w * x);
  rot[8] = 1.0f - 2.0f * (x * x + y * y);
}

}  // namespace

extern "C" {

int splat_colmap_read(const char* dir, splat_colmap_camera* cameras, uint64_t camera_capacity,
                      uint64_t* num_cameras, splat_colmap_image* images, uint64_t image_capacity,
                      uint64_t* num_images, splat_colmap_point* points, uint64_t point_capacity,
                      uint64_t* num_points, char* err, size_t err_len) {
  if (!dir || !num_cameras || !num_images || !num_points)
    return set_err(err, err_len, SPLAT_ERR_INVALID, "null argument");
  try {
    const std::string d(dir);
    const auto pick = [&](const char* stem) {
      const std::string bin = d + "/" + stem + ".bin", txt = d + "/" + stem + ".txt";
      if (is_file(bin)) return std::make_pair(bin, true);
      if (is_file(txt)) return std::make_pair(txt, false);
      throw Fail{SPLAT_ERR_INVALID,
                 std::string("neither ") + stem + ".bin nor " + stem + ".txt found in " + d};
    };
    const auto pc = pick("cameras"), pi = pick("images"), pp = pick("points3D");
    auto cams = pc.second ? read_cameras_bin(pc.first) : read_cameras_txt(pc.first);
    // The reference collects the cameras into a map by id (src/colmap/mod.rs:234): a
    // repeated id keeps one record, the last one read.  Here it keeps the position of the
    // id's first occurrence.
    {
      std::vector<splat_colmap_camera> unique;
      for (const auto& c : cams) {
        bool seen = false;
        for (auto& u : unique)
          if (u.id == c.id) {
            u = c;
            seen = true;
            break;
          }
        if (!seen) unique.push_back(c);
      }
      cams.swap(unique);
    }
    const auto imgs = pi.second ? read_images_bin(pi.first) : read_images_txt(pi.first);
    const auto pts = pp.second ? read_points_bin(pp.first) : read_points_txt(pp.first);
    *num_cameras = cams.size();
    *num_images = imgs.size();
    *num_points = pts.size();
    if (camera_capacity < cams.size() || image_capacity < imgs.size() ||
        point_capacity < pts.size())
      return set_err(err, err_len, SPLAT_ERR_CAPACITY, "output capacity too small");
    if (!cams.empty()) std::memcpy(cameras, cams.data(), cams.size() * sizeof(cams[0]));
    if (!imgs.empty()) std::memcpy(images, imgs.data(), imgs.size() * sizeof(imgs[0]));
    if (!pts.empty()) std::memcpy(points, pts.data(), pts.size() * sizeof(pts[0]));
    return SPLAT_OK;
  } catch (const Fail& f) {
    return set_err(err, err_len, f.code, f.msg);
  } catch (const std::exception& e) {
    return set_err(err, err_len, SPLAT_ERR_INVALID, e.what());
  }
}

int splat_colmap_write(const char* dir, int32_t binary, const splat_colmap_camera* cameras,
                       uint64_t num_cameras, const splat_colmap_image* images,
                       uint64_t num_images, const splat_colmap_point* points,
                       uint64_t num_points, char* err, size_t err_len) {
  if (!dir || (num_cameras && !cameras) || (num_images && !images) || (num_points && !points))
    return set_err(err, err_len, SPLAT_ERR_INVALID, "null argument");
  try {
    const std::string d(dir);
    const auto cams = sorted_cameras(cameras, num_cameras);
    std::string c, i, p;
    if (binary) {  // src/colmap/binary.rs:165-224
      put<uint64_t>(c, num_cameras);
      for (const auto* cam : cams) {
        put(c, cam->id);
        put(c, cam->model);
        put(c, cam->width);
        put(c, cam->height);
        for (uint32_t k = 0; k < cam->num_params; k++) put(c, cam->params[k]);
      }
      put<uint64_t>(i, num_images);
      for (uint64_t k = 0; k < num_images; k++) {
        const auto& im = images[k];
        put(i, im.id);
        for (double q : im.qvec) put(i, q);
        for (double t : im.tvec) put(i, t);
        put(i, im.camera_id);
        i.append(im.name, strnlen(im.name, sizeof(im.name)));
        i.push_back('\0');
        put<uint64_t>(i, 0);
      }
      put<uint64_t>(p, num_points);
      for (uint64_t k = 0; k < num_points; k++) {
        const auto& pt = points[k];
        put(p, pt.id);
        for (double x : pt.xyz) put(p, x);
        for (uint8_t ch : pt.rgb) put(p, ch);
        put(p, pt.error);
        put<uint64_t>(p, 0);
      }
    } else {  // src/colmap/text.rs:132-190
      c = "# Camera list with one line of data per camera:\n"
          "#   CAMERA_ID, MODEL, WIDTH, HEIGHT, PARAMS[]\n";
      for (const auto* cam : cams) {
        c += std::to_string(cam->id) + " " + kModelName[cam->model] + " " +
             std::to_string(cam->width) + " " + std::to_string(cam->height) + " ";
        for (uint32_t k = 0; k < cam->num_params; k++)
          c += (k ? " " : "") + fmt_f64(cam->params[k]);
        c += "\n";
      }
      i = "# Image list with two lines of data per image:\n"
          "#   IMAGE_ID, QW, QX, QY, QZ, TX, TY, TZ, CAMERA_ID, NAME\n"
          "#   POINTS2D[] as (X, Y, POINT3D_ID)\n";
      for (uint64_t k = 0; k < num_images; k++) {
        const auto& im = images[k];
        i += std::to_string(im.id);
        for (double q : im.qvec) i += " " + fmt_f64(q);
        for (double t : im.tvec) i += " " + fmt_f64(t);
        i += " " + std::to_string(im.camera_id) + " " +
             std::string(im.name, strnlen(im.name, sizeof(im.name))) + "\n\n";
      }
      p = "# 3D point list with one line of data per point:\n"
          "#   POINT3D_ID, X, Y, Z, R, G, B, ERROR, TRACK[] as (IMAGE_ID, POINT2D_IDX)\n";
      for (uint64_t k = 0; k < num_points; k++) {
        const auto& pt = points[k];
        p += std::to_string(pt.id);
        for (double x : pt.xyz) p += " " + fmt_f64(x);
        for (uint8_t ch : pt.rgb) p += " " + std::to_string(ch);
        p += " " + fmt_f64(pt.error) + "\n";
      }
    }
    const char* ext = binary ? ".bin" : ".txt";
    for (const auto& f : {std::make_pair("cameras", &c), std::make_pair("images", &i),
                          std::make_pair("points3D", &p)}) {
      const std::string path = d + "/" + f.first + ext;
      if (!write_file(path, *f.second)) throw Fail{SPLAT_ERR_IO, "cannot write " + path};
    }
    return SPLAT_OK;
  } catch (const Fail& f) {
    return set_err(err, err_len, f.code, f.msg);
  } catch (const std::exception& e) {
    return set_err(err, err_len, SPLAT_ERR_INVALID, e.what());
  }
}

int splat_camera_from_colmap(const splat_colmap_camera* cameras, uint64_t num_cameras,
                             const splat_colmap_image* image, splat_camera* out, char* err,
                             size_t err_len) {
  if ((num_cameras && !cameras) || !image || !out)
    return set_err(err, err_len, SPLAT_ERR_INVALID, "null argument");
  const splat_colmap_camera* cam = nullptr;
  for (uint64_t k = 0; k < num_cameras; k++)
    if (cameras[k].id == image->camera_id) cam = &cameras[k];
  const std::string name(image->name, strnlen(image->name, sizeof(image->name)));
  if (!cam)
    return set_err(err, err_len, SPLAT_ERR_INVALID,
                   "image " + name + " refers to unknown camera " +
                       std::to_string(image->camera_id));
  double fx, fy, cx, cy;
  if (cam->model == 0) {  // SIMPLE_PINHOLE: f, cx, cy
    fx = fy = cam->params[0];
    cx = cam->params[1];
    cy = cam->params[2];
  } else if (cam->model == 1) {  // PINHOLE: fx, fy, cx, cy
    fx = cam->params[0];
    fy = cam->params[1];
    cx = cam->params[2];
    cy = cam->params[3];
  } else {
    const char* model = cam->model >= 0 && cam->model < kModels ? kModelName[cam->model] : "?";
    return set_err(err, err_len, SPLAT_ERR_INVALID,
                   "camera " + std::to_string(cam->id) + " uses the " + model +
                       " model; only undistorted PINHOLE and SIMPLE_PINHOLE cameras are supported");
  }
  splat_camera c;
  c.width = (uint32_t)cam->width;
  c.height = (uint32_t)cam->height;
  c.fx = (float)fx;
  c.fy = (float)fy;
  c.cx = (float)cx;
  c.cy = (float)cy;
  quat_to_rot((float)image->qvec[0], (float)image->qvec[1], (float)image->qvec[2],
              (float)image->qvec[3], c.rot);
  for (int k = 0; k < 3; k++) c.trans[k] = (float)image->tvec[k];
  *out = c;
  return SPLAT_OK;
}

int splat_camera_match_image_size(const splat_camera* camera, uint32_t width, uint32_t height,
                                  splat_camera* out) {
  if (!camera || !out) return SPLAT_ERR_INVALID;
  if (width == camera->width && height == camera->height) {
    *out = *camera;
    return SPLAT_OK;
  }
  const double cw = camera->width, ch = camera->height, w = width, h = height;
  const double width_from_height = cw * h / ch;
  const double height_from_width = ch * w / cw;
  if (width == 0 || height == 0 || std::fabs(width_from_height - w) > 1.0 ||
      std::fabs(height_from_width - h) > 1.0 || !(camera->width && camera->height))
    return SPLAT_ERR_INVALID;
  // Camera::resized, src/camera.rs:48-62
  const float sx = (float)width / (float)camera->width;
  const float sy = (float)height / (float)camera->height;
  splat_camera c = *camera;
  c.width = width;
  c.height = height;
  c.fx = camera->fx * sx;
  c.fy = camera->fy * sy;
  c.cx = camera->cx * sx;
  c.cy = camera->cy * sy;
  *out = c;
  return SPLAT_OK;
}

}  // extern "C"
