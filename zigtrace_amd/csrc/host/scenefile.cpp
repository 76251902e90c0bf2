// Scene-file reader (the format of scenefile.zig:1-18): one directive per
// line, '#' starts a comment, tokens separated by spaces, tabs or CRs,
// materials named before use. Validation order and the diagnostic strings
// are the reference's (scenefile.zig:35-203), so a file rejected there is
// rejected here with the same line and message.
#include <cctype>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <unordered_map>
#include <vector>

#include "host_scene.hpp"

namespace {

struct Reject {
  const char* what;
};

// Numbers follow std.fmt.parseFloat(f64): optional sign, then inf/infinity/
// nan or decimal/hex digits with single '_' separators between digits.
bool readFloat(const std::string& tok, double* out) {
  if (tok.empty()) return false;
  size_t i = 0;
  std::string clean;
  if (tok[0] == '+' || tok[0] == '-') clean.push_back(tok[i++]);
  std::string body = tok.substr(i), low;
  for (char c : body) low.push_back((char)std::tolower((unsigned char)c));
  if (low == "inf" || low == "infinity") {
    *out = clean == "-" ? -std::numeric_limits<double>::infinity() : std::numeric_limits<double>::infinity();
    return true;
  }
  if (low == "nan") {
    *out = std::numeric_limits<double>::quiet_NaN();
    return true;
  }
  const bool hex = low.rfind("0x", 0) == 0;
  auto isd = [hex](char c) { return hex ? std::isxdigit((unsigned char)c) != 0 : std::isdigit((unsigned char)c) != 0; };
  bool digit = false;
  for (size_t k = 0; k < body.size(); ++k) {
    const char c = body[k];
    if (c == '_') {
      if (k == 0 || k + 1 == body.size() || !isd(body[k - 1]) || !isd(body[k + 1])) return false;
      continue;
    }
    digit = digit || std::isdigit((unsigned char)c);
    clean.push_back(c);
  }
  if (!digit) return false;
  char* end = nullptr;
  const double v = std::strtod(clean.c_str(), &end);
  if (!end || *end) return false;
  *out = v;
  return true;
}

// std.fmt.parseInt(usize, tok, 10), restricted to a positive result.
bool readPositive(const std::string& tok, uint64_t* out) {
  size_t k = (!tok.empty() && tok[0] == '+') ? 1 : 0;
  if (k >= tok.size()) return false;
  unsigned __int128 v = 0;
  for (; k < tok.size(); ++k) {
    const char c = tok[k];
    if (c == '_') {
      if (k == 0 || k + 1 == tok.size() || !std::isdigit((unsigned char)tok[k - 1]) ||
          !std::isdigit((unsigned char)tok[k + 1]))
        return false;
      continue;
    }
    if (!std::isdigit((unsigned char)c)) return false;
    v = v * 10 + (unsigned)(c - '0');
    if (v > std::numeric_limits<uint64_t>::max()) return false;
  }
  if (v == 0) return false;
  *out = (uint64_t)v;
  return true;
}

class Line {
 public:
  explicit Line(const std::string& text) {
    size_t i = 0;
    while (i < text.size()) {
      while (i < text.size() && std::strchr(" \t\r", text[i]) && text[i]) ++i;
      const size_t j0 = i;
      while (i < text.size() && !(text[i] == ' ' || text[i] == '\t' || text[i] == '\r')) ++i;
      if (i > j0) toks_.push_back(text.substr(j0, i - j0));
    }
  }
  bool empty() const { return toks_.empty(); }
  bool more() const { return at_ < toks_.size(); }
  std::string word() {
    if (!more()) throw Reject{"unexpected end of line"};
    return toks_[at_++];
  }
  double number() {
    double v;
    if (!readFloat(word(), &v)) throw Reject{"expected a number"};
    if (!std::isfinite(v)) throw Reject{"number must be finite"};
    return v;
  }
  uint64_t positive() {
    uint64_t v;
    if (!readPositive(word(), &v)) throw Reject{"expected a positive integer"};
    return v;
  }
  void vec(double out[3]) {
    for (int i = 0; i < 3; ++i) out[i] = number();
  }
  void color(double out[3]) {
    vec(out);
    if (out[0] < 0 || out[1] < 0 || out[2] < 0) throw Reject{"colour components must be non-negative"};
  }
  uint32_t material(const std::unordered_map<std::string, uint32_t>& names) {
    const auto it = names.find(word());
    if (it == names.end()) throw Reject{"undefined material"};
    return it->second;
  }
  void finish() {
    if (more()) throw Reject{"unexpected extra tokens"};
  }

 private:
  std::vector<std::string> toks_;
  size_t at_ = 0;
};

bool tiny(const double v[3]) {  // Vec3.nearZero, vec.zig:91-94
  return std::fabs(v[0]) < 1e-8 && std::fabs(v[1]) < 1e-8 && std::fabs(v[2]) < 1e-8;
}
void crossv(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

void readCamera(Line& ln, zt_scene* s) {  // scenefile.zig:107-138
  zt_camera_config c = s->camera;
  bool from = false, to = false;
  while (ln.more()) {
    const std::string key = ln.word();
    if (key == "from") {
      ln.vec(c.look_from);
      from = true;
    } else if (key == "to") {
      ln.vec(c.look_at);
      to = true;
    } else if (key == "up") {
      ln.vec(c.up);
    } else if (key == "fov") {
      c.vfov = ln.number();
      if (!(c.vfov > 0 && c.vfov < 180)) throw Reject{"fov must be between 0 and 180 degrees"};
    } else if (key == "aperture") {
      c.aperture = ln.number();
      if (c.aperture < 0) throw Reject{"aperture must be non-negative"};
    } else if (key == "focus") {
      const double f = ln.number();
      if (!(f > 0)) throw Reject{"focus distance must be positive"};
      c.focus_dist = f;
      c.has_focus_dist = 1;
    } else {
      throw Reject{"unknown camera parameter"};
    }
  }
  if (!from || !to) throw Reject{"camera needs 'from' and 'to'"};
  const double view[3] = {c.look_from[0] - c.look_at[0], c.look_from[1] - c.look_at[1], c.look_from[2] - c.look_at[2]};
  if (tiny(view)) throw Reject{"camera 'from' and 'to' must differ"};
  double x[3];
  crossv(c.up, view, x);
  if (tiny(x)) throw Reject{"camera 'up' is parallel to the view direction"};
  s->camera = c;
}

zt_material readMaterial(Line& ln) {  // scenefile.zig:140-156
  zt_material m{};
  const std::string kind = ln.word();
  if (kind == "lambertian") {
    m.kind = ZT_LAMBERTIAN;
    ln.color(m.color);
  } else if (kind == "metal") {
    m.kind = ZT_METAL;
    ln.color(m.color);
    m.param = ln.more() ? ln.number() : 0;
    if (m.param < 0 || m.param > 1) throw Reject{"metal fuzz must be in [0, 1]"};
  } else if (kind == "dielectric") {
    m.kind = ZT_DIELECTRIC;
    m.param = ln.number();
    if (!(m.param > 0)) throw Reject{"index of refraction must be positive"};
  } else if (kind == "emissive") {
    m.kind = ZT_EMISSIVE;
    ln.color(m.color);
  } else {
    throw Reject{"unknown material type"};
  }
  return m;
}

void readDirective(const std::string& kw, Line& ln, zt_scene* s, std::unordered_map<std::string, uint32_t>& names) {
  if (kw == "image") {
    s->width = ln.positive();
    s->height = ln.positive();
  } else if (kw == "samples") {
    s->spp = ln.positive();
  } else if (kw == "depth") {
    s->max_depth = ln.positive();
  } else if (kw == "camera") {
    readCamera(ln, s);
  } else if (kw == "background") {
    const std::string kind = ln.word();
    zt_background bg = s->background;
    if (kind == "color") {
      bg.kind = ZT_BG_SOLID;
      ln.vec(bg.a);
    } else if (kind == "sky") {
      bg.kind = ZT_BG_SKY;
      ln.vec(bg.a);
      ln.vec(bg.b);
    } else {
      throw Reject{"background must be 'color' or 'sky'"};
    }
    s->background = bg;
  } else if (kw == "material") {
    const std::string name = ln.word();
    if (names.count(name)) throw Reject{"material already defined"};
    const zt_material m = readMaterial(ln);
    uint32_t id = 0;
    zt_scene_add_material(s, &m, &id);
    names.emplace(name, id);
  } else if (kw == "sphere") {
    zt_shape sh{};
    sh.kind = ZT_SPHERE;
    ln.vec(sh.v);
    sh.v[3] = ln.number();
    if (sh.v[3] == 0) throw Reject{"sphere radius must be non-zero"};
    sh.material = ln.material(names);
    zt_scene_add(s, &sh);
  } else if (kw == "plane") {
    zt_shape sh{};
    sh.kind = ZT_PLANE;
    ln.vec(sh.v);
    double n[3];
    ln.vec(n);
    if (tiny(n)) throw Reject{"plane normal must be non-zero"};
    const double inv = 1.0 / std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);  // Vec3.normalize
    for (int i = 0; i < 3; ++i) sh.v[3 + i] = n[i] * inv;
    sh.material = ln.material(names);
    zt_scene_add(s, &sh);
  } else if (kw == "triangle") {
    zt_shape sh{};
    sh.kind = ZT_TRIANGLE;
    ln.vec(sh.v);
    ln.vec(sh.v + 3);
    ln.vec(sh.v + 6);
    const double e1[3] = {sh.v[3] - sh.v[0], sh.v[4] - sh.v[1], sh.v[5] - sh.v[2]};
    const double e2[3] = {sh.v[6] - sh.v[0], sh.v[7] - sh.v[1], sh.v[8] - sh.v[2]};
    double x[3];
    crossv(e1, e2, x);
    if (tiny(x)) throw Reject{"degenerate triangle"};
    sh.material = ln.material(names);
    zt_scene_add(s, &sh);
  } else if (kw == "quad") {
    double c[3], u[3], v[3], x[3];
    ln.vec(c);
    ln.vec(u);
    ln.vec(v);
    crossv(u, v, x);
    if (tiny(x)) throw Reject{"degenerate quad"};
    const uint32_t m = ln.material(names);
    zt_scene_add_quad(s, c, u, v, m);
  } else {
    throw Reject{"unknown directive"};
  }
}

}  // namespace

extern "C" int zt_scene_parse(const char* src, uint64_t len, zt_scene** out, uint64_t* diag_line, char* diag_msg,
                              uint64_t diag_cap) {
  if (!out || (!src && len)) return ZT_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  zt_scene* s = zt_scene_create();
  if (!s) return ZT_ERR_OUT_OF_MEMORY;
  std::unordered_map<std::string, uint32_t> names;
  uint64_t line_no = 0;
  size_t pos = 0;
  auto report = [&](uint64_t line, const char* msg) {
    if (diag_line) *diag_line = line;
    if (diag_msg && diag_cap) {
      std::strncpy(diag_msg, msg, diag_cap - 1);
      diag_msg[diag_cap - 1] = '\0';
    }
  };
  try {
    for (;;) {
      size_t end = pos;
      while (end < len && src[end] != '\n') ++end;
      ++line_no;
      std::string text(src + pos, end - pos);
      const size_t hash = text.find('#');
      if (hash != std::string::npos) text.resize(hash);
      Line ln(text);
      if (!ln.empty()) {
        const std::string kw = ln.word();
        readDirective(kw, ln, s, names);
        ln.finish();
      }
      if (end >= len) break;
      pos = end + 1;
    }
  } catch (const Reject& r) {
    report(line_no, r.what);
    zt_scene_destroy(s);
    return ZT_ERR_SYNTAX;
  }
  report(0, "");
  const int rc = zt_scene_build(s);
  if (rc != ZT_OK) {
    zt_scene_destroy(s);
    return rc;
  }
  *out = s;
  return ZT_OK;
}
