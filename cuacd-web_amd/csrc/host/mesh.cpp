// mesh.cpp — host mesh helpers (see mesh.hpp). Restates src/mesh.ts.
#include "mesh.hpp"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <regex>
#include <tuple>

namespace cuacd {

Mesh makeBox(const float lo[3], const float hi[3]) {
  const float x0 = lo[0], y0 = lo[1], z0 = lo[2], x1 = hi[0], y1 = hi[1], z1 = hi[2];
  Mesh m;
  m.verts = {x0, y0, z0, x1, y0, z0, x1, y1, z0, x0, y1, z0, x0, y0, z1, x1, y0, z1, x1, y1, z1, x0, y1, z1};
  // bottom, top, front, back, right, left — two triangles each, outward
  m.tris = {0, 2, 1, 0, 3, 2, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4,
            2, 3, 7, 2, 7, 6, 1, 2, 6, 1, 6, 5, 0, 4, 7, 0, 7, 3};
  return m;
}

Mesh unitCube() {
  const float lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  return makeBox(lo, hi);
}

Mesh lShape() {
  const float a0[3] = {0, 0, 0}, a1[3] = {1, 1, 1};
  const float b0[3] = {1, 0, 0}, b1[3] = {2, 1, 0.5f};
  return mergeMeshes(makeBox(a0, a1), makeBox(b0, b1));
}

Mesh mergeMeshes(const Mesh& a, const Mesh& b) {
  Mesh m;
  m.verts = a.verts;
  m.verts.insert(m.verts.end(), b.verts.begin(), b.verts.end());
  m.tris = a.tris;
  const uint32_t off = a.nv();
  m.tris.reserve(a.tris.size() + b.tris.size());
  for (uint32_t t : b.tris) m.tris.push_back(t + off);
  return m;
}

double signedVolume(const Mesh& m) {
  const auto& v = m.verts;
  const auto& t = m.tris;
  if (t.empty()) return 0.0;
  const double ox = v[0], oy = v[1], oz = v[2];
  double s = 0.0;
  for (size_t i = 0; i + 2 < t.size(); i += 3) {
    const size_t a = 3 * (size_t)t[i], b = 3 * (size_t)t[i + 1], c = 3 * (size_t)t[i + 2];
    const double ax = v[a] - ox, ay = v[a + 1] - oy, az = v[a + 2] - oz;
    const double bx = v[b] - ox, by = v[b + 1] - oy, bz = v[b + 2] - oz;
    const double cx = v[c] - ox, cy = v[c + 1] - oy, cz = v[c + 2] - oz;
    s += ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
  }
  return s / 6.0;
}

double volume(const Mesh& m) { return std::fabs(signedVolume(m)); }

EdgeCSR edgeCSR(const Mesh& m) {
  EdgeCSR r;
  const uint32_t n = m.nv();
  const auto& T = m.tris;
  r.off.assign(n + 1, 0);
  for (uint32_t x : T) r.off[x + 1]++;
  for (uint32_t v = 0; v < n; v++) r.off[v + 1] += r.off[v];
  std::vector<uint32_t> fill(r.off.begin(), r.off.begin() + n);
  r.dst.assign(T.size(), 0);
  r.face.assign(T.size(), 0);
  for (size_t i = 0; i + 2 < T.size(); i += 3) {
    for (int e = 0; e < 3; e++) {
      const uint32_t a = T[i + e], b = T[i + (e + 1) % 3];
      const uint32_t k = fill[a]++;
      r.dst[k] = b;
      r.face[k] = (uint32_t)(i / 3);
    }
  }
  return r;
}

bool isClosed(const Mesh& m) {
  const uint32_t n = m.nv();
  for (uint32_t x : m.tris)
    if (x >= n) return false;
  const EdgeCSR c = edgeCSR(m);
  auto count = [&](uint32_t a, uint32_t b) {
    uint32_t k = 0;
    for (uint32_t i = c.off[a]; i < c.off[a + 1]; i++) k += (c.dst[i] == b);
    return k;
  };
  for (uint32_t a = 0; a < n; a++) {
    for (uint32_t i = c.off[a]; i < c.off[a + 1]; i++) {
      const uint32_t b = c.dst[i];
      if (count(a, b) != 1 || count(b, a) != 1) return false;
    }
  }
  return true;
}

Box3 bboxPoints(const float* v, uint32_t nv) {
  Box3 b;
  for (int a = 0; a < 3; a++) {
    b.lo[a] = INFINITY;
    b.hi[a] = -INFINITY;
  }
  for (uint32_t i = 0; i < nv; i++) {
    for (int a = 0; a < 3; a++) {
      const float x = v[3 * i + a];
      if (x < b.lo[a]) b.lo[a] = x;
      if (x > b.hi[a]) b.hi[a] = x;
    }
  }
  return b;
}

Box3 bbox(const Mesh& m) { return bboxPoints(m.verts.data(), m.nv()); }

Mesh normalize(const Mesh& m, Normalization* norm) {
  const Box3 b = bbox(m);
  double center[3], extent = 0;
  for (int a = 0; a < 3; a++) {
    center[a] = ((double)b.lo[a] + (double)b.hi[a]) / 2;
    extent = std::max(extent, (double)b.hi[a] - (double)b.lo[a]);
  }
  const double scale = extent > 0 ? extent / 2 : 1.0;
  Mesh out;
  out.verts.resize(m.verts.size());
  for (size_t i = 0; i + 2 < m.verts.size(); i += 3)
    for (int a = 0; a < 3; a++) out.verts[i + a] = (float)(((double)m.verts[i + a] - center[a]) / scale);
  out.tris = m.tris;
  if (norm) {
    norm->scale = scale;
    for (int a = 0; a < 3; a++) norm->center[a] = center[a];
  }
  return out;
}

Mesh denormalize(const Mesh& m, const Normalization& n) {
  Mesh out;
  out.verts.resize(m.verts.size());
  for (size_t i = 0; i + 2 < m.verts.size(); i += 3)
    for (int a = 0; a < 3; a++) out.verts[i + a] = (float)((double)m.verts[i + a] * n.scale + n.center[a]);
  out.tris = m.tris;
  return out;
}

Mesh weld(const std::vector<float>& soup) {
  std::map<std::tuple<uint32_t, uint32_t, uint32_t>, uint32_t> idx;
  Mesh m;
  for (size_t i = 0; i + 2 < soup.size(); i += 3) {
    uint32_t k[3];
    std::memcpy(k, &soup[i], 12);
    // the reference keys vertices by their decimal text, where -0 and 0 coincide
    for (int a = 0; a < 3; a++)
      if (k[a] == 0x80000000u) k[a] = 0;
    auto key = std::make_tuple(k[0], k[1], k[2]);
    auto it = idx.find(key);
    uint32_t id;
    if (it == idx.end()) {
      id = m.nv();
      idx.emplace(key, id);
      m.verts.push_back(soup[i]);
      m.verts.push_back(soup[i + 1]);
      m.verts.push_back(soup[i + 2]);
    } else {
      id = it->second;
    }
    m.tris.push_back(id);
  }
  return m;
}

bool parseBinarySTL(const std::vector<uint8_t>& buf, Mesh* out) {
  if (buf.size() < 84) return false;
  uint32_t n;
  std::memcpy(&n, buf.data() + 80, 4);
  if ((size_t)84 + (size_t)n * 50 != buf.size()) return false;
  std::vector<float> soup;
  soup.reserve((size_t)n * 9);
  size_t off = 84;
  for (uint32_t i = 0; i < n; i++) {
    off += 12;  // normal
    for (int k = 0; k < 9; k++) {
      float f;
      std::memcpy(&f, buf.data() + off + 4 * k, 4);
      soup.push_back(f);
    }
    off += 38;
  }
  *out = weld(soup);
  return true;
}

namespace {

bool jsSpace(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

// Number(token) of a token with no white space (ECMAScript StringToNumber): "" is 0; decimal
// literals with optional sign, exponent, leading or trailing point; Infinity; 0x/0o/0b integers
double jsNumber(const std::string& t) {
  if (t.empty()) return 0.0;
  if (t == "Infinity" || t == "+Infinity") return INFINITY;
  if (t == "-Infinity") return -INFINITY;
  if (t.size() > 2 && t[0] == '0' && (t[1] == 'x' || t[1] == 'X' || t[1] == 'o' || t[1] == 'O' || t[1] == 'b' || t[1] == 'B')) {
    const int base = (t[1] == 'x' || t[1] == 'X') ? 16 : ((t[1] == 'o' || t[1] == 'O') ? 8 : 2);
    double v = 0;
    for (size_t i = 2; i < t.size(); i++) {
      const char c = t[i];
      int d = c >= '0' && c <= '9' ? c - '0' : (c >= 'a' && c <= 'f' ? c - 'a' + 10 : (c >= 'A' && c <= 'F' ? c - 'A' + 10 : 99));
      if (d >= base) return NAN;
      v = v * base + d;
    }
    return v;
  }
  static const std::regex dec("[+-]?(([0-9]+\\.?[0-9]*)|(\\.[0-9]+))([eE][+-]?[0-9]+)?");
  if (!std::regex_match(t, dec)) return NAN;
  return std::strtod(t.c_str(), nullptr);
}

// parseInt(s, 10): leading white space, an optional sign, the longest run of digits; NaN without digits
double jsParseInt(const std::string& s) {
  size_t i = 0;
  while (i < s.size() && jsSpace(s[i])) i++;
  bool neg = false;
  if (i < s.size() && (s[i] == '+' || s[i] == '-')) neg = s[i++] == '-';
  double v = 0;
  size_t d0 = i;
  while (i < s.size() && s[i] >= '0' && s[i] <= '9') v = v * 10 + (s[i++] - '0');
  if (i == d0) return NAN;
  return neg ? -v : v;
}

// ToUint32: what a Uint32Array stores for a number
uint32_t jsToUint32(double x) {
  if (!std::isfinite(x)) return 0;
  const double t = std::trunc(x);
  const double m = std::fmod(t, 4294967296.0);
  return (uint32_t)(int64_t)(m < 0 ? m + 4294967296.0 : m);
}

std::vector<std::string> splitWs(const std::string& s) {  // s.split(/\s+/)
  std::vector<std::string> out(1);
  size_t i = 0;
  while (i < s.size()) {
    if (jsSpace(s[i])) {
      while (i < s.size() && jsSpace(s[i])) i++;
      out.emplace_back();
    } else {
      out.back().push_back(s[i++]);
    }
  }
  return out;
}

std::string jsTrim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && jsSpace(s[a])) a++;
  while (b > a && jsSpace(s[b - 1])) b--;
  return s.substr(a, b - a);
}

Mesh weldDoubles(const std::vector<double>& pos) {  // src/mesh.ts:196-212 over number[] positions
  std::map<std::tuple<uint64_t, uint64_t, uint64_t>, uint32_t> idx;
  Mesh m;
  for (size_t i = 0; i + 2 < pos.size(); i += 3) {
    uint64_t k[3];
    for (int a = 0; a < 3; a++) {
      const double x = pos[i + a] == 0.0 ? 0.0 : pos[i + a];  // "-0" and "0" print alike
      std::memcpy(&k[a], &x, 8);
    }
    const auto key = std::make_tuple(k[0], k[1], k[2]);
    auto it = idx.find(key);
    uint32_t id;
    if (it == idx.end()) {
      id = m.nv();
      idx.emplace(key, id);
      for (int a = 0; a < 3; a++) m.verts.push_back((float)pos[i + a]);
    } else {
      id = it->second;
    }
    m.tris.push_back(id);
  }
  return m;
}

}  // namespace

bool parseSTL(const std::vector<uint8_t>& buf, Mesh* out) {
  const size_t hn = std::min<size_t>(80, buf.size());
  std::string head(buf.begin(), buf.begin() + hn);
  size_t h0 = 0;
  while (h0 < head.size() && jsSpace(head[h0])) h0++;
  bool looksBinary = false;
  if (buf.size() >= 84) {
    uint32_t n;
    std::memcpy(&n, buf.data() + 80, 4);
    looksBinary = (uint64_t)84 + (uint64_t)n * 50 == buf.size();
  }
  const bool isBinary = !(head.compare(h0, 5, "solid") == 0 && !looksBinary);
  if (isBinary) {
    if (buf.size() < 84) return false;
    uint32_t n;
    std::memcpy(&n, buf.data() + 80, 4);
    if (n && (uint64_t)84 + (uint64_t)(n - 1) * 50 + 48 > buf.size()) return false;  // a triangle past the end
    std::vector<float> soup;
    soup.reserve((size_t)n * 9);
    size_t off = 84;
    for (uint32_t i = 0; i < n; i++) {
      off += 12;
      for (int k = 0; k < 9; k++) {
        float f;
        std::memcpy(&f, buf.data() + off + 4 * k, 4);
        soup.push_back(f);
      }
      off += 38;
    }
    *out = weld(soup);
    return true;
  }
  const std::string text(buf.begin(), buf.end());
  static const std::regex re("vertex\\s+([-+0-9.eE]+)\\s+([-+0-9.eE]+)\\s+([-+0-9.eE]+)");
  std::vector<double> pos;
  for (auto it = std::sregex_iterator(text.begin(), text.end(), re); it != std::sregex_iterator(); ++it)
    for (int k = 1; k <= 3; k++) pos.push_back(jsNumber((*it)[k].str()));
  *out = weldDoubles(pos);
  return true;
}

Mesh parseOBJ(const std::string& text) {
  Mesh m;
  std::vector<uint32_t> tris;
  size_t at = 0;
  while (at <= text.size()) {
    size_t e = text.find('\n', at);
    if (e == std::string::npos) e = text.size();
    const std::string line = text.substr(at, e - at);
    at = e + 1;
    if (line.compare(0, 2, "v ") == 0) {
      const std::vector<std::string> p = splitWs(line);
      for (int k = 1; k <= 3; k++) m.verts.push_back((float)(k < (int)p.size() ? jsNumber(p[k]) : NAN));
    } else if (line.compare(0, 2, "f ") == 0) {
      const std::vector<std::string> tok = splitWs(jsTrim(line));
      std::vector<double> p;
      for (size_t i = 1; i < tok.size(); i++) p.push_back(jsParseInt(tok[i].substr(0, tok[i].find('/'))) - 1);
      for (size_t i = 1; i + 1 < p.size(); i++) {
        m.tris.push_back(jsToUint32(p[0]));
        m.tris.push_back(jsToUint32(p[i]));
        m.tris.push_back(jsToUint32(p[i + 1]));
      }
    }
  }
  return m;
}

std::string jsNumberString(double x) {
  if (std::isnan(x)) return "NaN";
  if (x == 0) return "0";
  if (std::isinf(x)) return x > 0 ? "Infinity" : "-Infinity";
  std::string sign = x < 0 ? "-" : "";
  char buf[64];
  const auto r = std::to_chars(buf, buf + sizeof(buf), std::fabs(x), std::chars_format::scientific);
  const std::string sci(buf, r.ptr);  // d[.ddd]e[+-]xx: the shortest round-trip digits
  const size_t ep = sci.find('e');
  std::string digits;
  for (size_t i = 0; i < ep; i++)
    if (sci[i] != '.') digits.push_back(sci[i]);
  const int k = (int)digits.size();
  const int n = std::atoi(sci.c_str() + ep + 1) + 1;  // the value is 0.digits x 10^n
  std::string s;
  if (k <= n && n <= 21) {
    s = digits + std::string(n - k, '0');
  } else if (0 < n && n <= 21) {
    s = digits.substr(0, n) + "." + digits.substr(n);
  } else if (-6 < n && n <= 0) {
    s = "0." + std::string(-n, '0') + digits;
  } else {
    const int e = n - 1;
    s = digits.substr(0, 1);
    if (k > 1) s += "." + digits.substr(1);
    s += std::string("e") + (e >= 0 ? "+" : "-") + std::to_string(std::abs(e));
  }
  return sign + s;
}

std::string toOBJ(const std::vector<const Mesh*>& parts) {
  if (parts.empty()) return "\n";  // [].join("\n") + "\n"
  std::string out;
  uint64_t base = 1;
  for (size_t i = 0; i < parts.size(); i++) {
    const Mesh& m = *parts[i];
    out += "o part_" + std::to_string(i) + "\n";
    for (size_t k = 0; k + 2 < m.verts.size(); k += 3)
      out += "v " + jsNumberString(m.verts[k]) + " " + jsNumberString(m.verts[k + 1]) + " " +
             jsNumberString(m.verts[k + 2]) + "\n";
    for (size_t k = 0; k + 2 < m.tris.size(); k += 3)
      out += "f " + std::to_string(m.tris[k] + base) + " " + std::to_string(m.tris[k + 1] + base) + " " +
             std::to_string(m.tris[k + 2] + base) + "\n";
    base += m.nv();
  }
  return out;
}

}  // namespace cuacd
