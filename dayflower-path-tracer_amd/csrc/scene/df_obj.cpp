// Wavefront OBJ ingest: ObjectLoader.load (scene/wavefront/ObjectLoader.java:118-212)
// and its ObjectModel / IndexedModel (:330-637), producing the triangle meshes
// the scene compiler turns into TriangleMesh primitives.
//
// Behaviour kept from the reference, because it decides which triangles,
// vertex attributes and mesh splits a scene gets:
//   * a line is split on single spaces only, empty tokens dropped; lines whose
//     first token is not g / usemtl / v / vt / vn / f are ignored; numbers
//     parse as Java's Float.valueOf / Integer.parseInt; indices are 1-based
//     (no negative indices);
//   * faces are fans (v0, vi+1, vi+2);
//   * whether texture coordinates / normals are used is decided for the whole
//     file (any face index that carries one); a face index without one then
//     uses entry 0;
//   * vertices are shared per (position, uv, normal) index triple; missing
//     normals are the normalized sums of the unit face normals around each
//     position; tangents are accumulated per position with the reference's
//     formula, whose z component repeats the y terms (:465-467), then
//     normalized (a zero sum gives NaN, as Vector3F.normalize does);
//   * v is flipped to 1 - v (Triangle.flipTextureCoordinatesForY);
//   * a mesh ends where the group of a triangle's first vertex changes; its
//     material name is that of the last triangle before the split;
//   * the files named aphroditegirl.obj and smoothMonkey2.obj are scaled by
//     100 in object space (the reference's per-file hack, :163-178).
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "df_scene_model.h"

namespace dfs {

namespace {

// Java's Float.valueOf: decimal (or hex) to the nearest float, optional
// trailing type suffix, surrounding whitespace allowed.
bool parse_float(const std::string& s, float* out) {
  size_t b = 0, e = s.size();
  while (b < e && static_cast<unsigned char>(s[b]) <= ' ') ++b;
  while (e > b && static_cast<unsigned char>(s[e - 1]) <= ' ') --e;
  std::string t = s.substr(b, e - b);
  if (!t.empty() && (t.back() == 'f' || t.back() == 'F' || t.back() == 'd' || t.back() == 'D')) {
    const bool hex = t.size() > 2 && (t.find("0x") != std::string::npos || t.find("0X") != std::string::npos);
    if (!hex || t.find_first_of("pP") != std::string::npos) t.pop_back();
  }
  if (t.empty()) return false;
  if (t == "NaN" || t == "+NaN" || t == "-NaN") {
    *out = std::nanf("");
    return true;
  }
  if (t == "Infinity" || t == "+Infinity" || t == "-Infinity") {
    *out = t[0] == '-' ? -INFINITY : INFINITY;
    return true;
  }
  for (char c : t)  // strtof also takes "inf", "nan(...)": Java does not
    if (!(std::isdigit(static_cast<unsigned char>(c)) || c == '.' || c == '+' || c == '-' || c == 'e' || c == 'E' ||
          c == 'x' || c == 'X' || c == 'p' || c == 'P' || (c >= 'a' && c <= 'f') || (c >= 'A' && c <= 'F')))
      return false;
  errno = 0;
  char* end = nullptr;
  const float v = std::strtof(t.c_str(), &end);
  if (end != t.c_str() + t.size()) return false;
  *out = v;  // overflow gives +-Infinity and underflow zero, as in Java
  return true;
}

// Java's Integer.parseInt (decimal, optional sign, 32-bit range).
bool parse_int(const std::string& s, int* out) {
  if (s.empty()) return false;
  size_t i = 0;
  bool neg = false;
  if (s[0] == '+' || s[0] == '-') {
    neg = s[0] == '-';
    i = 1;
    if (s.size() == 1) return false;
  }
  int64_t v = 0;
  for (; i < s.size(); ++i) {
    if (s[i] < '0' || s[i] > '9') return false;
    v = v * 10 + (s[i] - '0');
    if (v > 2147483648LL) return false;
  }
  if (neg) v = -v;
  if (v > 2147483647LL || v < -2147483648LL) return false;
  *out = static_cast<int>(v);
  return true;
}

// String.split(String) for a one-character separator: the whole string when
// the separator does not occur, else the pieces with trailing empty strings
// removed.
std::vector<std::string> java_split(const std::string& s, char sep) {
  if (s.find(sep) == std::string::npos) return {s};
  std::vector<std::string> parts;
  size_t start = 0;
  for (;;) {
    const size_t p = s.find(sep, start);
    if (p == std::string::npos) {
      parts.push_back(s.substr(start));
      break;
    }
    parts.push_back(s.substr(start, p - start));
    start = p + 1;
  }
  while (!parts.empty() && parts.back().empty()) parts.pop_back();
  return parts;
}

struct ObjIndex {
  int v = 0, t = 0, n = 0;
  bool operator<(const ObjIndex& o) const { return std::tie(v, t, n) < std::tie(o.v, o.t, o.n); }
};

Vec3 sub(const Vec3& a, const Vec3& b) { return Vec3{a.x - b.x, a.y - b.y, a.z - b.z}; }
Vec3 add(const Vec3& a, const Vec3& b) { return Vec3{a.x + b.x, a.y + b.y, a.z + b.z}; }
Vec3 cross(const Vec3& a, const Vec3& b) {
  return Vec3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
Vec3 normalize(const Vec3& a) {  // Vector3F.normalize: divide by the (float) length
  const float l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  return Vec3{a.x / l, a.y / l, a.z / l};
}

std::string base_name(const std::string& path) {
  const size_t p = path.find_last_of('/');
  return p == std::string::npos ? path : path.substr(p + 1);
}

}  // namespace

bool load_obj(const std::string& path, std::vector<ObjMesh>* out, std::string* error) {
  auto fail = [&](const std::string& what) {
    if (error) *error = path + ": " + what;
    return false;
  };
  std::ifstream in(path, std::ios::binary);
  if (!in) return fail("cannot open");
  std::stringstream ss;
  ss << in.rdbuf();
  const std::string text = ss.str();

  // ---- ObjectModel (:396-440): the file's statements
  std::vector<Vec3> positions, normals;
  std::vector<Vec2> uvs;
  std::vector<ObjIndex> corners;
  std::vector<std::string> corner_group, corner_material;
  bool has_uv = false, has_normals = false;
  std::string group, material;
  size_t line_no = 0;
  size_t pos = 0;
  while (pos < text.size()) {
    // BufferedReader.readLine: lines end at \n, \r or \r\n
    size_t e = text.find_first_of("\r\n", pos);
    if (e == std::string::npos) e = text.size();
    const std::string line = text.substr(pos, e - pos);
    pos = e;
    if (pos < text.size() && text[pos] == '\r') {
      ++pos;
      if (pos < text.size() && text[pos] == '\n') ++pos;
    } else if (pos < text.size()) {
      ++pos;  // '\n'
    }
    ++line_no;
    std::vector<std::string> tokens;
    for (std::string& t : java_split(line, ' '))
      if (!t.empty()) tokens.push_back(std::move(t));
    if (tokens.empty() || tokens[0] == "#") continue;
    const std::string& k = tokens[0];
    const std::string where = "line " + std::to_string(line_no);
    auto need = [&](size_t n) { return tokens.size() >= n; };
    if (k == "g") {
      if (!need(2)) return fail(where + ": g without a name");
      group = tokens[1];
    } else if (k == "usemtl") {
      if (!need(2)) return fail(where + ": usemtl without a name");
      material = tokens[1];
    } else if (k == "v" || k == "vn") {
      Vec3 v;
      if (!need(4) || !parse_float(tokens[1], &v.x) || !parse_float(tokens[2], &v.y) || !parse_float(tokens[3], &v.z))
        return fail(where + ": bad " + k);
      (k == "v" ? positions : normals).push_back(v);
    } else if (k == "vt") {
      Vec2 v;
      if (!need(3) || !parse_float(tokens[1], &v.x) || !parse_float(tokens[2], &v.y)) return fail(where + ": bad vt");
      uvs.push_back(v);
    } else if (k == "f") {
      // doParseObjectIndex (:538-561)
      auto parse_index = [&](const std::string& tok, ObjIndex* oi) {
        const std::vector<std::string> parts = java_split(tok, '/');
        if (parts.empty() || !parse_int(parts[0], &oi->v)) return false;
        oi->v -= 1;
        if (parts.size() > 1) {
          if (!parts[1].empty()) {
            has_uv = true;
            if (!parse_int(parts[1], &oi->t)) return false;
            oi->t -= 1;
          }
          if (parts.size() > 2) {
            has_normals = true;
            if (!parse_int(parts[2], &oi->n)) return false;
            oi->n -= 1;
          }
        }
        return true;
      };
      for (size_t i = 0; i + 3 < tokens.size(); ++i) {
        ObjIndex tri[3];
        if (!parse_index(tokens[1], &tri[0]) || !parse_index(tokens[2 + i], &tri[1]) ||
            !parse_index(tokens[3 + i], &tri[2]))
          return fail(where + ": bad face index");
        for (const ObjIndex& oi : tri) {
          corners.push_back(oi);
          corner_group.push_back(group);
          corner_material.push_back(material);
        }
      }
    }
  }

  // ---- toIndexedModel (:458-533): model 0 shares vertices per index triple,
  // model 1 per position (for the normal and tangent sums)
  struct Model {
    std::vector<int> indices;
    std::vector<Vec2> uv;
    std::vector<Vec3> pos, nrm, tan;
    std::vector<std::string> grp, mtl;
  } m0, m1;
  std::map<ObjIndex, int> vertex_of;
  std::unordered_map<int, int> position_of;
  std::vector<int> m0_to_m1;
  for (size_t i = 0; i < corners.size(); ++i) {
    const ObjIndex& oi = corners[i];
    if (oi.v < 0 || oi.v >= static_cast<int>(positions.size())) return fail("face refers to a missing vertex");
    Vec2 uv;
    if (has_uv) {
      if (oi.t < 0 || oi.t >= static_cast<int>(uvs.size())) return fail("face refers to a missing texture coordinate");
      uv = uvs[oi.t];
    }
    Vec3 n;
    if (has_normals) {
      if (oi.n < 0 || oi.n >= static_cast<int>(normals.size())) return fail("face refers to a missing normal");
      n = normals[oi.n];
    }
    auto it = vertex_of.find(oi);
    int mv;
    if (it == vertex_of.end()) {
      mv = static_cast<int>(m0.pos.size());
      vertex_of.emplace(oi, mv);
      m0.uv.push_back(uv);
      m0.pos.push_back(positions[oi.v]);
      m0.grp.push_back(corner_group[i]);
      m0.mtl.push_back(corner_material[i]);
      if (has_normals) m0.nrm.push_back(n);
      m0_to_m1.push_back(-1);
    } else {
      mv = it->second;
    }
    auto jt = position_of.find(oi.v);
    int np;
    if (jt == position_of.end()) {
      np = static_cast<int>(m1.pos.size());
      position_of.emplace(oi.v, np);
      m1.uv.push_back(uv);
      m1.pos.push_back(positions[oi.v]);
      m1.nrm.push_back(n);
      m1.tan.push_back(Vec3{});
    } else {
      np = jt->second;
    }
    m0.indices.push_back(mv);
    m1.indices.push_back(np);
    m0_to_m1[mv] = np;
  }
  if (!has_normals) {  // calculateNormals (:380-400)
    for (size_t i = 0; i + 2 < m1.indices.size(); i += 3) {
      const int a = m1.indices[i], b = m1.indices[i + 1], c = m1.indices[i + 2];
      const Vec3 n = normalize(cross(sub(m1.pos[b], m1.pos[a]), sub(m1.pos[c], m1.pos[a])));
      m1.nrm[a] = add(m1.nrm[a], n);
      m1.nrm[b] = add(m1.nrm[b], n);
      m1.nrm[c] = add(m1.nrm[c], n);
    }
    for (Vec3& n : m1.nrm) n = normalize(n);
    for (size_t i = 0; i < m0.pos.size(); ++i) m0.nrm.push_back(m1.nrm[m0_to_m1[i]]);
  }
  // calculateTangents (:402-430)
  for (size_t i = 0; i + 2 < m1.indices.size(); i += 3) {
    const int a = m1.indices[i], b = m1.indices[i + 1], c = m1.indices[i + 2];
    const Vec3 e0 = sub(m1.pos[b], m1.pos[a]);
    const Vec3 e1 = sub(m1.pos[c], m1.pos[a]);
    const float du0 = m1.uv[b].x - m1.uv[a].x;
    const float dv0 = m1.uv[b].y - m1.uv[a].y;
    const float du1 = m1.uv[c].x - m1.uv[a].x;
    const float dv1 = m1.uv[c].y - m1.uv[a].y;
    const float dividend = du0 * dv1 - du1 * dv0;
    // Float.compare(dividend, 0.0F) == 0: +0 only (-0 compares below +0)
    const bool zero = dividend == 0.0f && !std::signbit(dividend);
    const float fraction = zero ? 0.0f : 1.0f / dividend;
    const Vec3 t{fraction * (dv1 * e0.x - dv0 * e1.x), fraction * (dv1 * e0.y - dv0 * e1.y),
                 fraction * (dv1 * e0.y - dv0 * e1.y)};
    m1.tan[a] = add(m1.tan[a], t);
    m1.tan[b] = add(m1.tan[b], t);
    m1.tan[c] = add(m1.tan[c], t);
  }
  for (Vec3& t : m1.tan) t = normalize(t);
  for (size_t i = 0; i < m0.pos.size(); ++i) m0.tan.push_back(m1.tan[m0_to_m1[i]]);

  // ---- load (:118-212): triangles, split where the group changes
  const std::string name = base_name(path);
  const bool scaled = name == "aphroditegirl.obj" || name == "smoothMonkey2.obj";
  float o2w[16], w2o[16];
  if (scaled) {
    Primitive p;
    p.scale = Vec3{100.0f, 100.0f, 100.0f};
    if (!transform_matrices(p, o2w, w2o)) return fail("singular scaling");
  }
  if (m0.indices.empty()) return fail("no faces");
  out->clear();
  std::vector<Triangle> tris;
  std::string prev_group, prev_material;
  for (size_t i = 0; i + 2 < m0.indices.size(); i += 3) {
    const int ia = m0.indices[i], ib = m0.indices[i + 1], ic = m0.indices[i + 2];
    if (prev_group != m0.grp[ia]) {
      if (!tris.empty()) {
        out->push_back(ObjMesh{prev_group, prev_material, std::move(tris)});
        tris.clear();
      }
      prev_group = m0.grp[ia];
    }
    prev_material = m0.mtl[ia];
    Triangle t;
    Vertex* vs[3] = {&t.a, &t.b, &t.c};
    const int ids[3] = {ia, ib, ic};
    for (int k = 0; k < 3; ++k) {
      Vertex& v = *vs[k];
      v.position = m0.pos[ids[k]];
      v.normal = m0.nrm[ids[k]];
      v.tangent = m0.tan[ids[k]];
      v.uv = Vec2{m0.uv[ids[k]].x, 1.0f - m0.uv[ids[k]].y};
      if (scaled) transform_vertex(o2w, w2o, &v);
    }
    tris.push_back(t);
  }
  out->push_back(ObjMesh{prev_group, prev_material, std::move(tris)});
  return true;
}

}  // namespace dfs
