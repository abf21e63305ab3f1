// Deterministic generators for the benchmark workloads (procedural scenes
// expressed in the reference's object model) and the .scene file format.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>

#include "df_scene_model.h"

namespace dfs {

namespace {

// SplitMix64 (Steele, Lea, Flood 2014) - the generator for every seeded
// scene parameter in this file.
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform in [0, 1)
  double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }
  float range(float a, float b) { return static_cast<float>(a + (b - a) * uniform()); }
  int irange(int a, int b) { return a + static_cast<int>(uniform() * (b - a + 1)); }  // inclusive
};

// Seeded 3-D gradient noise (Perlin's improved noise construction with a
// shuffled permutation) summed over octaves: the fBm used to perturb the
// icosphere radii.
class GradientNoise {
 public:
  explicit GradientNoise(uint64_t seed) {
    Rng r(seed);
    for (int i = 0; i < 256; ++i) perm_[i] = i;
    for (int i = 255; i > 0; --i) std::swap(perm_[i], perm_[static_cast<int>(r.uniform() * (i + 1))]);
    for (int i = 0; i < 256; ++i) perm_[256 + i] = perm_[i];
  }
  double noise(double x, double y, double z) const {
    const int X = static_cast<int>(std::floor(x)) & 255;
    const int Y = static_cast<int>(std::floor(y)) & 255;
    const int Z = static_cast<int>(std::floor(z)) & 255;
    x -= std::floor(x);
    y -= std::floor(y);
    z -= std::floor(z);
    const double u = fade(x), v = fade(y), w = fade(z);
    const int A = perm_[X] + Y, AA = perm_[A] + Z, AB = perm_[A + 1] + Z;
    const int B = perm_[X + 1] + Y, BA = perm_[B] + Z, BB = perm_[B + 1] + Z;
    return lerp(w,
                lerp(v, lerp(u, grad(perm_[AA], x, y, z), grad(perm_[BA], x - 1, y, z)),
                     lerp(u, grad(perm_[AB], x, y - 1, z), grad(perm_[BB], x - 1, y - 1, z))),
                lerp(v, lerp(u, grad(perm_[AA + 1], x, y, z - 1), grad(perm_[BA + 1], x - 1, y, z - 1)),
                     lerp(u, grad(perm_[AB + 1], x, y - 1, z - 1), grad(perm_[BB + 1], x - 1, y - 1, z - 1))));
  }
  double fbm(double x, double y, double z, int octaves) const {
    double sum = 0.0, amp = 1.0, freq = 1.0, norm = 0.0;
    for (int i = 0; i < octaves; ++i) {
      sum += amp * noise(x * freq, y * freq, z * freq);
      norm += amp;
      amp *= 0.5;
      freq *= 2.0;
    }
    return sum / norm;
  }

 private:
  static double fade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
  static double lerp(double t, double a, double b) { return a + t * (b - a); }
  static double grad(int hash, double x, double y, double z) {
    const int h = hash & 15;
    const double u = h < 8 ? x : y;
    const double v = h < 4 ? y : (h == 12 || h == 14) ? x : z;
    return ((h & 1) == 0 ? u : -u) + ((h & 2) == 0 ? v : -v);
  }
  int perm_[512];
};

Vec3 v3(double x, double y, double z) { return Vec3{static_cast<float>(x), static_cast<float>(y), static_cast<float>(z)}; }

Color gray(float v) { return Color{v, v, v}; }

Color random_color(Rng& r, float lo, float hi) { return Color{r.range(lo, hi), r.range(lo, hi), r.range(lo, hi)}; }

// Unit icosphere: an icosahedron whose faces are split into four by edge
// midpoints `levels` times, vertices projected to the unit sphere.  Returns
// positions and faces in face order (the faces of base face f occupy the
// contiguous range [f * 4^levels, (f + 1) * 4^levels)).
void icosphere(int levels, std::vector<std::array<double, 3>>* verts, std::vector<std::array<int, 3>>* faces) {
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  std::vector<std::array<double, 3>>& V = *verts;
  V = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
       {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  for (auto& p : V) {
    const double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    p = {p[0] / l, p[1] / l, p[2] / l};
  }
  std::vector<std::array<int, 3>> F = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                                       {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                       {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                                       {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
  for (int l = 0; l < levels; ++l) {
    std::map<std::pair<int, int>, int> mid;
    auto midpoint = [&](int a, int b) {
      const auto key = std::make_pair(std::min(a, b), std::max(a, b));
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      std::array<double, 3> m = {(V[a][0] + V[b][0]) * 0.5, (V[a][1] + V[b][1]) * 0.5, (V[a][2] + V[b][2]) * 0.5};
      const double len = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
      m = {m[0] / len, m[1] / len, m[2] / len};
      V.push_back(m);
      const int idx = static_cast<int>(V.size()) - 1;
      mid.emplace(key, idx);
      return idx;
    };
    std::vector<std::array<int, 3>> G;
    G.reserve(F.size() * 4);
    for (const auto& f : F) {
      const int a = midpoint(f[0], f[1]);
      const int b = midpoint(f[1], f[2]);
      const int c = midpoint(f[2], f[0]);
      G.push_back({f[0], a, c});
      G.push_back({f[1], b, a});
      G.push_back({f[2], c, b});
      G.push_back({a, b, c});
    }
    F.swap(G);
  }
  *faces = std::move(F);
}

// Displaced icosphere (radius * (1 + amplitude * fbm(p))) with smooth
// per-vertex normals, spherical UVs and longitude tangents.
std::vector<Triangle> noisy_icosphere(int levels, double radius, Vec3 center, double amplitude, int octaves,
                                      uint64_t seed) {
  std::vector<std::array<double, 3>> V;
  std::vector<std::array<int, 3>> F;
  icosphere(levels, &V, &F);
  GradientNoise noise(seed);
  std::vector<std::array<double, 3>> P(V.size()), N(V.size(), {0, 0, 0});
  for (size_t i = 0; i < V.size(); ++i) {
    const double r = radius * (1.0 + amplitude * noise.fbm(V[i][0] * 2.0 + 11.3, V[i][1] * 2.0 + 7.1,
                                                           V[i][2] * 2.0 + 3.7, octaves));
    P[i] = {V[i][0] * r, V[i][1] * r, V[i][2] * r};
  }
  for (const auto& f : F) {
    const auto& a = P[f[0]];
    const auto& b = P[f[1]];
    const auto& c = P[f[2]];
    const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    for (int k = 0; k < 3; ++k)
      for (int j = 0; j < 3; ++j) N[f[k]][j] += n[j];
  }
  std::vector<Triangle> tris;
  tris.reserve(F.size());
  auto vertex = [&](int i) {
    Vertex v;
    v.position = v3(P[i][0] + center.x, P[i][1] + center.y, P[i][2] + center.z);
    const double l = std::sqrt(N[i][0] * N[i][0] + N[i][1] * N[i][1] + N[i][2] * N[i][2]);
    v.normal = v3(N[i][0] / l, N[i][1] / l, N[i][2] / l);
    const double phi = std::atan2(V[i][2], V[i][0]);
    v.uv = Vec2{static_cast<float>(0.5 + phi / (2.0 * M_PI)), static_cast<float>(0.5 - std::asin(V[i][1]) / M_PI)};
    double tx = -std::sin(phi), tz = std::cos(phi);
    v.tangent = v3(tx, 0.0, tz);
    return v;
  };
  for (const auto& f : F) tris.push_back(Triangle{vertex(f[0]), vertex(f[1]), vertex(f[2])});
  return tris;
}

// Random flat-shaded triangles: centre uniform in the box, edge size
// uniform in [smin, smax], random orientation.
std::vector<Triangle> scattered_triangles(int count, Vec3 box_min, Vec3 box_max, float smin, float smax, Rng& r) {
  std::vector<Triangle> tris;
  tris.reserve(count);
  for (int i = 0; i < count; ++i) {
    const double cx = r.range(box_min.x, box_max.x);
    const double cy = r.range(box_min.y, box_max.y);
    const double cz = r.range(box_min.z, box_max.z);
    const double size = r.range(smin, smax);
    double pts[3][3];
    for (int k = 0; k < 3; ++k) {
      // random direction on the sphere
      const double z = 2.0 * r.uniform() - 1.0;
      const double a = 2.0 * M_PI * r.uniform();
      const double s = std::sqrt(std::max(0.0, 1.0 - z * z));
      pts[k][0] = cx + size * 0.5 * s * std::cos(a);
      pts[k][1] = cy + size * 0.5 * s * std::sin(a);
      pts[k][2] = cz + size * 0.5 * z;
    }
    const double e1[3] = {pts[1][0] - pts[0][0], pts[1][1] - pts[0][1], pts[1][2] - pts[0][2]};
    const double e2[3] = {pts[2][0] - pts[0][0], pts[2][1] - pts[0][1], pts[2][2] - pts[0][2]};
    double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    double nl = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (nl < 1e-12) {
      n[0] = 0;
      n[1] = 1;
      n[2] = 0;
      nl = 1;
    }
    const double el = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    Triangle t;
    Vertex* vs[3] = {&t.a, &t.b, &t.c};
    const float uvs[3][2] = {{0, 0}, {1, 0}, {0, 1}};
    for (int k = 0; k < 3; ++k) {
      vs[k]->position = v3(pts[k][0], pts[k][1], pts[k][2]);
      vs[k]->normal = v3(n[0] / nl, n[1] / nl, n[2] / nl);
      vs[k]->uv = Vec2{uvs[k][0], uvs[k][1]};
      vs[k]->tangent = el > 0 ? v3(e1[0] / el, e1[1] / el, e1[2] / el) : v3(1, 0, 0);
    }
    tris.push_back(t);
  }
  return tris;
}

void add_ground(Scene& s, int material, int albedo) {
  const int black = constant_texture(s, gray(0.0f));
  const int surf = add_surface(s, material, albedo, black, black);
  // Plane(a=(0,0,0), b=(0,0,1), c=(1,0,0)) -> normal (0,1,0), as in the
  // reference's demo scenes (Scenes.java:378).
  add_primitive(s, make_plane(Vec3{0, 0, 0}, Vec3{0, 0, 1}, Vec3{1, 0, 0}), surf);
}

void finish_camera(Scene& s, int w, int h, float fov_x, Vec3 eye, Vec3 target, float aperture, float focal) {
  camera_default(s.camera);
  camera_set_resolution(s.camera, static_cast<float>(w), static_cast<float>(h));
  camera_set_field_of_view_x(s.camera, fov_x);
  camera_set_aperture(s.camera, aperture);
  camera_set_focal_distance(s.camera, focal);
  camera_look_at(s.camera, eye, target);
}

Vec3 sun_direction(float elevation_deg, float azimuth_deg) {
  const double e = elevation_deg * M_PI / 180.0, a = azimuth_deg * M_PI / 180.0;
  return v3(std::cos(e) * std::sin(a), std::sin(e), std::cos(e) * std::cos(a));
}

// The ~100k-triangle mesh scene: a noisy icosphere of 81,920 triangles
// (20 * 4^6; fBm amplitude 0.15, 5 octaves, seed 7) in 4 groups plus 20,000
// scattered triangles of edge 0.01-0.2 m in a 10 m box (seed 8) in 4
// groups, each group one TriangleMesh primitive with a material drawn
// uniformly from {Lambert, Phong, ClearCoat, Glass}.
void add_mesh_groups(Scene& s, int levels, uint64_t noise_seed, uint64_t scatter_seed, int scatter_count,
                     Vec3 sphere_center, float sphere_radius, Rng& mat_rng) {
  static const int kMats[4] = {kMaterialLambertian, kMaterialPhong, kMaterialClearCoat, kMaterialGlass};
  const int black = constant_texture(s, gray(0.0f));
  std::vector<Triangle> ico = noisy_icosphere(levels, sphere_radius, sphere_center, 0.15, 5, noise_seed);
  const size_t per = ico.size() / 4;
  for (int g = 0; g < 4; ++g) {
    std::vector<Triangle> part(ico.begin() + g * per, g == 3 ? ico.end() : ico.begin() + (g + 1) * per);
    const int mat = kMats[mat_rng.irange(0, 3)];
    const int alb = constant_texture(s, random_color(mat_rng, 0.2f, 0.9f));
    add_primitive(s, make_mesh(std::move(part)), add_surface(s, mat, alb, black, black));
  }
  Rng sr(scatter_seed);
  std::vector<Triangle> sc = scattered_triangles(scatter_count, Vec3{-5, 0, -5}, Vec3{5, 10, 5}, 0.01f, 0.2f, sr);
  // group by x quadrant so that groups stay spatially compact
  std::vector<Triangle> groups[4];
  for (const Triangle& t : sc) {
    const float cx = (t.a.position.x + t.b.position.x + t.c.position.x) / 3.0f;
    const int g = std::min(3, std::max(0, static_cast<int>((cx + 5.0f) / 2.5f)));
    groups[g].push_back(t);
  }
  for (int g = 0; g < 4; ++g) {
    if (groups[g].empty()) continue;
    const int mat = kMats[mat_rng.irange(0, 3)];
    const int alb = constant_texture(s, random_color(mat_rng, 0.2f, 0.9f));
    add_primitive(s, make_mesh(std::move(groups[g])), add_surface(s, mat, alb, black, black));
  }
}

Compiled config1(uint64_t seed) {
  // 256x256, 16 spp, depth 5; checkerboard ground (1 m tiles: scale 0.5
  // because a tile is half a period of frac(u * scale)); three spheres with
  // centres from seed 1; Perez sky turbidity 2, sun 45 deg; pinhole FOV 40;
  // Reinhard.
  Scene s;
  s.name = "config1_spheres_256";
  const int checker = checkerboard_texture(s, gray(0.8f), gray(0.2f), 0.0f, 0.5f, 0.5f);
  add_ground(s, kMaterialLambertian, checker);
  Rng r(seed ? seed : 1);
  const int black = constant_texture(s, gray(0.0f));
  struct SphereSpec {
    float radius;
    int material;
    Color albedo;
  } specs[3] = {{1.0f, kMaterialLambertian, Color{0.7f, 0.3f, 0.3f}},
                {0.75f, kMaterialGlass, gray(1.0f)},
                {0.5f, kMaterialReflection, gray(0.9f)}};
  for (const SphereSpec& sp : specs) {
    const float x = r.range(-3.0f, 3.0f);
    const float z = r.range(-3.0f, 3.0f);
    const int alb = constant_texture(s, sp.albedo);
    // Sphere at the object origin placed by the primitive's transform, as the
    // reference's scenes do (Scenes.java:380-384).
    add_primitive(s, make_sphere(Vec3{0, 0, 0}, sp.radius), add_surface(s, sp.material, alb, black, black),
                  Vec3{x, sp.radius, z});
  }
  finish_camera(s, 256, 256, 40.0f, Vec3{0.0f, 3.0f, 11.0f}, Vec3{0.0f, 0.5f, 0.0f}, 0.0f, 11.0f);
  sky_set(&s.sky, sun_direction(45.0f, 30.0f), 2.0f);
  Compiled c = compile(s);
  c.width = 256;
  c.height = 256;
  c.spp = 16;
  c.params = default_params();
  c.params.pt_depth_maximum = 5;
  c.params.tone_mapper_type = DF_TONE_MAPPER_REINHARD;
  return c;
}

Scene mesh_scene_base(const char* name, int w, int h, Rng& mat_rng) {
  Scene s;
  s.name = name;
  const int checker = checkerboard_texture(s, gray(0.8f), gray(0.2f), 0.0f, 0.5f, 0.5f);
  add_ground(s, kMaterialLambertian, checker);
  add_mesh_groups(s, 6, 7, 8, 20000, Vec3{0.0f, 3.2f, 0.0f}, 2.5f, mat_rng);
  (void)w;
  (void)h;
  return s;
}

Compiled config2(uint64_t seed) {
  // 1024x768, 64 spp, depth 8, ~100k triangles, turbidity 3, ACES.
  Rng mat_rng(seed ? seed : 9);
  Scene s = mesh_scene_base("config2_mesh_100k_1024x768", 1024, 768, mat_rng);
  finish_camera(s, 1024, 768, 40.0f, Vec3{0.0f, 5.0f, 17.0f}, Vec3{0.0f, 3.0f, 0.0f}, 0.0f, 17.0f);
  sky_set(&s.sky, sun_direction(45.0f, 30.0f), 3.0f);
  Compiled c = compile(s);
  c.width = 1024;
  c.height = 768;
  c.spp = 64;
  c.params = default_params();
  c.params.pt_depth_maximum = 8;
  c.params.tone_mapper_type = DF_TONE_MAPPER_FILMIC_CURVE_ACES_MODIFIED;
  return c;
}

Compiled config3(uint64_t seed) {
  // 1920x1080, 128 spp, depth 8: the mesh scene + terrain + 50 spheres with
  // noise normal mapping and fBm/UV/blend textures + one emissive triangle.
  // The terrain record carries the configured fBm parameters (frequency
  // 0.25, gain 0.5, -2..2, 8 octaves); the reference kernel's march ignores
  // them and uses sin(x) * sin(z) at scale 10 (GPURendererKernel.java:1470-1538).
  Rng mat_rng(seed ? seed : 9);
  Scene s = mesh_scene_base("config3_mesh_terrain_1080p", 1920, 1080, mat_rng);
  const int black = constant_texture(s, gray(0.0f));
  {
    const int alb = constant_texture(s, Color{0.35f, 0.55f, 0.25f});
    add_primitive(s, make_terrain(0.25f, 0.5f, -2.0f, 2.0f, 8), add_surface(s, kMaterialLambertian, alb, black, black));
  }
  Rng r(12);
  static const int kMats[5] = {kMaterialClearCoat, kMaterialGlass, kMaterialLambertian, kMaterialPhong,
                               kMaterialReflection};
  for (int i = 0; i < 50; ++i) {
    const float radius = r.range(0.2f, 1.0f);
    const float x = r.range(-9.0f, 9.0f);
    const float z = r.range(-9.0f, 9.0f);
    int alb;
    const int kind = r.irange(0, 2);
    if (kind == 0) {
      alb = fbm_texture(s, random_color(r, 0.05f, 0.5f), gray(1.0f), r.range(0.5f, 2.0f), 0.5f, 8);
    } else if (kind == 1) {
      alb = uv_texture(s);
    } else {
      const int a = fbm_texture(s, random_color(r, 0.05f, 0.3f), gray(1.0f), 0.5f, 0.5f, 8);
      const int b = checkerboard_texture(s, random_color(r, 0.2f, 0.9f), random_color(r, 0.2f, 0.9f), 0.0f, 4.0f, 4.0f);
      alb = blend_texture(s, a, b, 0.5f);
    }
    const int mat = kMats[r.irange(0, 4)];
    const float noise_amount = r.range(0.1f, 0.4f);
    const float noise_scale = r.range(0.5f, 2.0f);
    add_primitive(s, make_sphere(Vec3{0, 0, 0}, radius), add_surface(s, mat, alb, black, black, noise_amount, noise_scale),
                  Vec3{x, radius, z});
  }
  {
    // Emissive triangle area light: radiance 10 is not representable in the
    // 8-bit ARGB emission texture; it saturates to 1.0.
    Triangle t;
    Vertex* vs[3] = {&t.a, &t.b, &t.c};
    const Vec3 p[3] = {Vec3{-3.0f, 7.5f, -4.0f}, Vec3{3.0f, 7.5f, -4.0f}, Vec3{0.0f, 9.5f, -5.0f}};
    for (int k = 0; k < 3; ++k) {
      vs[k]->position = p[k];
      vs[k]->normal = Vec3{0.0f, -0.447214f, 0.894427f};
      vs[k]->uv = Vec2{k == 1 ? 1.0f : 0.0f, k == 2 ? 1.0f : 0.0f};
      vs[k]->tangent = Vec3{1, 0, 0};
    }
    const int alb = constant_texture(s, gray(0.8f));
    const int emi = constant_texture(s, gray(10.0f));
    add_primitive(s, make_triangle(t), add_surface(s, kMaterialLambertian, alb, emi, black));
  }
  finish_camera(s, 1920, 1080, 40.0f, Vec3{0.0f, 4.0f, 17.0f}, Vec3{0.0f, 2.5f, 0.0f}, 0.05f, 8.0f);
  sky_set(&s.sky, sun_direction(20.0f, 30.0f), 2.5f);
  Compiled c = compile(s);
  c.width = 1920;
  c.height = 1080;
  c.spp = 128;
  c.params = default_params();
  c.params.pt_depth_maximum = 8;
  c.params.tone_mapper_type = DF_TONE_MAPPER_FILMIC_CURVE_ACES_MODIFIED;
  return c;
}

}  // namespace

Compiled build_batch_scene(uint64_t seed) {
  // One scene of the uneven batch mix: 4-64 spheres, 0-50,000 triangles
  // (log-uniform), one plane, resolution from {64, 128, 256, 512}^2, spp
  // from {8, 32, 128}, depth 4-8, materials uniform over the five types,
  // turbidity uniform in 2-6.
  Rng r(seed * 0x2545F4914F6CDD1Dull + 0x1234567ull);
  Scene s;
  s.name = "batch_" + std::to_string(seed);
  static const int kMats[5] = {kMaterialClearCoat, kMaterialGlass, kMaterialLambertian, kMaterialPhong,
                               kMaterialReflection};
  const int black = constant_texture(s, gray(0.0f));
  {
    const int checker = checkerboard_texture(s, random_color(r, 0.2f, 0.9f), random_color(r, 0.2f, 0.9f), 0.0f, 0.5f, 0.5f);
    add_primitive(s, make_plane(Vec3{0, 0, 0}, Vec3{0, 0, 1}, Vec3{1, 0, 0}),
                  add_surface(s, kMaterialLambertian, checker, black, black));
  }
  const int spheres = r.irange(4, 64);
  for (int i = 0; i < spheres; ++i) {
    const float radius = r.range(0.2f, 1.0f);
    const int alb = constant_texture(s, random_color(r, 0.1f, 0.9f));
    add_primitive(s, make_sphere(Vec3{0, 0, 0}, radius), add_surface(s, kMats[r.irange(0, 4)], alb, black, black),
                  Vec3{r.range(-8.0f, 8.0f), radius, r.range(-8.0f, 8.0f)});
  }
  // log-uniform in [1, 50001) minus 1 -> [0, 50000]
  const int tris = static_cast<int>(std::exp(r.uniform() * std::log(50001.0))) - 1;
  if (tris > 0) {
    std::vector<Triangle> sc = scattered_triangles(tris, Vec3{-6, 0, -6}, Vec3{6, 6, 6}, 0.05f, 0.6f, r);
    const int alb = constant_texture(s, random_color(r, 0.1f, 0.9f));
    add_primitive(s, make_mesh(std::move(sc)), add_surface(s, kMats[r.irange(0, 4)], alb, black, black));
  }
  static const int kRes[4] = {64, 128, 256, 512};
  static const int kSpp[3] = {8, 32, 128};
  const int res = kRes[r.irange(0, 3)];
  const int spp = kSpp[r.irange(0, 2)];
  const int depth = r.irange(4, 8);
  const float turbidity = r.range(2.0f, 6.0f);
  finish_camera(s, res, res, 40.0f, Vec3{0.0f, 5.0f, 18.0f}, Vec3{0.0f, 1.0f, 0.0f}, 0.0f, 18.0f);
  sky_set(&s.sky, sun_direction(r.range(15.0f, 75.0f), r.range(0.0f, 360.0f)), turbidity);
  Compiled c = compile(s);
  c.width = res;
  c.height = res;
  c.spp = spp;
  c.params = default_params();
  c.params.pt_depth_maximum = depth;
  return c;
}

// The differentiation scene: 512x512, 32 spp, depth 6 (Russian roulette off:
// depth 7).  A noisy icosphere subdivided one level further than the mesh
// scene (20 * 4^7 = 327,680 triangles, fBm amplitude 0.15, 5 octaves, seed 7)
// in 48 groups of consecutive triangles, plus 20,000 scattered triangles
// (seed 8) in 16 groups by x/z quadrant: 64 TriangleMesh primitives, each
// with its own surface (albedo group k = the k-th mesh primitive), materials
// uniform over {Lambert, Phong, ClearCoat, Glass}.  Four emissive area-light
// triangles (emission j = the j-th triangle primitive) above the scene, a
// checkerboard ground.  The albedo, emission, turbidity and sun elevation
// that a render uses come from the differentiable-mode parameters; the
// textures here are the initial values.
Compiled config5(uint64_t seed) {
  Rng mat_rng(seed ? seed : 9);
  Scene s;
  s.name = "config5_diff_mesh_512";
  const int checker = checkerboard_texture(s, gray(0.8f), gray(0.2f), 0.0f, 0.5f, 0.5f);
  add_ground(s, kMaterialLambertian, checker);
  static const int kMats[4] = {kMaterialLambertian, kMaterialPhong, kMaterialClearCoat, kMaterialGlass};
  const int black = constant_texture(s, gray(0.0f));
  std::vector<Triangle> ico = noisy_icosphere(7, 2.5f, Vec3{0.0f, 3.2f, 0.0f}, 0.15, 5, 7);
  const size_t per = ico.size() / 48;
  for (int g = 0; g < 48; ++g) {
    std::vector<Triangle> part(ico.begin() + g * per, g == 47 ? ico.end() : ico.begin() + (g + 1) * per);
    const int alb = constant_texture(s, random_color(mat_rng, 0.2f, 0.9f));
    add_primitive(s, make_mesh(std::move(part)), add_surface(s, kMats[mat_rng.irange(0, 3)], alb, black, black));
  }
  Rng sr(8);
  std::vector<Triangle> sc = scattered_triangles(20000, Vec3{-5, 0, -5}, Vec3{5, 10, 5}, 0.01f, 0.2f, sr);
  std::vector<Triangle> groups[16];
  for (const Triangle& t : sc) {
    const float cx = (t.a.position.x + t.b.position.x + t.c.position.x) / 3.0f;
    const float cz = (t.a.position.z + t.b.position.z + t.c.position.z) / 3.0f;
    const int gx = std::min(3, std::max(0, static_cast<int>((cx + 5.0f) / 2.5f)));
    const int gz = std::min(3, std::max(0, static_cast<int>((cz + 5.0f) / 2.5f)));
    groups[gx * 4 + gz].push_back(t);
  }
  for (int g = 0; g < 16; ++g) {
    const int alb = constant_texture(s, random_color(mat_rng, 0.2f, 0.9f));
    const int mat = kMats[mat_rng.irange(0, 3)];
    if (groups[g].empty()) continue;
    add_primitive(s, make_mesh(std::move(groups[g])), add_surface(s, mat, alb, black, black));
  }
  for (int j = 0; j < 4; ++j) {  // area lights around and above the sphere, facing down
    const float a = static_cast<float>(j) * 1.5707964f + 0.4f;
    const float cx = 6.0f * std::cos(a), cz = 6.0f * std::sin(a);
    Triangle t;
    Vertex* vs[3] = {&t.a, &t.b, &t.c};
    const Vec3 p[3] = {Vec3{cx - 1.0f, 9.0f, cz - 1.0f}, Vec3{cx + 1.0f, 9.0f, cz - 1.0f}, Vec3{cx, 9.0f, cz + 1.0f}};
    for (int k = 0; k < 3; ++k) {
      vs[k]->position = p[k];
      vs[k]->normal = Vec3{0.0f, -1.0f, 0.0f};
      vs[k]->uv = Vec2{k == 1 ? 1.0f : 0.0f, k == 2 ? 1.0f : 0.0f};
      vs[k]->tangent = Vec3{1, 0, 0};
    }
    // distinct albedo textures keep the four surfaces distinct records
    const int alb = constant_texture(s, gray(0.80f + 0.01f * static_cast<float>(j)));
    const int emi = constant_texture(s, gray(1.0f));
    add_primitive(s, make_triangle(t), add_surface(s, kMaterialLambertian, alb, emi, black));
  }
  finish_camera(s, 512, 512, 40.0f, Vec3{0.0f, 5.0f, 17.0f}, Vec3{0.0f, 3.0f, 0.0f}, 0.0f, 17.0f);
  sky_set(&s.sky, sun_direction(45.0f, 30.0f), 3.0f);
  Compiled c = compile(s);
  c.width = 512;
  c.height = 512;
  c.spp = 32;
  c.params = default_params();
  c.params.pt_depth_maximum = 6;
  c.params.pt_depth_russian_roulette = 7;
  c.params.tone_mapper_type = DF_TONE_MAPPER_FILMIC_CURVE_ACES_MODIFIED;
  return c;
}

Compiled build_config(int config_id, uint64_t seed) {
  switch (config_id) {
    case 1:
      return config1(seed);
    case 2:
      return config2(seed);
    case 3:
      return config3(seed);
    case 5:
      return config5(seed);
    default:
      return Compiled{};
  }
}

// ---------------------------------------------------------------- .scene I/O

namespace {

void put_u32(std::string& out, uint32_t v) {
  const char b[4] = {static_cast<char>(v >> 24), static_cast<char>(v >> 16), static_cast<char>(v >> 8),
                     static_cast<char>(v)};
  out.append(b, 4);
}

template <typename T>
void put_array(std::string& out, const std::vector<T>& a) {
  put_u32(out, static_cast<uint32_t>(a.size()));
  for (const T& x : a) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    put_u32(out, u);
  }
}

struct Reader {
  const std::string& s;
  size_t pos = 0;
  bool ok = true;
  uint32_t u32() {
    if (pos + 4 > s.size()) {
      ok = false;
      return 0;
    }
    const unsigned char* p = reinterpret_cast<const unsigned char*>(s.data() + pos);
    pos += 4;
    return (static_cast<uint32_t>(p[0]) << 24) | (static_cast<uint32_t>(p[1]) << 16) |
           (static_cast<uint32_t>(p[2]) << 8) | p[3];
  }
  template <typename T>
  std::vector<T> array() {
    const uint32_t n = u32();
    std::vector<T> a;
    if (!ok || n > (s.size() - pos) / 4) {
      ok = false;
      return a;
    }
    a.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t u = u32();
      std::memcpy(&a[i], &u, 4);
    }
    return a;
  }
};

}  // namespace

bool write_scene_file(const Compiled& c, const std::string& path) {
  // DataOutputStream.writeUTF (2-byte big-endian length + modified UTF-8;
  // identical to plain bytes for ASCII names) then each array as int32 length
  // + big-endian elements (CompiledScene.java:273-295, Arrays2.java:247-273).
  std::string out;
  const std::string& name = c.name;
  out.push_back(static_cast<char>((name.size() >> 8) & 0xFF));
  out.push_back(static_cast<char>(name.size() & 0xFF));
  out += name;
  put_array(out, c.point2fs);
  put_array(out, c.point3fs);
  put_array(out, c.vector3fs);
  put_array(out, c.planes);
  put_array(out, c.spheres);
  put_array(out, c.terrains);
  put_array(out, c.triangles);
  put_array(out, c.bvh);
  put_array(out, c.textures);
  put_array(out, c.surfaces);
  put_array(out, c.primitives);
  put_array(out, c.object_to_world);
  put_array(out, c.world_to_object);
  put_array(out, c.camera);
  std::ofstream f(path, std::ios::binary);
  if (!f) return false;
  f.write(out.data(), static_cast<std::streamsize>(out.size()));
  return static_cast<bool>(f);
}

bool read_scene_file(const std::string& path, Compiled* out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::string s((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  if (s.size() < 2) return false;
  const size_t n = (static_cast<unsigned char>(s[0]) << 8) | static_cast<unsigned char>(s[1]);
  if (2 + n > s.size()) return false;
  Compiled c;
  c.name = s.substr(2, n);
  Reader r{s, 2 + n};
  c.point2fs = r.array<float>();
  c.point3fs = r.array<float>();
  c.vector3fs = r.array<float>();
  c.planes = r.array<int32_t>();
  c.spheres = r.array<float>();
  c.terrains = r.array<float>();
  c.triangles = r.array<int32_t>();
  c.bvh = r.array<int32_t>();
  c.textures = r.array<float>();
  c.surfaces = r.array<float>();
  c.primitives = r.array<int32_t>();
  c.object_to_world = r.array<float>();
  c.world_to_object = r.array<float>();
  c.camera = r.array<float>();
  if (!r.ok || c.camera.size() != 19) return false;
  c.params = default_params();
  c.width = static_cast<int>(c.camera[16]);
  c.height = static_cast<int>(c.camera[17]);
  c.spp = 1;
  sky_set(&c.sky, Vec3{1.0f, 1.0f, 1.0f}, 2.0f);  // Sky() default: set(1,1,1) (Sky.java:86-88, 268-270)
  *out = std::move(c);
  return true;
}

}  // namespace dfs
