// Scene object model -> CompiledScene arrays, camera and sky host math, and
// the SAH-sweep BVH builder.  Host C++ only (built with g++).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <unordered_map>

#include "df_scene_model.h"

namespace dfs {

namespace {

inline float fsat(float v) { return v < 0.0f ? 0.0f : v > 1.0f ? 1.0f : v; }

// Java (int) cast of a float: truncation toward zero, saturating, NaN -> 0.
inline int32_t jf2i(float v) {
  if (v != v) return 0;
  if (v >= 2147483648.0f) return INT32_MAX;
  if (v <= -2147483648.0f) return INT32_MIN;
  return static_cast<int32_t>(v);
}

inline float jsqrt(float v) { return static_cast<float>(std::sqrt(static_cast<double>(v))); }

struct KeyHash {
  size_t operator()(const std::vector<uint32_t>& k) const {
    uint64_t h = 1469598103934665603ull;
    for (uint32_t x : k) {
      h ^= x;
      h *= 1099511628211ull;
    }
    return static_cast<size_t>(h);
  }
};

inline uint32_t fbits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

// Deduplicating pool of float tuples (Point2F / Point3F / Vector3F in the
// reference, deduplicated by equals() in encounter order,
// SceneCompiler.java:806-818).  Returns element offsets.
class TuplePool {
 public:
  explicit TuplePool(int arity) : arity_(arity) {}
  int add(const float* v) {
    std::vector<uint32_t> key(arity_);
    for (int i = 0; i < arity_; ++i) key[i] = fbits(v[i]);
    auto it = map_.find(key);
    if (it != map_.end()) return it->second;
    const int off = static_cast<int>(data_.size());
    data_.insert(data_.end(), v, v + arity_);
    map_.emplace(std::move(key), off);
    return off;
  }
  std::vector<float> take() {
    if (data_.empty()) return std::vector<float>(1, 0.0f);  // Arrays2 minimum length 1
    return std::move(data_);
  }

 private:
  int arity_;
  std::vector<float> data_;
  std::unordered_map<std::vector<uint32_t>, int, KeyHash> map_;
};

}  // namespace

int32_t pack_argb(const Color& c) {
  const int r = jf2i(fsat(c.r) * 255.0f + 0.5f);
  const int g = jf2i(fsat(c.g) * 255.0f + 0.5f);
  const int b = jf2i(fsat(c.b) * 255.0f + 0.5f);
  const int a = 255;
  const uint32_t u = (static_cast<uint32_t>(a & 0xFF) << 24) | (static_cast<uint32_t>(r & 0xFF) << 16) |
                     (static_cast<uint32_t>(g & 0xFF) << 8) | static_cast<uint32_t>(b & 0xFF);
  return static_cast<int32_t>(u);
}

// ---------------------------------------------------------------- textures

int add_texture(Scene& s, const Texture& t) {
  s.textures.push_back(t);
  return static_cast<int>(s.textures.size()) - 1;
}

int constant_texture(Scene& s, Color c) {
  Texture t;
  t.type = kTextureConstant;
  t.payload = {static_cast<float>(pack_argb(c))};
  return add_texture(s, t);
}

int checkerboard_texture(Scene& s, Color c0, Color c1, float radians, float scale_u, float scale_v) {
  // CheckerboardTexture.toArray (CheckerboardTexture.java:353-363)
  Texture t;
  t.type = kTextureCheckerboard;
  t.payload = {static_cast<float>(pack_argb(c0)), static_cast<float>(pack_argb(c1)),
               static_cast<float>(std::cos(static_cast<double>(radians))),
               static_cast<float>(std::sin(static_cast<double>(radians))), scale_u, scale_v};
  return add_texture(s, t);
}

int fbm_texture(Scene& s, Color addend, Color multiplier, float frequency, float gain, int octaves) {
  // FractionalBrownianMotionTexture.toArray (:208-218)
  Texture t;
  t.type = kTextureFBM;
  t.payload = {static_cast<float>(pack_argb(addend)), static_cast<float>(pack_argb(multiplier)), frequency, gain,
               static_cast<float>(octaves)};
  return add_texture(s, t);
}

int uv_texture(Scene& s) {
  Texture t;
  t.type = kTextureUV;
  return add_texture(s, t);
}

int surface_normal_texture(Scene& s, bool tangent_space) {
  Texture t;
  t.type = kTextureSurfaceNormal;
  t.payload = {tangent_space ? 1.0f : 0.0f};
  return add_texture(s, t);
}

int bullseye_texture(Scene& s, Color a, Color b) {
  Texture t;
  t.type = kTextureBullseye;
  t.payload = {static_cast<float>(pack_argb(a)), static_cast<float>(pack_argb(b))};
  return add_texture(s, t);
}

int blend_texture(Scene& s, int a, int b, float factor) {
  Texture t;
  t.type = kTextureBlend;
  t.payload = {0.0f, 0.0f, factor};
  t.blend_a = a;
  t.blend_b = b;
  return add_texture(s, t);
}

int image_texture(Scene& s, float radians, float scale_u, float scale_v, int w, int h,
                  const std::vector<int32_t>& argb) {
  // ImageTexture layout (ImageTexture.java RELATIVE_OFFSET_*): cos, sin, W, H,
  // scaleU, scaleV, then W*H packed ARGB ints stored as floats.
  Texture t;
  t.type = kTextureImage;
  t.payload = {static_cast<float>(std::cos(static_cast<double>(radians))),
               static_cast<float>(std::sin(static_cast<double>(radians))),
               static_cast<float>(w),
               static_cast<float>(h),
               scale_u,
               scale_v};
  for (int32_t v : argb) t.payload.push_back(static_cast<float>(v));
  return add_texture(s, t);
}

int add_surface(Scene& s, int material, int albedo, int emission, int normal, float noise_amount, float noise_scale) {
  Surface f;
  f.material = material;
  f.albedo = albedo;
  f.emission = emission;
  f.normal = normal;
  f.noise_amount = noise_amount;
  f.noise_scale = noise_scale;
  s.surfaces.push_back(f);
  return static_cast<int>(s.surfaces.size()) - 1;
}

// ---------------------------------------------------------------- shapes

Shape make_plane(Vec3 a, Vec3 b, Vec3 c) {
  // Plane(a, b, c): normal = normalize(cross(b - a, c - a)) (Plane.java:109-116,
  // Vector3F.normalNormalized).
  Shape s;
  s.type = kShapePlane;
  s.a = a;
  s.b = b;
  s.c = c;
  const float abx = b.x - a.x, aby = b.y - a.y, abz = b.z - a.z;
  const float acx = c.x - a.x, acy = c.y - a.y, acz = c.z - a.z;
  const float nx = aby * acz - abz * acy;
  const float ny = abz * acx - abx * acz;
  const float nz = abx * acy - aby * acx;
  const float len = jsqrt(nx * nx + ny * ny + nz * nz);
  s.plane_normal = Vec3{nx / len, ny / len, nz / len};
  return s;
}

Shape make_sphere(Vec3 position, float radius) {
  Shape s;
  s.type = kShapeSphere;
  s.position = position;
  s.radius = radius;
  return s;
}

Shape make_terrain(float frequency, float gain, float minimum, float maximum, int octaves) {
  Shape s;
  s.type = kShapeTerrain;
  s.frequency = frequency;
  s.gain = gain;
  s.minimum = minimum;
  s.maximum = maximum;
  s.octaves = octaves;
  return s;
}

Shape make_triangle(const Triangle& t) {
  Shape s;
  s.type = kShapeTriangle;
  s.triangles.push_back(t);
  return s;
}

Shape make_mesh(std::vector<Triangle> tris) {
  Shape s;
  s.type = kShapeTriangleMesh;
  s.triangles = std::move(tris);
  return s;
}

int add_primitive(Scene& s, Shape shape, int surface, Vec3 translation) {
  Primitive p;
  p.surface = surface;
  p.translation = translation;
  return add_primitive(s, std::move(shape), p);
}

int add_primitive(Scene& s, Shape shape, const Primitive& transform_and_surface) {
  s.shapes.push_back(std::move(shape));
  Primitive p = transform_and_surface;
  p.shape = static_cast<int>(s.shapes.size()) - 1;
  s.primitives.push_back(p);
  return static_cast<int>(s.primitives.size()) - 1;
}

// ---------------------------------------------------------------- transforms

namespace {

// Matrix44F.multiply (math4j Matrix44F.java): row-major, each element a
// left-to-right sum of four products.
void mat_mul(const float a[16], const float b[16], float out[16]) {
  float r[16];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      r[i * 4 + j] = a[i * 4 + 0] * b[0 * 4 + j] + a[i * 4 + 1] * b[1 * 4 + j] + a[i * 4 + 2] * b[2 * 4 + j] +
                     a[i * 4 + 3] * b[3 * 4 + j];
  std::memcpy(out, r, sizeof r);
}

// Matrix44F.inverse: cofactors from the 2x2 minors of rows 1-2 and 3-4.
bool mat_inverse(const float m[16], float out[16]) {
  auto e = [&](int r, int c) { return m[(r - 1) * 4 + (c - 1)]; };
  const float a = e(1, 1) * e(2, 2) - e(1, 2) * e(2, 1);
  const float b = e(1, 1) * e(2, 3) - e(1, 3) * e(2, 1);
  const float c = e(1, 1) * e(2, 4) - e(1, 4) * e(2, 1);
  const float d = e(1, 2) * e(2, 3) - e(1, 3) * e(2, 2);
  const float f0 = e(1, 2) * e(2, 4) - e(1, 4) * e(2, 2);
  const float f = e(1, 3) * e(2, 4) - e(1, 4) * e(2, 3);
  const float g = e(3, 1) * e(4, 2) - e(3, 2) * e(4, 1);
  const float h = e(3, 1) * e(4, 3) - e(3, 3) * e(4, 1);
  const float i = e(3, 1) * e(4, 4) - e(3, 4) * e(4, 1);
  const float j = e(3, 2) * e(4, 3) - e(3, 3) * e(4, 2);
  const float k = e(3, 2) * e(4, 4) - e(3, 4) * e(4, 2);
  const float l = e(3, 3) * e(4, 4) - e(3, 4) * e(4, 3);
  const float det = a * l - b * k + c * j + d * i - f0 * h + f * g;
  if (std::fabs(det) < 1.0e-12f) return false;
  const float r = 1.0f / det;
  const float o[16] = {
      (+e(2, 2) * l - e(2, 3) * k + e(2, 4) * j) * r, (-e(1, 2) * l + e(1, 3) * k - e(1, 4) * j) * r,
      (+e(4, 2) * f - e(4, 3) * f0 + e(4, 4) * d) * r, (-e(3, 2) * f + e(3, 3) * f0 - e(3, 4) * d) * r,
      (-e(2, 1) * l + e(2, 3) * i - e(2, 4) * h) * r, (+e(1, 1) * l - e(1, 3) * i + e(1, 4) * h) * r,
      (-e(4, 1) * f + e(4, 3) * c - e(4, 4) * b) * r, (+e(3, 1) * f - e(3, 3) * c + e(3, 4) * b) * r,
      (+e(2, 1) * k - e(2, 2) * i + e(2, 4) * g) * r, (-e(1, 1) * k + e(1, 2) * i - e(1, 4) * g) * r,
      (+e(4, 1) * f0 - e(4, 2) * c + e(4, 4) * a) * r, (-e(3, 1) * f0 + e(3, 2) * c - e(3, 4) * a) * r,
      (-e(2, 1) * j + e(2, 2) * h - e(2, 3) * g) * r, (+e(1, 1) * j - e(1, 2) * h + e(1, 3) * g) * r,
      (-e(4, 1) * d + e(4, 2) * b - e(4, 3) * a) * r, (+e(3, 1) * d - e(3, 2) * b + e(3, 3) * a) * r};
  std::memcpy(out, o, sizeof o);
  return true;
}

}  // namespace

bool transform_matrices(const Primitive& p, float o2w[16], float w2o[16]) {
  const float qx = p.rotation[0], qy = p.rotation[1], qz = p.rotation[2], qw = p.rotation[3];
  // Matrix44F.rotation(QuaternionF): rows u, v, w
  const float wv[3] = {2.0f * (qx * qz - qw * qy), 2.0f * (qy * qz + qw * qx), 1.0f - 2.0f * (qx * qx + qy * qy)};
  const float vv[3] = {2.0f * (qx * qy + qw * qz), 1.0f - 2.0f * (qx * qx + qz * qz), 2.0f * (qy * qz - qw * qx)};
  const float uv[3] = {1.0f - 2.0f * (qy * qy + qz * qz), 2.0f * (qx * qy - qw * qz), 2.0f * (qx * qz + qw * qy)};
  const float t[16] = {1, 0, 0, p.translation.x, 0, 1, 0, p.translation.y, 0, 0, 1, p.translation.z, 0, 0, 0, 1};
  const float rot[16] = {uv[0], uv[1], uv[2], 0, vv[0], vv[1], vv[2], 0, wv[0], wv[1], wv[2], 0, 0, 0, 0, 1};
  const float sc[16] = {p.scale.x, 0, 0, 0, 0, p.scale.y, 0, 0, 0, 0, p.scale.z, 0, 0, 0, 0, 1};
  float tr[16];
  mat_mul(t, rot, tr);
  mat_mul(tr, sc, o2w);
  return mat_inverse(o2w, w2o);
}

void transform_vertex(const float m[16], const float w[16], Vertex* v) {
  // Point3F.transformAndDivide
  const Vec3 q = v->position;
  const float x = m[0] * q.x + m[1] * q.y + m[2] * q.z + m[3];
  const float y = m[4] * q.x + m[5] * q.y + m[6] * q.z + m[7];
  const float z = m[8] * q.x + m[9] * q.y + m[10] * q.z + m[11];
  const float h = m[12] * q.x + m[13] * q.y + m[14] * q.z + m[15];
  // MathF.equals(w, 1): Float.compare, so NaN never equals and -0 != +0 (irrelevant at 1)
  v->position = h == 1.0f ? Vec3{x, y, z} : Vec3{x / h, y / h, z / h};
  // Vector3F.transformTranspose
  auto tt = [&](const Vec3& n) {
    return Vec3{w[0] * n.x + w[4] * n.y + w[8] * n.z, w[1] * n.x + w[5] * n.y + w[9] * n.z,
                w[2] * n.x + w[6] * n.y + w[10] * n.z};
  };
  v->normal = tt(v->normal);
  v->tangent = tt(v->tangent);
}

float to_radians(float degrees) { return static_cast<float>(static_cast<double>(degrees) / 180.0 * M_PI); }

// ---------------------------------------------------------------- camera

namespace {
constexpr int kCamAperture = 0, kCamEye = 1, kCamFovX = 4, kCamFovY = 5, kCamFocal = 6, kCamU = 7, kCamV = 10,
              kCamW = 13, kCamResX = 16, kCamResY = 17, kCamLens = 18;

inline float to_degrees(float rad) { return static_cast<float>(static_cast<double>(rad) * 180.0 / M_PI); }
}  // namespace

void camera_set_field_of_view_x(float cam[19], float degrees) {
  // Camera.setFieldOfViewX (Camera.java:743-748)
  cam[kCamFovX] = degrees;
  const float half = static_cast<float>(std::tan(static_cast<double>(to_radians(degrees) * 0.5f)));
  const float ratio = cam[kCamResY] / cam[kCamResX];
  const float at = static_cast<float>(std::atan(static_cast<double>(half * ratio)));
  cam[kCamFovY] = to_degrees(2.0f * at);
}

void camera_set_resolution(float cam[19], float w, float h) {
  cam[kCamResX] = w;
  cam[kCamResY] = h;
  camera_set_field_of_view_x(cam, cam[kCamFovX]);
}

void camera_set_eye(float cam[19], float x, float y, float z) {
  cam[kCamEye + 0] = x;
  cam[kCamEye + 1] = y;
  cam[kCamEye + 2] = z;
}

void camera_set_aperture(float cam[19], float radius) {
  cam[kCamAperture] = std::max(std::min(radius, 25.0f), 0.0f);  // Camera.java:694-697
}

void camera_set_focal_distance(float cam[19], float d) {
  cam[kCamFocal] = std::max(std::min(d, 100.0f), 0.2f);  // Camera.java:779-782
}

void camera_update(float cam[19], float yaw, float pitch) {
  // Camera.update (Camera.java:877-946): W = -(sin(yaw)cos(pitch), sin(pitch),
  // cos(yaw)cos(pitch)), U = normalize(up x W), V = normalize(W x U).
  auto fsin = [](float a) { return static_cast<float>(std::sin(static_cast<double>(a))); };
  auto fcos = [](float a) { return static_cast<float>(std::cos(static_cast<double>(a))); };
  const float w0x = -(fsin(yaw) * fcos(pitch));
  const float w0y = -fsin(pitch);
  const float w0z = -(fcos(yaw) * fcos(pitch));
  const float wr = 1.0f / jsqrt(w0x * w0x + w0y * w0y + w0z * w0z);
  const float wx = w0x * wr, wy = w0y * wr, wz = w0z * wr;
  cam[kCamW + 0] = wx;
  cam[kCamW + 1] = wy;
  cam[kCamW + 2] = wz;
  const float upx = 0.0f, upy = 1.0f, upz = 0.0f;
  const float u0x = upy * wz - upz * wy;
  const float u0y = upz * wx - upx * wz;
  const float u0z = upx * wy - upy * wx;
  const float ur = 1.0f / jsqrt(u0x * u0x + u0y * u0y + u0z * u0z);
  const float ux = u0x * ur, uy = u0y * ur, uz = u0z * ur;
  cam[kCamU + 0] = ux;
  cam[kCamU + 1] = uy;
  cam[kCamU + 2] = uz;
  const float v0x = wy * uz - wz * uy;
  const float v0y = wz * ux - wx * uz;
  const float v0z = wx * uy - wy * ux;
  const float vr = 1.0f / jsqrt(v0x * v0x + v0y * v0y + v0z * v0z);
  cam[kCamV + 0] = v0x * vr;
  cam[kCamV + 1] = v0y * vr;
  cam[kCamV + 2] = v0z * vr;
}

void camera_default(float cam[19]) {
  // Camera() defaults (Camera.java:239-254).
  std::memset(cam, 0, sizeof(float) * 19);
  camera_set_aperture(cam, 0.0f);
  cam[kCamLens] = 1.0f;  // CAMERA_LENS_THIN
  camera_set_eye(cam, 55.0f, 42.0f, 155.6f);
  cam[kCamResX] = 800.0f;
  cam[kCamResY] = 800.0f;
  camera_set_field_of_view_x(cam, 90.0f);
  camera_set_focal_distance(cam, 30.0f);
  camera_update(cam, 0.0f, 0.0f);
}

void camera_look_at(float cam[19], Vec3 eye, Vec3 target) {
  camera_set_eye(cam, eye.x, eye.y, eye.z);
  const double dx = static_cast<double>(target.x) - eye.x;
  const double dy = static_cast<double>(target.y) - eye.y;
  const double dz = static_cast<double>(target.z) - eye.z;
  const double horiz = std::sqrt(dx * dx + dz * dz);
  // The primary ray through the image centre is W (GPURendererKernel.java:659-674),
  // and W = -(sin(yaw)cos(pitch), sin(pitch), cos(yaw)cos(pitch)).
  const float yaw = static_cast<float>(std::atan2(-dx, -dz));
  const float pitch = static_cast<float>(std::atan2(-dy, horiz));
  camera_update(cam, yaw, pitch);
}

// ---------------------------------------------------------------- sky

void sky_set(df_sky* sky, Vec3 d, float turbidity) {
  // Sky.set(Vector3F, float) (Sky.java:295-333).  Only the quantities the
  // kernel reads are produced; the sun spectrum and sampling histograms are
  // not read by the render kernel.
  const float onb_u[3] = {1.0f, 0.0f, 0.0f};  // OrthoNormalBasis33F(w=y, v=z, u=x), Sky.java:61
  const float onb_v[3] = {0.0f, 0.0f, 1.0f};
  const float onb_w[3] = {0.0f, 1.0f, 0.0f};
  std::memcpy(sky->onb_u, onb_u, sizeof onb_u);
  std::memcpy(sky->onb_v, onb_v, sizeof onb_v);
  std::memcpy(sky->onb_w, onb_w, sizeof onb_w);

  const float len = jsqrt(d.x * d.x + d.y * d.y + d.z * d.z);
  const float wx = d.x / len, wy = d.y / len, wz = d.z / len;
  sky->sun_direction_world[0] = wx;
  sky->sun_direction_world[1] = wy;
  sky->sun_direction_world[2] = wz;

  // transformReverse(onb) = (d.u, d.v, d.w), then normalize.
  const float sx = wx * onb_u[0] + wy * onb_u[1] + wz * onb_u[2];
  const float sy = wx * onb_v[0] + wy * onb_v[1] + wz * onb_v[2];
  const float sz = wx * onb_w[0] + wy * onb_w[1] + wz * onb_w[2];
  const float sl = jsqrt(sx * sx + sy * sy + sz * sz);
  sky->sun_direction[0] = sx / sl;
  sky->sun_direction[1] = sy / sl;
  sky->sun_direction[2] = sz / sl;

  sky->sun_origin[0] = 0.0f + wx * 10000.0f;
  sky->sun_origin[1] = 0.0f + wy * 10000.0f;
  sky->sun_origin[2] = 0.0f + wz * 10000.0f;

  float z = sky->sun_direction[2];
  z = z < -1.0f ? -1.0f : z > 1.0f ? 1.0f : z;
  sky->theta = static_cast<float>(std::acos(static_cast<double>(z)));
  sky->turbidity = turbidity;

  const double theta = sky->theta;
  const double theta2 = theta * theta;
  const double theta3 = theta * theta * theta;
  const double t = turbidity;
  const double t2 = t * t;
  const double pi_f = static_cast<double>(static_cast<float>(M_PI));  // MathF.PI
  const double chi = (4.0 / 9.0 - t / 120.0) * (pi_f - 2.0 * theta);

  sky->zenith_relative_luminance = ((4.0453 * t - 4.9710) * std::tan(chi) - 0.2155 * t + 2.4192) * 1000.0;
  sky->zenith_x = (0.00165 * theta3 - 0.00374 * theta2 + 0.00208 * theta + 0.0) * t2 +
                  (-0.02902 * theta3 + 0.06377 * theta2 - 0.03202 * theta + 0.00394) * t +
                  (0.11693 * theta3 - 0.21196 * theta2 + 0.06052 * theta + 0.25885);
  sky->zenith_y = (0.00275 * theta3 - 0.00610 * theta2 + 0.00316 * theta + 0.0) * t2 +
                  (-0.04212 * theta3 + 0.08970 * theta2 - 0.04153 * theta + 0.00515) * t +
                  (0.15346 * theta3 - 0.26756 * theta2 + 0.06669 * theta + 0.26688);
  // Preetham et al. 1999 Perez coefficient fits as tabulated in Sky.java:320-334.
  sky->perez_relative_luminance[0] = 0.17872 * t - 1.46303;
  sky->perez_relative_luminance[1] = -0.35540 * t + 0.42749;
  sky->perez_relative_luminance[2] = -0.02266 * t + 5.32505;
  sky->perez_relative_luminance[3] = 0.12064 * t - 2.57705;
  sky->perez_relative_luminance[4] = -0.06696 * t + 0.37027;
  sky->perez_x[0] = -0.01925 * t - 0.25922;
  sky->perez_x[1] = -0.06651 * t + 0.00081;
  sky->perez_x[2] = -0.00041 * t + 0.21247;
  sky->perez_x[3] = -0.06409 * t - 0.89887;
  sky->perez_x[4] = -0.00325 * t + 0.04517;
  sky->perez_y[0] = -0.01669 * t - 0.26078;
  sky->perez_y[1] = -0.09495 * t + 0.00921;
  sky->perez_y[2] = -0.00792 * t + 0.21023;
  sky->perez_y[3] = -0.04405 * t - 1.65369;
  sky->perez_y[4] = -0.01092 * t + 0.05291;
  sky->is_sky_active = 1;
  sky->is_sun_active = 1;
}

// ---------------------------------------------------------------- BVH

namespace {

struct LeafItem {
  float mn[3], mx[3];
  float center[3];
  int tri;
};

struct Builder {
  const std::vector<LeafItem>* items;
  std::vector<BvhNode> nodes;

  // Restates doCreateBoundingVolumeHierarchy (BoundingVolumeHierarchy.java:805-986):
  // leaf when fewer than 4 triangles; otherwise sweep split planes
  // start+step, start+2step, ... (step = extent / (1024 / (depth + 1))) on each
  // axis, classify leaf centers with "center < split", skip candidates with at
  // most one item on a side, and keep the lowest SAH-style cost
  // |L|*area(L) + |R|*area(R) (half-areas).  The candidate sweep uses sorted
  // centers and prefix/suffix bounds, which selects the same split as the
  // reference's per-candidate rescan because min/max are exact.
  int build(std::vector<int>& idx, int depth, const float bmn[3], const float bmx[3]) {
    const int size0 = static_cast<int>(idx.size());
    const int node_index = static_cast<int>(nodes.size());
    nodes.emplace_back();
    nodes[node_index].depth = depth;
    std::copy(bmn, bmn + 3, nodes[node_index].minimum);
    std::copy(bmx, bmx + 3, nodes[node_index].maximum);
    auto make_leaf = [&]() {
      BvhNode& n = nodes[node_index];
      n.leaf = true;
      for (int i : idx) n.tris.push_back((*items)[i].tri);
    };
    if (size0 < 4) {
      make_leaf();
      return node_index;
    }
    const float side_x = bmx[0] - bmn[0];
    const float side_y = bmx[1] - bmn[1];
    const float side_z = bmx[2] - bmn[2];
    float minimum_cost = static_cast<float>(size0) * (side_x * side_y + side_y * side_z + side_z * side_x);
    float best_split = 3.402823466e38f;
    int best_axis = -1;

    std::vector<int> order(idx);
    std::vector<float> pre_mn, pre_mx, suf_mn, suf_mx;
    for (int axis = 0; axis < 3; ++axis) {
      const float start = bmn[axis];
      const float stop = bmx[axis];
      if (std::fabs(stop - start) < 1.0e-4f) continue;
      const float step = (stop - start) / (1024.0f / (static_cast<float>(depth) + 1.0f));
      // Sort by center along the axis (stable, so equal centers keep list order).
      std::stable_sort(order.begin(), order.end(),
                       [&](int a, int b) { return (*items)[a].center[axis] < (*items)[b].center[axis]; });
      const int n = size0;
      pre_mn.assign(static_cast<size_t>(n) * 3, 0.0f);
      pre_mx.assign(static_cast<size_t>(n) * 3, 0.0f);
      suf_mn.assign(static_cast<size_t>(n) * 3, 0.0f);
      suf_mx.assign(static_cast<size_t>(n) * 3, 0.0f);
      float amn[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
      float amx[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
      for (int i = 0; i < n; ++i) {
        const LeafItem& it = (*items)[order[i]];
        for (int k = 0; k < 3; ++k) {
          amn[k] = std::min(amn[k], it.mn[k]);
          amx[k] = std::max(amx[k], it.mx[k]);
          pre_mn[i * 3 + k] = amn[k];
          pre_mx[i * 3 + k] = amx[k];
        }
      }
      for (int k = 0; k < 3; ++k) {
        amn[k] = 3.402823466e38f;
        amx[k] = -3.402823466e38f;
      }
      for (int i = n - 1; i >= 0; --i) {
        const LeafItem& it = (*items)[order[i]];
        for (int k = 0; k < 3; ++k) {
          amn[k] = std::min(amn[k], it.mn[k]);
          amx[k] = std::max(amx[k], it.mx[k]);
          suf_mn[i * 3 + k] = amn[k];
          suf_mx[i * 3 + k] = amx[k];
        }
      }
      int cursor = 0;  // number of items with center < split (monotone in split)
      for (float split = start + step; split < stop - step; split += step) {
        // A step below half an ulp of the split would never advance; the
        // reference would spin forever there, this build stops the sweep.
        if (split + step == split) break;
        while (cursor < n && (*items)[order[cursor]].center[axis] < split) ++cursor;
        const int count_left = cursor;
        const int count_right = n - cursor;
        if (count_left <= 1 || count_right <= 1) continue;
        const float* lmn = &pre_mn[(count_left - 1) * 3];
        const float* lmx = &pre_mx[(count_left - 1) * 3];
        const float* rmn = &suf_mn[count_left * 3];
        const float* rmx = &suf_mx[count_left * 3];
        const float slx = lmx[0] - lmn[0], sly = lmx[1] - lmn[1], slz = lmx[2] - lmn[2];
        const float srx = rmx[0] - rmn[0], sry = rmx[1] - rmn[1], srz = rmx[2] - rmn[2];
        const float surface_left = slx * sly + sly * slz + slz * slx;
        const float surface_right = srx * sry + sry * srz + srz * srx;
        const float cost = surface_left * static_cast<float>(count_left) + surface_right * static_cast<float>(count_right);
        if (cost < minimum_cost) {
          minimum_cost = cost;
          best_split = split;
          best_axis = axis;
        }
      }
    }
    if (best_axis == -1) {
      make_leaf();
      return node_index;
    }
    std::vector<int> left, right;
    left.reserve(size0 / 2);
    right.reserve(size0 / 2);
    float lmn[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
    float lmx[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
    float rmn[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
    float rmx[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
    for (int i : idx) {
      const LeafItem& it = (*items)[i];
      if (it.center[best_axis] < best_split) {
        left.push_back(i);
        for (int k = 0; k < 3; ++k) {
          lmn[k] = std::min(lmn[k], it.mn[k]);
          lmx[k] = std::max(lmx[k], it.mx[k]);
        }
      } else {
        right.push_back(i);
        for (int k = 0; k < 3; ++k) {
          rmn[k] = std::min(rmn[k], it.mn[k]);
          rmx[k] = std::max(rmx[k], it.mx[k]);
        }
      }
    }
    idx.clear();
    idx.shrink_to_fit();
    const int l = build(left, depth + 1, lmn, lmx);
    const int r = build(right, depth + 1, rmn, rmx);
    nodes[node_index].left = l;
    nodes[node_index].right = r;
    return node_index;
  }
};

}  // namespace

std::vector<BvhNode> build_bvh(const std::vector<Triangle>& tris) {
  std::vector<LeafItem> items(tris.size());
  float mn[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
  float mx[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
  for (size_t i = 0; i < tris.size(); ++i) {
    const Vec3& p0 = tris[i].a.position;
    const Vec3& p1 = tris[i].b.position;
    const Vec3& p2 = tris[i].c.position;
    LeafItem& it = items[i];
    it.mx[0] = std::max(std::max(p0.x, p1.x), p2.x);
    it.mx[1] = std::max(std::max(p0.y, p1.y), p2.y);
    it.mx[2] = std::max(std::max(p0.z, p1.z), p2.z);
    it.mn[0] = std::min(std::min(p0.x, p1.x), p2.x);
    it.mn[1] = std::min(std::min(p0.y, p1.y), p2.y);
    it.mn[2] = std::min(std::min(p0.z, p1.z), p2.z);
    for (int k = 0; k < 3; ++k) {
      it.center[k] = (it.mx[k] + it.mn[k]) * 0.5f;
      mn[k] = std::min(mn[k], it.mn[k]);
      mx[k] = std::max(mx[k], it.mx[k]);
    }
    it.tri = static_cast<int>(i);
  }
  Builder b;
  b.items = &items;
  std::vector<int> idx(items.size());
  std::iota(idx.begin(), idx.end(), 0);
  b.build(idx, 0, mn, mx);
  return std::move(b.nodes);
}

// ---------------------------------------------------------------- compiler

namespace {

// objectToWorld / worldToObject rows of one primitive (SceneCompiler.java:695-747).
void push_matrices(std::vector<float>& o2w, std::vector<float>& w2o, const Primitive& p) {
  float a[16], b[16];
  if (!transform_matrices(p, a, b)) throw std::invalid_argument("a primitive's transform cannot be inverted");
  o2w.insert(o2w.end(), a, a + 16);
  w2o.insert(w2o.end(), b, b + 16);
}

// Pre-order list of node indices (Node.toList, BoundingVolumeHierarchy.java:622-637).
void preorder(const std::vector<BvhNode>& nodes, int i, std::vector<int>& out) {
  out.push_back(i);
  if (!nodes[i].leaf) {
    preorder(nodes, nodes[i].left, out);
    preorder(nodes, nodes[i].right, out);
  }
}

}  // namespace

Compiled compile(const Scene& s) {
  Compiled c;
  c.name = s.name;
  TuplePool p2(2), p3(3), v3(3);

  // Textures: records [type, size, payload] at cumulative offsets
  // (SceneCompiler.java:634-646); blend children rewritten to offsets.
  std::vector<int> tex_off(s.textures.size());
  {
    int off = 0;
    for (size_t i = 0; i < s.textures.size(); ++i) {
      tex_off[i] = off;
      off += 2 + static_cast<int>(s.textures[i].payload.size());
    }
    for (size_t i = 0; i < s.textures.size(); ++i) {
      const Texture& t = s.textures[i];
      c.textures.push_back(static_cast<float>(t.type));
      c.textures.push_back(static_cast<float>(2 + t.payload.size()));
      std::vector<float> pl = t.payload;
      if (t.type == kTextureBlend) {
        pl[0] = static_cast<float>(tex_off[t.blend_a]);
        pl[1] = static_cast<float>(tex_off[t.blend_b]);
      }
      c.textures.insert(c.textures.end(), pl.begin(), pl.end());
    }
    if (c.textures.empty()) c.textures.push_back(0.0f);
  }
  // Surfaces: [material, albedo offset, emission offset, normal offset, noise amount, noise scale]
  for (const Surface& f : s.surfaces) {
    c.surfaces.push_back(static_cast<float>(f.material));
    c.surfaces.push_back(static_cast<float>(tex_off[f.albedo]));
    c.surfaces.push_back(static_cast<float>(tex_off[f.emission]));
    c.surfaces.push_back(static_cast<float>(tex_off[f.normal]));
    c.surfaces.push_back(f.noise_amount);
    c.surfaces.push_back(f.noise_scale);
  }
  if (c.surfaces.empty()) c.surfaces.push_back(0.0f);

  auto add_triangle = [&](const Triangle& t) -> int {
    const int off = static_cast<int>(c.triangles.size());
    const Vertex* vs[3] = {&t.a, &t.b, &t.c};
    int32_t rec[12];
    for (int k = 0; k < 3; ++k) {
      const float pos[3] = {vs[k]->position.x, vs[k]->position.y, vs[k]->position.z};
      const float nrm[3] = {vs[k]->normal.x, vs[k]->normal.y, vs[k]->normal.z};
      const float uv[2] = {vs[k]->uv.x, vs[k]->uv.y};
      const float tan[3] = {vs[k]->tangent.x, vs[k]->tangent.y, vs[k]->tangent.z};
      rec[0 + k] = p3.add(pos);
      rec[3 + k] = v3.add(nrm);
      rec[6 + k] = p2.add(uv);
      rec[9 + k] = v3.add(tan);
    }
    c.triangles.insert(c.triangles.end(), rec, rec + 12);
    return off;
  };

  for (const Primitive& prim : s.primitives) {
    const Shape& sh = s.shapes[prim.shape];
    int shape_offset = 0;
    switch (sh.type) {
      case kShapePlane: {
        shape_offset = static_cast<int>(c.planes.size());
        const float a[3] = {sh.a.x, sh.a.y, sh.a.z};
        const float b[3] = {sh.b.x, sh.b.y, sh.b.z};
        const float cc[3] = {sh.c.x, sh.c.y, sh.c.z};
        const float n[3] = {sh.plane_normal.x, sh.plane_normal.y, sh.plane_normal.z};
        c.planes.push_back(p3.add(a));
        c.planes.push_back(p3.add(b));
        c.planes.push_back(p3.add(cc));
        c.planes.push_back(v3.add(n));
        break;
      }
      case kShapeSphere: {
        shape_offset = static_cast<int>(c.spheres.size());
        const float pos[3] = {sh.position.x, sh.position.y, sh.position.z};
        c.spheres.push_back(static_cast<float>(p3.add(pos)));
        c.spheres.push_back(sh.radius);
        break;
      }
      case kShapeTerrain: {
        shape_offset = static_cast<int>(c.terrains.size());
        c.terrains.push_back(sh.frequency);
        c.terrains.push_back(sh.gain);
        c.terrains.push_back(sh.minimum);
        c.terrains.push_back(sh.maximum);
        c.terrains.push_back(static_cast<float>(sh.octaves));
        break;
      }
      case kShapeTriangle: {
        shape_offset = add_triangle(sh.triangles[0]);
        break;
      }
      case kShapeTriangleMesh: {
        // Mesh triangles are emitted in mesh order, then the BVH segment
        // (SceneCompiler.java:892-995): tree = [1, next, min, max, left],
        // leaf = [2, next, min, max, n, tri offsets...], offsets relative to
        // the segment start, -1 = none.
        std::vector<int> tri_off(sh.triangles.size());
        for (size_t i = 0; i < sh.triangles.size(); ++i) tri_off[i] = add_triangle(sh.triangles[i]);
        const std::vector<BvhNode> nodes = build_bvh(sh.triangles);
        std::vector<int> order;
        preorder(nodes, 0, order);
        std::vector<int> offsets(order.size());
        int j = 0;
        for (size_t i = 0; i < order.size(); ++i) {
          offsets[i] = j;
          const BvhNode& n = nodes[order[i]];
          j += n.leaf ? 5 + static_cast<int>(n.tris.size()) : 5;
        }
        // next = first later node (pre-order) whose depth <= this depth; left =
        // the immediately following node (depth + 1).
        std::vector<int> next(order.size(), -1);
        {
          std::vector<std::pair<int, int>> stack;  // (depth, position) awaiting next
          for (size_t i = 0; i < order.size(); ++i) {
            const int d = nodes[order[i]].depth;
            while (!stack.empty() && stack.back().first >= d) {
              next[stack.back().second] = offsets[i];
              stack.pop_back();
            }
            stack.emplace_back(d, static_cast<int>(i));
          }
        }
        shape_offset = static_cast<int>(c.bvh.size());
        for (size_t i = 0; i < order.size(); ++i) {
          const BvhNode& n = nodes[order[i]];
          const int mn = p3.add(n.minimum);
          const int mx = p3.add(n.maximum);
          if (n.leaf) {
            c.bvh.push_back(2);
            c.bvh.push_back(next[i]);
            c.bvh.push_back(mn);
            c.bvh.push_back(mx);
            c.bvh.push_back(static_cast<int>(n.tris.size()));
            for (int t : n.tris) c.bvh.push_back(tri_off[t]);
          } else {
            c.bvh.push_back(1);
            c.bvh.push_back(next[i]);
            c.bvh.push_back(mn);
            c.bvh.push_back(mx);
            c.bvh.push_back(offsets[i + 1]);
          }
        }
        break;
      }
      default:
        break;
    }
    c.primitives.push_back(sh.type);
    c.primitives.push_back(shape_offset);
    c.primitives.push_back(prim.surface * 6);
    push_matrices(c.object_to_world, c.world_to_object, prim);
  }

  c.point2fs = p2.take();
  c.point3fs = p3.take();
  c.vector3fs = v3.take();
  if (c.planes.empty()) c.planes.push_back(0);
  if (c.spheres.empty()) c.spheres.push_back(0.0f);
  if (c.terrains.empty()) c.terrains.push_back(0.0f);
  if (c.triangles.empty()) c.triangles.push_back(0);
  if (c.bvh.empty()) c.bvh.push_back(0);
  c.camera.assign(s.camera, s.camera + 19);
  c.sky = s.sky;
  return c;
}

df_params default_params() {
  // AbstractRendererKernel.java:171-184 and AbstractKernel.java:196-205.
  df_params p{};
  p.renderer_type = DF_RENDERER_PATH_TRACER;
  p.tone_mapper_type = DF_TONE_MAPPER_FILMIC_CURVE_ACES_MODIFIED;
  p.tone_mapper_exposure = 1.0f;
  p.pt_depth_maximum = 5;
  p.pt_depth_russian_roulette = 5;
  p.ao_maximum_distance = 200.0f;
  p.normal_mapping = 1;
  p.wireframes = 0;
  p.shader_type = DF_SHADER_GOURAUD;
  p.selected_primitive_offset = -1;
  p.global_amplitude = 0.5f;
  p.global_frequency = 0.2f;
  p.global_lacunarity = 2.0f;
  p.global_gain = 1.0f / 2.0f;
  p.global_octaves = 2;
  p.math_mode = DF_MATH_FAST;
  return p;
}

}  // namespace dfs
