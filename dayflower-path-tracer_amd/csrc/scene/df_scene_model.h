// Host-side scene object model and compiler for the Dayflower flat layouts.
//
// This is the producer side of the drop-in boundary: it builds the same
// CompiledScene arrays the reference's SceneCompiler emits
// (scene/compiler/SceneCompiler.java:79-188) from a small object model that
// mirrors the reference's Primitive/Shape/Surface/Texture classes
// (scene/Primitive.java, scene/shape/*.java, scene/Surface.java,
// scene/texture/*.java).  Layout constants are those of the reference's
// TYPE/SIZE/RELATIVE_OFFSET_* fields.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/dayflower_types.h"

namespace dfs {

// ---- layout constants (reference scene/*.java TYPE / SIZE fields) ----
constexpr int kMaterialClearCoat = 1;   // material/ClearCoatMaterial.java:35
constexpr int kMaterialGlass = 2;       // material/GlassMaterial.java:35
constexpr int kMaterialLambertian = 3;  // material/LambertianMaterial.java:35
constexpr int kMaterialPhong = 4;       // material/PhongMaterial.java:35
constexpr int kMaterialReflection = 5;  // material/ReflectionMaterial.java:35

constexpr int kShapePlane = 1;         // shape/Plane.java (TYPE)
constexpr int kShapeSphere = 2;        // shape/Sphere.java
constexpr int kShapeTerrain = 3;       // shape/Terrain.java
constexpr int kShapeTriangle = 4;      // shape/Triangle.java
constexpr int kShapeTriangleMesh = 5;  // shape/TriangleMesh.java

constexpr int kTextureBlend = 1;
constexpr int kTextureBullseye = 2;
constexpr int kTextureCheckerboard = 3;
constexpr int kTextureConstant = 4;
constexpr int kTextureFBM = 5;
constexpr int kTextureImage = 6;
constexpr int kTextureSurfaceNormal = 7;
constexpr int kTextureUV = 8;

struct Color {
  float r = 0.0f, g = 0.0f, b = 0.0f;
};

// Color.pack() of image4j (Color.java:1493-1515,1709-1711): ARGB with
// alpha 255, each component saturate(c) * 255 + 0.5 truncated.
int32_t pack_argb(const Color& c);

struct Vec3 {
  float x = 0.0f, y = 0.0f, z = 0.0f;
};
struct Vec2 {
  float x = 0.0f, y = 0.0f;
};

struct Vertex {
  Vec3 position;
  Vec3 normal;
  Vec2 uv;
  Vec3 tangent;
};

struct Triangle {
  Vertex a, b, c;
};

// A texture record is [type, size, payload...] (scene/Texture.java:33-38).
// Blend textures reference two other textures by index into the scene's
// texture list; the compiler rewrites those to element offsets
// (SceneCompiler.java:779-790).
struct Texture {
  int type = kTextureConstant;
  std::vector<float> payload;  // excludes type and size
  int blend_a = -1, blend_b = -1;
};

struct Surface {
  int material = kMaterialLambertian;
  int albedo = -1, emission = -1, normal = -1;  // texture indices
  float noise_amount = 0.0f, noise_scale = 0.0f;
};

struct Shape {
  int type = kShapeSphere;
  // plane
  Vec3 a, b, c, plane_normal;
  // sphere
  Vec3 position;
  float radius = 1.0f;
  // terrain (stored but not read by the reference kernel's march)
  float frequency = 0.0f, gain = 0.0f, minimum = 0.0f, maximum = 0.0f;
  int octaves = 0;
  // triangle / mesh
  std::vector<Triangle> triangles;
};

struct Primitive {
  int shape = -1;    // index into Scene::shapes
  int surface = -1;  // index into Scene::surfaces
  // Transform (scene/Transform.java): position, rotation quaternion (x, y, z, w), scale
  Vec3 translation;
  float rotation[4] = {0.0f, 0.0f, 0.0f, 1.0f};
  Vec3 scale{1.0f, 1.0f, 1.0f};
};

struct Scene {
  std::string name;
  std::vector<Texture> textures;
  std::vector<Surface> surfaces;
  std::vector<Shape> shapes;
  std::vector<Primitive> primitives;
  float camera[19] = {0};
  df_sky sky{};
};

// Flat arrays in CompiledScene order.
struct Compiled {
  std::string name;
  std::vector<float> point2fs, point3fs, vector3fs;
  std::vector<int32_t> planes;
  std::vector<float> spheres, terrains;
  std::vector<int32_t> triangles, bvh;
  std::vector<float> textures, surfaces;
  std::vector<int32_t> primitives;
  std::vector<float> object_to_world, world_to_object;
  std::vector<float> camera;
  df_sky sky{};
  // render configuration attached by the generators
  int width = 0, height = 0, spp = 0;
  df_params params{};
};

// ---- texture helpers ----
int add_texture(Scene& s, const Texture& t);
int constant_texture(Scene& s, Color c);
int checkerboard_texture(Scene& s, Color c0, Color c1, float radians, float scale_u, float scale_v);
int fbm_texture(Scene& s, Color addend, Color multiplier, float frequency, float gain, int octaves);
int uv_texture(Scene& s);
int surface_normal_texture(Scene& s, bool tangent_space);
int bullseye_texture(Scene& s, Color a, Color b);
int blend_texture(Scene& s, int a, int b, float factor);
int image_texture(Scene& s, float radians, float scale_u, float scale_v, int w, int h, const std::vector<int32_t>& argb);
int add_surface(Scene& s, int material, int albedo, int emission, int normal, float noise_amount = 0.0f,
                float noise_scale = 0.0f);

// ---- shapes ----
Shape make_plane(Vec3 a, Vec3 b, Vec3 c);
Shape make_sphere(Vec3 position, float radius);
Shape make_terrain(float frequency, float gain, float minimum, float maximum, int octaves);
Shape make_triangle(const Triangle& t);
Shape make_mesh(std::vector<Triangle> tris);
int add_primitive(Scene& s, Shape shape, int surface, Vec3 translation = Vec3{});
// As above with the full Transform of the primitive (p.shape is set here).
int add_primitive(Scene& s, Shape shape, const Primitive& transform_and_surface);

// Transform.doCalculateObjectToWorldAndWorldToObject (Transform.java): object
// to world = translation * rotation(quaternion) * scaling, world to object its
// inverse, row-major, with math4j Matrix44F's float arithmetic.  Returns
// false where Matrix44F.inverse refuses (|determinant| < 1e-12).
bool transform_matrices(const Primitive& p, float o2w[16], float w2o[16]);
// Vertex.transformToWorldSpace (shape/Triangle.java:796-798): the position
// by o2w (divided by w unless w is exactly 1), normal and tangent by the
// transpose of w2o's upper 3x3.
void transform_vertex(const float o2w[16], const float w2o[16], Vertex* v);
// MathF.toRadians: (float) Math.toRadians(degrees)
float to_radians(float degrees);

// ---- Wavefront OBJ (scene/wavefront/ObjectLoader.java:118-212, 330-637) ----
struct ObjMesh {
  std::string group, material;
  std::vector<Triangle> triangles;
};
// One mesh per run of triangles in the same group, in file order.  Returns
// false with a message on I/O or syntax errors, out-of-range indices or a
// file without faces.
bool load_obj(const std::string& path, std::vector<ObjMesh>* out, std::string* error);

// ---- camera (scene/Camera.java:239-264, 743-748, 877-946) ----
void camera_default(float cam[19]);
void camera_set_resolution(float cam[19], float w, float h);
void camera_set_field_of_view_x(float cam[19], float degrees);
void camera_set_eye(float cam[19], float x, float y, float z);
void camera_set_aperture(float cam[19], float radius);
void camera_set_focal_distance(float cam[19], float d);
void camera_update(float cam[19], float yaw_radians, float pitch_radians);
void camera_look_at(float cam[19], Vec3 eye, Vec3 target);

// ---- sky (scene/Sky.java:295-374) ----
void sky_set(df_sky* sky, Vec3 sun_direction_world, float turbidity);

// ---- BVH (scene/bvh/BoundingVolumeHierarchy.java:118-157, 805-986) ----
struct BvhNode {
  bool leaf = false;
  int depth = 0;
  float minimum[3], maximum[3];
  int left = -1, right = -1;       // indices into the node vector
  std::vector<int> tris;           // triangle indices (into the mesh) for leaves
};
// Builds the tree; node 0 is the root.
std::vector<BvhNode> build_bvh(const std::vector<Triangle>& tris);

// ---- compiler ----
Compiled compile(const Scene& s);

// ---- .scene file (CompiledScene.java:273-377, Arrays2.java:56-273) ----
bool write_scene_file(const Compiled& c, const std::string& path);
bool read_scene_file(const std::string& path, Compiled* out);

// ---- workload generators ----
Compiled build_config(int config_id, uint64_t seed);
Compiled build_batch_scene(uint64_t seed);

df_params default_params();

}  // namespace dfs
