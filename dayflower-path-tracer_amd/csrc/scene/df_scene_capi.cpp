// extern "C" surface of the scene compiler (include/dayflower_scene.h).
#include <algorithm>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../../include/dayflower_scene.h"
#include "df_scene_model.h"

struct dfs_scene {
  dfs::Compiled c;
};

struct dfs_builder {
  dfs::Scene s;
  std::string error;
  int default_surface = -1;  // Surface(Lambertian, GRAY, BLACK, BLACK), made on first use
};

namespace {

template <typename T>
std::vector<T> copy_array(const T* p, size_t n) {
  if (!p || n == 0) return std::vector<T>();
  return std::vector<T>(p, p + n);
}

int wrap(dfs::Compiled&& c, dfs_scene** out) {
  dfs_scene* s = new (std::nothrow) dfs_scene;
  if (!s) return DF_ENOMEM;
  s->c = std::move(c);
  *out = s;
  return DF_OK;
}

}  // namespace

extern "C" {

int dfs_build_config(int config_id, uint64_t seed, dfs_scene** out) {
  if (!out || (!(config_id >= 1 && config_id <= 3) && config_id != 5)) return DF_EINVAL;
  try {
    return wrap(dfs::build_config(config_id, seed), out);
  } catch (const std::bad_alloc&) {
    return DF_ENOMEM;
  } catch (...) {
    return DF_EINVAL;
  }
}

int dfs_build_batch_scene(uint64_t seed, dfs_scene** out) {
  if (!out) return DF_EINVAL;
  try {
    return wrap(dfs::build_batch_scene(seed), out);
  } catch (const std::bad_alloc&) {
    return DF_ENOMEM;
  } catch (...) {
    return DF_EINVAL;
  }
}

int dfs_read_scene(const char* path, dfs_scene** out) {
  if (!path || !out) return DF_EINVAL;
  try {
    dfs::Compiled c;
    if (!dfs::read_scene_file(path, &c)) return DF_EIO;
    return wrap(std::move(c), out);
  } catch (const std::bad_alloc&) {
    return DF_ENOMEM;
  }
}

int dfs_write_scene(const dfs_scene* scene, const char* path) {
  if (!scene || !path) return DF_EINVAL;
  return dfs::write_scene_file(scene->c, path) ? DF_OK : DF_EIO;
}

int dfs_scene_from_arrays(const df_compiled_scene* a, const char* name, dfs_scene** out) {
  if (!a || !out || !a->camera || a->camera_len != 19) return DF_EINVAL;
  try {
    dfs::Compiled c;
    c.name = name ? name : "";
    c.point2fs = copy_array(a->point2fs, a->point2fs_len);
    c.point3fs = copy_array(a->point3fs, a->point3fs_len);
    c.vector3fs = copy_array(a->vector3fs, a->vector3fs_len);
    c.planes = copy_array(a->planes, a->planes_len);
    c.spheres = copy_array(a->spheres, a->spheres_len);
    c.terrains = copy_array(a->terrains, a->terrains_len);
    c.triangles = copy_array(a->triangles, a->triangles_len);
    c.bvh = copy_array(a->bvh, a->bvh_len);
    c.textures = copy_array(a->textures, a->textures_len);
    c.surfaces = copy_array(a->surfaces, a->surfaces_len);
    c.primitives = copy_array(a->primitives, a->primitives_len);
    c.object_to_world = copy_array(a->object_to_world, a->object_to_world_len);
    c.world_to_object = copy_array(a->world_to_object, a->world_to_object_len);
    c.camera = copy_array(a->camera, a->camera_len);
    c.params = dfs::default_params();
    c.width = static_cast<int>(c.camera[16]);
    c.height = static_cast<int>(c.camera[17]);
    c.spp = 1;
    dfs::sky_set(&c.sky, dfs::Vec3{1.0f, 1.0f, 1.0f}, 2.0f);
    return wrap(std::move(c), out);
  } catch (const std::bad_alloc&) {
    return DF_ENOMEM;
  }
}

void dfs_free(dfs_scene* scene) { delete scene; }

int dfs_get_compiled(const dfs_scene* s, df_compiled_scene* o) {
  if (!s || !o) return DF_EINVAL;
  const dfs::Compiled& c = s->c;
  o->point2fs = c.point2fs.data();
  o->point2fs_len = c.point2fs.size();
  o->point3fs = c.point3fs.data();
  o->point3fs_len = c.point3fs.size();
  o->vector3fs = c.vector3fs.data();
  o->vector3fs_len = c.vector3fs.size();
  o->planes = c.planes.data();
  o->planes_len = c.planes.size();
  o->spheres = c.spheres.data();
  o->spheres_len = c.spheres.size();
  o->terrains = c.terrains.data();
  o->terrains_len = c.terrains.size();
  o->triangles = c.triangles.data();
  o->triangles_len = c.triangles.size();
  o->bvh = c.bvh.data();
  o->bvh_len = c.bvh.size();
  o->textures = c.textures.data();
  o->textures_len = c.textures.size();
  o->surfaces = c.surfaces.data();
  o->surfaces_len = c.surfaces.size();
  o->primitives = c.primitives.data();
  o->primitives_len = c.primitives.size();
  o->object_to_world = c.object_to_world.data();
  o->object_to_world_len = c.object_to_world.size();
  o->world_to_object = c.world_to_object.data();
  o->world_to_object_len = c.world_to_object.size();
  o->camera = c.camera.data();
  o->camera_len = c.camera.size();
  return DF_OK;
}

int dfs_get_sky(const dfs_scene* s, df_sky* o) {
  if (!s || !o) return DF_EINVAL;
  *o = s->c.sky;
  return DF_OK;
}

int dfs_get_render_config(const dfs_scene* s, int32_t* w, int32_t* h, int32_t* spp, df_params* p) {
  if (!s) return DF_EINVAL;
  if (w) *w = s->c.width;
  if (h) *h = s->c.height;
  if (spp) *spp = s->c.spp;
  if (p) *p = s->c.params;
  return DF_OK;
}

int dfs_get_stats(const dfs_scene* s, dfs_stats* o) {
  if (!s || !o) return DF_EINVAL;
  const dfs::Compiled& c = s->c;
  std::memset(o, 0, sizeof *o);
  o->primitives = static_cast<int32_t>(c.primitives.size() / 3);
  for (size_t i = 0; i + 2 < c.primitives.size(); i += 3) {
    switch (c.primitives[i]) {
      case dfs::kShapePlane:
        o->planes++;
        break;
      case dfs::kShapeSphere:
        o->spheres++;
        break;
      case dfs::kShapeTerrain:
        o->terrains++;
        break;
      default:
        break;
    }
  }
  o->triangles = c.triangles.size() >= 12 ? static_cast<int32_t>(c.triangles.size() / 12) : 0;
  for (size_t i = 0; i < c.bvh.size();) {
    const int type = c.bvh[i];
    if (type == 1) {
      o->bvh_nodes++;
      i += 5;
    } else if (type == 2 && i + 4 < c.bvh.size()) {
      o->bvh_nodes++;
      i += 5 + static_cast<size_t>(c.bvh[i + 4]);
    } else {
      break;
    }
  }
  o->textures_len = static_cast<int32_t>(c.textures.size());
  return DF_OK;
}

int dfs_get_name(const dfs_scene* s, char* buf, size_t cap) {
  if (!s || !buf || cap == 0) return DF_EINVAL;
  const size_t n = std::min(cap - 1, s->c.name.size());
  std::memcpy(buf, s->c.name.data(), n);
  buf[n] = '\0';
  return DF_OK;
}

int dfs_sky_set(df_sky* out, float x, float y, float z, float turbidity) {
  if (!out) return DF_EINVAL;
  if (x == 0.0f && y == 0.0f && z == 0.0f) return DF_EINVAL;
  dfs::sky_set(out, dfs::Vec3{x, y, z}, turbidity);
  return DF_OK;
}

int dfs_camera_setup(float* cam, float w, float h, float fov_x, float ex, float ey, float ez, float tx, float ty,
                     float tz, float aperture, float focal) {
  if (!cam || w <= 0.0f || h <= 0.0f) return DF_EINVAL;
  dfs::camera_default(cam);
  dfs::camera_set_resolution(cam, w, h);
  dfs::camera_set_field_of_view_x(cam, fov_x);
  dfs::camera_set_aperture(cam, aperture);
  dfs::camera_set_focal_distance(cam, focal);
  dfs::camera_look_at(cam, dfs::Vec3{ex, ey, ez}, dfs::Vec3{tx, ty, tz});
  return DF_OK;
}

void dfs_default_params(df_params* out) {
  if (out) *out = dfs::default_params();
}

int dfs_camera_default(float* cam) {
  if (!cam) return DF_EINVAL;
  dfs::camera_default(cam);
  return DF_OK;
}

int dfs_camera_set_eye(float* cam, float x, float y, float z) {
  if (!cam) return DF_EINVAL;
  dfs::camera_set_eye(cam, x, y, z);
  return DF_OK;
}

int dfs_camera_set_resolution(float* cam, float w, float h) {
  if (!cam || !(w > 0.0f) || !(h > 0.0f)) return DF_EINVAL;
  dfs::camera_set_resolution(cam, w, h);
  return DF_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- builder

namespace {

int bfail(dfs_builder* b, int code, const std::string& why) {
  b->error = why;
  return code;
}

dfs::Color rgb3(const float* c) { return dfs::Color{c[0], c[1], c[2]}; }
dfs::Vec3 vec3(const float* v) { return dfs::Vec3{v[0], v[1], v[2]}; }

bool texture_ok(const dfs_builder* b, int32_t i) { return i >= 0 && i < static_cast<int32_t>(b->s.textures.size()); }
bool surface_ok(const dfs_builder* b, int32_t i) { return i >= 0 && i < static_cast<int32_t>(b->s.surfaces.size()); }

int texture_out(dfs_builder* b, int index, int32_t* out) {
  if (out) *out = index;
  return DF_OK;
}

// The primitive's transform and surface, checked for an invertible matrix.
int placement(dfs_builder* b, int32_t surface, const dfs_transform* t, dfs::Primitive* p) {
  if (!surface_ok(b, surface)) return bfail(b, DF_EINVAL, "surface index out of range");
  p->surface = surface;
  if (t) {
    p->translation = vec3(t->position);
    for (int i = 0; i < 4; ++i) p->rotation[i] = t->rotation[i];
    p->scale = vec3(t->scale);
  }
  float o2w[16], w2o[16];
  if (!dfs::transform_matrices(*p, o2w, w2o)) return bfail(b, DF_EINVAL, "transform cannot be inverted");
  return DF_OK;
}

dfs::Triangle triangle_from(const float* v) {
  dfs::Triangle t;
  dfs::Vertex* vs[3] = {&t.a, &t.b, &t.c};
  for (int k = 0; k < 3; ++k, v += 11) {
    vs[k]->position = vec3(v);
    vs[k]->normal = vec3(v + 3);
    vs[k]->uv = dfs::Vec2{v[6], v[7]};
    vs[k]->tangent = vec3(v + 8);
  }
  return t;
}

template <typename F>
int guarded(dfs_builder* b, F&& f) {
  if (!b) return DF_EINVAL;
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return bfail(b, DF_ENOMEM, "out of memory");
  } catch (const std::exception& e) {
    return bfail(b, DF_EINVAL, e.what());
  }
}

}  // namespace

extern "C" {

int dfs_builder_create(const char* name, dfs_builder** out) {
  if (!out) return DF_EINVAL;
  dfs_builder* b = new (std::nothrow) dfs_builder;
  if (!b) return DF_ENOMEM;
  b->s.name = name ? name : "";
  dfs::camera_default(b->s.camera);
  dfs::sky_set(&b->s.sky, dfs::Vec3{1.0f, 1.0f, 1.0f}, 2.0f);  // Sky(): set(1, 1, 1) (Sky.java:86-88)
  *out = b;
  return DF_OK;
}

void dfs_builder_free(dfs_builder* b) { delete b; }

int dfs_builder_last_error(const dfs_builder* b, char* buf, size_t cap) {
  if (!b || !buf || cap == 0) return DF_EINVAL;
  const size_t n = std::min(cap - 1, b->error.size());
  std::memcpy(buf, b->error.data(), n);
  buf[n] = '\0';
  return DF_OK;
}

int dfs_builder_constant_texture(dfs_builder* b, const float rgb[3], int32_t* index) {
  return guarded(b, [&]() -> int {
    if (!rgb) return bfail(b, DF_EINVAL, "colour is NULL");
    return texture_out(b, dfs::constant_texture(b->s, rgb3(rgb)), index);
  });
}

int dfs_builder_checkerboard_texture(dfs_builder* b, const float rgb0[3], const float rgb1[3], float scale_u,
                                     float scale_v, float degrees, int32_t* index) {
  return guarded(b, [&]() -> int {
    if (!rgb0 || !rgb1) return bfail(b, DF_EINVAL, "colour is NULL");
    return texture_out(
        b, dfs::checkerboard_texture(b->s, rgb3(rgb0), rgb3(rgb1), dfs::to_radians(degrees), scale_u, scale_v), index);
  });
}

int dfs_builder_fbm_texture(dfs_builder* b, const float addend[3], const float multiplier[3], float frequency,
                            float gain, int32_t octaves, int32_t* index) {
  return guarded(b, [&]() -> int {
    if (!addend || !multiplier) return bfail(b, DF_EINVAL, "colour is NULL");
    return texture_out(b, dfs::fbm_texture(b->s, rgb3(addend), rgb3(multiplier), frequency, gain, octaves), index);
  });
}

int dfs_builder_blend_texture(dfs_builder* b, int32_t ta, int32_t tb, float factor, int32_t* index) {
  return guarded(b, [&]() -> int {
    if (!texture_ok(b, ta) || !texture_ok(b, tb)) return bfail(b, DF_EINVAL, "texture index out of range");
    return texture_out(b, dfs::blend_texture(b->s, ta, tb, factor), index);
  });
}

int dfs_builder_bullseye_texture(dfs_builder* b, const float rgb0[3], const float rgb1[3], int32_t* index) {
  return guarded(b, [&]() -> int {
    if (!rgb0 || !rgb1) return bfail(b, DF_EINVAL, "colour is NULL");
    return texture_out(b, dfs::bullseye_texture(b->s, rgb3(rgb0), rgb3(rgb1)), index);
  });
}

int dfs_builder_uv_texture(dfs_builder* b, int32_t* index) {
  return guarded(b, [&]() -> int { return texture_out(b, dfs::uv_texture(b->s), index); });
}

int dfs_builder_surface_normal_texture(dfs_builder* b, int32_t tangent_space, int32_t* index) {
  return guarded(b, [&]() -> int { return texture_out(b, dfs::surface_normal_texture(b->s, tangent_space != 0), index); });
}

int dfs_builder_image_texture(dfs_builder* b, const int32_t* rgb, int32_t w, int32_t h, float degrees, float scale_u,
                              float scale_v, int32_t* index) {
  return guarded(b, [&]() -> int {
    if (!rgb || w <= 0 || h <= 0) return bfail(b, DF_EINVAL, "image is empty");
    if (static_cast<int64_t>(w) * h > (int64_t{1} << 26)) return bfail(b, DF_EINVAL, "image larger than 2^26 pixels");
    const std::vector<int32_t> data(rgb, rgb + static_cast<size_t>(w) * h);
    return texture_out(b, dfs::image_texture(b->s, dfs::to_radians(degrees), scale_u, scale_v, w, h, data), index);
  });
}

int dfs_builder_surface(dfs_builder* b, int32_t material, int32_t albedo, int32_t emission, int32_t normal,
                        float noise_amount, float noise_scale, int32_t* index) {
  return guarded(b, [&]() -> int {
    if (material < dfs::kMaterialClearCoat || material > dfs::kMaterialReflection)
      return bfail(b, DF_EINVAL, "material must be 1..5");
    if (!texture_ok(b, albedo) || !texture_ok(b, emission) || !texture_ok(b, normal))
      return bfail(b, DF_EINVAL, "texture index out of range");
    const int i = dfs::add_surface(b->s, material, albedo, emission, normal, noise_amount, noise_scale);
    if (index) *index = i;
    return DF_OK;
  });
}

int dfs_builder_plane(dfs_builder* b, const float a[3], const float b3[3], const float c[3], int32_t surface,
                      const dfs_transform* t) {
  return guarded(b, [&]() -> int {
    if (!a || !b3 || !c) return bfail(b, DF_EINVAL, "point is NULL");
    dfs::Primitive p;
    if (const int rc = placement(b, surface, t, &p)) return rc;
    dfs::add_primitive(b->s, dfs::make_plane(vec3(a), vec3(b3), vec3(c)), p);
    return DF_OK;
  });
}

int dfs_builder_sphere(dfs_builder* b, const float center[3], float radius, int32_t surface, const dfs_transform* t) {
  return guarded(b, [&]() -> int {
    if (!center) return bfail(b, DF_EINVAL, "centre is NULL");
    dfs::Primitive p;
    if (const int rc = placement(b, surface, t, &p)) return rc;
    dfs::add_primitive(b->s, dfs::make_sphere(vec3(center), radius), p);
    return DF_OK;
  });
}

int dfs_builder_terrain(dfs_builder* b, float frequency, float gain, float minimum, float maximum, int32_t octaves,
                        int32_t surface, const dfs_transform* t) {
  return guarded(b, [&]() -> int {
    dfs::Primitive p;
    if (const int rc = placement(b, surface, t, &p)) return rc;
    dfs::add_primitive(b->s, dfs::make_terrain(frequency, gain, minimum, maximum, octaves), p);
    return DF_OK;
  });
}

int dfs_builder_triangle(dfs_builder* b, const float v[33], int32_t surface, const dfs_transform* t) {
  return guarded(b, [&]() -> int {
    if (!v) return bfail(b, DF_EINVAL, "vertices are NULL");
    dfs::Primitive p;
    if (const int rc = placement(b, surface, t, &p)) return rc;
    dfs::add_primitive(b->s, dfs::make_triangle(triangle_from(v)), p);
    return DF_OK;
  });
}

int dfs_builder_mesh(dfs_builder* b, const float* v, int64_t n, int32_t surface, const dfs_transform* t) {
  return guarded(b, [&]() -> int {
    if (!v || n <= 0) return bfail(b, DF_EINVAL, "a mesh needs at least one triangle");
    dfs::Primitive p;
    if (const int rc = placement(b, surface, t, &p)) return rc;
    std::vector<dfs::Triangle> tris;
    tris.reserve(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) tris.push_back(triangle_from(v + i * 33));
    dfs::add_primitive(b->s, dfs::make_mesh(std::move(tris)), p);
    return DF_OK;
  });
}

int dfs_builder_load_obj(dfs_builder* b, const char* path, dfs_surface_mapper mapper, void* user, int32_t surface,
                         const dfs_transform* t, int32_t* primitives) {
  return guarded(b, [&]() -> int {
    if (!path) return bfail(b, DF_EINVAL, "path is NULL");
    std::vector<dfs::ObjMesh> meshes;
    std::string err;
    if (!dfs::load_obj(path, &meshes, &err)) return bfail(b, DF_EIO, err);
    if (!mapper && surface < 0) {
      if (b->default_surface < 0) {  // ObjectLoader.load(file): Surface(Lambertian, GRAY, BLACK, BLACK)
        const int gray = dfs::constant_texture(b->s, dfs::Color{0.5f, 0.5f, 0.5f});
        const int black0 = dfs::constant_texture(b->s, dfs::Color{});
        const int black1 = dfs::constant_texture(b->s, dfs::Color{});
        b->default_surface = dfs::add_surface(b->s, dfs::kMaterialLambertian, gray, black0, black1);
      }
      surface = b->default_surface;
    }
    std::vector<int32_t> surfaces;
    for (const dfs::ObjMesh& m : meshes) {
      const int32_t sidx = mapper ? mapper(m.group.c_str(), m.material.c_str(), user) : surface;
      if (!surface_ok(b, sidx))
        return bfail(b, DF_EINVAL, "no valid surface for group '" + m.group + "', material '" + m.material + "'");
      surfaces.push_back(sidx);
    }
    for (size_t i = 0; i < meshes.size(); ++i) {
      dfs::Primitive p;
      if (const int rc = placement(b, surfaces[i], t, &p)) return rc;
      dfs::add_primitive(b->s, dfs::make_mesh(std::move(meshes[i].triangles)), p);
    }
    if (primitives) *primitives = static_cast<int32_t>(meshes.size());
    return DF_OK;
  });
}

int dfs_builder_set_camera(dfs_builder* b, const float cam[19]) {
  return guarded(b, [&]() -> int {
    if (!cam) return bfail(b, DF_EINVAL, "camera is NULL");
    std::memcpy(b->s.camera, cam, sizeof(float) * 19);
    return DF_OK;
  });
}

int dfs_builder_set_sky(dfs_builder* b, const df_sky* sky) {
  return guarded(b, [&]() -> int {
    if (!sky) return bfail(b, DF_EINVAL, "sky is NULL");
    b->s.sky = *sky;
    return DF_OK;
  });
}

int dfs_builder_compile(dfs_builder* b, dfs_scene** out) {
  return guarded(b, [&]() -> int {
    if (!out) return bfail(b, DF_EINVAL, "out is NULL");
    if (b->s.primitives.empty()) return bfail(b, DF_EINVAL, "the scene has no primitives");
    dfs::Compiled c = dfs::compile(b->s);
    c.params = dfs::default_params();
    c.width = static_cast<int>(c.camera[16]);
    c.height = static_cast<int>(c.camera[17]);
    c.spp = 1;
    return wrap(std::move(c), out);
  });
}

}  // extern "C"
