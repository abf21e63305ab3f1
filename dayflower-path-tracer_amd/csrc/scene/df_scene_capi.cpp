// extern "C" surface of the scene compiler (include/dayflower_scene.h).
#include <algorithm>
#include <cstring>
#include <exception>
#include <new>

#include "../../../include/dayflower_scene.h"
#include "df_scene_model.h"

struct dfs_scene {
  dfs::Compiled c;
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
  if (!out || !(config_id >= 1 && config_id <= 3) && config_id != 5) return DF_EINVAL;
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

}  // extern "C"
