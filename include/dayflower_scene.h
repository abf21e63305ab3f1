/*
 * dayflower_scene.h - C ABI of the host-side scene compiler
 * (libdayflower_scene.so, plain C++ built with g++, no GPU needed).
 *
 * It produces the reference's flat CompiledScene arrays:
 *   - from the procedural benchmark workloads (dfs_build_config,
 *     dfs_build_batch_scene), which stand in for the reference's
 *     SceneLoader.loadScene + SceneCompiler.compile
 *     (scene/loader/SceneLoader.java:113-154, scene/compiler/SceneCompiler.java:79-188);
 *   - from a reference .scene file (dfs_read_scene; CompiledScene.read,
 *     scene/compiler/CompiledScene.java:323-377) and back (dfs_write_scene;
 *     CompiledScene.write :273-320);
 * plus the Camera and Sky host math that fills the camera array and the sun
 * and sky scalars (scene/Camera.java:743-748, 877-946; scene/Sky.java:295-374).
 */
#ifndef DAYFLOWER_SCENE_H
#define DAYFLOWER_SCENE_H

#include "dayflower_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dfs_scene dfs_scene;

/* Scene statistics (element counts of the compiled arrays). */
typedef struct dfs_stats {
  int32_t primitives;
  int32_t triangles;
  int32_t spheres;
  int32_t planes;
  int32_t terrains;
  int32_t bvh_nodes;
  int32_t textures_len;
} dfs_stats;

/* Builds benchmark workload `config_id` (1: 256x256 spheres + checkerboard
 * plane; 2: 1024x768 ~100k-triangle mesh; 3: 1920x1080 mesh + terrain + 50
 * spheres + emissive triangle; 5: 512x512 differentiation scene, 347,680
 * triangles in 64 mesh groups + 4 area lights).  seed 0 selects the
 * workload's documented seed.  Replaces SceneLoader.loadCompiledScene (SceneLoader.java:113-133). */
int dfs_build_config(int config_id, uint64_t seed, dfs_scene** out);

/* Builds one scene of the uneven batch mix from its seed (100..355). */
int dfs_build_batch_scene(uint64_t seed, dfs_scene** out);

/* Reads a .scene file (CompiledScene.read, CompiledScene.java:323-377); the
 * sky is the reference Sky() default (sun direction (1,1,1), turbidity 2). */
int dfs_read_scene(const char* path, dfs_scene** out);

/* Writes the scene's arrays as a .scene file (CompiledScene.write :273-320). */
int dfs_write_scene(const dfs_scene* scene, const char* path);

/* Copies externally supplied arrays into a new handle (e.g. to write them). */
int dfs_scene_from_arrays(const df_compiled_scene* arrays, const char* name, dfs_scene** out);

void dfs_free(dfs_scene* scene);

/* Borrowed views of the handle's arrays; valid until dfs_free. */
int dfs_get_compiled(const dfs_scene* scene, df_compiled_scene* out);
int dfs_get_sky(const dfs_scene* scene, df_sky* out);

/* Render configuration attached to a workload: width, height, samples per
 * pixel and renderer parameters (may be NULL individually). */
int dfs_get_render_config(const dfs_scene* scene, int32_t* width, int32_t* height, int32_t* spp, df_params* params);

int dfs_get_stats(const dfs_scene* scene, dfs_stats* out);

/* Writes the scene's name (NUL-terminated, truncated to cap) into buf. */
int dfs_get_name(const dfs_scene* scene, char* buf, size_t cap);

/* Sky.set(sunDirectionWorld, turbidity) (Sky.java:295-333). */
int dfs_sky_set(df_sky* out, float sun_x, float sun_y, float sun_z, float turbidity);

/* Camera host math: resolution/FOV (Camera.java:743-748, 806-818) and
 * Camera.update from a look-at (Camera.java:877-946).  cam has 19 floats. */
int dfs_camera_setup(float* cam, float width, float height, float fov_x_degrees, float eye_x, float eye_y,
                     float eye_z, float target_x, float target_y, float target_z, float aperture_radius,
                     float focal_distance);

/* Reference defaults of the renderer scalars
 * (AbstractRendererKernel.java:171-184, AbstractKernel.java:196-205). */
void dfs_default_params(df_params* out);

#ifdef __cplusplus
}
#endif

#endif /* DAYFLOWER_SCENE_H */
