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

/* Camera() defaults (Camera.java:239-254), Camera.setEye and
 * Camera.setResolution on a 19-float camera array. */
int dfs_camera_default(float* cam);
int dfs_camera_set_eye(float* cam, float x, float y, float z);
int dfs_camera_set_resolution(float* cam, float width, float height);

/* ---- Scene construction and asset ingest ----
 * Scenes built the way the reference builds them in Java: Scene,
 * Primitive(shape, surface, transform), Surface and the texture classes
 * (scene/Scene.java, scene/Primitive.java, scene/Surface.java, the classes
 * of scene/texture and scene/shape), Wavefront OBJ models
 * (ObjectLoader.load, scene/wavefront/ObjectLoader.java:118-212) and image
 * textures (ImageTexture, scene/texture/ImageTexture.java:337-356), then
 * SceneCompiler.compile (scene/compiler/SceneCompiler.java:79-188).  Every
 * call returns DF_OK or an error code; dfs_builder_last_error says why. */
typedef struct dfs_builder dfs_builder;

/* Transform(position, rotation, scale) (scene/Transform.java): rotation is a
 * quaternion (x, y, z, w).  A NULL transform is Transform(): the identity. */
typedef struct dfs_transform {
  float position[3];
  float rotation[4];
  float scale[3];
} dfs_transform;

/* new Scene(name): Camera() and Sky() defaults, no primitives. */
int dfs_builder_create(const char* name, dfs_builder** out);
void dfs_builder_free(dfs_builder* b);
int dfs_builder_last_error(const dfs_builder* b, char* buf, size_t cap);

/* Textures: *index receives the texture's index for surfaces and blends.
 * Colours are linear RGB triples, packed as image4j Color.pack. */
int dfs_builder_constant_texture(dfs_builder* b, const float rgb[3], int32_t* index);
/* CheckerboardTexture(color0, color1, scaleU, scaleV, degrees) */
int dfs_builder_checkerboard_texture(dfs_builder* b, const float rgb0[3], const float rgb1[3], float scale_u,
                                     float scale_v, float degrees, int32_t* index);
/* FractionalBrownianMotionTexture(addend, multiplier, frequency, gain, octaves) */
int dfs_builder_fbm_texture(dfs_builder* b, const float addend[3], const float multiplier[3], float frequency,
                            float gain, int32_t octaves, int32_t* index);
/* BlendTexture(textureA, textureB, factor) */
int dfs_builder_blend_texture(dfs_builder* b, int32_t texture_a, int32_t texture_b, float factor, int32_t* index);
/* BullseyeTexture(colorA, colorB) */
int dfs_builder_bullseye_texture(dfs_builder* b, const float rgb0[3], const float rgb1[3], int32_t* index);
int dfs_builder_uv_texture(dfs_builder* b, int32_t* index);
int dfs_builder_surface_normal_texture(dfs_builder* b, int32_t tangent_space, int32_t* index);
/* ImageTexture(degrees, width, height, scaleU, scaleV, data): data is the
 * decoded image as the reference's TYPE_INT_RGB raster, width x height packed
 * 0x00RRGGBB ints, rows from the top (ImageTexture.load, :470-548). */
int dfs_builder_image_texture(dfs_builder* b, const int32_t* rgb, int32_t width, int32_t height, float degrees,
                              float scale_u, float scale_v, int32_t* index);

/* Surface(material DF material id 1..5, albedo, emission, normal textures,
 * noiseAmount, noiseScale). */
int dfs_builder_surface(dfs_builder* b, int32_t material, int32_t albedo, int32_t emission, int32_t normal,
                        float noise_amount, float noise_scale, int32_t* index);

/* Primitives: a shape with a surface and a transform. */
int dfs_builder_plane(dfs_builder* b, const float a[3], const float b3[3], const float c[3], int32_t surface,
                      const dfs_transform* t);
int dfs_builder_sphere(dfs_builder* b, const float center[3], float radius, int32_t surface, const dfs_transform* t);
int dfs_builder_terrain(dfs_builder* b, float frequency, float gain, float minimum, float maximum, int32_t octaves,
                        int32_t surface, const dfs_transform* t);
/* Triangle(a, b, c) or TriangleMesh(triangles): 33 floats per triangle, per
 * vertex position[3], normal[3], texture coordinates[2], tangent[3]. */
int dfs_builder_triangle(dfs_builder* b, const float vertices[33], int32_t surface, const dfs_transform* t);
int dfs_builder_mesh(dfs_builder* b, const float* vertices, int64_t triangles, int32_t surface,
                     const dfs_transform* t);

/* ObjectLoader.load(file, surfaceMapper, transform): one TriangleMesh
 * primitive per run of triangles in one group, in file order.  The mapper
 * receives (group, material) names and returns a surface index; with no
 * mapper every mesh takes `surface`, and with surface < 0 as well the
 * reference's default Surface(Lambertian, GRAY, BLACK, BLACK).  *primitives
 * (may be NULL) receives the number of primitives added. */
typedef int32_t (*dfs_surface_mapper)(const char* group, const char* material, void* user);
int dfs_builder_load_obj(dfs_builder* b, const char* path, dfs_surface_mapper mapper, void* user, int32_t surface,
                         const dfs_transform* t, int32_t* primitives);

int dfs_builder_set_camera(dfs_builder* b, const float cam[19]);
int dfs_builder_set_sky(dfs_builder* b, const df_sky* sky);

/* SceneCompiler.compile: a new scene handle (render configuration: the
 * camera's resolution, 1 sample per pixel, reference default parameters). */
int dfs_builder_compile(dfs_builder* b, dfs_scene** out);

#ifdef __cplusplus
}
#endif

#endif /* DAYFLOWER_SCENE_H */
