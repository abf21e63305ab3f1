/*
 * dayflower_types.h - plain-old-data types shared by the renderer C ABI
 * (dayflower_hip.h) and the scene compiler C ABI (dayflower_scene.h).
 *
 * Every array keeps the flattened CompiledScene layout of the reference
 * unchanged (src/main/java/org/dayflower/pathtracer/scene/compiler/CompiledScene.java:55-71,
 * written in this order by CompiledScene.write :273-295).  Offsets stored in
 * the arrays are element offsets exactly as the reference's SceneCompiler
 * emits them (SceneCompiler.java:634-1053).
 */
#ifndef DAYFLOWER_TYPES_H
#define DAYFLOWER_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: 0 = OK, negative = error.  There is no silent CPU fallback. */
enum {
  DF_OK = 0,
  DF_EINVAL = -1,      /* bad argument / size mismatch                        */
  DF_ENODEVICE = -2,   /* no HIP device or the requested ordinal is absent    */
  DF_EHIP = -3,        /* a HIP runtime call failed (message in df_last_error) */
  DF_ESTATE = -4,      /* call order violated (e.g. execute before update)    */
  DF_ENOMEM = -5,      /* device or host allocation failed                    */
  DF_EIO = -6          /* file read/write failed                               */
};

/* Renderer types (AbstractRendererKernel.java:40-52). */
enum {
  DF_RENDERER_AMBIENT_OCCLUSION = 1,
  DF_RENDERER_PATH_TRACER = 2,
  DF_RENDERER_RAY_CASTER = 3,
  DF_RENDERER_RAY_MARCHER = 4,
  DF_RENDERER_RAY_TRACER = 5,
  DF_RENDERER_SURFACE_NORMALS = 6
};

/* Tone mappers (AbstractRendererKernel.java:58-64). */
enum {
  DF_TONE_MAPPER_FILMIC_CURVE_ACES_MODIFIED = 1,
  DF_TONE_MAPPER_REINHARD = 2,
  DF_TONE_MAPPER_REINHARD_MODIFIED_1 = 3,
  DF_TONE_MAPPER_REINHARD_MODIFIED_2 = 4
};

/* Shader types (AbstractRendererKernel.java:54-56). */
enum { DF_SHADER_FLAT = 1, DF_SHADER_GOURAUD = 2 };

/* Transcendental-function rounding mode of the GPU kernels.
 * DF_MATH_FAST  : single-precision device library (sinf, cosf, powf, ...).
 * DF_MATH_EXACT : each transcendental evaluated in double and rounded once
 *                 to float, i.e. the rounding of java.lang.Math under the
 *                 reference's thread-pool execution.  Used for bit-level
 *                 parity runs; every other operation is identical in both. */
enum { DF_MATH_FAST = 0, DF_MATH_EXACT = 1 };

/* How the path tracer is scheduled on the GPU; every choice gives the same
 * bits.  AUTO times the candidate schedules (df_schedule: the megakernel with
 * strip or square waves, the wavefront with one shade queue or material
 * bins) on the first passes of a frame, one pass at a time, two passes each,
 * and keeps the fastest until the scene, resolution or parameters change.
 * DF_PT_MEGAKERNEL : one work-item runs a pixel's whole path.
 * DF_PT_WAVEFRONT  : a kernel per stage (closest-hit, shade, shadow) over
 *                    queues of live paths. */
enum { DF_PT_AUTO = 0, DF_PT_MEGAKERNEL = 1, DF_PT_WAVEFRONT = 2 };

/* A launch schedule of the path tracer (no counterpart in the reference):
 * what the autotuner chooses among for DF_PT_AUTO (df_get_schedule), or what
 * df_set_schedule fixes.  Every schedule gives the same bits.
 *   engine       DF_PT_MEGAKERNEL or DF_PT_WAVEFRONT;
 *   wave_tile    megakernel: the pixels of a wave within its 16x16 tile,
 *                DF_TILE_ROWS (a 16x4 strip) or DF_TILE_SQUARE (8x8);
 *   shade_binned wavefront: 1 = hits are binned by material and each bin
 *                is shaded by code for that material only, 0 = one queue;
 *   bvh_pairs    wavefront: 1 = the bounce traces walk the mesh trees over
 *                64-byte two-child records with a per-lane stack (a scene
 *                whose trees are not all binary, or need a stack deeper than
 *                16, is walked over the nodes as with 0);
 *   tail_bounce  wavefront: the first bounce whose trace launch runs the
 *                trace kernel built for short queues (four waves per SIMD,
 *                no spilled registers) instead of the one for full queues
 *                (seven waves per SIMD); 0 = every bounce, at least the
 *                maximum depth = none; ignored with bvh_pairs.  The
 *                autotuner's wavefront candidates and the default wavefront
 *                schedule use DF_TAIL_BOUNCE_DEFAULT. */
enum { DF_TILE_ROWS = 0, DF_TILE_SQUARE = 1 };
#define DF_TAIL_BOUNCE_DEFAULT 4
typedef struct df_schedule {
  int32_t engine;
  int32_t wave_tile;
  int32_t shade_binned;
  int32_t bvh_pairs;
  int32_t tail_bounce;
} df_schedule;

/* Kernel stages, for per-stage counters and timings (df_get_stage_stats).
 * MEGAKERNEL covers every renderer run as one kernel; the others are the
 * wavefront path tracer's kernels. */
enum {
  DF_STAGE_MEGAKERNEL = 0,
  DF_STAGE_RAYGEN = 1,
  DF_STAGE_TRACE = 2,     /* sun any-hit queries of a bounce, then the next closest-hit queries */
  DF_STAGE_SHADE = 3,     /* hit record, textures, sun sample, BSDF        */
  DF_STAGE_ESCAPE = 4,    /* any-hit + sky after the maximum depth         */
  DF_STAGE_FILM = 5,      /* film update, wireframes, image pass           */
  DF_STAGE_COUNT = 6
};

/* The 14 flat arrays of a compiled scene plus the camera, as (pointer,
 * element count) pairs, in CompiledScene file order.  The data is borrowed
 * for the duration of the call that receives it and copied. */
typedef struct df_compiled_scene {
  const float* point2fs;        size_t point2fs_len;
  const float* point3fs;        size_t point3fs_len;
  const float* vector3fs;       size_t vector3fs_len;
  const int32_t* planes;        size_t planes_len;
  const float* spheres;         size_t spheres_len;
  const float* terrains;        size_t terrains_len;
  const int32_t* triangles;     size_t triangles_len;
  const int32_t* bvh;           size_t bvh_len;
  const float* textures;        size_t textures_len;
  const float* surfaces;        size_t surfaces_len;
  const int32_t* primitives;    size_t primitives_len;
  const float* object_to_world; size_t object_to_world_len;
  const float* world_to_object; size_t world_to_object_len;
  const float* camera;          size_t camera_len;   /* 19 floats */
} df_compiled_scene;

/* Sun and sky scalars read by the kernel (GPURendererKernel.java:100-127,
 * produced by Sky.set, Sky.java:295-374).  The sky frame is the reference's
 * fixed ONB u=(1,0,0) v=(0,0,1) w=(0,1,0) unless a caller supplies another. */
typedef struct df_sky {
  float onb_u[3];
  float onb_v[3];
  float onb_w[3];
  float sun_direction_world[3];
  float sun_direction[3];       /* sun direction in the sky frame */
  float sun_origin[3];          /* 10000 * sun_direction_world      */
  float theta;                  /* acos(sun_direction.z)            */
  float turbidity;              /* informational (host chain rule)  */
  double zenith_relative_luminance;
  double zenith_x;
  double zenith_y;
  double perez_relative_luminance[5];
  double perez_x[5];
  double perez_y[5];
  int32_t is_sky_active;
  int32_t is_sun_active;
} df_sky;

/* Renderer scalars (AbstractRendererKernel.java:129-184 defaults,
 * AbstractKernel.java:196-205 global noise defaults). */
typedef struct df_params {
  int32_t renderer_type;
  int32_t tone_mapper_type;
  float tone_mapper_exposure;
  int32_t pt_depth_maximum;
  int32_t pt_depth_russian_roulette;
  float ao_maximum_distance;
  int32_t normal_mapping;
  int32_t wireframes;
  int32_t shader_type;
  int32_t selected_primitive_offset;
  float global_amplitude;
  float global_frequency;
  float global_gain;
  float global_lacunarity;
  int32_t global_octaves;
  int32_t math_mode;
  int32_t pt_engine;        /* DF_PT_* */
} df_params;

/* Differentiable mode of the path tracer (no counterpart in the reference).
 * Continuous parameters replace the 8-bit albedo / emission textures of the
 * surfaces mapped to them: an albedo group's colour replaces the surface's
 * albedo texture, a checkerboard group's two colours replace those of the
 * surface's checkerboard albedo texture (whose tile rule stays, without its
 * single-colour shading quirk), an emission replaces the emission texture.
 * The sky is differentiated with respect to its turbidity and the sun's
 * elevation and azimuth: the renderer's current sky must be Sky.set of those
 * parameters, the sun towards (cos e sin a, sin e, cos e cos a) in the world,
 * and their derivatives are taken in dual numbers through Sky.set, the sky's
 * Perez evaluation and the sun cone sample.  Gradient layout (df_loss_gradient):
 * albedo groups x RGB, emissions, checkerboard groups x 2 colours x RGB,
 * turbidity, sun elevation, sun azimuth (per degree). */
typedef struct df_diff {
  int32_t n_albedo;
  int32_t n_emission;
  const int32_t* surface_albedo;    /* per surface record (surfaces_len / 6): albedo group or -1 */
  const int32_t* surface_emission;  /* per surface record: emission index or -1 */
  size_t n_surfaces;
  const float* albedo;              /* n_albedo * 3, linear RGB */
  const float* emission;            /* n_emission, white radiance */
  int32_t n_checker;                /* checkerboard groups; 0: none (surface_checker may be null) */
  int32_t reserved;
  const int32_t* surface_checker;   /* per surface record: checkerboard group or -1 */
  const float* checker;             /* n_checker * 2 * 3: the texture's first colour (dark tiles), its second */
  double turbidity;
  double elevation_deg;
  double azimuth_deg;
} df_diff;

/* Work counters.  A ray is one intersection query (closest-hit or any-hit);
 * a sample is one film update of one pixel.  The remaining fields count the
 * algorithmic units of the intersection and sky work: BVH box tests,
 * triangle tests, primitive ray transforms, sphere tests, terrain height
 * evaluations and sky evaluations (the Perez model of an escaped ray).
 * The ray count is the reference algorithm's; the unit counts are the work
 * this implementation did (it skips terrain heights and queries whose
 * outcome is already decided). */
typedef struct df_counters {
  uint64_t rays;
  uint64_t samples;
  uint64_t box_tests;
  uint64_t tri_tests;
  uint64_t prim_tests;
  uint64_t sphere_tests;
  uint64_t terrain_steps;
  uint64_t sky_evals;
} df_counters;

#ifdef __cplusplus
}
#endif

#endif /* DAYFLOWER_TYPES_H */
