/*
 * dayflower_hip.h - C ABI of the MI355X renderer (libdayflower_hip.so).
 *
 * Drop-in replacement for the reference's renderer-kernel surface:
 * AbstractRendererKernel (src/main/java/org/dayflower/pathtracer/kernel/
 * AbstractRendererKernel.java:36-727) plus the Aparapi Kernel calls the
 * application makes on it (execute(Range), put, get, dispose;
 * main/DayflowerApplication.java:319-343, 832-876).  Every entry point
 * returns 0 on success or a negative DF_E* code (dayflower_types.h); the
 * message of the last failure is available from df_last_error.  There is no
 * CPU fallback: without a HIP device df_create fails with DF_ENODEVICE.
 *
 * Threading: a handle is not thread-safe; callers serialise calls on one
 * handle, as the reference serialises execute/get/update with
 * synchronized(pixels1) (DayflowerApplication.java:645-654, 852-860).
 * Scalar parameters take effect at the next df_execute.
 */
#ifndef DAYFLOWER_HIP_H
#define DAYFLOWER_HIP_H

#include "dayflower_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct df_ctx df_ctx;

/* Number of visible HIP devices (0 when none). */
int df_device_count(int32_t* count);

/* Creates a renderer on HIP device `device`.  Replaces
 * new GPURendererKernel(sceneLoader) (GPURendererKernel.java:166-234) minus
 * the scene, which df_load_scene supplies. */
int df_create(df_ctx** out, int32_t device);

/* Validates a compiled scene on the host without touching a device: every
 * shape, surface, texture, triangle and BVH offset must lie inside its array
 * and every BVH pointer must name a node further along the pre-order.  The
 * same check guards df_load_scene.  Writes "ok" or the first problem to msg. */
int df_validate_scene(const df_compiled_scene* scene, char* msg, size_t cap);

/* Releases every device and host resource (Kernel.dispose). */
void df_destroy(df_ctx* ctx);

/* Uploads the compiled scene (the put() calls of GPURendererKernel.update,
 * :426-447).  The arrays are validated (every offset in range), copied, and
 * repacked device-side; the camera array is taken from scene->camera. */
int df_load_scene(df_ctx* ctx, const df_compiled_scene* scene);

/* Allocates the per-pixel film, sample counts, image, seed and primitive
 * offset buffers for width x height and fills the primitive offsets with -1
 * (update(W, H, byte[], localSize), GPURendererKernel.java:411-449 and
 * AbstractImageKernel.java:145-165).  local_size_hint is accepted for
 * signature parity and ignored (the workgroup shape is fixed at 16x16).
 * Seeds are set from master seed 0; call df_seed or df_set_seeds after. */
int df_update(df_ctx* ctx, int32_t width, int32_t height, int32_t local_size_hint);

/* Fills the per-pixel 48-bit LCG seeds deterministically from a master
 * seed: seed[i] = SplitMix64(master + (i + 1) * 0x9E3779B97F4A7C15).
 * Replaces ThreadLocalRandom seeding (AbstractKernel.java:1148-1156). */
int df_seed(df_ctx* ctx, uint64_t master);

/* Sets / reads the per-pixel seeds exactly (n must equal width * height). */
int df_set_seeds(df_ctx* ctx, const int64_t* seeds, size_t n);
int df_get_seeds(df_ctx* ctx, int64_t* seeds, size_t n);

/* updateCamera() (GPURendererKernel.java:455-461): re-uploads the 19-float
 * camera array (Camera.update is host math, dfs_camera_setup). */
int df_upload_camera(df_ctx* ctx, const float* camera, size_t n);

/* updatePrimitives() (GPURendererKernel.java:467-521): new objectToWorld and
 * worldToObject matrices, 16 floats per primitive, n_prim primitives. */
int df_upload_transforms(df_ctx* ctx, const float* object_to_world, const float* world_to_object, int32_t n_prim);

/* put(sceneSurfaces) after a material change (GPURendererKernel.java:314-342). */
int df_upload_surfaces(df_ctx* ctx, const float* surfaces, size_t n);

/* updateSunAndSky() (GPURendererKernel.java:527-569). */
int df_upload_sky(df_ctx* ctx, const df_sky* sky);

/* toggleSky() / toggleSun() (GPURendererKernel.java:374-397). */
int df_toggle_sky(df_ctx* ctx);
int df_toggle_sun(df_ctx* ctx);

/* Renderer setters (AbstractRendererKernel.java:392-575).  Invalid enum
 * values are rejected with DF_EINVAL (the reference ignores them). */
int df_set_params(df_ctx* ctx, const df_params* params);
int df_get_params(df_ctx* ctx, df_params* params);

/* togglePrimitiveSelection(x, y) (GPURendererKernel.java:351-369): reads the
 * primitive offset under pixel (x, y) and toggles the selection tint. */
int df_toggle_primitive_selection(df_ctx* ctx, int32_t x, int32_t y);

/* togglePrimitiveMaterial() (GPURendererKernel.java:314-342): cycles the
 * material of the selected primitive's surface. */
int df_toggle_primitive_material(df_ctx* ctx);

/* clear() (AbstractImageKernel.java:119-121): the next pass overwrites the
 * film instead of averaging into it. */
int df_clear(df_ctx* ctx);

/* execute(range) followed by clearFilmFlags(), `passes` times: one sample per
 * pixel per pass.  Asynchronous on the handle's stream. */
int df_execute(df_ctx* ctx, int32_t passes);

/* The path tracer's engine for the next df_execute: DF_PT_MEGAKERNEL or
 * DF_PT_WAVEFRONT, or DF_PT_AUTO while the autotuner is still timing the
 * schedules on the current frame (no counterpart in the reference). */
int df_get_engine(df_ctx* ctx, int32_t* engine);

/* The full schedule of the next passes (engine DF_PT_AUTO while the
 * autotuner is still timing its candidates), and a fixed schedule for every
 * later df_execute (NULL: back to params.pt_engine, re-tuning DF_PT_AUTO).
 * No counterpart in the reference; every schedule gives the same bits.  In a
 * batch, a path-tracer job whose schedule is the wavefront is rendered on it
 * by df_batch_execute instead of by the batch kernel. */
int df_get_schedule(df_ctx* ctx, df_schedule* schedule);
int df_set_schedule(df_ctx* ctx, const df_schedule* schedule);

/* The per-lane stack depth the loaded scene's two-child walk uses (schedule
 * bvh_pairs), or 0 when the scene is walked over the nodes with either
 * setting (a tree node with other than two children, or a stack deeper than
 * 16).  No counterpart in the reference. */
int df_get_bvh_pair_stack(df_ctx* ctx, int32_t* depth);

/* Frame-loop graphs (no counterpart in the reference): with graphs on
 * (default off: on an MI355X replaying measured 1-7 % slower per pass than
 * launching the same kernels eagerly), df_execute captures the launches of
 * one wavefront pass (one graph
 * per combination of the film-clear and image-pass flags) or the megakernel
 * launch into a HIP graph and replays it while the scene, camera, sky,
 * parameters and frame are unchanged; any change is captured anew on the
 * next df_execute.  Results are the same bits either way.  Launches timed by
 * df_set_stage_timing are not captured. */
int df_set_graphs(df_ctx* ctx, int32_t enabled);

/* Waits for the handle's stream. */
int df_synchronize(df_ctx* ctx);

/* Device time of the last df_execute in milliseconds (HIP events on the
 * handle's stream; waits for it). */
int df_last_execute_ms(df_ctx* ctx, float* ms);

/* Readbacks (get(...)); sizes must match exactly. */
int df_get_image_bgra(df_ctx* ctx, uint8_t* bgra, size_t n);        /* 4 * W * H bytes */
int df_get_image_float(df_ctx* ctx, float* rgb, size_t n);          /* 3 * W * H */
int df_get_primitive_offsets(df_ctx* ctx, int32_t* offsets, size_t n);  /* W * H */
int df_get_film(df_ctx* ctx, float* rgb, size_t rgb_n, int32_t* samples, size_t samples_n);

/* Restores film and sample counts (with df_set_seeds this resumes a render
 * bit-exactly). */
int df_set_film(df_ctx* ctx, const float* rgb, size_t rgb_n, const int32_t* samples, size_t samples_n);

/* Cumulative work counters since the last reset (rays = intersection
 * queries, samples = film updates). */
int df_get_counters(df_ctx* ctx, df_counters* out);
int df_reset_counters(df_ctx* ctx);

/* Unit counters (no counterpart in the reference): with them off (default
 * on), the wavefront schedule's trace kernels leave box_tests, tri_tests,
 * prim_tests, sphere_tests and terrain_steps at zero (rays, samples and
 * sky_evals are still counted), which saves about 3 % of their time.  The
 * counts are a function of the scene, parameters and seeds, so a caller can
 * time passes without them and count the same passes again (df_set_seeds)
 * with them.  Results are the same bits either way. */
int df_set_unit_counters(df_ctx* ctx, int32_t enabled);

/* Per-stage instrumentation (no counterpart in the reference, whose kernel
 * is one Aparapi launch).  With timing on, every kernel launch is bracketed
 * by HIP events on the handle's stream; df_get_stage_stats returns, for one
 * DF_STAGE_*, the counters its kernels accumulated, the summed kernel time
 * in milliseconds and the number of launches since the last
 * df_reset_counters.  It synchronizes the stream. */
int df_set_stage_timing(df_ctx* ctx, int32_t enabled);
int df_get_stage_stats(df_ctx* ctx, int32_t stage, df_counters* counters, double* milliseconds,
                       int64_t* launches);

/* Message of the last failed call on this handle (never NULL). */
const char* df_last_error(df_ctx* ctx);

/* ---- Differentiable path tracer -------------------------------------------
 * (no counterpart in the reference; planned with the ABI as df_adjoint).
 * df_set_diff switches the path tracer to the parameters of *diff (copied);
 * the current sky (df_upload_sky) must be Sky.set of diff's turbidity,
 * elevation and azimuth (DF_EINVAL otherwise), a checkerboard group may map
 * only surfaces with a checkerboard albedo texture and no albedo group, and
 * at most 1024 parameters fit.  df_clear_diff restores the textures, and so
 * does df_load_scene; df_upload_surfaces refuses to take a checkerboard
 * group's checkerboard texture away.  df_loss_gradient renders `passes`
 * samples per pixel into a cleared film (the forward pass), returns
 * L = sum (film - target)^2 over the W x H x 3 linear film, and, when grad
 * is not NULL, dL/dparameters (n_grad = 3 n_albedo + n_emission +
 * 6 n_checker + 3, laid out as df_diff describes) by replaying every path
 * from the seeds the forward pass started from.  It needs the path tracer,
 * Russian roulette off (pt_depth_russian_roulette > pt_depth_maximum),
 * pt_depth_maximum <= 8 and the sky df_set_diff was given (DF_ESTATE after
 * another df_upload_sky).  forward_ms / adjoint_ms (may be NULL) receive the
 * device times of the two passes.  On return the seeds are those after the
 * forward pass. */
int df_set_diff(df_ctx* ctx, const df_diff* diff);
int df_clear_diff(df_ctx* ctx);
int df_loss_gradient(df_ctx* ctx, const float* target, size_t n, int32_t passes, double* loss, double* grad,
                     size_t n_grad, float* forward_ms, float* adjoint_ms);

/* ---- Batches of independent scenes ---------------------------------------
 * The reference renders one scene per kernel; these entry points (planned
 * with the ABI as df_batch_*) render many scenes in one persistent launch:
 * workgroups take (scene, 16x16 tile) items from a device queue, most
 * expensive first, and render all of a tile's passes.  Each scene keeps its
 * own renderer handle, so its result equals rendering it alone with
 * df_execute(passes), bit for bit. */
typedef struct df_batch df_batch;

int df_batch_create(df_batch** out, int32_t device);
void df_batch_destroy(df_batch* batch);

/* Adds a scene: load, sky, parameters, width x height frame, df_seed(seed);
 * `passes` samples per pixel per df_batch_execute.  *index = its position. */
int df_batch_add(df_batch* batch, const df_compiled_scene* scene, const df_sky* sky, const df_params* params,
                 int32_t width, int32_t height, int32_t passes, uint64_t seed, int32_t* index);
int df_batch_count(df_batch* batch, int32_t* count);

/* The scene's renderer handle, owned by the batch (readbacks, counters, or a
 * df_execute of that scene alone).  Do not df_destroy it. */
int df_batch_job(df_batch* batch, int32_t index, df_ctx** ctx);

/* Renders every scene's passes; returns when the batch is done. */
int df_batch_execute(df_batch* batch);

/* df_batch_execute without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (serving loops,
 * code around HIP graph capture, harnesses that forbid hipMalloc): the
 * batch's job table, tile work list and work heads live in a static device
 * array of about 4.7 MiB per GPU, set up by df_batch_create and held until
 * the process ends.  Limits: at most 1024 scenes and 262144 16x16 tiles
 * (their sum over the scenes) per call, else DF_EINVAL with a message and
 * nothing rendered.  Results equal df_batch_execute's bit for bit (every job
 * through the batch kernel, whatever its own schedule).  Calls
 * share their GPU's array: this process's calls are serialised per device
 * and each returns after its batch is done; calls that overlapped on one GPU
 * would race. */
int df_batch_execute_noalloc(df_batch* batch);

/* Graphs for batches (default off): the group launches of a batch are
 * captured once and replayed while the batch's work tables keep their size. */
int df_batch_set_graphs(df_batch* batch, int32_t enabled);

/* Device time of the last df_batch_execute in milliseconds. */
int df_batch_last_execute_ms(df_batch* batch, float* ms);
const char* df_batch_last_error(df_batch* batch);

/* ---- Terrain march check --------------------------------------------------
 * (no counterpart in the reference: a self-check for tests.)  Marches n rays
 * (6 floats each: object-space origin, then direction) against the terrain
 * height field on `device` twice: with the kernels' march, which skips the
 * positions a crossing provably cannot lie at (t_skip, and the heights it
 * evaluated in evals), and with the reference's march over every position
 * (GPURendererKernel.java:1470-1506, t_plain), both in math_mode and with no
 * closer hit.  The two distances must be equal bit for bit.  n <= 2^26;
 * allocates and frees its own device buffers. */
int df_check_terrain_march(int32_t device, int32_t math_mode, const float* rays, size_t n, float* t_skip,
                           float* t_plain, uint32_t* evals);

/* ---- Reciprocal check -----------------------------------------------------
 * (no counterpart in the reference: a self-check for tests.)  The triangle
 * test takes 1 / det from the hardware reciprocal and one fused correction
 * when 2^-14 < |det| < 2^126; this compares that with the correctly rounded
 * division 1.0f / x for every float in that range, both signs, on `device`:
 * the values tested, the mismatches (must be 0) and the smallest mismatching
 * bit pattern (0xFFFFFFFF when none). */
int df_check_reciprocal(int32_t device, uint64_t* tested, uint64_t* mismatches, uint32_t* first_mismatch);

/* ---- Measured roofs -------------------------------------------------------
 * (no counterpart in the reference: measurement only.)  Runs one
 * microbenchmark on `device` and returns its sustained rate: FLOP/s of
 * single (unpacked) FP32 fused multiply-adds, or bytes/s read from the L2
 * caches, from the Infinity Cache (MALL), or from HBM.  work_per_launch (may
 * be NULL) receives the FLOP or bytes of one timed launch.  Allocates and
 * frees its own buffers (at most 2 GiB, for DF_ROOF_HBM_READ). */
enum { DF_ROOF_VALU_FP32 = 1, DF_ROOF_L2_READ = 2, DF_ROOF_MALL_READ = 3, DF_ROOF_HBM_READ = 4 };
int df_measure_roof(int32_t device, int32_t kind, double* rate, double* work_per_launch);

#ifdef __cplusplus
}
#endif

#endif /* DAYFLOWER_HIP_H */
