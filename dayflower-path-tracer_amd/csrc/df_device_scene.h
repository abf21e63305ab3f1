// Device-side scene representation shared by the host runtime (which builds
// it) and the HIP kernels (which read it).  The reference's flat arrays are
// uploaded unchanged; next to them the runtime builds a private repack:
//   * one DPrim record per primitive (type, offsets, both matrices and the
//     shape's geometry gathered out of point3Fs / vector3Fs), read with
//     wave-uniform scalar loads in the primitive loop;
//   * every TriangleMesh BVH re-laid as 32-byte nodes in the reference's own
//     pre-order (box, skip pointer, leaf triangle range), so the traversal
//     visits boxes and triangles in exactly the reference's order
//     (GPURendererKernel.java:1097-1194) with one 32-byte fetch per node
//     instead of three dependent gathers;
//   * the triangles of every leaf stored contiguously as (A, B - A, C - A)
//     in 48 bytes, with the original triangle offset for surface properties.
#pragma once

#include <stdint.h>

namespace dfk {

// One primitive, 46 words (184 bytes).
struct DPrim {
  int32_t type;          // shape type (1 plane, 2 sphere, 3 terrain, 4 triangle, 5 mesh)
  int32_t shape_offset;  // reference shape offset (planes / spheres / triangles / BVH segment)
  int32_t surface_offset;
  int32_t node_root;     // mesh: index of the root in the repacked node array
  float w2o[12];         // world -> object rows 1..3
  float o2w[12];         // object -> world rows 1..3
  float geo[8];          // sphere: pos.xyz, radius; plane: a.xyz, n.xyz; triangle: unused
  int32_t run_end;       // mesh: one past the last of the consecutive meshes starting here; else index + 1
  int32_t xform;         // index of the first primitive with the same w2o (bit for bit)
  int32_t pad[8];
};
static_assert(sizeof(DPrim) == 184, "DPrim layout");

// World-space bounding sphere of what a primitive's test can hit, for
// skipping primitives a ray cannot reach.  on = 0: never skipped (planes,
// terrains, single triangles).  A ray skips the primitive only when the
// sphere of radius R = r + k |c - o|^2 lies off its line or wholly behind
// its origin.  R is at least the sphere's radius enlarged by 1e-4 |c - o| +
// 1e-6 s / r' |c - o|^2 (r' the sphere's own radius, s <= 16 the squared
// stretch ratio of the primitive's transform), which exceeds the float
// error of the object-space sphere and box tests (about 1e-7 s |c - o|^2 /
// r' and 1e-7 |c - o|) and of this test by an order of magnitude, so a
// skipped primitive is one whose own test would report no hit.
struct DCull {
  float c[3];
  float r;  // R = r + k |c - o|^2
  float k;
  int32_t on;
  int32_t pad[2];
};

// Repacked BVH node (pre-order).  box = (min.xyz, max.xyz); next = index of
// the skip target or -1; tri >= 0: first packed triangle of a leaf;
// tri < 0: tree node whose left child is node -tri - 2 (a leaf without
// triangles is stored as a tree node whose left child is its skip target).
struct DNode {
  float mn[3];
  float mx[3];
  int32_t next;
  int32_t tri;
};

// Packed triangle: a.xyz, e0 = b - a, e1 = c - a (computed in float exactly
// as the reference computes them per test), the reference triangle offset and
// a last-in-leaf flag.
struct DTri {
  float a[3];
  float e0[3];
  float e1[3];
  int32_t tri_offset;
  int32_t last;
  int32_t pad;
};

// Two-child BVH records (64 bytes, one cache line), for the walk that tests
// both children of a tree node from one fetch and keeps the second child's
// test result on a short per-lane stack: the boxes of both children of tree
// node i of the DNode array, component-major (min x of child 0 and child 1,
// min y, ...), and the children: >= 0 a tree node (its record), -1 a leaf
// without triangles, <= -2 a leaf whose first packed triangle is -ref - 2.
// Indexed like DNode (records of leaf nodes unused).
constexpr int kPairStackMax = 16;  // deepest per-lane stack the kernels provide
struct alignas(64) DPair {
  float mn[3][2];
  float mx[3][2];
  int32_t ref[2];
  int32_t pad[2];
};
static_assert(sizeof(DPair) == 64, "DPair layout");

struct DScene {
  // unchanged reference arrays
  const float* point2fs;
  const float* point3fs;
  const float* vector3fs;
  const int32_t* planes;
  const float* spheres;
  const float* terrains;
  const int32_t* triangles;
  const float* textures;
  const float* surfaces;
  const int32_t* primitives;
  // repack
  const DPrim* prims;
  const DNode* nodes;
  const DTri* tris;
  const DCull* cull;
  // Terrain march positions tc_0 = 0.001, tc_{k+1} = tc_k + 0.01 * tc_k (float
  // arithmetic, every tc_k < 10): the sequence every terrain query visits.
  const float* march_tc;
  // two-child records (null when some mesh's tree is not binary or needs a
  // deeper stack than kPairStackMax) and the stack depth the scene needs
  const DPair* pairs;
  int32_t pair_stack;
  int32_t prim_count;
  int32_t march_n;
};

// Copies of the work counters per stage, one 64-byte line each: a wave adds
// into the copy its global index selects and the host sums the copies.
// Atomics on one line serialize (about 11 ns apiece), so a single copy made
// the counter flush a large part of every launch.
constexpr int kCounterSlots = 64;

struct DCamera {
  float c[19];
  float fov_x_tan;  // tan(fovX * 0.5 * PI/180) rounded once to float (host)
  float fov_y_tan;  // tan(-fovY * 0.5 * PI/180)
  int32_t pad;
};

struct DSky {
  float onb_u[3], onb_v[3], onb_w[3];
  float sun_dir_world[3];
  float sun_dir[3];
  float sun_origin[3];
  float theta;
  int32_t sky_active;
  int32_t sun_active;
  double zenith_l, zenith_x, zenith_y;
  double pl[5], px[5], py[5];
  // Per-frame constants hoisted out of the per-ray sky evaluation
  // (GPURendererKernel.java:2193-2205), each computed with the same
  // operation order as the kernel expression it replaces:
  double l_den;           // whole luminance denominator
  double x_den0, y_den0;  // (1 + A * exp(B)) of the x / y denominators
  double x_c, y_c;        // E * cos(theta0) * cos(theta0) of the x / y denominators
};

struct DParams {
  int32_t renderer_type;
  int32_t tone_mapper_type;
  float exposure;
  int32_t depth_max;
  int32_t depth_rr;
  float ao_max;
  int32_t normal_mapping;
  int32_t wireframes;
  int32_t shader_type;
  int32_t selected;
  float ga, gf, gg, gl;
  int32_t go;
  int32_t width, height;
  int32_t passes;
  int32_t clear_first;
  // wavefront schedule: 1 = the trace kernel bins each closest-hit result by
  // its hit surface's material and the shade kernel drains the bins one by one
  int32_t shade_binned;
  // megakernel schedule: 0 = a wave covers a 16x4 strip of its 16x16 tile, 1 = an 8x8 quadrant
  int32_t wave_tile;
  // wavefront schedule: 0 = its trace kernels leave the unit counters (box,
  // triangle, primitive and sphere tests, terrain heights) at zero
  int32_t count_units;
  // wavefront schedule: 1 = the bounce traces walk the meshes over the
  // two-child records (DPair) with a per-lane stack
  int32_t bvh_pairs;
  // wavefront schedule: the first bounce whose trace launch runs the kernel
  // for short queues (four waves per SIMD, no spilled registers)
  int32_t tail_bounce;
  // sRGB constants (image4j RGBColorSpace.SRGB)
  float srgb_break, srgb_gamma_inv, srgb_slope, srgb_slope_match, srgb_segment;
  // differentiable mode (diff_on = 1): continuous albedo / emission
  // parameters replace the 8-bit textures of the surfaces mapped to them.
  // Albedo slots: k < checker_base is albedo group k; checkerboard group g
  // has slots checker_base + 2 g + t, t the tile (0: the texture's first
  // colour); diff_albedo holds every slot's RGB.
  int32_t diff_on;
  int32_t n_surfaces;               // surface records (6 floats each)
  const int32_t* surface_albedo;    // per surface record: albedo group or -1
  const int32_t* surface_emission;  // per surface record: emission index or -1
  const int32_t* surface_checker;   // per surface record: checkerboard group or -1
  const float* diff_albedo;         // albedo slots x RGB
  const float* diff_emission;       // white radiance per index
  int32_t checker_base;
};

// A double with its derivatives with respect to the sky's parameters:
// turbidity, sun elevation and sun azimuth (degrees).
struct Dual3 {
  double v;
  double d[3];
};

// Sky.set of the sky's parameters in dual numbers: what the adjoint
// differentiates the sky radiance and the sun sample through.
struct DSkyDual {
  float onb_u[3], onb_v[3], onb_w[3];
  Dual3 world[3];  // unit vector towards the sun, world frame
  Dual3 dir[3];    // the same in the sky frame
  Dual3 theta;     // the sun's zenith angle
  Dual3 zl, zx, zy;
  Dual3 pl[5], px[5], py[5];
};

// Inputs of the adjoint (path-replay) kernel beyond those of a render.
// Gradient layout: albedo groups x RGB, emissions, checkerboard groups x 2
// colours x RGB, turbidity, sun elevation, sun azimuth (per degree).
struct DAdjoint {
  const DSkyDual* sky;  // device copy of the dual Sky.set
  const float* weight;  // dLoss/dfilm per pixel and channel, divided by the passes
  double* grad;
  int32_t n_albedo;
  int32_t n_emission;
  int32_t n_checker;
};

// Parameters of the adjoint that one launch handles: LDS holds them per block.
constexpr int kMaxDiffParams = 1024;
constexpr int kMaxAdjointDepth = 8;

// Per-pixel frame buffers.
struct DFrame {
  float* film;             // 3 * W * H
  int32_t* film_samples;   // W * H
  float* image_float;      // 3 * W * H
  uint8_t* image_bgra;     // 4 * W * H
  int32_t* primitive_offsets;
  uint64_t* seeds;
  unsigned long long* counters;  // per stage: rays, samples, box, tri, prim, sphere, terrain, sky

};

// Everything one scene of a batch needs: its scene, camera, sky, parameters
// and frame (the arguments of render_kernel).
struct DJob {
  DScene S;
  DCamera C;
  DSky K;
  DParams P;
  DFrame F;
};

// One work item of a batch launch: a 16x16 tile of a job, all its passes.
struct DBatchItem {
  int32_t job;
  int32_t tile_x;
  int32_t tile_y;
  int32_t pad;
};

// Path state of the wavefront path tracer, one slot per pixel (y * W + x),
// and its work queues.  float4 records keep each slot one 16-byte access.
struct DPaths {
  float4* ray_o;      // current ray origin
  float4* ray_d;      // current ray direction
  float4* thru;       // path throughput
  float4* rad;        // radiance so far (without the sun term while its shadow query is pending)
  float4* hit;        // closest hit: t, primitive index, shape type, shape offset (int bits)
  float4* sh_o;       // shadow query origin
  float4* sh_d;       // shadow query direction
  float4* rad_lit;    // radiance with the sun term, taken when the shadow query hits nothing
  float4* sample;     // the pass's sample colour; w = 1: set the film instead of accumulating
  float4* ray0_o;     // primary ray, kept for the wireframe test
  float4* ray0_d;
  int32_t* queue[2];  // live paths of even / odd bounces (kQueueStripes stripes of n entries)
  int32_t* shadow_queue;
  int32_t* noise_queue;  // a bounce's paths whose hit surface has noise normal mapping
  int32_t* sky_queue;    // the pass's paths that escaped, their sky term pending (every bounce)
  int32_t* shade_bins;   // kShadeBins bins of kQueueStripes stripes of n entries (binned shading)
  float4* noise;         // their three normal-map fBm values (x, y, z)
  int32_t* ctl;       // per-pass queue counts and work heads (groups of kCtlGroup ints)
  int64_t n;          // slots; also the capacity of one queue stripe
  int32_t ctl_len;
  int32_t pad;
};

// Work queues are striped: a wave appends to, and first drains, the stripe its
// global index selects, so the atomics on a queue's count and work head spread
// over kQueueStripes 64-byte lines (atomics on one line serialize).  A stripe
// can receive every entry, so each holds n.
constexpr int kQueueStripes = 16;
constexpr int kLineInts = 16;                          // one 64-byte line
constexpr int kCtlGroup = kQueueStripes * kLineInts;  // one count or head per stripe

// Shade bins of the binned schedule: 0 the paths that hit nothing, 1..5 the
// hits on a surface of that material (ClearCoat, Glass, Lambertian, Phong,
// Reflection), 6 the hits on a surface with any other material value.
constexpr int kShadeBins = 7;

// Length of DPaths::ctl for a maximum depth, in ints: B + 1 live-path counts,
// B + 1 shadow counts (the last stays zero), 3B + 4 work heads, per bounce the
// noise queue's count and work head, then per bounce the shade bins' counts
// and work heads, B = max(1, depth), each a group of per-stripe lines.
__host__ __device__ inline int wavefront_ctl_len(int depth_max) {
  const int b = depth_max > 1 ? depth_max : 1;
  return (7 * b + 8 + 2 * kShadeBins * b) * kCtlGroup;
}

}  // namespace dfk
