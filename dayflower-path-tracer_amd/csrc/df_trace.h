// Per-path tracing state and the renderer bodies of the megakernel.
//
// Tracer<M> holds one pixel's state (RNG stream, hit record, scratch
// colour) and restates doIntersectPrimitives, the surface-property,
// texture, normal-mapping, sun/sky and renderer functions of
// GPURendererKernel.java.  The primitive loop is wave-uniform (every lane
// tests primitive i at the same time), so the DPrim records are read with
// scalar loads; only the BVH walk and the shading diverge.
#pragma once

#include "df_device.h"
#include "df_dual.h"

namespace dfk {

// What one path-tracer bounce contributed, for the adjoint's replay: the
// albedo slot and emission parameter it used, the sun term's albedo factor
// and its derivatives with respect to the sun's elevation and azimuth, and
// how the throughput changed (m *= s, times the albedo when uses_albedo).
// Input: the dual Sky.set the derivatives are taken through.
struct Vtx {
  const DSkyDual* sky;   // in
  bool miss;
  bool uses_albedo;
  bool rr;               // Russian roulette scaled the albedo
  int aslot, eslot;
  float a[3];
  float e[3];            // emission added at the vertex
  float s;
  float sun_w;
  float dsun_w[2];       // d sun_w / d elevation, d azimuth (per degree)
};

// The mesh walk tests leaves' triangles once this fraction of the lanes
// still walking wait at a leaf (measured best on the 1080p workload: 1/8,
// 1/6, 1/2, 3/4 and all slower).
constexpr int kLeafWaitNum = 1, kLeafWaitDen = 4;

// Node visits per step of the mesh walk loop (measured on the 1080p
// workload: trace 7.79 ms per pass with one, 7.35 with two, 7.18-7.28 with
// four, 7.27 with six, 7.31 with eight, 7.40 with an inner loop of up to
// eight; measured again once the last bounces had their own kernel: three
// 7.05, four 7.02-7.08, five 6.98-7.04, six 7.05).
constexpr int kVisitsPerStep = 5;

// kUnits: count the unit counters (box, triangle, primitive and sphere
// tests, terrain height evaluations); the wavefront's trace kernels also come
// without them, for callers that need only rays and samples (about 3 % of
// those kernels' time).
// kPairs: the mesh walks go over the two-child records (walk_meshes_pairs),
// with the per-lane stack in LDS at pstack_key / pstack_ref.
template <int M, bool kUnits = true, bool kPairs = false>
struct Tracer {
  const DScene& S;
  const DCamera& C;
  const DSky& K;
  const DParams& P;
  Noise noise;
  // Values precomputed per path (wavefront shading): when noise_in is set,
  // a hit on a surface with noise normals takes its three noise values
  // (x, y, z) and its albedo texture's colour (w, int bits) from
  // noise_in[noise_index]; with noise_skip, normal_map_noise leaves the
  // normal unperturbed (the pass computing the values).
  const float4* noise_in = nullptr;
  int noise_index = 0;
  bool noise_skip = false;
  // the lane's stack of walk_meshes_pairs (entry k at [k * kPairStride]; LDS)
  typedef __attribute__((address_space(3))) float lds_f;
  typedef __attribute__((address_space(3))) int lds_i;
  static constexpr int kPairStride = 256;
  lds_f* pstack_key = nullptr;
  lds_i* pstack_ref = nullptr;
  Rng rng;
  Hit h;
  float tmp[3];
  // pt_bounce<true>: set when the path missed every primitive and its sky
  // term is left to a pass of its own (wf_sky)
  bool sky_pending = false;
  unsigned rays;
  unsigned samples;
  unsigned box_tests;
  unsigned tri_tests;
  unsigned prim_tests;
  unsigned sphere_tests;
  unsigned terrain_steps;
  unsigned sky_evals;

  __device__ __forceinline__ static void unit(unsigned& counter) {
    if constexpr (kUnits) ++counter;
  }

  __device__ Tracer(const DScene& s, const DCamera& c, const DSky& k, const DParams& p, const int* perm)
      : S(s), C(c), K(k), P(p), rays(0), samples(0), box_tests(0), tri_tests(0), prim_tests(0),
        sphere_tests(0), terrain_steps(0), sky_evals(0) {
    noise.perm = perm;
    h.t = DF_INF;
    h.prim = -1;
    h.shape_type = -1;
    h.shape_off = -1;
    h.p[0] = h.p[1] = h.p[2] = 0.0f;
    h.ns[0] = h.ns[1] = h.ns[2] = 0.0f;
    h.uv[0] = h.uv[1] = 0.0f;
    h.tan[0] = h.tan[1] = h.tan[2] = 0.0f;
  }

  // ---------------------------------------------------------- primary ray (:573-724)
  __device__ __forceinline__ bool primary_ray(int x, int y, float ray[6]) {
    const float* c = C.c;
    const float eyeX = c[1], eyeY = c[2], eyeZ = c[3];
    const float uX = c[7], uY = c[8], uZ = c[9];
    const float vX = c[10], vY = c[11], vZ = c[12];
    const float wX = c[13], wY = c[14], wZ = c[15];
    const float fx1 = C.fov_x_tan, fy1 = C.fov_y_tan;
    const float hX = uX * fx1, hY = uY * fx1, hZ = uZ * fx1;
    const float vtX = vX * fy1, vtY = vY * fy1, vtZ = vZ * fy1;
    float sampleX = 0.5f, sampleY = 0.5f;
    if (P.renderer_type == 2) {
      sampleX = rng.next();
      sampleY = rng.next();
      sampleX = sampleX < 0.5f ? sqrtf(2.0f * sampleX) - 1.0f : 1.0f - sqrtf(2.0f - 2.0f * sampleX);
      sampleY = sampleY < 0.5f ? sqrtf(2.0f * sampleY) - 1.0f : 1.0f - sqrtf(2.0f - 2.0f * sampleY);
    }
    const float sx = (sampleX + static_cast<float>(x)) / (c[16] - 1.0f);
    const float sy = (sampleY + static_cast<float>(y)) / (c[17] - 1.0f);
    const float sx0 = 2.0f * sx - 1.0f;
    const float sy0 = 2.0f * sy - 1.0f;
    float w = 1.0f;
    if (c[18] == 2.0f) {
      const float dp = sx0 * sx0 + sy0 * sy0;
      if (dp > 1.0f) return false;
      w = sqrtf(1.0f - dp);
    }
    const float mX = eyeX + wX * w, mY = eyeY + wY * w, mZ = eyeZ + wZ * w;
    const float ppX = mX + (hX * sx0) + (vtX * sy0);
    const float ppY = mY + (hY * sx0) + (vtY * sy0);
    const float ppZ = mZ + (hZ * sx0) + (vtZ * sy0);
    const float focal = c[6];
    const float piX = eyeX + (ppX - eyeX) * focal;
    const float piY = eyeY + (ppY - eyeY) * focal;
    const float piZ = eyeZ + (ppZ - eyeZ) * focal;
    float apX = eyeX, apY = eyeY, apZ = eyeZ;
    if (c[0] > 0.00001f) {
      const float r1 = rng.next();
      const float r2 = rng.next();
      const float angle = DF_PI_2 * r1;
      const float dist = c[0] * sqrtf(r2);
      const float ax = Math<M>::cos(angle) * dist;
      const float ay = Math<M>::sin(angle) * dist;
      apX = eyeX + uX * ax + vX * ay;
      apY = eyeY + uY * ax + vY * ay;
      apZ = eyeZ + uZ * ax + vZ * ay;
    }
    const float d0X = piX - apX, d0Y = piY - apY, d0Z = piZ - apZ;
    const float lr = jrsqrt(d0X * d0X + d0Y * d0Y + d0Z * d0Z);
    ray[0] = apX;
    ray[1] = apY;
    ray[2] = apZ;
    ray[3] = d0X * lr;
    ray[4] = d0Y * lr;
    ray[5] = d0Z * lr;
    return true;
  }

  // ---------------------------------------------------------- ray query (:1048-1411)
  // Testing = true is the any-hit query (isTesting), false the closest-hit
  // query that also fills the hit record.
  __device__ __forceinline__ float intersect(float ox, float oy, float oz, float dx, float dy, float dz, bool testing) {
    return testing ? intersect_t<true>(ox, oy, oz, dx, dy, dz) : intersect_t<false>(ox, oy, oz, dx, dy, dz);
  }

  template <bool testing>
  __device__ __forceinline__ float intersect_t(float ox, float oy, float oz, float dx, float dy, float dz) {
    int ci, cst, cso;
    const float closest = query<testing>(ox, oy, oz, dx, dy, dz, ci, cst, cso);
    if (!testing) fill_hit(ox, oy, oz, dx, dy, dz, closest, ci, cst, cso);
    return closest;
  }

  // The BVH walks (:1097-1194) of the marked meshes `mine` (bits relative to
  // `base`), in index order, as one loop of node visits.  Updates the closest
  // hit; returns true when an any-hit query has its hit.
  __device__ __forceinline__ bool walk_meshes(uint64_t mine, int base, float ox, float oy, float oz, float dx,
                                              float dy, float dz, float& closest, int& ci, int& cst, int& cso,
                                              bool testing) {
    float oxo = 0.0f, oyo = 0.0f, ozo = 0.0f, dxo = 0.0f, dyo = 0.0f, dzo = 0.0f;
    float rdx = 0.0f, rdy = 0.0f, rdz = 0.0f;
    int idx = -1, pi = 0, xform = -1;
    int leaf = -1;  // first triangle of the leaf this lane has reached, its tests pending
    for (;;) {
      // A leaf's triangles are tested by the lanes that have reached one
      // together, once kLeafWaitNum / kLeafWaitDen of the lanes still walking
      // have (or all of them): a lane waits at its leaf, so its own order of
      // tests is unchanged, and the wave runs the triangle loop for many lanes
      // at once instead of for one or two in almost every step.
      const uint64_t waiting = __ballot(leaf >= 0);
      if (waiting != 0ull && __popcll(waiting) * kLeafWaitDen >= __popcll(__ballot(true)) * kLeafWaitNum) {
        if (leaf >= 0) {
          for (int k = leaf;; ++k) {
            const DTri tr = at(S.tris, k);
            unit(tri_tests);
            const float cur = isect_triangle(oxo, oyo, ozo, dxo, dyo, dzo, tr, closest);
            if (cur < closest) {
              closest = cur;
              ci = pi;
              cst = 4;
              cso = tr.tri_offset;
            }
            if (testing && ci != -1) return true;
            if (tr.last) break;
          }
          leaf = -1;
        }
        continue;
      }
      if (leaf >= 0) continue;
      if (idx == -1) {  // this lane's next mesh
        if (!mine) break;
        pi = base + __builtin_ctzll(mine);
        mine &= mine - 1;
        unit(prim_tests);
        const DPrim& pr = S.prims[pi];
        idx = pr.node_root;
        // The object-space ray and its reciprocals, unless the previous mesh
        // had the same transform (then they are these bits already).
        if (pr.xform != xform) {
          xform = pr.xform;
          const float m0 = pr.w2o[0], m1 = pr.w2o[1], m2 = pr.w2o[2], m3 = pr.w2o[3];
          const float m4 = pr.w2o[4], m5 = pr.w2o[5], m6 = pr.w2o[6], m7 = pr.w2o[7];
          const float m8 = pr.w2o[8], m9 = pr.w2o[9], m10 = pr.w2o[10], m11 = pr.w2o[11];
          oxo = m0 * ox + m1 * oy + m2 * oz + m3;
          oyo = m4 * ox + m5 * oy + m6 * oz + m7;
          ozo = m8 * ox + m9 * oy + m10 * oz + m11;
          dxo = m0 * dx + m1 * dy + m2 * dz;
          dyo = m4 * dx + m5 * dy + m6 * dz;
          dzo = m8 * dx + m9 * dy + m10 * dz;
          rdx = recip_rn(dxo);
          rdy = recip_rn(dyo);
          rdz = recip_rn(dzo);
        }
      }
      // One node visit: the box test and the next node (or the leaf whose
      // triangles come first).
      auto visit = [&]() {
        const DNode nd = at(S.nodes, idx);
        unit(box_tests);
        const float t0x = (nd.mn[0] - oxo) * rdx, t0y = (nd.mn[1] - oyo) * rdy, t0z = (nd.mn[2] - ozo) * rdz;
        const float t1x = (nd.mx[0] - oxo) * rdx, t1y = (nd.mx[1] - oyo) * rdy, t1z = (nd.mx[2] - ozo) * rdz;
        const float tmax = jmin(jmax(t0x, t1x), jmin(jmax(t0y, t1y), jmax(t0z, t1z)));
        const float tmin = jmax(jmin(t0x, t1x), jmax(jmin(t0y, t1y), jmin(t0z, t1z)));
        // The box is entered when (:1122-1134) its entry or exit lies in
        // (0.001, closest), unless tmin > tmax (false for NaN); written without
        // branches, as selects.
        const int in_min = static_cast<int>(tmin > 0.001f) & static_cast<int>(tmin < closest);
        const int in_max = static_cast<int>(tmax > 0.001f) & static_cast<int>(tmax < closest);
        const bool hit = static_cast<bool>(static_cast<int>(!(tmin > tmax)) & (in_min | in_max));
        const bool tree = nd.tri < 0;
        leaf = hit && !tree ? nd.tri : -1;                   // a leaf's triangles come before the next visit
        idx = hit && tree ? -nd.tri - 2 : nd.next;           // tree node: descend to the left child
      };
      // Up to kVisitsPerStep visits per step of the loop while the lane meets
      // neither a leaf nor its mesh's end: the loop's own work (the leaf
      // batch ballots, the mesh switch) is paid once per few visits.  Visits
      // only happen between leaf tests, which alone change `closest`, so each
      // lane's sequence of tests is unchanged.
      visit();
#pragma unroll
      for (int v = 1; v < kVisitsPerStep; ++v)
        if (leaf < 0 && idx != -1) visit();
    }
    return false;
  }

  // walk_meshes over the two-child records (DPair): one 64-byte fetch tests
  // both children of a tree node.  Each child's slab test does not depend on
  // the closest hit, only the decision does: the box is entered when its
  // entry or exit lies in (0.001, closest) and tmin <= tmax (:1122-1134),
  // which is `key < closest` for key = tmin if tmin > 0.001, else tmax if
  // tmax > 0.001, else never (and never if tmin > tmax; NaN compares false
  // throughout).  The first child is decided at once; the second child's key
  // waits on the lane's stack (when it can still be entered: closest only
  // decreases) until the first child's subtree is done and is then compared
  // with the closest hit of that moment, so every box is decided where the
  // pre-order skip walk tests it, against the same closest hit, and the
  // triangles are tested in the same order.  The mesh's root is tested from
  // its DNode.  The host provides these records only when every tree node has
  // two children and the stack never holds more than S.pair_stack entries.
  __device__ __forceinline__ static float pair_key(float tmin, float tmax) {
    const float never = __builtin_inff();
    const float k = tmin > 0.001f ? tmin : (tmax > 0.001f ? tmax : never);
    return tmin > tmax ? never : k;
  }

  __device__ __forceinline__ bool walk_meshes_pairs(uint64_t mine, int base, float ox, float oy, float oz,
                                                    float dx, float dy, float dz, float& closest, int& ci, int& cst,
                                                    int& cso, bool testing) {
    float oxo = 0.0f, oyo = 0.0f, ozo = 0.0f, dxo = 0.0f, dyo = 0.0f, dzo = 0.0f;
    float rdx = 0.0f, rdy = 0.0f, rdz = 0.0f;
    int cur = -1;  // a tree node whose children are tested next
    int sp = 0;    // entries on the lane's stack
    int pi = 0, xform = -1;
    int leaf = -1;  // first triangle of the leaf this lane has reached, its tests pending
    for (;;) {
      // leaves' triangles in batches, as in walk_meshes
      const uint64_t waiting = __ballot(leaf >= 0);
      if (waiting != 0ull && __popcll(waiting) * kLeafWaitDen >= __popcll(__ballot(true)) * kLeafWaitNum) {
        if (leaf >= 0) {
          for (int k = leaf;; ++k) {
            const DTri tr = at(S.tris, k);
            unit(tri_tests);
            const float cur_t = isect_triangle(oxo, oyo, ozo, dxo, dyo, dzo, tr, closest);
            if (cur_t < closest) {
              closest = cur_t;
              ci = pi;
              cst = 4;
              cso = tr.tri_offset;
            }
            if (testing && ci != -1) return true;
            if (tr.last) break;
          }
          leaf = -1;
        }
        continue;
      }
      if (leaf >= 0) continue;
      if (cur < 0) {
        if (sp > 0) {  // the deepest pending second child
          --sp;
          const float k = pstack_key[sp * kPairStride];
          const int r = pstack_ref[sp * kPairStride];
          if (k < closest) {
            if (r >= 0) cur = r;
            else if (r <= -2) leaf = -r - 2;
          }
          continue;
        }
        // this lane's next mesh: its root box from the DNode
        if (!mine) break;
        pi = base + __builtin_ctzll(mine);
        mine &= mine - 1;
        unit(prim_tests);
        const DPrim& pr = S.prims[pi];
        if (pr.xform != xform) {
          xform = pr.xform;
          const float m0 = pr.w2o[0], m1 = pr.w2o[1], m2 = pr.w2o[2], m3 = pr.w2o[3];
          const float m4 = pr.w2o[4], m5 = pr.w2o[5], m6 = pr.w2o[6], m7 = pr.w2o[7];
          const float m8 = pr.w2o[8], m9 = pr.w2o[9], m10 = pr.w2o[10], m11 = pr.w2o[11];
          oxo = m0 * ox + m1 * oy + m2 * oz + m3;
          oyo = m4 * ox + m5 * oy + m6 * oz + m7;
          ozo = m8 * ox + m9 * oy + m10 * oz + m11;
          dxo = m0 * dx + m1 * dy + m2 * dz;
          dyo = m4 * dx + m5 * dy + m6 * dz;
          dzo = m8 * dx + m9 * dy + m10 * dz;
          rdx = recip_rn(dxo);
          rdy = recip_rn(dyo);
          rdz = recip_rn(dzo);
        }
        const int root = pr.node_root;
        const DNode nd = at(S.nodes, root);
        unit(box_tests);
        const float t0x = (nd.mn[0] - oxo) * rdx, t0y = (nd.mn[1] - oyo) * rdy, t0z = (nd.mn[2] - ozo) * rdz;
        const float t1x = (nd.mx[0] - oxo) * rdx, t1y = (nd.mx[1] - oyo) * rdy, t1z = (nd.mx[2] - ozo) * rdz;
        const float tmax = jmin(jmax(t0x, t1x), jmin(jmax(t0y, t1y), jmax(t0z, t1z)));
        const float tmin = jmax(jmin(t0x, t1x), jmax(jmin(t0y, t1y), jmin(t0z, t1z)));
        if (pair_key(tmin, tmax) < closest) {
          if (nd.tri >= 0) leaf = nd.tri;
          else if (-nd.tri - 2 != nd.next) cur = root;  // (a tree-coded node whose left is its skip target is empty)
        }
        continue;
      }
      // Both children of `cur`; as in walk_meshes, up to kVisitsPerStep of
      // these per step of the loop while the lane descends into tree nodes.
      auto expand = [&]() {
        const DPair P = at(S.pairs, cur);
        unit(box_tests);
        unit(box_tests);
        float key[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const float t0x = (P.mn[0][c] - oxo) * rdx, t0y = (P.mn[1][c] - oyo) * rdy, t0z = (P.mn[2][c] - ozo) * rdz;
          const float t1x = (P.mx[0][c] - oxo) * rdx, t1y = (P.mx[1][c] - oyo) * rdy, t1z = (P.mx[2][c] - ozo) * rdz;
          const float tmax = jmin(jmax(t0x, t1x), jmin(jmax(t0y, t1y), jmax(t0z, t1z)));
          const float tmin = jmax(jmin(t0x, t1x), jmax(jmin(t0y, t1y), jmin(t0z, t1z)));
          key[c] = pair_key(tmin, tmax);
        }
        int r = -1;
        if (key[0] < closest) {
          if (key[1] < closest) {  // the second child waits for the first child's subtree
            pstack_key[sp * kPairStride] = key[1];
            pstack_ref[sp * kPairStride] = P.ref[1];
            ++sp;
          }
          r = P.ref[0];
        } else if (key[1] < closest) {
          r = P.ref[1];
        }
        cur = r >= 0 ? r : -1;
        leaf = r <= -2 ? -r - 2 : -1;
      };
      expand();
#pragma unroll
      for (int v = 1; v < kVisitsPerStep; ++v)
        if (cur >= 0) expand();
    }
    return false;
  }

  // The primitive loop of doIntersectPrimitives: the closest (or, testing,
  // the first) hit's distance, primitive index, shape type and shape offset.
  //
  // An any-hit query (testing) keeps `closest` at the no-hit value until it
  // ends, so no test in it depends on what was tested before: each shape's
  // test compares with `closest` only to reject (the box cull's
  // entry-or-exit-before-closest rule then reduces to "entry or exit
  // ahead"), and the query ends at the first hit.  Whether some primitive is
  // hit is therefore the same in any order, and that is all its callers read
  // (the returned distance is only compared with the no-hit value).  So an
  // any-hit query tests a block's primitives other than meshes first, in
  // index order, and walks its mesh runs after them: a ray blocked by the
  // terrain, a sphere, a plane or a single triangle skips the BVH walks
  // (trace 9.93 -> 9.25 ms per 1080p pass, box tests -27 %, the film's bits
  // unchanged).  A closest-hit query runs the reference's index order.
  template <bool testing>
  __device__ __forceinline__ float query(float ox, float oy, float oz, float dx, float dy, float dz, int& ci,
                                         int& cst, int& cso) {
    ++rays;
    float closest = DF_INF;
    ci = -1;
    cst = -1;
    cso = -1;
    const int n = S.prim_count;
    // Primitives in blocks of 64: first a branch-free pass over the block's
    // bounding spheres marks, per lane, the primitives the ray can reach
    // (approximate reciprocal and square root: the spheres' enlargement
    // covers their error), then the marked primitives run (in index order,
    // or for an any-hit query in the two phases described above).
    const float dd = dx * dx + dy * dy + dz * dz;
    const float ia = __builtin_amdgcn_rcpf(dd), dl = __builtin_amdgcn_sqrtf(dd);
    for (int base = 0; base < n; base += 64) {
      const int m = n - base < 64 ? n - base : 64;
      uint64_t cand = 0, any = 0;  // this lane's marks; the wave's union
      for (int j0 = 0; j0 < m; j0 += 8) {  // the sphere array is padded to a multiple of 8
        unsigned c8 = 0;                     // marks of primitives j0 .. j0 + 7
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int j = j0 + u;
          const DCull cu = load_uniform(S.cull, base + j);  // wave-uniform: scalar loads
          // This test is no arithmetic of the reference's, only a conservative
          // filter, so it takes fused multiply-adds: each rounds once where the
          // products and sums it replaces round twice, so its error stays
          // within what the DCull allowance was sized for.
          const float lx = cu.c[0] - ox, ly = cu.c[1] - oy, lz = cu.c[2] - oz;
          const float l2 = __builtin_fmaf(lx, lx, __builtin_fmaf(ly, ly, lz * lz));
          const float tca = __builtin_fmaf(lx, dx, __builtin_fmaf(ly, dy, lz * dz));
          // the enlarged radius r + 1e-4 |c - o| + k |c - o|^2, bounded above
          // without a square root as a + b |c - o|^2 (DCull)
          const float R = __builtin_fmaf(cu.k, l2, cu.r);
          const float d2 = __builtin_fmaf(-(tca * tca), ia, l2);
          // comparisons are false for NaN: a ray with non-finite terms is never skipped
          const int miss = static_cast<int>(d2 > R * R) | static_cast<int>(tca < -R * dl);
          const int keep = static_cast<int>(j < m) & ((cu.on == 0) | (miss ^ 1));
          c8 |= static_cast<unsigned>(keep) << u;
          any |= __ballot(keep) != 0 ? uint64_t{1} << j : uint64_t{0};
        }
        cand |= static_cast<uint64_t>(c8) << j0;
      }
      // testing: phase 0 the primitives other than meshes, phase 1 the mesh runs
      const uint64_t marked = any;
      for (int phase = testing ? 0 : 1; phase < 2; ++phase) {
        any = marked;
        while (any) {
          const int j = __builtin_ctzll(any);
          const DPrim pr = load_uniform(S.prims, base + j);
          if (testing && phase == 0 && pr.type == 5) {  // a mesh run: left for phase 1
            const int end = (pr.run_end < base + m ? pr.run_end : base + m) - base;
            any &= ~((end >= 64 ? ~uint64_t{0} : ((uint64_t{1} << end) - 1)) & ~((uint64_t{1} << j) - 1));
            continue;
          }
          if (testing && phase == 1 && pr.type != 5) {  // tested in phase 0
            any &= any - 1;
            continue;
          }
          if (pr.type == 5) {
            // A run of consecutive meshes (a single mesh is a run of one): each
            // lane walks its own marked meshes of the run, in index order, in one
            // loop whose steps are single node visits, so a lane whose walk ends
            // moves straight on to its next mesh instead of waiting for the wave.
            // Every lane's own sequence of tests, and so every result, is the
            // reference's.
            const int end = (pr.run_end < base + m ? pr.run_end : base + m) - base;
            const uint64_t run = (end >= 64 ? ~uint64_t{0} : ((uint64_t{1} << end) - 1)) & ~((uint64_t{1} << j) - 1);
            any &= ~run;
            bool done;
            if constexpr (kPairs)
              done = walk_meshes_pairs(cand & run, base, ox, oy, oz, dx, dy, dz, closest, ci, cst, cso, testing);
            else
              done = walk_meshes(cand & run, base, ox, oy, oz, dx, dy, dz, closest, ci, cst, cso, testing);
            if (done) return closest;
            continue;
          }
          any &= any - 1;
          if (!((cand >> j) & 1)) continue;
          const int i = base + j;
          const int st = pr.type;
          unit(prim_tests);
          const float* m = pr.w2o;
          const float oxo = m[0] * ox + m[1] * oy + m[2] * oz + m[3];
          const float oyo = m[4] * ox + m[5] * oy + m[6] * oz + m[7];
          const float ozo = m[8] * ox + m[9] * oy + m[10] * oz + m[11];
          const float dxo = m[0] * dx + m[1] * dy + m[2] * dz;
          const float dyo = m[4] * dx + m[5] * dy + m[6] * dz;
          const float dzo = m[8] * dx + m[9] * dy + m[10] * dz;
          float cur = DF_INF;
          if (st == 1) {
            cur = isect_plane(oxo, oyo, ozo, dxo, dyo, dzo, pr.geo[0], pr.geo[1], pr.geo[2], pr.geo[3], pr.geo[4],
                              pr.geo[5], closest);
          } else if (st == 2) {
            unit(sphere_tests);
            cur = isect_sphere(oxo, oyo, ozo, dxo, dyo, dzo, pr.geo[0], pr.geo[1], pr.geo[2], pr.geo[3], closest);
          } else if (st == 3) {
            cur = isect_terrain<M>(oxo, oyo, ozo, dxo, dyo, dzo, closest, S.march_tc, S.march_n,
                                   kUnits ? &terrain_steps : nullptr);
          } else if (st == 4) {
            const DTri tr = S.tris[pr.node_root];
            unit(tri_tests);
            cur = isect_triangle(oxo, oyo, ozo, dxo, dyo, dzo, tr, closest);
          }
          if (cur < closest) {
            closest = cur;
            ci = i;
            cst = st;
            cso = pr.shape_offset;
          }
          if (testing && ci != -1) return closest;
        }
      }
    }
    return closest;
  }

  // The hit record of the query's result (the tail of doIntersectPrimitives):
  // surface point, normals, texture coordinates, normal mapping.
  __device__ __forceinline__ void fill_hit(float ox, float oy, float oz, float dx, float dy, float dz,
                                           float closest, int ci, int cst, int cso) {
    {
      if (ci != -1) {
        const DPrim& pr = S.prims[ci];
        const float* m = pr.w2o;
        const float oxo = m[0] * ox + m[1] * oy + m[2] * oz + m[3];
        const float oyo = m[4] * ox + m[5] * oy + m[6] * oz + m[7];
        const float ozo = m[8] * ox + m[9] * oy + m[10] * oz + m[11];
        const float dxo = m[0] * dx + m[1] * dy + m[2] * dz;
        const float dyo = m[4] * dx + m[5] * dy + m[6] * dz;
        const float dzo = m[8] * dx + m[9] * dy + m[10] * dz;
        h.t = closest;
        h.prim = ci * 3;
        h.shape_type = cst;
        h.shape_off = cso;
        if (cst == 1) {
          props_plane(oxo, oyo, ozo, dxo, dyo, dzo, closest, cso);
        } else if (cst == 2) {
          props_sphere(oxo, oyo, ozo, dxo, dyo, dzo, closest, pr);
        } else if (cst == 3) {
          props_terrain(oxo, oyo, ozo, dxo, dyo, dzo, closest);
        } else if (cst == 4) {
          props_triangle(oxo, oyo, ozo, dxo, dyo, dzo, closest, cso);
        }
        to_world(pr);
        normal_map_image(pr);
        normal_map_noise(pr);
      } else {
        h.t = DF_INF;
        h.prim = -1;
        h.shape_type = -1;
        h.shape_off = -1;
      }
    }
  }

  // ---------------------------------------------------------- surface properties

  __device__ __forceinline__ void props_plane(float ox, float oy, float oz, float dx, float dy, float dz, float t,
                              int so) {  // :2349-2450 (UV part; the plane ONB is never read)
    const int* pl = S.planes + so;
    const float* A = S.point3fs + pl[0];
    const float* B = S.point3fs + pl[1];
    const float* Cc = S.point3fs + pl[2];
    const float* N = S.vector3fs + pl[3];
    const float aX = A[0], aY = A[1], aZ = A[2], bX = B[0], bY = B[1], bZ = B[2], cX = Cc[0], cY = Cc[1], cZ = Cc[2];
    const float nX = N[0], nY = N[1], nZ = N[2];
    const float pX = ox + dx * t, pY = oy + dy * t, pZ = oz + dz * t;
    const float anX = fabsf(nX), anY = fabsf(nY), anZ = fabsf(nZ);
    const bool isX = anX > anY && anX > anZ;
    const bool isY = anY > anZ;
    const float a1X = isX ? aY : isY ? aZ : aX;
    const float a1Y = isX ? aZ : isY ? aX : aY;
    const float b1X = isX ? cY - aX : isY ? cZ - aX : cX - aX;
    const float b1Y = isX ? cZ - aY : isY ? cX - aY : cY - aY;
    const float c1X = isX ? bY - aX : isY ? bZ - aX : bX - aX;
    const float c1Y = isX ? bZ - aY : isY ? bX - aY : bY - aY;
    const float det = b1X * c1Y - b1Y * c1X;
    const float dr = recip_rn(det);
    const float bNU = -b1Y * dr, bNV = b1X * dr, bND = (b1Y * a1X - b1X * a1Y) * dr;
    const float cNU = c1Y * dr, cNV = -c1X * dr, cND = (c1X * a1Y - c1Y * a1X) * dr;
    const float hU = isX ? pY : isY ? pZ : pX;
    const float hV = isX ? pZ : isY ? pX : pY;
    h.uv[0] = hU * bNU + hV * bNV + bND;
    h.uv[1] = hU * cNU + hV * cNV + cND;
    h.p[0] = pX;
    h.p[1] = pY;
    h.p[2] = pZ;
    h.ns[0] = nX;
    h.ns[1] = nY;
    h.ns[2] = nZ;
  }

  __device__ __forceinline__ void props_sphere(float ox, float oy, float oz, float dx, float dy, float dz, float t,
                               const DPrim& pr) {  // :2452-2492
    const float pX = ox + dx * t, pY = oy + dy * t, pZ = oz + dz * t;
    const float n0X = pX - pr.geo[0], n0Y = pY - pr.geo[1], n0Z = pZ - pr.geo[2];
    const float r = jrsqrt(n0X * n0X + n0Y * n0Y + n0Z * n0Z);
    const float nX = n0X * r, nY = n0Y * r, nZ = n0Z * r;
    h.uv[0] = 0.5f + Math<M>::atan2(-nZ, -nX) * DF_PI_2_RECIP;
    h.uv[1] = 0.5f - Math<M>::asinpi(-nY);
    h.p[0] = pX;
    h.p[1] = pY;
    h.p[2] = pZ;
    h.ns[0] = nX;
    h.ns[1] = nY;
    h.ns[2] = nZ;
  }

  __device__ __forceinline__ void props_terrain(float ox, float oy, float oz, float dx, float dy, float dz, float t) {  // :2495-2539
    const float sr = 1.0f / 10.0f;
    const float pX = ox * sr + dx * t, pY = oy * sr + dy * t, pZ = oz * sr + dz * t;
    const float e = 0.02f;
    const float nX = (Math<M>::sin(pX - e) * Math<M>::sin(pZ)) - (Math<M>::sin(pX + e) * Math<M>::sin(pZ));
    const float nY = 2.0f * e;
    const float nZ = (Math<M>::sin(pX) * Math<M>::sin(pZ - e)) - (Math<M>::sin(pX) * Math<M>::sin(pZ + e));
    const float r = jrsqrt(nX * nX + nY * nY + nZ * nZ);
    h.p[0] = pX;
    h.p[1] = pY;
    h.p[2] = pZ;
    h.uv[0] = pX;
    h.uv[1] = pZ;
    h.ns[0] = nX * r;
    h.ns[1] = nY * r;
    h.ns[2] = nZ * r;
  }

  __device__ __forceinline__ void props_triangle(float ox, float oy, float oz, float dx, float dy, float dz, float t,
                                 int to) {  // :2541-2650
    const int32_t* T = S.triangles + to;
    const float* A = S.point3fs + T[0];
    const float* B = S.point3fs + T[1];
    const float* Cc = S.point3fs + T[2];
    const float pX = ox + dx * t, pY = oy + dy * t, pZ = oz + dz * t;
    const float abX = B[0] - A[0], abY = B[1] - A[1], abZ = B[2] - A[2];
    const float acX = Cc[0] - A[0], acY = Cc[1] - A[1], acZ = Cc[2] - A[2];
    const float v0X = dy * acZ - dz * acY, v0Y = dz * acX - dx * acZ, v0Z = dx * acY - dy * acX;
    const float v1X = ox - A[0], v1Y = oy - A[1], v1Z = oz - A[2];
    const float v2X = v1Y * abZ - v1Z * abY, v2Y = v1Z * abX - v1X * abZ, v2Z = v1X * abY - v1Y * abX;
    const float det = abX * v0X + abY * v0Y + abZ * v0Z;
    const float dr = recip_rn(det);
    const float bu = (v1X * v0X + v1Y * v0Y + v1Z * v0Z) * dr;
    const float bv = (dx * v2X + dy * v2Y + dz * v2Z) * dr;
    const float bw = 1.0f - bu - bv;
    const float* uvA = S.point2fs + T[6];
    const float* uvB = S.point2fs + T[7];
    const float* uvC = S.point2fs + T[8];
    h.uv[0] = bw * uvA[0] + bu * uvB[0] + bv * uvC[0];
    h.uv[1] = bw * uvA[1] + bu * uvB[1] + bv * uvC[1];
    h.t = t;
    h.p[0] = pX;
    h.p[1] = pY;
    h.p[2] = pZ;
    float n0X = 0.0f, n0Y = 0.0f, n0Z = 0.0f, t0X = 0.0f, t0Y = 0.0f, t0Z = 0.0f;
    if (P.shader_type == 1) {
      n0X = abY * acZ - abZ * acY;
      n0Y = abZ * acX - abX * acZ;
      n0Z = abX * acY - abY * acX;
    } else if (P.shader_type == 2) {
      const float* nA = S.vector3fs + T[3];
      const float* nB = S.vector3fs + T[4];
      const float* nC = S.vector3fs + T[5];
      const float* tA = S.vector3fs + T[9];
      const float* tB = S.vector3fs + T[10];
      const float* tC = S.vector3fs + T[11];
      n0X = nA[0] * bw + nB[0] * bu + nC[0] * bv;
      n0Y = nA[1] * bw + nB[1] * bu + nC[1] * bv;
      n0Z = nA[2] * bw + nB[2] * bu + nC[2] * bv;
      t0X = tA[0] * bw + tB[0] * bu + tC[0] * bv;
      t0Y = tA[1] * bw + tB[1] * bu + tC[1] * bv;
      t0Z = tA[2] * bw + tB[2] * bu + tC[2] * bv;
    }
    const float nr = jrsqrt(n0X * n0X + n0Y * n0Y + n0Z * n0Z);
    const float tr = jrsqrt(t0X * t0X + t0Y * t0Y + t0Z * t0Z);
    h.ns[0] = n0X * nr;
    h.ns[1] = n0Y * nr;
    h.ns[2] = n0Z * nr;
    h.tan[0] = t0X * tr;
    h.tan[1] = t0Y * tr;
    h.tan[2] = t0Z * tr;
  }

  __device__ __forceinline__ void to_world(const DPrim& pr) {  // :3998-4073 (the fields that are read)
    const float* w = pr.w2o;
    const float* o = pr.o2w;
    const float x = h.p[0], y = h.p[1], z = h.p[2];
    h.p[0] = o[0] * x + o[1] * y + o[2] * z + o[3];
    h.p[1] = o[4] * x + o[5] * y + o[6] * z + o[7];
    h.p[2] = o[8] * x + o[9] * y + o[10] * z + o[11];
    const float nx = h.ns[0], ny = h.ns[1], nz = h.ns[2];
    h.ns[0] = w[0] * nx + w[4] * ny + w[8] * nz;
    h.ns[1] = w[1] * nx + w[5] * ny + w[9] * nz;
    h.ns[2] = w[2] * nx + w[6] * ny + w[10] * nz;
    const float tx = h.tan[0], ty = h.tan[1], tz = h.tan[2];
    h.tan[0] = w[0] * tx + w[4] * ty + w[8] * tz;
    h.tan[1] = w[1] * tx + w[5] * ty + w[9] * tz;
    h.tan[2] = w[2] * tx + w[6] * ty + w[10] * tz;
  }

  // Whether the surface of primitive ci perturbs the shading normal with
  // noise (the test of normal_map_noise).
  // The shade bin (kShadeBins) of a closest-hit result: 0 for a miss, the hit
  // surface's material for 1..5, 6 for any other material value.
  __device__ __forceinline__ int shade_bin(int ci) const {
    if (ci < 0) return 0;
    const int mat = jf2i(S.surfaces[S.prims[ci].surface_offset]);
    return mat >= 1 && mat <= 5 ? mat : 6;
  }

  __device__ __forceinline__ bool has_noise_normals(int ci) const {
    if (P.normal_mapping != 1 || ci < 0) return false;
    const int so = S.prims[ci].surface_offset;
    return S.surfaces[so + 4] > 0.0f && S.surfaces[so + 5] > 0.0f;
  }

  // The three fBm values of the noise normal map at the hit point, for a
  // surface of noise scale `scale`.
  __device__ __forceinline__ void noise_normal_values(float scale, float n[3]) {
    const float sr = 1.0f / scale;
    const float x1 = h.p[0] * sr, y1 = h.p[1] * sr, z1 = h.p[2] * sr;
    n[0] = noise.fbm3(scale, 0.5f, -0.26f, 0.26f, 16, x1, y1, z1);
    n[1] = noise.fbm3(scale, 0.5f, -0.26f, 0.26f, 16, y1, z1, x1);
    n[2] = noise.fbm3(scale, 0.5f, -0.26f, 0.26f, 16, z1, x1, y1);
  }

  __device__ __forceinline__ void normal_map_noise(const DPrim& pr) {  // :2658-2712
    const int so = pr.surface_offset;
    const float amount = S.surfaces[so + 4];
    const float scale = S.surfaces[so + 5];
    if (P.normal_mapping == 1 && amount > 0.0f && scale > 0.0f) {
      if (noise_skip) return;  // the caller computes the values itself
      float n[3];
      if (noise_in) {  // computed by an earlier pass, from the same hit point
        const float4 q = noise_in[noise_index];
        n[0] = q.x;
        n[1] = q.y;
        n[2] = q.z;
      } else {
        noise_normal_values(scale, n);
      }
      apply_noise_normal(amount, n);
    }
  }

  // The shading normal perturbed by the noise values n with weight amount.
  __device__ __forceinline__ void apply_noise_normal(float amount, const float n[3]) {
    const float nx = n[0], ny = n[1], nz = n[2];
    const float s1X = h.ns[0] + nx * amount;
    const float s1Y = h.ns[1] + ny * amount;
    const float s1Z = h.ns[2] + nz * amount;
    const float r = jrsqrt(s1X * s1X + s1Y * s1Y + s1Z * s1Z);
    h.ns[0] = s1X * r;
    h.ns[1] = s1Y * r;
    h.ns[2] = s1Z * r;
  }

  __device__ __forceinline__ void normal_map_image(const DPrim& pr) {  // :2714-2846
    const int so = pr.surface_offset;
    const int to = jf2i(S.surfaces[so + 3]);
    const int tt = jf2i(S.textures[to]);
    if (!(P.normal_mapping == 1 && tt == 6)) return;
    const int rgb = tex_image_simple(to);
    const float r = ((((rgb >> 16) & 0xFF) * DF_COLOR_RECIP) - 0.5f) * 2.0f;
    const float g = ((((rgb >> 8) & 0xFF) * DF_COLOR_RECIP) - 0.5f) * 2.0f;
    const float b = ((((rgb >> 0) & 0xFF) * DF_COLOR_RECIP) - 0.0f) * 2.0f;
    const float wX = h.ns[0], wY = h.ns[1], wZ = h.ns[2];
    const int type = pr.type;
    float sX, sY, sZ;
    if (type == 2) {
      const float v0X = -2.0f * DF_PI * wY, v0Y = 2.0f * DF_PI * wX, v0Z = 0.0f;
      const float u0X = v0Y * wZ - v0Z * wY, u0Y = v0Z * wX - v0X * wZ, u0Z = v0X * wY - v0Y * wX;
      const float ur = jrsqrt(u0X * u0X + u0Y * u0Y + u0Z * u0Z);
      const float u1X = u0X * ur, u1Y = u0Y * ur, u1Z = u0Z * ur;
      const float v1X = wY * u1Z - wZ * u1Y, v1Y = wZ * u1X - wX * u1Z, v1Z = wX * u1Y - wY * u1X;
      const float vr = jrsqrt(v1X * v1X + v1Y * v1Y + v1Z * v1Z);
      const float v2X = v1X * vr, v2Y = v1Y * vr, v2Z = v1Z * vr;
      sX = r * u1X + g * v2X + b * wX;
      sY = r * u1Y + g * v2Y + b * wY;
      sZ = r * u1Z + g * v2Z + b * wZ;
    } else if (type == 4) {
      const float uX = h.tan[0], uY = h.tan[1], uZ = h.tan[2];
      const float v0X = wY * uZ - wZ * uY, v0Y = wZ * uX - wX * uZ, v0Z = wX * uY - wY * uX;
      const float vr = jrsqrt(v0X * v0X + v0Y * v0Y + v0Z * v0Z);
      const float v1X = v0X * vr, v1Y = v0Y * vr, v1Z = v0Z * vr;
      sX = r * uX + g * v1X + b * wX;
      sY = r * uY + g * v1Y + b * wY;
      sZ = r * uZ + g * v1Z + b * wZ;
    } else {
      const float aX = fabsf(wX), aY = fabsf(wY), aZ = fabsf(wZ);
      const bool isWX = aX < aY && aX < aZ;
      const bool isWY = aY < aZ;
      const float v0X = isWX ? 0.0f : isWY ? wZ : wY;
      const float v0Y = isWX ? wZ : isWY ? 0.0f : -wX;
      const float v0Z = isWX ? -wY : isWY ? -wX : 0.0f;
      const float vr = jrsqrt(v0X * v0X + v0Y * v0Y + v0Z * v0Z);
      const float v1X = v0X * vr, v1Y = v0Y * vr, v1Z = v0Z * vr;
      const float u0X = v1Y * wZ - v1Z * wY, u0Y = v1Z * wX - v1X * wZ, u0Z = v1X * wY - v1Y * wX;
      const float ur = jrsqrt(u0X * u0X + u0Y * u0Y + u0Z * u0Z);
      const float u1X = u0X * ur, u1Y = u0Y * ur, u1Z = u0Z * ur;
      sX = r * u1X + g * v1X + b * wX;
      sY = r * u1Y + g * v1Y + b * wY;
      sZ = r * u1Z + g * v1Z + b * wZ;
    }
    const float sr = jrsqrt(sX * sX + sY * sY + sZ * sZ);
    h.ns[0] = sX * sr;
    h.ns[1] = sY * sr;
    h.ns[2] = sZ * sr;
  }

  // ---------------------------------------------------------- textures (:1596-1960)

  __device__ __forceinline__ int tex_color(int off) {
    const int type = jf2i(S.textures[off]);
    if (type == 1) {
      const float* T = S.textures;
      const int ao = jf2i(T[off + 2]);
      const int bo = jf2i(T[off + 3]);
      const float f = T[off + 4];
      const int a = tex_color2(ao);
      const int b = tex_color2(bo);
      return quant3((1.0f - f) * ch_r(a) + f * ch_r(b), (1.0f - f) * ch_g(a) + f * ch_g(b),
                    (1.0f - f) * ch_b(a) + f * ch_b(b));
    }
    return tex_color2(off);
  }

  __device__ __forceinline__ int tex_color2(int off) {
    const float* T = S.textures;
    const int type = jf2i(T[off]);
    switch (type) {
      case 2: {  // Bullseye :1668-1712
        const int a = jf2i(T[off + 2]);
        const int b = jf2i(T[off + 3]);
        const float length = sqrtf(h.p[0] * h.p[0] + h.p[1] * h.p[1] + h.p[2] * h.p[2]);
        const float value = remainder_j(length * 0.25f, 1.0f);
        float cR = value > 0.5f ? ch_r(a) : ch_r(b);
        float cG = value > 0.5f ? ch_g(a) : ch_g(b);
        float cB = value > 0.5f ? ch_b(a) : ch_b(b);
        const float m = value > 0.5f ? 0.8f : 1.2f;
        if (ch_r(a) == ch_r(b) && ch_g(a) == ch_g(b) && ch_b(a) == ch_b(b)) {
          cR *= m;
          cG *= m;
          cB *= m;
        }
        return quant3(cR, cG, cB);
      }
      case 3: {  // Checkerboard :1714-1767
        const int c0 = jf2i(T[off + 2]);
        const int c1 = jf2i(T[off + 3]);
        const bool dark = checker_tile(off) == 0;
        const float m = dark ? 0.8f : 1.2f;
        float cR = dark ? ch_r(c0) : ch_r(c1);
        float cG = dark ? ch_g(c0) : ch_g(c1);
        float cB = dark ? ch_b(c0) : ch_b(c1);
        if (ch_r(c0) == ch_r(c1) && ch_g(c0) == ch_g(c1) && ch_b(c0) == ch_b(c1)) {
          cR *= m;
          cG *= m;
          cB *= m;
        }
        return quant3(cR, cG, cB);
      }
      case 4:  // Constant :1769-1774
        return jf2i(T[off + 2]);
      case 5: {  // fBm :1776-1815
        const int ad = jf2i(T[off + 2]);
        const int mu = jf2i(T[off + 3]);
        const float n = noise.fbm3(T[off + 4], T[off + 5], 0.0f, 1.0f, jf2i(T[off + 6]), h.p[0], h.p[1], h.p[2]);
        return quant3(n * ch_r(mu) + ch_r(ad), n * ch_g(mu) + ch_g(ad), n * ch_b(mu) + ch_b(ad));
      }
      case 6:
        return tex_image(off);
      case 7:  // SurfaceNormal :1931-1945
        return quant3((h.ns[0] + 1.0f) * 0.5f, (h.ns[1] + 1.0f) * 0.5f, (h.ns[2] + 1.0f) * 0.5f);
      case 8:  // UV :1947-1960
        return pack_rgb(saturate(h.uv[0], 0.0f, 1.0f) * 255.0f + 0.5f, saturate(h.uv[1], 0.0f, 1.0f) * 255.0f + 0.5f,
                        0.0f);
      default:
        return 0;
    }
  }

  // The tile of the checkerboard texture at `off` under the hit's uv
  // (:1741-1747): 0 where the texture takes its first colour, 1 the second.
  __device__ __forceinline__ int checker_tile(int off) {
    const float* T = S.textures;
    const float u = h.uv[0], v = h.uv[1];
    const float ca = T[off + 4], sa = T[off + 5];
    const float tu = modulo((u * ca - v * sa) * T[off + 6]);
    const float tv = modulo((v * ca + u * sa) * T[off + 7]);
    return (tu > 0.5f) != (tv > 0.5f) ? 0 : 1;
  }

  __device__ __forceinline__ void image_coords(int off, float* u2, float* v2, int* x, int* y, int* w, int* res) {
    const float* T = S.textures;
    const float u = h.uv[0], v = h.uv[1];
    const float width = T[off + 4], height = T[off + 5];
    const float ca = T[off + 2], sa = T[off + 3];
    const float u0 = (u * ca - v * sa);
    const float v0 = (v * ca + u * sa);
    const float u1 = remainder_j(u0 * T[off + 6] * width, width);
    const float v1 = remainder_j(v0 * T[off + 7] * height, height);
    *u2 = u1 >= 0.0f ? u1 : width - fabsf(u1);
    *v2 = v1 >= 0.0f ? v1 : height - fabsf(v1);
    *x = jf2i(*u2);
    *y = jf2i(*v2);
    *w = jf2i(width);
    *res = jf2i(width * height);
  }

  __device__ __forceinline__ int tex_image(int off) {  // :1817-1892
    float u2, v2;
    int x, y, w, res;
    image_coords(off, &u2, &v2, &x, &y, &w, &res);
    const float* D = S.textures + off + 8;
    const int i00 = y * w + x, i01 = y * w + x + 1, i10 = (y + 1) * w + x, i11 = (y + 1) * w + x + 1;
    const int c00 = i00 >= 0 && i00 < res ? jf2i(D[i00]) : 0;
    const int c01 = i01 >= 0 && i01 < res ? jf2i(D[i01]) : 0;
    const int c10 = i10 >= 0 && i10 < res ? jf2i(D[i10]) : 0;
    const int c11 = i11 >= 0 && i11 < res ? jf2i(D[i11]) : 0;
    const float fx = u2 - static_cast<float>(x);
    const float fy = v2 - static_cast<float>(y);
    const float r = lerp_j(lerp_j(ch_r(c00), ch_r(c01), fx), lerp_j(ch_r(c10), ch_r(c11), fx), fy);
    const float g = lerp_j(lerp_j(ch_g(c00), ch_g(c01), fx), lerp_j(ch_g(c10), ch_g(c11), fx), fy);
    const float b = lerp_j(lerp_j(ch_b(c00), ch_b(c01), fx), lerp_j(ch_b(c10), ch_b(c11), fx), fy);
    return quant3(r, g, b);
  }

  __device__ __forceinline__ int tex_image_simple(int off) {  // :1894-1929
    float u2, v2;
    int x, y, w, res;
    image_coords(off, &u2, &v2, &x, &y, &w, &res);
    const int i = y * w + x;
    return i >= 0 && i < res ? jf2i(S.textures[off + 8 + i]) : 0;
  }

  // ---------------------------------------------------------- lighting

  __device__ __forceinline__ int phong_shader0(bool check, float pX, float pY, float pZ, float nX, float nY, float nZ, float vX,
                               float vY, float vZ, float aR, float aG, float aB) {  // :1962-2039 (ka .2, kd .5, ks .5, ns 250)
    float r = 0.0f, g = 0.0f, b = 0.0f;
    r += aR * 0.2f * 1.0f;
    g += aG * 0.2f * 1.0f;
    b += aB * 0.2f * 1.0f;
    const float lX = K.sun_origin[0] - pX, lY = K.sun_origin[1] - pY, lZ = K.sun_origin[2] - pZ;
    const float l2 = lX * lX + lY * lY + lZ * lZ;
    const float lr = recip_rn(sqrtf(l2));
    const float lnX = lX * lr, lnY = lY * lr, lnZ = lZ * lr;
    if (check) intersect_t<false>(pX, pY, pZ, lnX, lnY, lnZ);
    const float t = check ? h.t : DF_INF;
    if (t >= DF_INF || t > l2) {
      const float ldn = lnX * nX + lnY * nY + lnZ * nZ;
      r += aR * 0.5f * jmax(ldn, 0.0f) * 1.0f;
      g += aG * 0.5f * jmax(ldn, 0.0f) * 1.0f;
      b += aB * 0.5f * jmax(ldn, 0.0f) * 1.0f;
      if (ldn > 0.0f) {
        const float rX = 2.0f * ldn * nX - lnX, rY = 2.0f * ldn * nY - lnY, rZ = 2.0f * ldn * nZ - lnZ;
        const float rr = jrsqrt(rX * rX + rY * rY + rZ * rZ);
        const float rdv = rX * rr * vX + rY * rr * vY + rZ * rr * vZ;
        const float sh = Math<M>::pow(jmax(rdv, 0.0f), 250.0f);
        r += 0.5f * sh * 1.0f;
        g += 0.5f * sh * 1.0f;
        b += 0.5f * sh * 1.0f;
      }
    }
    r = saturate(r, 0.0f, 1.0f);
    g = saturate(g, 0.0f, 1.0f);
    b = saturate(b, 0.0f, 1.0f);
    return ((jf2i(r * 255.0f + 0.5f) & 0xFF) << 16) | ((jf2i(g * 255.0f + 0.5f) & 0xFF) << 8) |
           (jf2i(b * 255.0f + 0.5f) & 0xFF);
  }

  __device__ __forceinline__ void sky(bool active, float dx, float dy, float dz) {  // :2127-2236
    float d0X = dx * K.onb_u[0] + dy * K.onb_u[1] + dz * K.onb_u[2];
    float d0Y = dx * K.onb_v[0] + dy * K.onb_v[1] + dz * K.onb_v[2];
    float d0Z = dx * K.onb_w[0] + dy * K.onb_w[1] + dz * K.onb_w[2];
    float r = 0.01f, g = 0.01f, b = 0.01f;
    if (d0Z >= 0.0f && active) {
      ++sky_evals;
      d0Z = jmax(d0Z, 0.001f);
      const float lr = jrsqrt(d0X * d0X + d0Y * d0Y + d0Z * d0Z);
      const float d1X = d0X * lr, d1Y = d0Y * lr, d1Z = d0Z * lr;
      const float dp = d1X * K.sun_dir[0] + d1Y * K.sun_dir[1] + d1Z * K.sun_dir[2];
      const double theta1 = ::acos(jmaxd(jmind(static_cast<double>(d1Z), 1.0), -1.0));
      const double cosTheta1 = ::cos(theta1);
      const double cosTheta1R = 1.0 / (cosTheta1 + 0.01);
      const double gamma = ::acos(jmaxd(jmind(static_cast<double>(dp), 1.0), -1.0));
      const double cosGamma = ::cos(gamma);
      const double* L = K.pl;
      const double* X = K.px;
      const double* Y = K.py;
      const double lnum = ((1.0 + L[0] * ::exp(L[1] * cosTheta1R)) * (1.0 + L[2] * ::exp(L[3] * gamma) + L[4] * cosGamma * cosGamma));
      const double lum = K.zenith_l * lnum / K.l_den * 1.0e-4;
      const double xden = (K.x_den0 * (1.0 + X[2] * ::exp(X[3] * theta1) + K.x_c));
      const double xnum = ((1.0 + X[0] * ::exp(X[1] * cosTheta1R)) * (1.0 + X[2] * ::exp(X[3] * gamma) + X[4] * cosGamma * cosGamma));
      const double x = K.zenith_x * xnum / xden;
      const double yden = (K.y_den0 * (1.0 + Y[2] * ::exp(Y[3] * theta1) + K.y_c));
      const double ynum = ((1.0 + Y[0] * ::exp(Y[1] * cosTheta1R)) * (1.0 + Y[2] * ::exp(Y[3] * gamma) + Y[4] * cosGamma * cosGamma));
      const double y = K.zenith_y * ynum / yden;
      const float fx = static_cast<float>(x), fy = static_cast<float>(y);
      const float cie0 = 1.0f / (0.0241f + 0.2562f * fx - 0.7341f * fy);
      const float cie1 = (-1.3515f - 1.7703f * fx + 5.9114f * fy) * cie0;
      const float cie2 = (0.03f - 31.4424f * fx + 30.0717f * fy) * cie0;
      const float cX = 10246.121f + cie1 * 187.75537f + cie2 * 213.14803f;
      const float cY = 10676.695f + cie1 * 192.59653f + cie2 * 76.29494f;
      const float cZ = 12372.504f + cie1 * 3482.8765f + cie2 * -235.71611f;
      const float cYR = 1.0f / cY;
      const float cR = static_cast<float>(static_cast<double>(cX) * lum * static_cast<double>(cYR));
      const float cG = static_cast<float>(lum);
      const float cB = static_cast<float>(static_cast<double>(cZ) * lum * static_cast<double>(cYR));
      r = 3.2410042f * cR + -1.5373994f * cG + -0.49861607f * cB;
      g = -0.9692241f * cR + 1.8759298f * cG + 0.041554242f * cB;
      b = 0.05563942f * cR + -0.20401107f * cG + 1.0571486f * cB;
      const float w = jmax(0.0f, -jmin(0.0f, jmin(r, jmin(g, b))));
      r += w;
      g += w;
      b += w;
    }
    tmp[0] = r;
    tmp[1] = g;
    tmp[2] = b;
  }

  // Samples a direction towards the sun.  Returns true when a shadow query
  // towards `sd` decides the result; tmp then holds the contribution when
  // nothing is hit (and 0 otherwise, as it also does on false).
  // A direction in the cone towards the sun of sky k for the draws (u, v);
  // returns dp0 = d . sun direction (:2250-2329).
  __device__ __forceinline__ float sun_cone_dir(const DSky& k, float pX, float pY, float pZ, float u, float v,
                                                float d[3]) {
    const float sdX = k.sun_dir_world[0], sdY = k.sun_dir_world[1], sdZ = k.sun_dir_world[2];
    const float sunR2 = 500.0f * 500.0f;
    const float toX = k.sun_origin[0] - pX, toY = k.sun_origin[1] - pY, toZ = k.sun_origin[2] - pZ;
    const float to2 = toX * toX + toY * toY + toZ * toZ;
    const float sinMax2 = sunR2 / to2;
    const float cosMax = sqrtf(jmax(0.0f, 1.0f - sinMax2));
    const float cosT = u * (cosMax - 1.0f) + 1.0f;
    const float sinT = sqrtf(jmax(0.0f, 1.0f - cosT * cosT));
    const float phi = DF_PI_2 * v;
    const float x = Math<M>::cos(phi) * sinT, y = Math<M>::sin(phi) * sinT, z = cosT;
    const float wX = sdX, wY = sdY, wZ = sdZ;
    const float aX = fabsf(wX), aY = fabsf(wY), aZ = fabsf(wZ);
    const bool c0 = aX < aY && aX < aZ;
    const bool c1 = aY < aZ;
    const float vX = c0 ? 0.0f : c1 ? wZ : wY;
    const float vY = c0 ? wZ : c1 ? 0.0f : -wX;
    const float vZ = c0 ? -wY : c1 ? -wX : 0.0f;
    const float vr = jrsqrt(vX * vX + vY * vY + vZ * vZ);
    const float vnX = vX * vr, vnY = vY * vr, vnZ = vZ * vr;
    const float unX = vnY * wZ - vnZ * wY, unY = vnZ * wX - vnX * wZ, unZ = vnX * wY - vnY * wX;
    const float rX = unX * x + vnX * y + wX * z;
    const float rY = unY * x + vnY * y + wY * z;
    const float rZ = unZ * x + vnZ * y + wZ * z;
    const float rr = jrsqrt(rX * rX + rY * rY + rZ * rZ);
    d[0] = rX * rr;
    d[1] = rY * rr;
    d[2] = rZ * rr;
    return d[0] * sdX + d[1] * sdY + d[2] * sdZ;
  }

  // Samples a direction towards the sun (two draws, returned in uv).
  // Returns true when a shadow query towards `sd` decides the result; tmp
  // then holds the contribution when nothing is hit (and 0 otherwise, as it
  // also does on false).
  __device__ __forceinline__ bool sun_sample(bool active, float pX, float pY, float pZ, float nX, float nY, float nZ,
                                             float aR, float aG, float aB, float sd[3],
                                             float uv[2]) {  // :2238-2347
    bool query_needed = false;
    float r = 0.0f, g = 0.0f, b = 0.0f;
    if (active) {
      const float u = rng.next();
      const float v = rng.next();
      uv[0] = u;
      uv[1] = v;
      float d[3];
      const float dp0 = sun_cone_dir(K, pX, pY, pZ, u, v, d);
      const float dp1 = d[0] * nX + d[1] * nY + d[2] * nZ;
      if (dp0 > 0.0f && dp1 > 0.0f) {
        query_needed = true;
        sd[0] = d[0];
        sd[1] = d[1];
        sd[2] = d[2];
        r = aR * 1.0f * dp1 * DF_PI_RECIP;
        g = aG * 1.0f * dp1 * DF_PI_RECIP;
        b = aB * 1.0f * dp1 * DF_PI_RECIP;
      }
    }
    tmp[0] = r;
    tmp[1] = g;
    tmp[2] = b;
    return query_needed;
  }

  // dp1 / pi of the sun sample of sky k at the draws uv (0 where the sample
  // is rejected): the factor of the albedo in the sun term.
  __device__ __forceinline__ float sun_weight(const DSky& k, float pX, float pY, float pZ, float nX, float nY,
                                              float nZ, const float uv[2]) {
    float d[3];
    const float dp0 = sun_cone_dir(k, pX, pY, pZ, uv[0], uv[1], d);
    const float dp1 = d[0] * nX + d[1] * nY + d[2] * nZ;
    return dp0 > 0.0f && dp1 > 0.0f ? dp1 * DF_PI_RECIP : 0.0f;
  }

  // The shadow query's verdict: nothing hit (:2331).
  __device__ __forceinline__ bool sun_visible(float pX, float pY, float pZ, const float sd[3]) {
    return intersect_t<true>(pX, pY, pZ, sd[0], sd[1], sd[2]) >= DF_INF - 0.0001f;
  }

  // Cosine-weighted direction about w (:3410-3444, :3466-3500).
  __device__ __forceinline__ void cosine_dir(float wX, float wY, float wZ, float* o) {
    const float u = rng.next();
    const float v = rng.next();
    const float phi = DF_PI_2 * u;
    const float ct = sqrtf(v), st = sqrtf(1.0f - v);
    const float x = Math<M>::cos(phi) * st, y = Math<M>::sin(phi) * st, z = ct;
    const bool isY = fabsf(wX) > 0.1f;
    const float uX = (isY ? 1.0f : 0.0f) * wZ;
    const float uY = -((isY ? 0.0f : 1.0f) * wZ);
    const float uZ = (isY ? 0.0f : 1.0f) * wY - (isY ? 1.0f : 0.0f) * wX;
    const float ur = jrsqrt(uX * uX + uY * uY + uZ * uZ);
    const float unX = uX * ur, unY = uY * ur, unZ = uZ * ur;
    const float vnX = wY * unZ - wZ * unY, vnY = wZ * unX - wX * unZ, vnZ = wX * unY - wY * unX;
    const float lX = unX * x + vnX * y + wX * z;
    const float lY = unY * x + vnY * y + wY * z;
    const float lZ = unZ * x + vnZ * y + wZ * z;
    const float lr = jrsqrt(lX * lX + lY * lY + lZ * lZ);
    o[0] = lX * lr;
    o[1] = lY * lr;
    o[2] = lZ * lr;
  }

  // ---------------------------------------------------------- renderers

  // One bounce of the path tracer (:3196-3688), after the closest-hit query
  // of the ray (o, d) has filled the hit record: sky on a miss, emission and
  // sun light, Russian roulette, then the material's scattered ray.
  //
  // The sun's shadow query is handed back instead of traced: with `shadow`
  // set the caller traces (h.p, sd) with the any-hit query and, when nothing
  // is hit, replaces p by pv, the radiance including the sun's contribution
  // (the two differ only in that one term, so either order of evaluation
  // gives the reference's bits).  The Russian-roulette exit re-tests the
  // path's current ray with the any-hit query (:3224-3231); that ray has just
  // hit a surface, the any-hit query runs the closest-hit query's tests up to
  // that surface's primitive, so it always hits and its sky term is never
  // added: the query is counted, not traced.
  //
  // Returns true while the path continues with the new ray in (o, d) and
  // throughput m; false when it has ended with radiance p (pending the shadow
  // query).  depth counts the bounces taken.
  // bin: the shade bin (kShadeBins) the path was binned in, or -1 (none).
  // A bin fixes whether the path missed and, for bins 1..5, the material of
  // the hit surface; a wave draining one bin then takes the same branches of
  // the material switch on every lane.  The arithmetic of every path is the
  // same for every bin.
  // kDeferSky: a missed path leaves its radiance without the sky term and
  // sets sky_pending; the caller adds the term m * sky(d) / pi to the
  // radiance later, as the same single float addition, and so carries no
  // sky code of its own.
  template <bool kDeferSky = false>
  __device__ __forceinline__ bool pt_bounce(float o[3], float d[3], float m[3], float p[3], int& depth,
                                            bool sky_active, bool sun_active, bool& shadow, float sd[3],
                                            float pv[3], Vtx* vtx = nullptr, int bin = -1) {
    const int depth_rr = P.depth_rr;
    float ox = o[0], oy = o[1], oz = o[2], dx = d[0], dy = d[1], dz = d[2];
    float pr = p[0], pg = p[1], pb = p[2];
    float mr = m[0], mg = m[1], mb = m[2];
    shadow = false;
    const int po = h.prim;
    const bool missed = bin == 0 ? true : bin > 0 ? false : (h.t == DF_INF || po == -1);
    if (missed) {
      if (vtx) vtx->miss = true;
      if constexpr (kDeferSky) {
        sky_pending = true;
        return false;
      }
      sky(sky_active, dx, dy, dz);
      pr += mr * tmp[0] * DF_PI_RECIP;
      pg += mg * tmp[1] * DF_PI_RECIP;
      pb += mb * tmp[2] * DF_PI_RECIP;
      p[0] = pr;
      p[1] = pg;
      p[2] = pb;
      return false;
    }
    const DPrim& prm = S.prims[po / 3];
    const int so = prm.surface_offset;
    const float pX = h.p[0], pY = h.p[1], pZ = h.p[2];
    const float sX = h.ns[0], sY = h.ns[1], sZ = h.ns[2];
    const float dp = sX * dx + sY * dy + sZ * dz;
    const float dp2 = dp * 2.0f;
    const bool ok = dp < 0.0f;
    const float wX = ok ? sX : -sX, wY = ok ? sY : -sY, wZ = ok ? sZ : -sZ;
    float aR, aG, aB, eR, eG, eB;
    int aslot = -1, eslot = -1;
    if (P.diff_on) {  // differentiable mode: parameters instead of 8-bit textures
      const int rec_i = so / 6;
      if (rec_i >= 0 && rec_i < P.n_surfaces) {
        aslot = P.surface_albedo[rec_i];
        eslot = P.surface_emission[rec_i];
        const int grp = aslot < 0 ? P.surface_checker[rec_i] : -1;
        if (grp >= 0) aslot = P.checker_base + 2 * grp + checker_tile(jf2i(S.surfaces[so + 1]));
      }
    }
    if (aslot >= 0) {
      aR = P.diff_albedo[aslot * 3 + 0];
      aG = P.diff_albedo[aslot * 3 + 1];
      aB = P.diff_albedo[aslot * 3 + 2];
    } else {
      // the albedo texture of a surface with noise normals: evaluated by the
      // pass that computed the noise values, at this hit record
      const int alb = noise_in && has_noise_normals(po / 3) ? __float_as_int(noise_in[noise_index].w)
                                                            : tex_color(jf2i(S.surfaces[so + 1]));
      aR = ch_r(alb);
      aG = ch_g(alb);
      aB = ch_b(alb);
    }
    if (eslot >= 0) {
      eR = eG = eB = P.diff_emission[eslot];
    } else {
      const int emi = tex_color(jf2i(S.surfaces[so + 2]));
      eR = ch_r(emi);
      eG = ch_g(emi);
      eB = ch_b(emi);
    }
    const int mat = bin >= 1 && bin <= 5 ? bin : jf2i(S.surfaces[so]);
    float uv[2] = {0.0f, 0.0f};
    if (mat == 1 || mat == 3 || mat == 4) {
      shadow = sun_sample(sun_active, pX, pY, pZ, wX, wY, wZ, aR, aG, aB, sd, uv);
    } else {
      tmp[0] = tmp[1] = tmp[2] = 0.0f;
    }
    if (vtx) {
      vtx->miss = false;
      vtx->aslot = aslot;
      vtx->eslot = eslot;
      vtx->a[0] = aR;
      vtx->a[1] = aG;
      vtx->a[2] = aB;
      vtx->e[0] = eR;
      vtx->e[1] = eG;
      vtx->e[2] = eB;
      vtx->sun_w = 0.0f;
      vtx->dsun_w[0] = vtx->dsun_w[1] = 0.0f;
      if (shadow) {
        // the sun term is a * dp1 / pi; dp1's derivatives at the same draws
        // (u, v) in dual numbers through the sun cone sample
        vtx->sun_w = sun_weight(K, pX, pY, pZ, wX, wY, wZ, uv);
        const Dual3 c = sun_cosine_dual(*vtx->sky, pX, pY, pZ, wX, wY, wZ, uv[0], uv[1]);
        vtx->dsun_w[0] = static_cast<float>(c.d[kDualElevation] * (1.0 / 3.141592653589793));
        vtx->dsun_w[1] = static_cast<float>(c.d[kDualAzimuth] * (1.0 / 3.141592653589793));
      }
      vtx->s = 1.0f;
      vtx->uses_albedo = false;
      vtx->rr = false;
    }
    if (shadow) {
      pv[0] = pr + mr * (eR + tmp[0]);
      pv[1] = pg + mg * (eG + tmp[1]);
      pv[2] = pb + mb * (eB + tmp[2]);
      pr += mr * (eR + 0.0f);
      pg += mg * (eG + 0.0f);
      pb += mb * (eB + 0.0f);
    } else {
      pr += mr * (eR + tmp[0]);
      pg += mg * (eG + tmp[1]);
      pb += mb * (eB + tmp[2]);
    }
    p[0] = pr;
    p[1] = pg;
    p[2] = pb;
    depth++;
    if (depth >= depth_rr) {
      const float pdf = jmax(aR, jmax(aG, aB));
      const float rnd = rng.next();
      if (rnd >= pdf) {
        ++rays;  // the any-hit re-test of (o, d): always a hit
        return false;
      }
      const float pdfr = recip_rn(pdf);
      if (vtx) vtx->rr = true;
      aR *= pdfr;
      aG *= pdfr;
      aB *= pdfr;
    }
      if (mat == 1) {  // ClearCoat :3323-3464
        const float a = 1.5f - 1.0f, b = 1.5f + 1.0f, c = 1.0f / 1.5f;
        const float e = (a * a) / (b * b);
        const float nnt = c;
        const float dp0 = wX * dx + wY * dy + wZ * dz;
        const float tir = 1.0f - nnt * nnt * (1.0f - dp0 * dp0);
        const float rX = dx - sX * dp2, rY = dy - sY * dp2, rZ = dz - sZ * dp2;
        if (tir < 0.0f) {
          ox = pX + wX * 0.02f;
          oy = pY + wY * 0.02f;
          oz = pZ + wZ * 0.02f;
          dx = rX;
          dy = rY;
          dz = rZ;
        } else {
          const float a1 = -dp0;
          const float a2 = 1.0f - a1;
          const float refl = e + (1.0f - e) * a2 * a2 * a2 * a2 * a2;
          const float trans = 1.0f - refl;
          const float prob = 0.25f + 0.5f * refl;
          const float rp = refl / prob;
          const float tp = trans / (1.0f - prob);
          const bool isR = rng.next() < prob;
          const float m = isR ? rp : tp;
          mr *= m;
          if (vtx) vtx->s = m;
          mg *= m;
          mb *= m;
          if (isR) {
            ox = pX + wX * 0.02f;
            oy = pY + wY * 0.02f;
            oz = pZ + wZ * 0.02f;
            dx = rX;
            dy = rY;
            dz = rZ;
          } else {
            float d[3];
            cosine_dir(wX, wY, wZ, d);
            ox = pX + wX * 0.01f;
            oy = pY + wY * 0.01f;
            oz = pZ + wZ * 0.01f;
            dx = d[0];
            dy = d[1];
            dz = d[2];
            mr *= aR;
            if (vtx) vtx->uses_albedo = true;
            mg *= aG;
            mb *= aB;
          }
        }
      } else if (mat == 3) {  // Lambertian :3465-3515
        float d[3];
        cosine_dir(wX, wY, wZ, d);
        ox = pX + wX * 0.01f;
        oy = pY + wY * 0.01f;
        oz = pZ + wZ * 0.01f;
        dx = d[0];
        dy = d[1];
        dz = d[2];
        mr *= aR;
        if (vtx) vtx->uses_albedo = true;
        mg *= aG;
        mb *= aB;
      } else if (mat == 4) {  // Phong :3516-3576
        const float u = rng.next();
        const float v = rng.next();
        const float phi = DF_PI_2 * u;
        const float ct = Math<M>::pow(1.0f - v, 1.0f / (50.0f + 1.0f));
        const float st = sqrtf(jmax(0.0f, 1.0f - ct * ct));
        const float x = Math<M>::cos(phi) * st, y = Math<M>::sin(phi) * st, z = ct;
        const float rwX = dx - sX * dp2, rwY = dy - sY * dp2, rwZ = dz - sZ * dp2;
        const float rwr = jrsqrt(rwX * rwX + rwY * rwY + rwZ * rwZ);
        const float rnX = rwX * rwr, rnY = rwY * rwr, rnZ = rwZ * rwr;
        const bool isY = fabsf(rnX) > 0.1f;
        const float uX = (isY ? 1.0f : 0.0f) * rnZ;
        const float uY = -((isY ? 0.0f : 1.0f) * rnZ);
        const float uZ = (isY ? 0.0f : 1.0f) * rnY - (isY ? 1.0f : 0.0f) * rnX;
        const float ur = jrsqrt(uX * uX + uY * uY + uZ * uZ);
        const float unX = uX * ur, unY = uY * ur, unZ = uZ * ur;
        const float vnX = rnY * unZ - rnZ * unY, vnY = rnZ * unX - rnX * unZ, vnZ = rnX * unY - rnY * unX;
        const float qX = unX * x + vnX * y + rwX * z;
        const float qY = unY * x + vnY * y + rwY * z;
        const float qZ = unZ * x + vnZ * y + rwZ * z;
        const float qr = jrsqrt(qX * qX + qY * qY + qZ * qZ);
        ox = pX + rnX * 0.01f;
        oy = pY + rnY * 0.01f;
        oz = pZ + rnZ * 0.01f;
        dx = qX * qr;
        dy = qY * qr;
        dz = qZ * qr;
        mr *= aR;
        if (vtx) vtx->uses_albedo = true;
        mg *= aG;
        mb *= aB;
      } else if (mat == 2) {  // Glass :3577-3670
        const float a = 1.5f - 1.0f, b = 1.5f + 1.0f, c = 1.0f / 1.5f, d = 1.5f / 1.0f;
        const float e = (a * a) / (b * b);
        const bool in = sX * wX + sY * wY + sZ * wZ > 0.0f;
        const float nnt = in ? c : d;
        const float dw = wX * dx + wY * dy + wZ * dz;
        const float tir = 1.0f - nnt * nnt * (1.0f - dw * dw);
        const float rX = dx - sX * dp2, rY = dy - sY * dp2, rZ = dz - sZ * dp2;
        if (tir < 0.0f) {
          ox = pX + wX * 0.01f;
          oy = pY + wY * 0.01f;
          oz = pZ + wZ * 0.01f;
          dx = rX;
          dy = rY;
          dz = rZ;
        } else {
          const float st = sqrtf(tir);
          const float sc = in ? dw * nnt + st : -(dw * nnt + st);
          const float t0X = dx * nnt - sX * sc, t0Y = dy * nnt - sY * sc, t0Z = dz * nnt - sZ * sc;
          const float tr = jrsqrt(t0X * t0X + t0Y * t0Y + t0Z * t0Z);
          const float tX = t0X * tr, tY = t0Y * tr, tZ = t0Z * tr;
          const float a1 = (in ? -dw : tX * sX + tY * sY + tZ * sZ);
          const float a2 = 1.0f - a1;
          const float refl = e + (1.0f - e) * a2 * a2 * a2 * a2 * a2;
          const float trans = 1.0f - refl;
          const float prob = 0.25f + 0.5f * refl;
          const float rp = refl / prob;
          const float tp = trans / (1.0f - prob);
          const bool isR = rng.next() < prob;
          const float m = isR ? rp : tp;
          mr *= m;
          if (vtx) vtx->s = m;
          mg *= m;
          mb *= m;
          const float eps = isR ? 0.01f : 0.000001f;
          ox = pX + wX * eps;
          oy = pY + wY * eps;
          oz = pZ + wZ * eps;
          dx = isR ? rX : tX;
          dy = isR ? rY : tY;
          dz = isR ? rZ : tZ;
        }
      } else if (mat == 5) {  // Reflection :3671-3686
        ox = pX + sX * 0.000001f;
        oy = pY + sY * 0.000001f;
        oz = pZ + sZ * 0.000001f;
        dx = dx - sX * dp2;
        dy = dy - sY * dp2;
        dz = dz - sZ * dp2;
        mr *= aR;
        if (vtx) vtx->uses_albedo = true;
        mg *= aG;
        mb *= aB;
      }
    o[0] = ox;
    o[1] = oy;
    o[2] = oz;
    d[0] = dx;
    d[1] = dy;
    d[2] = dz;
    m[0] = mr;
    m[1] = mg;
    m[2] = mb;
    return true;
  }

  // Path tracer (:3151-3705).  Returns the sample colour in out[3].
  __device__ __forceinline__ void render_pt(const float ray[6], bool sky_active, bool sun_active, float out[3],
                                            int* prim0) {
    const int depth_max = P.depth_max;
    int depth = 0;
    float o[3] = {ray[0], ray[1], ray[2]};
    float d[3] = {ray[3], ray[4], ray[5]};
    float m[3] = {1.0f, 1.0f, 1.0f};
    float p[3] = {0.0f, 0.0f, 0.0f};
    // The sky term of a path whose ray escapes (a miss, or the escape test
    // after the last bounce) is the path's last addition to its radiance, with
    // the escaping ray's d and m: it is evaluated after the loop, where the
    // wave's lanes are together again, instead of at whichever bounce each
    // lane's path escaped (the FP64 sky then ran once per bounce for the few
    // lanes escaping there).
    bool sky_due = false;
    for (;;) {
      intersect_t<false>(o[0], o[1], o[2], d[0], d[1], d[2]);
      if (depth == 0) *prim0 = h.prim;
      bool shadow;
      float sd[3], pv[3];
      sky_pending = false;
      const bool more = pt_bounce<true>(o, d, m, p, depth, sky_active, sun_active, shadow, sd, pv);
      if (shadow && sun_visible(h.p[0], h.p[1], h.p[2], sd)) {
        p[0] = pv[0];
        p[1] = pv[1];
        p[2] = pv[2];
      }
      if (!more) {
        sky_due = sky_pending;
        break;
      }
      if (depth >= depth_max) {  // the escape test (:3689-3700)
        sky_due = !(intersect_t<true>(o[0], o[1], o[2], d[0], d[1], d[2]) < DF_INF);
        break;
      }
    }
    if (sky_due) {
      sky(sky_active, d[0], d[1], d[2]);
      p[0] += m[0] * tmp[0] * DF_PI_RECIP;
      p[1] += m[1] * tmp[1] * DF_PI_RECIP;
      p[2] += m[2] * tmp[2] * DF_PI_RECIP;
    }
    out[0] = p[0];
    out[1] = p[1];
    out[2] = p[2];
  }

  // Ambient occlusion (:3029-3149).
  __device__ __forceinline__ void render_ao(const float ray[6], float out[3], int* prim0) {
    const float dx = ray[3], dy = ray[4], dz = ray[5];
    float pr = 0.0f, pg = 0.0f, pb = 0.0f;
    intersect_t<false>(ray[0], ray[1], ray[2], dx, dy, dz);
    const int po = h.prim;
    *prim0 = po;
    if (h.t != DF_INF && po != -1) {
      const float pX = h.p[0], pY = h.p[1], pZ = h.p[2];
      const float sX = h.ns[0], sY = h.ns[1], sZ = h.ns[2];
      const float dp = sX * dx + sY * dy + sZ * dz;
      const bool ok = dp < 0.0f;
      const float w0X = ok ? sX : -sX, w0Y = ok ? sY : -sY, w0Z = ok ? sZ : -sZ;
      const float wr = jrsqrt(w0X * w0X + w0Y * w0Y + w0Z * w0Z);
      const float w1X = w0X * wr, w1Y = w0Y * wr, w1Z = w0Z * wr;
      const bool isY = fabsf(w1X) > 0.1f;
      const float u0X = isY ? 0.0f : 1.0f;
      const float u0Y = isY ? 1.0f : 0.0f;
      const float u1X = u0Y * w1Z;
      const float u1Y = -(u0X * w1Z);
      const float u1Z = u0X * w1Y - u0Y * w1X;
      const float ur = jrsqrt(u1X * u1X + u1Y * u1Y + u1Z * u1Z);
      const float u2X = u1X * ur, u2Y = u1Y * ur, u2Z = u1Z * ur;
      const float v0X = w1Y * u2Z - w1Z * u2Y, v0Y = w1Z * u2X - w1X * u2Z, v0Z = w1X * u2Y - w1Y * u2X;
      const float xi = rng.next();
      const float xj = rng.next();
      const float phi = DF_PI_2 * xi;
      const float cp = Math<M>::cos(phi), sp = Math<M>::sin(phi);
      const float st = sqrtf(xj), ct = sqrtf(1.0f - xj);
      const float d0X = cp * st, d0Y = sp * st, d0Z = ct;
      const float d1X = d0X * u2X + d0Y * v0X + d0Z * w1X;
      const float d1Y = d0X * u2Y + d0Y * v0Y + d0Z * w1Y;
      const float d1Z = d0X * u2Z + d0Y * v0Z + d0Z * w1Z;
      const float dr = jrsqrt(d1X * d1X + d1Y * d1Y + d1Z * d1Z);
      const float t = intersect_t<false>(pX + w1X * 0.01f, pY + w1Y * 0.01f, pZ + w1Z * 0.01f, d1X * dr, d1Y * dr, d1Z * dr);
      const bool hit = t < P.ao_max;
      const float r = hit ? 1.0f : 0.0f;
      pr += (1.0f - r) * 1.0f + r * 0.0f;
      pg += (1.0f - r) * 1.0f + r * 0.0f;
      pb += (1.0f - r) * 1.0f + r * 0.0f;
    }
    out[0] = pr;
    out[1] = pg;
    out[2] = pb;
  }

  // Ray caster (:3707-3800).
  __device__ __forceinline__ void render_rc(const float ray[6], bool sky_active, float out[3], int* prim0) {
    const float dx = ray[3], dy = ray[4], dz = ray[5];
    intersect_t<false>(ray[0], ray[1], ray[2], dx, dy, dz);
    const int po = h.prim;
    *prim0 = po;
    if (h.t == DF_INF || po == -1) {
      sky(sky_active, dx, dy, dz);
      out[0] = tmp[0];
      out[1] = tmp[1];
      out[2] = tmp[2];
      return;
    }
    const int so = S.prims[po / 3].surface_offset;
    const int alb = tex_color(jf2i(S.surfaces[so + 1]));
    const float sX = h.ns[0], sY = h.ns[1], sZ = h.ns[2];
    const float dp = sX * dx + sY * dy + sZ * dz;
    const bool ok = dp < 0.0f;
    const int c = phong_shader0(true, h.p[0], h.p[1], h.p[2], ok ? sX : -sX, ok ? sY : -sY, ok ? sZ : -sZ, -dx, -dy,
                                -dz, ch_r(alb), ch_g(alb), ch_b(alb));
    out[0] = ch_r(c);
    out[1] = ch_g(c);
    out[2] = ch_b(c);
  }

  // Ray tracer (:3866-3996).
  __device__ __forceinline__ void render_rt(const float ray[6], bool sky_active, float out[3], int* prim0) {
    int depth = 0;
    float ox = ray[0], oy = ray[1], oz = ray[2], dx = ray[3], dy = ray[4], dz = ray[5];
    float pr = 0.0f, pg = 0.0f, pb = 0.0f;
    do {
      intersect_t<false>(ox, oy, oz, dx, dy, dz);
      const int po = h.prim;
      if (depth == 0) *prim0 = po;
      if (h.t == DF_INF || po == -1) {
        sky(sky_active, dx, dy, dz);
        out[0] = pr + tmp[0];
        out[1] = pg + tmp[1];
        out[2] = pb + tmp[2];
        return;
      }
      depth++;
      const int so = S.prims[po / 3].surface_offset;
      const int mat = jf2i(S.surfaces[so]);
      const int alb = tex_color(jf2i(S.surfaces[so + 1]));
      const float aR = ((alb >> 16) & 0xFF) / 255.0f;
      const float aG = ((alb >> 8) & 0xFF) / 255.0f;
      const float aB = (alb & 0xFF) / 255.0f;
      const float pX = h.p[0], pY = h.p[1], pZ = h.p[2];
      const float sX = h.ns[0], sY = h.ns[1], sZ = h.ns[2];
      const float dp = sX * dx + sY * dy + sZ * dz;
      const bool ok = dp < 0.0f;
      const int c = phong_shader0(true, pX, pY, pZ, ok ? sX : -sX, ok ? sY : -sY, ok ? sZ : -sZ, -dx, -dy, -dz, aR,
                                  aG, aB);
      pr += ch_r(c);
      pg += ch_g(c);
      pb += ch_b(c);
      if (mat == 5) {
        const float dp2 = dp * 2.0f;
        ox = pX + sX * 0.000001f;
        oy = pY + sY * 0.000001f;
        oz = pZ + sZ * 0.000001f;
        dx = dx - sX * dp2;
        dy = dy - sY * dp2;
        dz = dz - sZ * dp2;
      } else {
        depth = 5;
      }
    } while (depth < 5);
    out[0] = pr;
    out[1] = pg;
    out[2] = pb;
  }

  // Surface normals (:2848-2908).
  __device__ __forceinline__ void render_sn(const float ray[6], bool sky_active, float out[3], int* prim0) {
    const float dx = ray[3], dy = ray[4], dz = ray[5];
    intersect_t<false>(ray[0], ray[1], ray[2], dx, dy, dz);
    const int po = h.prim;
    *prim0 = po;
    if (h.t == DF_INF || po == -1) {
      sky(sky_active, dx, dy, dz);
      out[0] = tmp[0];
      out[1] = tmp[1];
      out[2] = tmp[2];
      return;
    }
    out[0] = (h.ns[0] + 1.0f) * 0.5f;
    out[1] = (h.ns[1] + 1.0f) * 0.5f;
    out[2] = (h.ns[2] + 1.0f) * 0.5f;
  }

  // Ray marcher over the global simplex fractal (:3802-3864).
  __device__ __forceinline__ float get_y(float x, float z) const { return noise.fractal2(P.ga, P.gf, P.gg, P.gl, P.go, x, z); }

  __device__ __forceinline__ void render_rm(const float ray[6], bool sky_active, float out[3]) {
    const float ox = ray[0], oz = ray[2];
    const float oy = get_y(ox, oz) + 0.1f;
    const float dx = ray[3], dy = ray[4], dz = ray[5];
    float dt = 0.01f;
    const float maxt = jmax(oy, 1.0f) * 20.0f;
    sky(sky_active, dx, dy, dz);
    float pr = tmp[0], pg = tmp[1], pb = tmp[2];
    for (float t = 0.001f; t < maxt; t += dt) {
      const float pX = ox + dx * t, pY = oy + dy * t, pZ = oz + dz * t;
      const float hh = get_y(pX, pZ);
      if (pY < hh) {
        const float nX = get_y(pX - DF_EPS, pZ) - get_y(pX + DF_EPS, pZ);
        const float nY = 2.0f * DF_EPS;
        const float nZ = get_y(pX, pZ - DF_EPS) - get_y(pX, pZ + DF_EPS);
        const float nr = recip_rn(sqrtf(nX * nX + nY * nY + nZ * nZ));
        const float n = noise.fbm3(8.0f, 0.5f, 0.0f, 1.0f, 16, pX, pY, pZ);
        const float a = saturate(n * 1.0f + 0.1f, 0.0f, 1.0f);
        const int c = phong_shader0(false, pX, pY, pZ, nX * nr, nY * nr, nZ * nr, -dx, -dy, -dz, a, a, a);
        pr = ch_r(c);
        pg = ch_g(c);
        pb = ch_b(c);
        t = maxt;
      }
      dt = 0.01f * t;
    }
    out[0] = pr;
    out[1] = pg;
    out[2] = pb;
  }

  // Wireframe overlay (:2910-3027); returns true when the pixel is on an edge.
  __device__ __forceinline__ bool wireframe_edge(const float ray[6]) {
    intersect_t<false>(ray[0], ray[1], ray[2], ray[3], ray[4], ray[5]);
    const float lwc = Math<M>::cos(DF_PI * 0.5f / 4096.0f);
    if (!(h.t != DF_INF && h.prim != -1) || h.shape_type != 4) return false;
    const float* c = C.c;
    const float uX = c[7], uY = c[8], uZ = c[9], vX = c[10], vY = c[11], vZ = c[12], wX = c[13], wY = c[14], wZ = c[15];
    const float pX = h.p[0], pY = h.p[1], pZ = h.p[2];
    const float pcX = pX * uX + pY * uY + pZ * uZ;
    const float pcY = pX * vX + pY * vY + pZ * vZ;
    const float pcZ = pX * wX + pY * wY + pZ * wZ;
    const float pcN = jrsqrt(pcX * pcX + pcY * pcY + pcZ * pcZ);
    const int so = h.shape_off;
    const float* A = S.point3fs + S.triangles[so + 0];
    const float* B = S.point3fs + S.triangles[so + 1];
    const float* Cc = S.point3fs + S.triangles[so + 2];
    const float aX = A[0] * uX + A[1] * uY + A[2] * uZ, aY = A[0] * vX + A[1] * vY + A[2] * vZ,
                aZ = A[0] * wX + A[1] * wY + A[2] * wZ;
    const float bX = B[0] * uX + B[1] * uY + B[2] * uZ, bY = B[0] * vX + B[1] * vY + B[2] * vZ,
                bZ = B[0] * wX + B[1] * wY + B[2] * wZ;
    const float cX = Cc[0] * uX + Cc[1] * uY + Cc[2] * uZ, cY = Cc[0] * vX + Cc[1] * vY + Cc[2] * vZ,
                cZ = Cc[0] * wX + Cc[1] * wY + Cc[2] * wZ;
    const float dAC = ((aX - cX) * (aX - cX)) + ((aY - cY) * (aY - cY)) + ((aZ - cZ) * (aZ - cZ));
    const float dBA = ((bX - aX) * (bX - aX)) + ((bY - aY) * (bY - aY)) + ((bZ - aZ) * (bZ - aZ));
    const float dCB = ((cX - bX) * (cX - bX)) + ((cY - bY) * (cY - bY)) + ((cZ - bZ) * (cZ - bZ));
    const float tAC = (((pcX - aX) * (cX - aX)) + ((pcY - aY) * (cY - aY)) + ((pcZ - aZ) * (cZ - aZ))) / dAC;
    const float tBA = (((pcX - bX) * (aX - bX)) + ((pcY - bY) * (aY - bY)) + ((pcZ - bZ) * (aZ - bZ))) / dBA;
    const float tCB = (((pcX - cX) * (bX - cX)) + ((pcY - cY) * (bY - cY)) + ((pcZ - cZ) * (bZ - cZ))) / dCB;
    const float prACX = (1.0f - tAC) * aX + tAC * cX, prACY = (1.0f - tAC) * aY + tAC * cY,
                prACZ = (1.0f - tAC) * aZ + tAC * cZ;
    const float prACN = jrsqrt(prACX * prACX + prACY * prACY + prACZ * prACZ);
    const float prBAX = (1.0f - tBA) * bX + tBA * aX, prBAY = (1.0f - tBA) * bY + tBA * aY,
                prBAZ = (1.0f - tBA) * bZ + tBA * aZ;
    const float prBAN = jrsqrt(prBAX * prBAX + prBAY * prBAY + prBAZ * prBAZ);
    const float prCBX = (1.0f - tCB) * cX + tCB * bX, prCBY = (1.0f - tCB) * cY + tCB * bY,
                prCBZ = (1.0f - tCB) * cZ + tCB * bZ;
    const float prCBN = jrsqrt(prCBX * prCBX + prCBY * prCBY + prCBZ * prCBZ);
    const float dpAC = prACX * pcX + prACY * pcY + prACZ * pcZ;
    const float dpBA = prBAX * pcX + prBAY * pcY + prBAZ * pcZ;
    const float dpCB = prCBX * pcX + prCBY * pcY + prCBZ * pcZ;
    return dpAC * prACN * pcN >= lwc || dpBA * prBAN * pcN >= lwc || dpCB * prCBN * pcN >= lwc;
  }
};

}  // namespace dfk
