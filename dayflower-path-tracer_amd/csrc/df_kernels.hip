// HIP kernels of the renderer for gfx950 and their launch wrappers.
//
// render_kernel<M>: one work-item per pixel, 16x16-pixel workgroups (each
// 64-wide wavefront covers a 16x4 pixel strip).  A launch runs `passes`
// consecutive passes of GPURendererKernel.run() for its pixel - each pass
// draws from the pixel's own 48-bit LCG stream in the reference's order - then
// writes film, seed, primitive offset and the tone-mapped pixel once.  The
// noise permutation table is staged in LDS per workgroup.
#include <hip/hip_runtime.h>

#include "df_kernels.h"
#include "df_pass_common.h"

namespace dfk {

// All `P.passes` passes of run() for pixel (x, y): film, sample count, seed
// and primitive offset stay in registers and are written once, followed by
// the tone-mapped pixel.  R is the renderer type (1 AO, 2 path tracer, 3 ray
// caster, 4 ray marcher, 5 ray tracer, 6 surface normals), a template
// parameter so that each instantiation holds only its own ray queries.
template <int M, int R>
__device__ __forceinline__ void render_pixel(const DScene& S, const DCamera& C, const DSky& K, const DParams& P,
                                             const DFrame& F, const int* perm, int x, int y,
                                             unsigned long long cnt[kCounterCount]) {
  const int gid = y * P.width + x;
  Tracer<M> tr(S, C, K, P, perm);
  tr.rng.s = F.seeds[gid];
  float film[3] = {F.film[gid * 3 + 0], F.film[gid * 3 + 1], F.film[gid * 3 + 2]};
  int nsamp = F.film_samples[gid];
  int prim_off = F.primitive_offsets[gid];
  const bool sky_active = K.sky_active == 1;
  const bool sun_active = K.sun_active == 1;
  for (int pass = 0; pass < P.passes; ++pass) {
    const bool clear = pass == 0 && P.clear_first;
    float ray[6];
    float col[3] = {0.0f, 0.0f, 0.0f};
    bool set = false;  // filmSetColor instead of filmAddColor
    if (tr.primary_ray(x, y, ray)) {
      if constexpr (R == 1) {
        tr.render_ao(ray, col, &prim_off);
      } else if constexpr (R == 3) {
        tr.render_rc(ray, sky_active, col, &prim_off);
      } else if constexpr (R == 4) {
        tr.render_rm(ray, sky_active, col);
      } else if constexpr (R == 5) {
        tr.render_rt(ray, sky_active, col, &prim_off);
      } else if constexpr (R == 6) {
        tr.render_sn(ray, sky_active, col, &prim_off);
      } else {
        tr.render_pt(ray, sky_active, sun_active, col, &prim_off);
      }
    } else {
      set = true;
    }
    // filmAddColor / filmSetColor (AbstractImageKernel.java:333-378)
    ++tr.samples;
    if (set || clear) {
      film[0] = col[0];
      film[1] = col[1];
      film[2] = col[2];
      nsamp = 1;
    } else {
      const int n = nsamp + 1;
      film[0] = film[0] + ((col[0] - film[0]) / static_cast<float>(n));
      film[1] = film[1] + ((col[1] - film[1]) / static_cast<float>(n));
      film[2] = film[2] + ((col[2] - film[2]) / static_cast<float>(n));
      nsamp = n;
    }
    if (!set && P.wireframes == 1 && tr.wireframe_edge(ray)) {
      ++tr.samples;
      film[0] = film[1] = film[2] = 0.0f;
      nsamp = 1;
    }
  }
  F.seeds[gid] = tr.rng.s;
  F.film[gid * 3 + 0] = film[0];
  F.film[gid * 3 + 1] = film[1];
  F.film[gid * 3 + 2] = film[2];
  F.film_samples[gid] = nsamp;
  F.primitive_offsets[gid] = prim_off;
  image_pass<M>(P, film, prim_off, F.image_float + gid * 3, F.image_bgra + gid * 4);
  take_counters(tr, cnt);
}

__device__ __forceinline__ void stage_perm_2d(int* perm) {
  load_perm(perm, threadIdx.y * 16 + threadIdx.x, 256);
  __syncthreads();
}

// Occupancy hint: four waves per SIMD (spilling some registers); without it
// the allocation leaves two, and the 512x512 mesh workload's pass took 4.55
// ms instead of 3.70 (five: 3.77, six: 3.88, eight: 4.42).
//
// P.wave_tile chooses the pixels of a workgroup's four waves: 0 = rows (a
// wave covers a 16x4 strip), 1 = squares (a wave covers an 8x8 quadrant, a
// more coherent set of primary rays); either way every pixel of the 16x16
// tile gets one work-item and the same arithmetic.
template <int M, int R>
__global__ __launch_bounds__(256, 4) void render_kernel(DScene S, DCamera C, DSky K, DParams P, DFrame F) {
  __shared__ int perm[kPermLen];
  stage_perm_2d(perm);
  int tx = threadIdx.x, ty = threadIdx.y;
  if (P.wave_tile == 1) {
    const int t = threadIdx.y * 16 + threadIdx.x, wave = t >> 6, lane = t & 63;
    tx = (wave & 1) * 8 + (lane & 7);
    ty = (wave >> 1) * 8 + (lane >> 3);
  }
  const int x = blockIdx.x * 16 + tx;
  const int y = blockIdx.y * 16 + ty;
  unsigned long long cnt[kCounterCount] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (x < P.width && y < P.height) render_pixel<M, R>(S, C, K, P, F, perm, x, y, cnt);
  flush_counters(cnt, F.counters, kStageMega);
}

// A batch of independent scenes in one persistent launch: each workgroup
// takes (job, tile) items from an atomic head - the items ordered by
// decreasing estimated cost - and renders all of the tile's passes.  A pixel's
// passes stay in one work-item, in its own RNG stream, so every job's result
// is the one render_kernel gives for that scene alone.
// Occupancy hint: seven waves per SIMD (spilling some registers) measured
// fastest on the 256-scene batch mix (1.527 s against 1.594 s at six, 1.554 s
// at eight, 1.79 s at five), far faster than the two an unconstrained
// allocation allows.
template <int M, int R>
__global__ __launch_bounds__(256, 7) void batch_kernel(const DJob* __restrict__ jobs,
                                                    const DBatchItem* __restrict__ items, int n_items, int* head) {
  __shared__ int perm[kPermLen];
  __shared__ int next;
  stage_perm_2d(perm);
  for (;;) {
    if (threadIdx.x == 0 && threadIdx.y == 0) next = atomicAdd(head, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(next);
    __syncthreads();  // every work-item has read `next` before it is replaced
    if (it >= n_items) break;
    const DBatchItem item = items[it];
    // jobs is read-only and restrict: its fields are scalar loads, as kernel arguments are
    const DJob& J = jobs[item.job];
    // a wave covers an 8x8 quadrant of the tile, a more coherent set of rays than a 16x4 strip (the first
    // 128 scenes of the batch mix 1.21 -> 1.13 s; all 256 unchanged)
    const int t = threadIdx.y * 16 + threadIdx.x, wave = t >> 6, lane = t & 63;
    const int x = item.tile_x * 16 + (wave & 1) * 8 + (lane & 7);
    const int y = item.tile_y * 16 + (wave >> 1) * 8 + (lane >> 3);
    unsigned long long cnt[kCounterCount] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (x < J.P.width && y < J.P.height) render_pixel<M, R>(J.S, J.C, J.K, J.P, J.F, perm, x, y, cnt);
    flush_counters(cnt, J.F.counters, kStageMega);
  }
}

// Per-pixel seeds from a master seed: SplitMix64 output for state
// master + (index + 1) * 0x9E3779B97F4A7C15 (Steele et al. 2014).  Replaces
// the reference's ThreadLocalRandom seeding (AbstractKernel.java:1148-1156).
__global__ void seed_kernel(uint64_t* seeds, uint64_t n, uint64_t master) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t z = master + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  seeds[i] = z ^ (z >> 31);
}

}  // namespace dfk

// ------------------------------------------------------------ launch wrappers

hipError_t dfk_launch_render(const dfk::DScene& s, const dfk::DCamera& c, const dfk::DSky& k, const dfk::DParams& p,
                             const dfk::DFrame& f, int math_mode, hipStream_t stream, const dfk_marks* marks) {
  const dim3 block(16, 16);
  const dim3 grid((p.width + 15) / 16, (p.height + 15) / 16);
  using Fn = void (*)(dfk::DScene, dfk::DCamera, dfk::DSky, dfk::DParams, dfk::DFrame);
  static const Fn kFast[6] = {dfk::render_kernel<0, 1>, dfk::render_kernel<0, 2>, dfk::render_kernel<0, 3>,
                              dfk::render_kernel<0, 4>, dfk::render_kernel<0, 5>, dfk::render_kernel<0, 6>};
  static const Fn kExact[6] = {dfk::render_kernel<1, 1>, dfk::render_kernel<1, 2>, dfk::render_kernel<1, 3>,
                               dfk::render_kernel<1, 4>, dfk::render_kernel<1, 5>, dfk::render_kernel<1, 6>};
  const int r = p.renderer_type >= 1 && p.renderer_type <= 6 ? p.renderer_type - 1 : 1;
  if (marks) marks->mark(marks->user, dfk::kStageMega, 1);
  hipLaunchKernelGGL((math_mode == 1 ? kExact : kFast)[r], grid, block, 0, stream, s, c, k, p, f);
  if (marks) marks->mark(marks->user, dfk::kStageMega, 0);
  return hipGetLastError();
}

hipError_t dfk_launch_batch(const dfk::DJob* jobs, const dfk::DBatchItem* items, int n_items, int* head,
                            int math_mode, int renderer_type, hipStream_t stream) {
  using Fn = void (*)(const dfk::DJob* __restrict__, const dfk::DBatchItem* __restrict__, int, int*);
  static const Fn kFast[6] = {dfk::batch_kernel<0, 1>, dfk::batch_kernel<0, 2>, dfk::batch_kernel<0, 3>,
                              dfk::batch_kernel<0, 4>, dfk::batch_kernel<0, 5>, dfk::batch_kernel<0, 6>};
  static const Fn kExact[6] = {dfk::batch_kernel<1, 1>, dfk::batch_kernel<1, 2>, dfk::batch_kernel<1, 3>,
                               dfk::batch_kernel<1, 4>, dfk::batch_kernel<1, 5>, dfk::batch_kernel<1, 6>};
  const int r = renderer_type >= 1 && renderer_type <= 6 ? renderer_type - 1 : 1;
  const Fn fn = (math_mode == 1 ? kExact : kFast)[r];
  static int grid[2][6] = {};
  int& g = grid[math_mode == 1 ? 1 : 0][r];
  if (g == 0) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    g = cus * per_cu;
  }
  if (n_items <= 0) return hipSuccess;
  const int blocks = n_items < g ? n_items : g;
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(16, 16), 0, stream, jobs, items, n_items, head);
  return hipGetLastError();
}

namespace dfk {

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// The reference's terrain march (GPURendererKernel.java:1470-1506) step by
// step over every position of the table, for checking isect_terrain.
template <int M>
__device__ float march_every_position(float ox, float oy, float oz, float dx, float dy, float dz, const float* tc_tab,
                                      int n) {
  const float sr = 1.0f / 10.0f;
  float t = 0.0f, td = 0.01f, lh = 0.0f, ly = 0.0f;
  for (int k = 0; k < n; ++k) {
    const float tc = tc_tab[k];
    const float px = ox * sr + dx * tc, py = oy * sr + dy * tc, pz = oz * sr + dz * tc;
    const float h = Math<M>::sin(px) * Math<M>::sin(pz);
    if (py < h) {
      t = tc - td + td * (lh - ly) / (py - ly - h + lh);
      break;
    }
    td = 0.01f * tc;
    lh = h;
    ly = py;
  }
  return t > DF_EPS ? t : DF_INF;
}

template <int M>
__global__ void terrain_check_kernel(const float* rays, int64_t n, const float* tc_tab, int tn, float* t_skip,
                                     float* t_plain, uint32_t* evals) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* r = rays + 6 * i;
  unsigned e = 0;
  t_skip[i] = isect_terrain<M>(r[0], r[1], r[2], r[3], r[4], r[5], DF_INF, tc_tab, tn, &e);
  evals[i] = e;
  t_plain[i] = march_every_position<M>(r[0], r[1], r[2], r[3], r[4], r[5], tc_tab, tn);
}

// recip_rn against the division for the float bit patterns [lo, hi) of both
// signs: counts the values tested and the mismatches, and keeps the smallest
// mismatching pattern.
__global__ void reciprocal_check_kernel(uint32_t lo, uint32_t hi, unsigned long long* tested,
                                        unsigned long long* mismatches, uint32_t* first) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  unsigned long long n = 0, bad = 0;
  for (uint64_t b = lo + static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; b < hi; b += stride) {
    for (uint32_t sign = 0; sign < 2; ++sign) {
      const uint32_t bits = static_cast<uint32_t>(b) | (sign << 31);
      const float x = __uint_as_float(bits);
      const float q = 1.0f / x;
      ++n;
      if (__float_as_uint(recip_rn(x)) != __float_as_uint(q)) {
        ++bad;
        atomicMin(first, bits);
      }
    }
  }
  n = wave_sum_u64(n);
  bad = wave_sum_u64(bad);
  if (__lane_id() == 0) {
    atomicAdd(tested, n);
    if (bad) atomicAdd(mismatches, bad);
  }
}

}  // namespace dfk

hipError_t dfk_launch_reciprocal_check(uint32_t lo, uint32_t hi, unsigned long long* tested,
                                       unsigned long long* mismatches, uint32_t* first, hipStream_t stream) {
  hipLaunchKernelGGL(dfk::reciprocal_check_kernel, dim3(8192), dim3(256), 0, stream, lo, hi, tested, mismatches,
                     first);
  return hipGetLastError();
}

hipError_t dfk_launch_terrain_check(const float* rays, int64_t n, const float* tc_tab, int tn, float* t_skip,
                                    float* t_plain, uint32_t* evals, int math_mode, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
  if (math_mode == 1)
    hipLaunchKernelGGL(dfk::terrain_check_kernel<1>, dim3(blocks), dim3(256), 0, stream, rays, n, tc_tab, tn, t_skip,
                       t_plain, evals);
  else
    hipLaunchKernelGGL(dfk::terrain_check_kernel<0>, dim3(blocks), dim3(256), 0, stream, rays, n, tc_tab, tn, t_skip,
                       t_plain, evals);
  return hipGetLastError();
}

hipError_t dfk_launch_seed(uint64_t* seeds, uint64_t n, uint64_t master, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
  hipLaunchKernelGGL(dfk::seed_kernel, dim3(blocks), dim3(256), 0, stream, seeds, n, master);
  return hipGetLastError();
}
