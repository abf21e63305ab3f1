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
#include "df_trace.h"

namespace dfk {

// Ken Perlin's improved-noise permutation (as tabulated in AbstractKernel.java:1130).
__constant__ int kPerm[256] = {
    151, 160, 137, 91,  90,  15,  131, 13,  201, 95,  96,  53,  194, 233, 7,   225, 140, 36,  103, 30,  69,  142,
    8,   99,  37,  240, 21,  10,  23,  190, 6,   148, 247, 120, 234, 75,  0,   26,  197, 62,  94,  252, 219, 203,
    117, 35,  11,  32,  57,  177, 33,  88,  237, 149, 56,  87,  174, 20,  125, 136, 171, 168, 68,  175, 74,  165,
    71,  134, 139, 48,  27,  166, 77,  146, 158, 231, 83,  111, 229, 122, 60,  211, 133, 230, 220, 105, 92,  41,
    55,  46,  245, 40,  244, 102, 143, 54,  65,  25,  63,  161, 1,   216, 80,  73,  209, 76,  132, 187, 208, 89,
    18,  169, 200, 196, 135, 130, 116, 188, 159, 86,  164, 100, 109, 198, 173, 186, 3,   64,  52,  217, 226, 250,
    124, 123, 5,   202, 38,  147, 118, 126, 255, 82,  85,  212, 207, 206, 59,  227, 47,  16,  58,  17,  182, 189,
    28,  42,  23,  183, 170, 213, 119, 248, 152, 2,   44,  154, 163, 70,  221, 153, 101, 155, 167, 43,  172, 9,
    129, 22,  39,  253, 19,  98,  108, 110, 79,  113, 224, 232, 178, 185, 112, 104, 218, 246, 97,  228, 251, 34,
    242, 193, 238, 210, 144, 12,  191, 179, 162, 241, 81,  51,  145, 235, 249, 14,  239, 107, 49,  192, 214, 31,
    181, 199, 106, 157, 184, 84,  204, 176, 115, 121, 50,  45,  127, 4,   150, 254, 138, 236, 205, 93,  222, 114,
    67,  29,  24,  72,  243, 141, 128, 195, 78,  66,  215, 61,  156, 180};

// Tone map + sRGB + BGRA (run() :282-307, AbstractImageKernel.java:410-472, 503-774).
template <int M>
__device__ void image_pass(const DParams& P, const float film[3], int prim_off, float* out_f, uint8_t* out_b) {
  const float tint_g = prim_off > -1 && prim_off == P.selected ? 1.0f : 0.0f;
  float r = film[0] + 0.0f, g = film[1] + tint_g, b = film[2] + 0.0f;
  const float ex = P.exposure;
  if (P.tone_mapper_type == 2) {  // Reinhard
    const float rm = r * ex, gm = g * ex, bm = b * ex;
    r = rm / (1.0f + rm);
    g = gm / (1.0f + gm);
    b = bm / (1.0f + bm);
  } else if (P.tone_mapper_type == 3) {  // Reinhard modified 1
    r = 1.0f - Math<M>::exp(-r * ex);
    g = 1.0f - Math<M>::exp(-g * ex);
    b = 1.0f - Math<M>::exp(-b * ex);
  } else if (P.tone_mapper_type == 4) {  // Reinhard modified 2 (L_white = 4)
    const float lw2r = 1.0f / (4.0f * 4.0f);
    const float rm = r * ex, gm = g * ex, bm = b * ex;
    r = rm * (1.0f + rm * lw2r) / (1.0f + rm);
    g = gm * (1.0f + gm * lw2r) / (1.0f + gm);
    b = bm * (1.0f + bm * lw2r) / (1.0f + bm);
  } else if (P.tone_mapper_type == 1) {  // filmic ACES (a 2.51, b 0.03, c 2.43, d 0.59, e 0.14)
    const float mr = jmax(r * ex - 0.0f, -3.402823466e38f);
    const float mg = jmax(g * ex - 0.0f, -3.402823466e38f);
    const float mb = jmax(b * ex - 0.0f, -3.402823466e38f);
    r = saturate((mr * (2.51f * mr + 0.03f)) / (mr * (2.43f * mr + 0.59f) + 0.14f), 0.0f, 1.0f);
    g = saturate((mg * (2.51f * mg + 0.03f)) / (mg * (2.43f * mg + 0.59f) + 0.14f), 0.0f, 1.0f);
    b = saturate((mb * (2.51f * mb + 0.03f)) / (mb * (2.43f * mb + 0.59f) + 0.14f), 0.0f, 1.0f);
  }
  const float bp = P.srgb_break, gi = P.srgb_gamma_inv, sl = P.srgb_slope, sm = P.srgb_slope_match,
              sg = P.srgb_segment;
  r = r <= 0.0f ? 0.0f : r >= 1.0f ? 1.0f : r <= bp ? r * sl : sm * Math<M>::pow(r, gi) - sg;
  g = g <= 0.0f ? 0.0f : g >= 1.0f ? 1.0f : g <= bp ? g * sl : sm * Math<M>::pow(g, gi) - sg;
  b = b <= 0.0f ? 0.0f : b >= 1.0f ? 1.0f : b <= bp ? b * sl : sm * Math<M>::pow(b, gi) - sg;
  out_f[0] = r;
  out_f[1] = g;
  out_f[2] = b;
  const int R = jf2i(jmax(jmin(r * 255.0f + 0.5f, 255.0f), 0.0f));
  const int G = jf2i(jmax(jmin(g * 255.0f + 0.5f, 255.0f), 0.0f));
  const int B = jf2i(jmax(jmin(b * 255.0f + 0.5f, 255.0f), 0.0f));
  const uint32_t px = static_cast<uint32_t>(B & 0xFF) | (static_cast<uint32_t>(G & 0xFF) << 8) |
                      (static_cast<uint32_t>(R & 0xFF) << 16) | 0xFF000000u;
  *reinterpret_cast<uint32_t*>(out_b) = px;
}

constexpr int kCounterCount = 8;

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// M: transcendental mode; R: renderer type (1 AO, 2 path tracer, 3 ray caster,
// 4 ray marcher, 5 ray tracer, 6 surface normals), a template parameter so
// that each instantiation holds only its own renderer's ray queries.
template <int M, int R>
__global__ __launch_bounds__(256) void render_kernel(DScene S, DCamera C, DSky K, DParams P, DFrame F) {
  __shared__ int perm[512];
  for (int i = threadIdx.y * 16 + threadIdx.x; i < 512; i += 256) perm[i] = kPerm[i & 255];
  __syncthreads();
  const int x = blockIdx.x * 16 + threadIdx.x;
  const int y = blockIdx.y * 16 + threadIdx.y;
  unsigned long long cnt[kCounterCount] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (x < P.width && y < P.height) {
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
    cnt[0] = tr.rays;
    cnt[1] = tr.samples;
    cnt[2] = tr.box_tests;
    cnt[3] = tr.tri_tests;
    cnt[4] = tr.prim_tests;
    cnt[5] = tr.sphere_tests;
    cnt[6] = tr.terrain_steps;
    cnt[7] = tr.sky_evals;
  }
  const bool leader = (threadIdx.x + threadIdx.y * 16) % 64 == 0;
  for (int i = 0; i < kCounterCount; ++i) {
    const unsigned long long v = wave_sum(cnt[i]);
    if (leader && v) atomicAdd(&F.counters[i], v);
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
                             const dfk::DFrame& f, int math_mode, hipStream_t stream) {
  const dim3 block(16, 16);
  const dim3 grid((p.width + 15) / 16, (p.height + 15) / 16);
  using Fn = void (*)(dfk::DScene, dfk::DCamera, dfk::DSky, dfk::DParams, dfk::DFrame);
  static const Fn kFast[6] = {dfk::render_kernel<0, 1>, dfk::render_kernel<0, 2>, dfk::render_kernel<0, 3>,
                              dfk::render_kernel<0, 4>, dfk::render_kernel<0, 5>, dfk::render_kernel<0, 6>};
  static const Fn kExact[6] = {dfk::render_kernel<1, 1>, dfk::render_kernel<1, 2>, dfk::render_kernel<1, 3>,
                               dfk::render_kernel<1, 4>, dfk::render_kernel<1, 5>, dfk::render_kernel<1, 6>};
  const int r = p.renderer_type >= 1 && p.renderer_type <= 6 ? p.renderer_type - 1 : 1;
  hipLaunchKernelGGL((math_mode == 1 ? kExact : kFast)[r], grid, block, 0, stream, s, c, k, p, f);
  return hipGetLastError();
}

hipError_t dfk_launch_seed(uint64_t* seeds, uint64_t n, uint64_t master, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
  hipLaunchKernelGGL(dfk::seed_kernel, dim3(blocks), dim3(256), 0, stream, seeds, n, master);
  return hipGetLastError();
}
