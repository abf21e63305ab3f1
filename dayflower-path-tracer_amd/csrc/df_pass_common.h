// Pieces shared by the megakernel and the wavefront kernels: the noise
// permutation, the per-pixel film-to-image pass and the counter reduction.
#pragma once

#include <hip/hip_runtime.h>

#include "df_trace.h"

namespace dfk {

// Ken Perlin's improved-noise permutation (as tabulated in AbstractKernel.java:1130).
static constexpr int kPerm[256] = {
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

// The noise lookup table: the permutation twice (512 entries, as the
// reference's noisePermutations), then each entry modulo 12 (512 more, as its
// noisePermutationsModulo12, AbstractKernel.java:1126-1141): simplex noise
// reads its gradient indices from the second half instead of computing
// `perm[i] % 12`.
struct PermTable {
  int v[kPermLen];
};
constexpr PermTable make_perm_table() {
  PermTable t{};
  for (int i = 0; i < 512; ++i) {
    t.v[i] = kPerm[i & 255];
    t.v[512 + i] = kPerm[i & 255] % 12;
  }
  return t;
}
static __constant__ PermTable kPermTable = make_perm_table();

// Copies the table into the block's LDS copy (every thread of the block calls it).
__device__ __forceinline__ void load_perm(int* perm, int tid, int nthreads) {
  for (int i = tid; i < kPermLen; i += nthreads) perm[i] = kPermTable.v[i];
}

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
// DFrame::counters holds, per stage (DF_STAGE_* order), kCounterSlots copies
// of the kCounterCount counters (df_device_scene.h).
enum Stage { kStageMega = 0, kStageRaygen, kStageTrace, kStageShade, kStageEscape, kStageFilm,
             kStageCount };

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Adds each lane's counters into the frame's counters of `stage`, one atomic
// per wave and non-zero counter, into the wave's copy.  Called by every lane
// of the wave.
__device__ __forceinline__ void flush_counters(const unsigned long long cnt[kCounterCount],
                                               unsigned long long* counters, int stage) {
  const unsigned block = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const unsigned waves_per_block = (blockDim.x * blockDim.y * blockDim.z + 63) / 64;
  const unsigned wave = block * waves_per_block + (threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z)) / 64;
  unsigned long long* out = counters + (stage * kCounterSlots + wave % kCounterSlots) * kCounterCount;
  const bool leader = __lane_id() == 0;
  for (int i = 0; i < kCounterCount; ++i) {
    const unsigned long long v = wave_sum(cnt[i]);
    if (leader && v) atomicAdd(&out[i], v);
  }
}

template <int M, bool U, bool P>
__device__ __forceinline__ void take_counters(const Tracer<M, U, P>& tr, unsigned long long cnt[kCounterCount]) {
  cnt[0] += tr.rays;
  cnt[1] += tr.samples;
  cnt[2] += tr.box_tests;
  cnt[3] += tr.tri_tests;
  cnt[4] += tr.prim_tests;
  cnt[5] += tr.sphere_tests;
  cnt[6] += tr.terrain_steps;
  cnt[7] += tr.sky_evals;
}

}  // namespace dfk
