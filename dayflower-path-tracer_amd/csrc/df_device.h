// Device functions of the Dayflower renderer kernel for gfx950.
//
// Every function restates the matching piece of the reference kernel
// (src/main/java/org/dayflower/pathtracer/kernel/GPURendererKernel.java,
// AbstractKernel.java, AbstractImageKernel.java) with the reference's float
// operation order.  The library is compiled with -ffp-contract=off and with
// HIP's correctly rounded fp32 division and square root, so every
// non-transcendental result is the one the reference's Java thread-pool
// execution produces.  Transcendentals go through Math<M>:
//   M = 0 (fast):  single-precision device library functions;
//   M = 1 (exact): evaluated in double and rounded once to float, which is
//                  what java.lang.Math gives the reference.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "df_device_scene.h"

namespace dfk {

// ------------------------------------------------------------ constants

#define DF_PI 3.14159274101257324f                 // (float) Math.PI
#define DF_PI_DIV_180 (DF_PI / 180.0f)
#define DF_PI_2 (DF_PI * 2.0f)
#define DF_PI_2_RECIP (1.0f / DF_PI_2)
#define DF_PI_RECIP (1.0f / DF_PI)
#define DF_INF 100000.0f                           // AbstractKernel.INFINITY
#define DF_EPS 0.02f                               // AbstractKernel.EPSILON
#define DF_COLOR_RECIP (1.0f / 255.0f)

// ------------------------------------------------------------ Java semantics

// p[i] of scene data, which no kernel writes: read through the constant
// address space, a wave-uniform index becomes scalar loads.  (The host
// compilation pass of this header has no such address space.)
template <class T>
__device__ __forceinline__ T load_uniform(const T* p, int i) {
#if defined(__HIP_DEVICE_COMPILE__)
  return ((const __attribute__((address_space(4))) T*)(p))[i];
#else
  return p[i];
#endif
}

// p[i] with a 32-bit byte offset (arrays here stay far below 4 GiB), so the
// load takes the scalar-base + 32-bit-offset form instead of 64-bit address
// arithmetic per access.
template <class T>
__device__ __forceinline__ T at(const T* p, int i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) +
                                     static_cast<uint32_t>(i) * static_cast<uint32_t>(sizeof(T)));
}

// Math.max / Math.min (float): NaN if either operand is NaN, -0 below +0 -
// the IEEE 754-2019 maximum / minimum, one v_maximum3_f32 / v_minimum3_f32
// on gfx950 (nested calls fuse into the three-operand form).
__device__ __forceinline__ float jmax(float a, float b) { return __builtin_elementwise_maximum(a, b); }
__device__ __forceinline__ float jmin(float a, float b) { return __builtin_elementwise_minimum(a, b); }
__device__ __forceinline__ double jmaxd(double a, double b) {
  if (a != a) return a;
  if (a == 0.0 && b == 0.0 && __builtin_signbit(a)) return b;
  return a >= b ? a : b;
}
__device__ __forceinline__ double jmind(double a, double b) {
  if (a != a) return a;
  if (a == 0.0 && b == 0.0 && __builtin_signbit(b)) return b;
  return a <= b ? a : b;
}
// Java (int): truncate, saturate, NaN -> 0.  Branch-free: the value is
// clamped into [-2^31, 2^31 - 128] (the largest float below 2^31; fmaxf takes
// -2^31 for NaN), converted, then the two cases the clamp cannot express are
// selected (2^31 and above, NaN).
__device__ __forceinline__ int jf2i(float v) {
  const int i = static_cast<int>(fminf(fmaxf(v, -2147483648.0f), 2147483520.0f));
  const int s = v >= 2147483648.0f ? 2147483647 : i;
  return v != v ? 0 : s;
}


template <int M>
struct Math;

template <>
struct Math<0> {
  static __device__ __forceinline__ float sin(float x) { return sinf(x); }
  static __device__ __forceinline__ float cos(float x) { return cosf(x); }
  static __device__ __forceinline__ float exp(float x) { return expf(x); }
  static __device__ __forceinline__ float pow(float a, float b) { return powf(a, b); }
  static __device__ __forceinline__ float atan2(float y, float x) { return atan2f(y, x); }
  static __device__ __forceinline__ float asinpi(float x) { return asinf(x) * DF_PI_RECIP; }
};

template <>
struct Math<1> {
  static __device__ __forceinline__ float sin(float x) { return static_cast<float>(::sin(static_cast<double>(x))); }
  static __device__ __forceinline__ float cos(float x) { return static_cast<float>(::cos(static_cast<double>(x))); }
  static __device__ __forceinline__ float exp(float x) { return static_cast<float>(::exp(static_cast<double>(x))); }
  static __device__ __forceinline__ float pow(float a, float b) {
    return static_cast<float>(::pow(static_cast<double>(a), static_cast<double>(b)));
  }
  static __device__ __forceinline__ float atan2(float y, float x) {
    return static_cast<float>(::atan2(static_cast<double>(y), static_cast<double>(x)));
  }
  static __device__ __forceinline__ float asinpi(float x) {
    return static_cast<float>(::asin(static_cast<double>(x)) / 3.141592653589793);
  }
};

// ------------------------------------------------------------ small helpers

__device__ __forceinline__ float saturate(float v, float a, float b) {  // AbstractKernel.java:553-558
  const float mn = jmin(a, b), mx = jmax(a, b);
  return v < mn ? mn : v > mx ? mx : v;
}
__device__ __forceinline__ float sat01(float v) { return v < 0.0f ? 0.0f : v > 1.0f ? 1.0f : v; }
__device__ __forceinline__ float modulo(float v) { return v - floorf(v); }
__device__ __forceinline__ float remainder_j(float x, float y) {  // AbstractKernel.java:533-537
  const int n = jf2i(x / y);
  return x - static_cast<float>(n) * y;
}
__device__ __forceinline__ float lerp_j(float a, float b, float t) { return (1.0f - t) * a + t * b; }
__device__ __forceinline__ float ch_r(int c) { return ((c >> 16) & 0xFF) * DF_COLOR_RECIP; }
__device__ __forceinline__ float ch_g(int c) { return ((c >> 8) & 0xFF) * DF_COLOR_RECIP; }
__device__ __forceinline__ float ch_b(int c) { return (c & 0xFF) * DF_COLOR_RECIP; }
__device__ __forceinline__ int pack_rgb(float r, float g, float b) {
  return ((jf2i(r) & 0xFF) << 16) | ((jf2i(g) & 0xFF) << 8) | (jf2i(b) & 0xFF);
}
__device__ __forceinline__ int quant3(float r, float g, float b) {
  return pack_rgb(saturate(r, 0.0f, 1.0f) * 255.0f + 0.5f, saturate(g, 0.0f, 1.0f) * 255.0f + 0.5f,
                  saturate(b, 0.0f, 1.0f) * 255.0f + 0.5f);
}

// ------------------------------------------------------------ RNG (AbstractKernel.java:438-440, 1101-1110)

struct Rng {
  uint64_t s;
  __device__ __forceinline__ float next() {
    s = (s * 0x5DEECE66Dull + 0xBull) & ((1ull << 48) - 1);
    return static_cast<float>(static_cast<int>(s >> 24)) * (1.0f / 16777216.0f);
  }
};

// ------------------------------------------------------------ noise (AbstractKernel.java:762-917)

// The permutation table (Ken Perlin's improved-noise table, as tabulated in
// AbstractKernel.java:1130) is staged in LDS by each block; gradient indices
// are perm % 12 and the 12 gradients of grad3 are computed from the index.
// Entries of the noise lookup table: the permutation twice, then the same modulo 12.
constexpr int kPermLen = 1024;

struct Noise {
  const int* perm;  // kPermLen entries (in LDS where noise is evaluated)

  static __device__ __forceinline__ float grad_dot3(int g, float x, float y, float z) {
    // simplexGradient3 rows: (1,1,0) (-1,1,0) (1,-1,0) (-1,-1,0) (1,0,1) (-1,0,1)
    // (1,0,-1) (-1,0,-1) (0,1,1) (0,-1,1) (0,1,-1) (0,-1,-1), selected without
    // branches; the products and sums are the reference's gx*x + gy*y + gz*z
    // with its zeros.
    const float s1 = (g & 1) ? -1.0f : 1.0f, s2 = (g & 2) ? -1.0f : 1.0f;
    const float gx = g < 8 ? s1 : 0.0f;
    const float gy = g < 4 ? s2 : (g < 8 ? 0.0f : s1);
    const float gz = g < 4 ? 0.0f : s2;
    return gx * x + gy * y + gz * z;
  }
  static __device__ __forceinline__ float grad_dot2(int g, float x, float y) {
    // the x and y components of the same twelve rows, selected without branches
    const float s1 = (g & 1) ? -1.0f : 1.0f, s2 = (g & 2) ? -1.0f : 1.0f;
    const float gx = g < 8 ? s1 : 0.0f;
    const float gy = g < 4 ? s2 : (g < 8 ? 0.0f : s1);
    return gx * x + gy * y;
  }
  static __device__ __forceinline__ int fast_floor(float v) {
    const int i = jf2i(v);
    return v < static_cast<float>(i) ? i - 1 : i;
  }
  __device__ __forceinline__ int p12(int i) const { return perm[512 + i]; }  // perm[i] % 12

  __device__ float simplex2(float x, float y) const {
    const float F2 = 0.3660254037844386f, G2 = 0.21132486540518713f;
    const float s = (x + y) * F2;
    const int i = fast_floor(x + s);
    const int j = fast_floor(y + s);
    const float t = static_cast<float>(i + j) * G2;
    const float x0 = x - (static_cast<float>(i) - t);
    const float y0 = y - (static_cast<float>(j) - t);
    const int i1 = x0 > y0 ? 1 : 0;
    const int j1 = x0 > y0 ? 0 : 1;
    const float x1 = x0 - static_cast<float>(i1) + G2;
    const float y1 = y0 - static_cast<float>(j1) + G2;
    const float x2 = x0 - 1.0f + 2.0f * G2;
    const float y2 = y0 - 1.0f + 2.0f * G2;
    const int ii = i & 0xFF, jj = j & 0xFF;
    const int gi0 = p12(ii + perm[jj]);
    const int gi1 = p12(ii + i1 + perm[jj + j1]);
    const int gi2 = p12(ii + 1 + perm[jj + 1]);
    // corner terms computed on every lane and selected, as in simplex3
    const float t0 = 0.5f - x0 * x0 - y0 * y0;
    const float v0 = (t0 * t0) * (t0 * t0) * grad_dot2(gi0, x0, y0);
    const float t1 = 0.5f - x1 * x1 - y1 * y1;
    const float v1 = (t1 * t1) * (t1 * t1) * grad_dot2(gi1, x1, y1);
    const float t2 = 0.5f - x2 * x2 - y2 * y2;
    const float v2 = (t2 * t2) * (t2 * t2) * grad_dot2(gi2, x2, y2);
    const float n0 = t0 < 0.0f ? 0.0f : v0, n1 = t1 < 0.0f ? 0.0f : v1, n2 = t2 < 0.0f ? 0.0f : v2;
    return 70.0f * (n0 + n1 + n2);
  }

  __device__ float simplex3(float x, float y, float z) const {
    const float F3 = 1.0f / 3.0f, G3 = 1.0f / 6.0f;
    const float s = (x + y + z) * F3;
    const int i = fast_floor(x + s);
    const int j = fast_floor(y + s);
    const int k = fast_floor(z + s);
    const float t = static_cast<float>(i + j + k) * G3;
    const float x0 = x - (static_cast<float>(i) - t);
    const float y0 = y - (static_cast<float>(j) - t);
    const float z0 = z - (static_cast<float>(k) - t);
    // The reference's six orderings (x0 >= y0 ? (y0 >= z0 ? XYZ : x0 >= z0 ?
    // XZY : ZXY) : (y0 < z0 ? ZYX : x0 < z0 ? YZX : YXZ), comparisons false
    // for NaN), selected without branches.
    const bool a = x0 >= y0, b = y0 >= z0, c = x0 >= z0, d = y0 < z0, e = x0 < z0;
    const bool xyz = a && b, xzy = a && !b && c, zxy = a && !b && !c;
    const bool zyx = !a && d, yzx = !a && !d && e, yxz = !a && !d && !e;
    const int i1 = xyz || xzy, j1 = yzx || yxz, k1 = zxy || zyx;
    const int i2 = xyz || xzy || zxy || yxz, j2 = xyz || zyx || yzx || yxz, k2 = xzy || zxy || zyx || yzx;
    const float x1 = x0 - static_cast<float>(i1) + G3;
    const float y1 = y0 - static_cast<float>(j1) + G3;
    const float z1 = z0 - static_cast<float>(k1) + G3;
    const float x2 = x0 - static_cast<float>(i2) + 2.0f * G3;
    const float y2 = y0 - static_cast<float>(j2) + 2.0f * G3;
    const float z2 = z0 - static_cast<float>(k2) + 2.0f * G3;
    const float x3 = x0 - 1.0f + 3.0f * G3;
    const float y3 = y0 - 1.0f + 3.0f * G3;
    const float z3 = z0 - 1.0f + 3.0f * G3;
    const int ii = i & 0xFF, jj = j & 0xFF, kk = k & 0xFF;
    const int gi0 = p12(ii + perm[jj + perm[kk]]);
    const int gi1 = p12(ii + i1 + perm[jj + j1 + perm[kk + k1]]);
    const int gi2 = p12(ii + i2 + perm[jj + j2 + perm[kk + k2]]);
    const int gi3 = p12(ii + 1 + perm[jj + 1 + perm[kk + 1]]);
    const float t0 = 0.6f - x0 * x0 - y0 * y0 - z0 * z0;
    const float n0 = t0 < 0.0f ? 0.0f : (t0 * t0) * (t0 * t0) * grad_dot3(gi0, x0, y0, z0);
    const float t1 = 0.6f - x1 * x1 - y1 * y1 - z1 * z1;
    const float n1 = t1 < 0.0f ? 0.0f : (t1 * t1) * (t1 * t1) * grad_dot3(gi1, x1, y1, z1);
    const float t2 = 0.6f - x2 * x2 - y2 * y2 - z2 * z2;
    const float n2 = t2 < 0.0f ? 0.0f : (t2 * t2) * (t2 * t2) * grad_dot3(gi2, x2, y2, z2);
    const float t3 = 0.6f - x3 * x3 - y3 * y3 - z3 * z3;
    const float n3 = t3 < 0.0f ? 0.0f : (t3 * t3) * (t3 * t3) * grad_dot3(gi3, x3, y3, z3);
    return 32.0f * (n0 + n1 + n2 + n3);
  }

  __device__ float fbm3(float frequency, float gain, float minimum, float maximum, int octaves, float x, float y,
                        float z) const {  // AbstractKernel.java:695-724
    float amp = 1.0f, max_amp = 0.0f, freq = frequency, noise = 0.0f;
    for (int i = 0; i < octaves; ++i) {
      noise += simplex3(x * freq, y * freq, z * freq) * amp;
      max_amp += amp;
      amp *= gain;
      freq *= 2.0f;
    }
    noise /= max_amp;
    noise = noise * (maximum - minimum) / 2.0f + (maximum + minimum) / 2.0f;
    return noise;
  }

  __device__ float fractal2(float amplitude, float frequency, float gain, float lacunarity, int octaves, float x,
                            float y) const {  // AbstractKernel.java:572-599
    float result = 0.0f, amp = amplitude, freq = frequency;
    for (int i = 0; i < octaves; ++i) {
      result += amp * simplex2(x * freq, y * freq);
      amp *= gain;
      freq *= lacunarity;
    }
    return result;
  }
};

// ------------------------------------------------------------ hit record

// The fields of the reference's 27-float intersection record that any
// renderer reads (GPURendererKernel.java:79-92); the geometric normal and the
// plane ONB are written there but never read, so they are not kept.
struct Hit {
  float t;
  int prim;      // primitive offset (index * 3), -1 on miss
  int shape_type;
  int shape_off;
  float p[3];
  float ns[3];
  float uv[2];
  float tan[3];
};

// ------------------------------------------------------------ intersection tests

__device__ __forceinline__ float isect_plane(float ox, float oy, float oz, float dx, float dy, float dz, float ax,
                                             float ay, float az, float nx, float ny, float nz,
                                             float tmax) {  // :1414-1431
  const float dp = nx * dx + ny * dy + nz * dz;
  if (dp < 0.0f || dp > 0.0f) {
    const float t = ((ax - ox) * nx + (ay - oy) * ny + (az - oz) * nz) / dp;
    if (t > 0.001f && t < tmax) return t;
  }
  return DF_INF;
}

__device__ __forceinline__ float isect_sphere(float ox, float oy, float oz, float dx, float dy, float dz, float px,
                                              float py, float pz, float radius, float tmax) {  // :1433-1467
  const float tox = ox - px, toy = oy - py, toz = oz - pz;
  const float r2 = radius * radius;
  const float a = dx * dx + dy * dy + dz * dz;
  const float b = 2.0f * (tox * dx + toy * dy + toz * dz);
  const float c = (tox * tox + toy * toy + toz * toz) - r2;
  const float disc2 = b * b - 4.0f * a * c;
  if (disc2 >= 0.0f) {
    const float disc = sqrtf(disc2);
    const float q = -0.5f * (b < 0.0f ? b - disc : b + disc);
    const float r0 = q / a;
    const float r1 = q == 0.0f ? r0 : c / q;
    const float t0 = jmin(r0, r1);
    const float t1 = jmax(r0, r1);
    if (t0 > 0.001f && t0 < tmax) return t0;
    if (t1 > 0.001f && t1 < tmax) return t1;
  }
  return DF_INF;
}

// 1 / x rounded to nearest, as the division 1.0f / x gives it: the hardware
// reciprocal (within one ulp) and one fused correction r + r (1 - x r),
// which is the correctly rounded quotient for every float with 2^-14 < |x| <
// 2^126 (df_check_reciprocal compares the two for every one of them on the
// GPU); outside that range, the division.  Four instructions against ten.
__device__ __forceinline__ float recip_rn(float x) {
  const float ax = fabsf(x);
  if (ax > 0x1p-14f && ax < 0x1p126f) {
    const float r = __builtin_amdgcn_rcpf(x);
    return __builtin_fmaf(__builtin_fmaf(-x, r, 1.0f), r, r);
  }
  return 1.0f / x;
}

// Java's 1 / sqrt(x): the correctly rounded square root, then the correctly
// rounded reciprocal.
__device__ __forceinline__ float jrsqrt(float x) { return recip_rn(sqrtf(x)); }

// Moller-Trumbore with the reference's epsilons (:1541-1594); e0 = b - a and
// e1 = c - a come precomputed (the same float subtraction).
__device__ __forceinline__ float isect_triangle(float ox, float oy, float oz, float dx, float dy, float dz,
                                                const DTri& T, float tmax) {
  const float e0x = T.e0[0], e0y = T.e0[1], e0z = T.e0[2];
  const float e1x = T.e1[0], e1y = T.e1[1], e1z = T.e1[2];
  const float v0x = dy * e1z - dz * e1y;
  const float v0y = dz * e1x - dx * e1z;
  const float v0z = dx * e1y - dy * e1x;
  const float det = e0x * v0x + e0y * v0y + e0z * v0z;
  float t = DF_INF;
  if (det < -0.0001f || det > 0.0001f) {
    const float dr = recip_rn(det);
    const float v1x = ox - T.a[0], v1y = oy - T.a[1], v1z = oz - T.a[2];
    const float u = (v1x * v0x + v1y * v0y + v1z * v0z) * dr;
    if (u >= 0.0f && u <= 1.0f) {
      const float v2x = v1y * e0z - v1z * e0y;
      const float v2y = v1z * e0x - v1x * e0z;
      const float v2z = v1x * e0y - v1y * e0x;
      const float v = (dx * v2x + dy * v2y + dz * v2z) * dr;
      t = v >= 0.0f && u + v <= 1.0f ? (e1x * v2x + e1y * v2y + e1z * v2z) * dr : 0.001f;
      t = t > 0.001f && t < tmax ? t : DF_INF;
    }
  }
  return t;
}

// Terrain height field y = sin(x) sin(z) at scale 10, marched from tc = 0.001
// to 10 with td = 0.01 * tc (:1470-1506).  Returns exactly what the
// reference's step-by-step march returns, evaluating far fewer heights:
//
// * the march positions tc_k are the same for every ray (`tc`, `n`);
// * g(tc) = py - h is Lipschitz in tc with constant L = |dy| + sqrt(dx^2 +
//   dz^2): dg/dtc = dy - (dx cos px sin pz + dz sin px cos pz), and by
//   Cauchy-Schwarz the bracket is at most sqrt(dx^2 + dz^2) times
//   sqrt(cos^2 px sin^2 pz + sin^2 px cos^2 pz) <= 1; L is rounded up by a
//   relative 2^-20.  So after a step with gap g no step with
//   tc_j <= tc + (g - margin) / L can cross; `margin` bounds the float
//   rounding of py and h at both steps (positions carry ~2^-23 relative
//   error, sin ~1 ulp; it keeps the coarser |dx|+|dy|+|dz| for the
//   positions' error), with a wide safety factor;
// * it also skips to the positive root of g + g' s - D2 s^2 / 2 (a lower
//   bound of the gap s past tc, from |g''| <= D2 = dx^2 + dz^2 and a lower
//   bound on g' at tc) when that is farther: rays leaving the surface or
//   skimming it, whose gap grows or stays small, otherwise advance by tiny
//   first-order steps (1080p pass of the mesh + terrain workload: 193 M ->
//   57 M height evaluations).  df_check_terrain_march compares this march
//   with the plain one on millions of rays;
// * a crossing at step j returns t >= tc_{j-1} (rounded), so once
//   tc_{k-1} > closest the terrain cannot win the closest-hit comparison
//   (tc_{k-1} >= 0.98 tc_k);
// * h <= 1, so a ray with dy >= 0 that is above py = 1 never crosses;
// * the interpolation needs (h, py) at the step before the crossing: they
//   are recomputed there when that step was skipped.
// `steps` (if not null) counts height evaluations.
template <int M>
__device__ __forceinline__ float isect_terrain(float ox, float oy, float oz, float dx, float dy, float dz,
                                               float closest, const float* __restrict__ tc_tab, int n,
                                               unsigned* steps) {
  const float sr = 1.0f / 10.0f;
  const float bx = ox * sr, by = oy * sr, bz = oz * sr;
  const float L1 = fabsf(dx) + fabsf(dy) + fabsf(dz);
  const float L = (fabsf(dy) + sqrtf(dx * dx + dz * dz)) * 1.00000095367431640625f;
  const float margin = 1.0e-5f + 9.5367431640625e-07f * (fabsf(bx) + fabsf(by) + fabsf(bz) + 10.0f * L1);
  const float inv_l = recip_rn(L);
  // second-order bound: |g''| <= D2 = dx^2 + dz^2 (rounded up, halved)
  const float qb = (dx * dx + dz * dz) * 0.500005f;
  const float gp_err = 1.0e-5f * L1;
  float lh = 0.0f, ly = 0.0f, t = 0.0f;
  int last = -1;
  unsigned evals = 0;
  int k = 0;
  float tc = n > 0 ? tc_tab[0] : 0.0f;  // tc_tab[k]
  while (k < n) {
    if (tc * 0.98f > closest) break;
    ++evals;
    const float px = bx + dx * tc;
    const float py = by + dy * tc;
    const float pz = bz + dz * tc;
    const float sx = Math<M>::sin(px), sz = Math<M>::sin(pz);
    const float h = sx * sz;
    if (py < h) {
      float td = 0.01f;
      if (k > 0) {
        const float tp = tc_tab[k - 1];
        td = 0.01f * tp;
        if (last != k - 1) {
          ++evals;
          const float qx = bx + dx * tp;
          const float qy = by + dy * tp;
          const float qz = bz + dz * tp;
          lh = Math<M>::sin(qx) * Math<M>::sin(qz);
          ly = qy;
        }
      }
      t = tc - td + td * (lh - ly) / (py - ly - h + lh);
      break;
    }
    if (dy >= 0.0f && py > 1.0f) break;
    const float g = py - h;
    lh = h;
    ly = py;
    last = k;
    // The next position to evaluate: k + 1, or after a step with g > margin
    // the first j > k with tc_j > thr (every position up to thr is skipped).
    int j = k + 1;
    if (g > margin) {
      const float gm = g - margin;
      float skip = gm * inv_l;
      if (qb > 0.0f) {
        // g(tc + s) >= g + g' s - D2 s^2 / 2, with g' = dy - (dx cos px sin pz
        // + dz sin px cos pz) >= dy - |dx| |cos px| |sin pz| - |dz| |sin px|
        // |cos pz| - gp_err (|cos| bounded above through 1 - sin^2 with room
        // for the sines' rounding, gp_err the rounding of this sum), and the
        // directional second derivative of sin x sin z is at most D2 in
        // magnitude (|sin x sin z| + |cos x cos z| <= 1).  Its positive root,
        // cancellation-free and shrunk past its own rounding, is a second
        // skip distance; rays climbing away from the surface or skimming it
        // take far longer skips than the first-order one.
        const float cx = __builtin_amdgcn_sqrtf(1.0f - sx * sx + 2.0e-6f);
        const float cz = __builtin_amdgcn_sqrtf(1.0f - sz * sz + 2.0e-6f);
        const float gp = dy - fabsf(dx) * cx * fabsf(sz) - fabsf(dz) * fabsf(sx) * cz - gp_err;
        const float sq = __builtin_amdgcn_sqrtf(gp * gp + 4.0f * qb * gm);
        const float s2 = gp > 0.0f ? (gp + sq) * __builtin_amdgcn_rcpf(2.0f * qb)
                                   : 2.0f * gm * __builtin_amdgcn_rcpf(sq - gp);
        skip = fmaxf(skip, s2 * 0.99999f);
      }
      const float thr = tc + skip;
      // tc_{k+m} ~ tc_k * 1.01^m: a first guess (approximate reciprocal and
      // logarithm; it only sets where the search starts), then the table
      // entries on both sides of it, loaded together, and a walk from there
      // when the guess was off.
      float mf = __log2f(thr * __builtin_amdgcn_rcpf(tc)) * 69.66071689357483f;
      mf = fminf(fmaxf(mf, 0.0f), static_cast<float>(n - 1 - k));
      j = k + 1 + static_cast<int>(mf);
      float lo = j - 1 > k ? tc_tab[j - 1] : tc;
      float hi = j < n ? tc_tab[j] : __builtin_inff();
      while (j - 1 > k && lo > thr) {
        --j;
        hi = lo;
        lo = j - 1 > k ? tc_tab[j - 1] : tc;
      }
      while (j < n && hi <= thr) {
        ++j;
        hi = j < n ? tc_tab[j] : __builtin_inff();
      }
      k = j;
      tc = hi;
      continue;
    }
    k = j;
    if (k < n) tc = tc_tab[k];
  }
  if (steps) *steps += evals;
  return t > DF_EPS ? t : DF_INF;
}

}  // namespace dfk
