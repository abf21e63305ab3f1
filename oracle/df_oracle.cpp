// df_oracle.cpp - CPU oracle for the Dayflower renderer kernel.
//
// TEST INFRASTRUCTURE ONLY.  This file restates GPURendererKernel.run() and
// its callees (src/main/java/org/dayflower/pathtracer/kernel/GPURendererKernel.java,
// AbstractImageKernel.java, AbstractKernel.java) as plain C++ with the
// semantics the reference has when Aparapi executes run() in its Java thread
// pool:
//   * every float operation rounds to float (compile with -ffp-contract=off,
//     no fast-math);
//   * java.lang.Math transcendentals evaluated in double and rounded once to
//     float; rsqrt(x) = 1 / sqrt(x);
//   * Math.min/Math.max NaN propagation; Java (int) casts (truncate,
//     saturate, NaN -> 0);
//   * the 48-bit LCG of java.util.Random per pixel.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// it, always as the checker.  Nothing in the product calls it.
//
// Functions follow the reference line by line; each carries the
// file:line range it restates.  The reference's per-work-item scratch
// arrays (intersections_$local$, rays_$private$6, colorTemporarySamples)
// become per-pixel locals here; fields of the hit record that a shape does
// not write start at zero for every pixel and pass.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <utility>
#include <vector>

#include "../include/dayflower_types.h"

namespace {

// ------------------------------------------------------------ planted mistakes
// The pin tests switch on one deliberate, plausible mistake at a time (oracle_plant) to show that each of
// their checks fails on it.  0 is the faithful restatement; nothing but those tests sets another value.
enum Plant {
  PLANT_NONE = 0,
  PLANT_CAMERA_RES = 1,        // pixel coordinates normalised by res instead of res - 1
  PLANT_CAMERA_BOX = 2,        // box filter in place of the tent
  PLANT_LENS_RADIUS = 3,       // lens radius aperture * r instead of aperture * sqrt(r)
  PLANT_FISHEYE_W = 4,         // fisheye w = 1 - r^2 instead of sqrt(1 - r^2)
  PLANT_LAMBERT_UNIFORM = 5,   // uniform hemisphere instead of cosine-weighted
  PLANT_PHONG_EXPONENT = 6,    // Phong lobe exponent 20 instead of 50
  PLANT_FRESNEL_PROB = 7,      // reflection probability R instead of 0.25 + 0.5 R
  PLANT_SNELL_RATIO = 8,       // index ratio inverted
  PLANT_SUN_CONE = 9,          // sin^2 of the half-angle r / d^2 instead of r^2 / d^2
  PLANT_AO_UNIFORM = 10,       // AO rays uniform over the hemisphere
  PLANT_PHONG_SHADER = 11,     // ray caster's specular exponent a tenth of 250
  PLANT_SPHERE_UV = 12,        // atan2 operands swapped
  PLANT_BARYCENTRIC = 13,      // barycentric weights of B and C swapped
  PLANT_NORMAL_MAP_B = 14,     // normal map blue decoded like red and green
  PLANT_NOISE_NORMAL = 15,     // noise normal's y component from the z rotation of the point
  PLANT_BULLSEYE = 16,         // bullseye ring frequency 0.5 instead of 0.25
  PLANT_PEREZ_THETA0 = 17,     // chromaticity denominators at the sun's zenith angle (no theta-1 quirk)
  PLANT_PEREZ_SWAP = 18,       // luminance Perez coefficients D and E exchanged
  PLANT_FBM_TEXTURE = 19,      // fBm texture noise mapped to [-1, 1] instead of [0, 1]
  PLANT_CLEARCOAT_OFFSET = 20,  // ClearCoat reflection leaves at 0.01 instead of 0.02 along the normal
  PLANT_CHECKER_TILE = 21       // differentiable checkerboard group: its two colours swapped on the tiles
};
int g_plant = PLANT_NONE;
inline bool planted(int k) { return g_plant == k; }

// ------------------------------------------------------------ Java semantics

inline float jmaxf(float a, float b) {  // Math.max(float, float)
  if (a != a) return a;
  if (a == 0.0f && b == 0.0f && std::signbit(a)) return b;
  return a >= b ? a : b;
}
inline float jminf(float a, float b) {  // Math.min(float, float)
  if (a != a) return a;
  if (a == 0.0f && b == 0.0f && std::signbit(b)) return b;
  return a <= b ? a : b;
}
inline double jmaxd(double a, double b) {
  if (a != a) return a;
  if (a == 0.0 && b == 0.0 && std::signbit(a)) return b;
  return a >= b ? a : b;
}
inline double jmind(double a, double b) {
  if (a != a) return a;
  if (a == 0.0 && b == 0.0 && std::signbit(b)) return b;
  return a <= b ? a : b;
}
inline int32_t jf2i(float v) {  // (int) float
  if (v != v) return 0;
  if (v >= 2147483648.0f) return INT32_MAX;
  if (v <= -2147483648.0f) return INT32_MIN;
  return static_cast<int32_t>(v);
}
inline float jsqrt(float x) { return std::sqrt(x); }  // correctly rounded, == (float)Math.sqrt
inline float jrsqrt(float x) { return 1.0f / std::sqrt(x); }
inline float jsin(float x) { return static_cast<float>(std::sin(static_cast<double>(x))); }
inline float jcos(float x) { return static_cast<float>(std::cos(static_cast<double>(x))); }
inline float jtan(float x) { return static_cast<float>(std::tan(static_cast<double>(x))); }
inline float jexp(float x) { return static_cast<float>(std::exp(static_cast<double>(x))); }
inline float jpow(float a, float b) {
  return static_cast<float>(std::pow(static_cast<double>(a), static_cast<double>(b)));
}
inline float jatan2(float y, float x) {
  return static_cast<float>(std::atan2(static_cast<double>(y), static_cast<double>(x)));
}
inline float jasinpi(float x) {  // asinpi: asin(x) / pi, one rounding
  return static_cast<float>(std::asin(static_cast<double>(x)) / M_PI);
}
inline float jfloor(float x) { return std::floor(x); }
inline float jabs(float x) { return std::fabs(x); }

// MathF constants (math4j MathF.java:43-64).
const float PI = static_cast<float>(M_PI);
const float PI_DIVIDED_BY_180 = PI / 180.0f;
const float PI_MULTIPLIED_BY_TWO = PI * 2.0f;
const float PI_MULTIPLIED_BY_TWO_RECIPROCAL = 1.0f / PI_MULTIPLIED_BY_TWO;
const float PI_RECIPROCAL = 1.0f / PI;

const float EPSILON = 0.02f;        // AbstractKernel.java:51
const float INFINITY_ = 100000.0f;  // AbstractKernel.java:56
const float COLOR_RECIPROCAL = 1.0f / 255.0f;

// record layout (GPURendererKernel.java:79-92)
enum {
  R_T = 0, R_PRIM = 1, R_P = 2, R_NG = 5, R_UV = 8, R_ONB_U = 10, R_ONB_V = 13, R_ONB_W = 16, R_NS = 19,
  R_SHAPE_TYPE = 22, R_SHAPE_OFF = 23, R_TAN = 24, R_SIZE = 27
};

// noise tables (AbstractKernel.java:1126-1141)
const float kGrad3[36] = {1, 1, 0, -1, 1, 0, 1, -1, 0, -1, -1, 0, 1, 0, 1, -1, 0, 1,
                          1, 0, -1, -1, 0, -1, 0, 1, 1, 0, -1, 1, 0, 1, -1, 0, -1, -1};
const int kPermBase[256] = {
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

struct Tables {
  int perm[512];
  int perm12[512];
  Tables() {
    for (int i = 0; i < 512; ++i) {
      perm[i] = kPermBase[i & 255];
      perm12[i] = perm[i] % 12;
    }
  }
};
const Tables kTables;

inline int fast_floor(float v) {  // doFastFloor (AbstractKernel.java:1184-1188)
  const int i = jf2i(v);
  return v < static_cast<float>(i) ? i - 1 : i;
}

// CIE daylight chromaticity (x, y) and luminance -> sRGB, lifted to
// non-negative (GPURendererKernel.java:2207-2230): the M1/M2 daylight
// weights, the S0/S1/S2 x CMF integrals, then the XYZ -> linear sRGB matrix.
void daylight_xyy_to_srgb(float fx, float fy, double lum, float* r_out, float* g_out, float* b_out) {
  const float cie0 = 1.0f / (0.0241f + 0.2562f * fx - 0.7341f * fy);
  const float cie1 = (-1.3515f - 1.7703f * fx + 5.9114f * fy) * cie0;
  const float cie2 = (0.03f - 31.4424f * fx + 30.0717f * fy) * cie0;
  const float cX = 10246.121f + cie1 * 187.75537f + cie2 * 213.14803f;
  const float cY = 10676.695f + cie1 * 192.59653f + cie2 * 76.29494f;
  const float cZ = 12372.504f + cie1 * 3482.8765f + cie2 * -235.71611f;
  const float cYR = 1.0f / cY;
  // float * double promotes to double (Java binary numeric promotion)
  const float cR = static_cast<float>(static_cast<double>(cX) * lum * static_cast<double>(cYR));
  const float cG = static_cast<float>(lum);
  const float cB = static_cast<float>(static_cast<double>(cZ) * lum * static_cast<double>(cYR));
  float r = 3.2410042f * cR + -1.5373994f * cG + -0.49861607f * cB;
  float g = -0.9692241f * cR + 1.8759298f * cG + 0.041554242f * cB;
  float b = 0.05563942f * cR + -0.20401107f * cG + 1.0571486f * cB;
  const float w = jmaxf(0.0f, -jminf(0.0f, jminf(r, jminf(g, b))));
  *r_out = r + w;
  *g_out = g + w;
  *b_out = b + w;
}

// ------------------------------------------------------------ FP64 adjoint
//
// Forward-mode differentiation of the path tracer in double precision, for
// the differentiable mode (continuous albedo per surface group, emission per
// light, the sky's turbidity and the sun's elevation).  Every path is traced
// by render_pt exactly as in the forward render (same floats, same random
// numbers, Russian roulette off); beside it a tangent carries the derivative
// of the path throughput and the sample with respect to every parameter,
// in double.  Geometry and visibility do not move with the parameters, so
// this is the interior term of the gradient.  The sky's derivatives come
// from dual numbers through Sky.set and the Perez evaluation, the sun's
// from dual numbers through the sun cone sample at the forward's random
// numbers: no finite differences anywhere.

// A value with its derivatives with respect to turbidity (t), sun elevation
// in degrees (e) and sun azimuth in degrees (a).
struct D3 {
  double v, t, e, a;
  D3(double value = 0.0, double dt = 0.0, double de = 0.0, double da = 0.0) : v(value), t(dt), e(de), a(da) {}
};
inline D3 operator+(D3 x, D3 y) { return D3(x.v + y.v, x.t + y.t, x.e + y.e, x.a + y.a); }
inline D3 operator-(D3 x, D3 y) { return D3(x.v - y.v, x.t - y.t, x.e - y.e, x.a - y.a); }
inline D3 operator-(D3 x) { return D3(-x.v, -x.t, -x.e, -x.a); }
inline D3 operator*(D3 x, D3 y) {
  return D3(x.v * y.v, x.t * y.v + x.v * y.t, x.e * y.v + x.v * y.e, x.a * y.v + x.v * y.a);
}
inline D3 operator/(D3 x, D3 y) {
  const double q = x.v / y.v;
  return D3(q, (x.t - q * y.t) / y.v, (x.e - q * y.e) / y.v, (x.a - q * y.a) / y.v);
}
inline D3 chain(D3 x, double f, double df) { return D3(f, df * x.t, df * x.e, df * x.a); }
inline D3 dexp(D3 a) {
  const double f = std::exp(a.v);
  return chain(a, f, f);
}
inline D3 dsin(D3 a) { return chain(a, std::sin(a.v), std::cos(a.v)); }
inline D3 dcos(D3 a) { return chain(a, std::cos(a.v), -std::sin(a.v)); }
inline D3 dtan(D3 a) {
  const double c = std::cos(a.v);
  return chain(a, std::tan(a.v), 1.0 / (c * c));
}
inline D3 dsqrt(D3 a) {  // at zero: derivative zero (the cone's centre, where 1 - cos^2 stays zero)
  if (!(a.v > 0.0)) return D3(0.0);
  const double f = std::sqrt(a.v);
  return chain(a, f, 0.5 / f);
}
inline D3 dacos_clamped(D3 a) {  // acos(max(min(a, 1), -1)): constant outside [-1, 1]
  if (a.v >= 1.0) return D3(0.0);
  if (a.v <= -1.0) return D3(M_PI);
  return chain(a, std::acos(a.v), -1.0 / std::sqrt(1.0 - a.v * a.v));
}
inline D3 dmax0(D3 a) { return a.v > 0.0 ? a : D3(0.0); }

// Sky.set (Sky.java:295-333) as a function of turbidity T, sun elevation e
// and sun azimuth a (degrees), the sun towards (cos e sin a, sin e,
// cos e cos a) in the world (y up, azimuth from +z towards +x).  The sky
// frame is the one the forward's sky holds (u = x, v = z, w = y).
struct DualSky {
  D3 world[3];  // unit vector towards the sun, world frame
  D3 dir[3];    // the same in the sky frame
  D3 theta;     // the sun's zenith angle
  D3 zl, zx, zy;
  D3 L[5], X[5], Y[5];
  float onb_u[3], onb_v[3], onb_w[3];
};

DualSky dual_sky(const df_sky* frame, double turbidity, double elevation_deg, double azimuth_deg) {
  DualSky K;
  std::memcpy(K.onb_u, frame->onb_u, sizeof K.onb_u);
  std::memcpy(K.onb_v, frame->onb_v, sizeof K.onb_v);
  std::memcpy(K.onb_w, frame->onb_w, sizeof K.onb_w);
  const double rad = M_PI / 180.0;
  const D3 el(elevation_deg * rad, 0.0, rad, 0.0);
  const D3 az(azimuth_deg * rad, 0.0, 0.0, rad);
  const D3 T(turbidity, 1.0, 0.0, 0.0);
  K.world[0] = dcos(el) * dsin(az);
  K.world[1] = dsin(el);
  K.world[2] = dcos(el) * dcos(az);
  const float* onb[3] = {K.onb_u, K.onb_v, K.onb_w};
  for (int i = 0; i < 3; ++i) K.dir[i] = K.world[0] * onb[i][0] + K.world[1] * onb[i][1] + K.world[2] * onb[i][2];
  const D3 len = dsqrt(K.dir[0] * K.dir[0] + K.dir[1] * K.dir[1] + K.dir[2] * K.dir[2]);
  for (D3& c : K.dir) c = c / len;
  K.theta = dacos_clamped(K.dir[2]);
  const D3 th = K.theta, th2 = th * th, th3 = th2 * th, T2 = T * T;
  const double pi_f = static_cast<double>(static_cast<float>(M_PI));  // MathF.PI
  const D3 chi = (D3(4.0 / 9.0) - T / D3(120.0)) * (D3(pi_f) - D3(2.0) * th);
  K.zl = ((D3(4.0453) * T - D3(4.9710)) * dtan(chi) - D3(0.2155) * T + D3(2.4192)) * D3(1000.0);
  auto poly = [&](const double c[12]) {  // (c0 th3 + c1 th2 + c2 th + c3) T2 + (...) T + (...)
    D3 r(0.0);
    const D3 coef[3] = {T2, T, D3(1.0)};
    for (int k = 0; k < 3; ++k)
      r = r + (D3(c[4 * k]) * th3 + D3(c[4 * k + 1]) * th2 + D3(c[4 * k + 2]) * th + D3(c[4 * k + 3])) * coef[k];
    return r;
  };
  const double cx[12] = {0.00165, -0.00374, 0.00208, 0.0, -0.02902, 0.06377, -0.03202, 0.00394,
                         0.11693, -0.21196, 0.06052, 0.25885};
  const double cy[12] = {0.00275, -0.00610, 0.00316, 0.0, -0.04212, 0.08970, -0.04153, 0.00515,
                         0.15346, -0.26756, 0.06669, 0.26688};
  K.zx = poly(cx);
  K.zy = poly(cy);
  // Perez coefficients, linear in T (Preetham et al. 1999 fits, Sky.java:320-334)
  const double pl[5][2] = {{0.17872, -1.46303}, {-0.35540, 0.42749}, {-0.02266, 5.32505}, {0.12064, -2.57705},
                           {-0.06696, 0.37027}};
  const double px[5][2] = {{-0.01925, -0.25922}, {-0.06651, 0.00081}, {-0.00041, 0.21247}, {-0.06409, -0.89887},
                           {-0.00325, 0.04517}};
  const double py[5][2] = {{-0.01669, -0.26078}, {-0.09495, 0.00921}, {-0.00792, 0.21023}, {-0.04405, -1.65369},
                           {-0.01092, 0.05291}};
  for (int i = 0; i < 5; ++i) {
    K.L[i] = D3(pl[i][0]) * T + D3(pl[i][1]);
    K.X[i] = D3(px[i][0]) * T + D3(px[i][1]);
    K.Y[i] = D3(py[i][0]) * T + D3(py[i][1]);
  }
  return K;
}

// doCalculateColorForSky (GPURendererKernel.java:2127-2236) in dual numbers
// for the direction (dx, dy, dz), which does not depend on the parameters.
// Returns false below the horizon, where the sky is a constant.
bool dual_sky_rgb(const DualSky& K, float dx, float dy, float dz, D3 out[3]) {
  const float d0X = dx * K.onb_u[0] + dy * K.onb_u[1] + dz * K.onb_u[2];
  const float d0Y = dx * K.onb_v[0] + dy * K.onb_v[1] + dz * K.onb_v[2];
  float d0Z = dx * K.onb_w[0] + dy * K.onb_w[1] + dz * K.onb_w[2];
  if (!(d0Z >= 0.0f)) return false;
  d0Z = jmaxf(d0Z, 0.001f);
  const double l = 1.0 / std::sqrt(static_cast<double>(d0X) * d0X + static_cast<double>(d0Y) * d0Y +
                                   static_cast<double>(d0Z) * d0Z);
  const double d1X = d0X * l, d1Y = d0Y * l, d1Z = d0Z * l;
  const D3 dp = K.dir[0] * D3(d1X) + K.dir[1] * D3(d1Y) + K.dir[2] * D3(d1Z);
  const double theta1 = std::acos(jmaxd(jmind(d1Z, 1.0), -1.0));
  const double cosTheta1R = 1.0 / (std::cos(theta1) + 0.01);
  const D3 gamma = dacos_clamped(dp);
  const D3 cosGamma = dcos(gamma);
  const D3 cosTheta0 = dcos(K.theta);
  const D3 one(1.0);
  auto f_sun = [&](const D3* C, D3 angle) {  // 1 + C2 exp(C3 angle) + C4 cos^2
    return one + C[2] * dexp(C[3] * angle) + C[4] * cosGamma * cosGamma;
  };
  const D3* L = K.L;
  const D3* X = K.X;
  const D3* Y = K.Y;
  const D3 lden = (one + L[0] * dexp(L[1])) * (one + L[2] * dexp(L[3] * K.theta) + L[4] * cosTheta0 * cosTheta0);
  const D3 lnum = (one + L[0] * dexp(L[1] * D3(cosTheta1R))) * f_sun(L, gamma);
  const D3 lum = K.zl * lnum / lden * D3(1.0e-4);
  const D3 xden = (one + X[0] * dexp(X[1])) * (one + X[2] * dexp(X[3] * D3(theta1)) + X[4] * cosTheta0 * cosTheta0);
  const D3 xnum = (one + X[0] * dexp(X[1] * D3(cosTheta1R))) * f_sun(X, gamma);
  const D3 x = K.zx * xnum / xden;
  const D3 yden = (one + Y[0] * dexp(Y[1])) * (one + Y[2] * dexp(Y[3] * D3(theta1)) + Y[4] * cosTheta0 * cosTheta0);
  const D3 ynum = (one + Y[0] * dexp(Y[1] * D3(cosTheta1R))) * f_sun(Y, gamma);
  const D3 y = K.zy * ynum / yden;
  // daylight chromaticity -> linear sRGB (GPURendererKernel.java:2207-2230)
  const D3 cie0 = one / (D3(0.0241) + D3(0.2562) * x - D3(0.7341) * y);
  const D3 cie1 = (D3(-1.3515) - D3(1.7703) * x + D3(5.9114) * y) * cie0;
  const D3 cie2 = (D3(0.03) - D3(31.4424) * x + D3(30.0717) * y) * cie0;
  const D3 cX = D3(10246.121) + cie1 * D3(187.75537) + cie2 * D3(213.14803);
  const D3 cY = D3(10676.695) + cie1 * D3(192.59653) + cie2 * D3(76.29494);
  const D3 cZ = D3(12372.504) + cie1 * D3(3482.8765) + cie2 * D3(-235.71611);
  const D3 cR = cX * lum / cY, cG = lum, cB = cZ * lum / cY;
  const D3 r = D3(3.2410042) * cR + D3(-1.5373994) * cG + D3(-0.49861607) * cB;
  const D3 g = D3(-0.9692241) * cR + D3(1.8759298) * cG + D3(0.041554242) * cB;
  const D3 b = D3(0.05563942) * cR + D3(-0.20401107) * cG + D3(1.0571486) * cB;
  // w = max(0, -min(0, min(r, g, b))): the smallest channel lifted to zero
  const D3* lo = &r;
  if (g.v < lo->v) lo = &g;
  if (b.v < lo->v) lo = &b;
  const D3 w = lo->v < 0.0 ? -*lo : D3(0.0);
  out[0] = r + w;
  out[1] = g + w;
  out[2] = b + w;
  return true;
}

// The cosine dp1 between the normal (nX, nY, nZ) and the sun direction that
// doCalculateColorForSun (GPURendererKernel.java:2238-2347) samples with the
// random numbers (u, v) at (pX, pY, pZ), as a function of the sun's
// elevation and azimuth (the sun's origin sits 10000 units along its
// direction).
D3 dual_sun_cosine(const DualSky& K, float pX, float pY, float pZ, float nX, float nY, float nZ, float u, float v) {
  const D3* w = K.world;
  const D3 toX = w[0] * D3(10000.0) - D3(pX), toY = w[1] * D3(10000.0) - D3(pY), toZ = w[2] * D3(10000.0) - D3(pZ);
  const D3 to2 = toX * toX + toY * toY + toZ * toZ;
  const D3 sinMax2 = D3(500.0 * 500.0) / to2;
  const D3 cosMax = dsqrt(dmax0(D3(1.0) - sinMax2));
  const D3 cosT = D3(u) * (cosMax - D3(1.0)) + D3(1.0);
  const D3 sinT = dsqrt(dmax0(D3(1.0) - cosT * cosT));
  const float phi = PI_MULTIPLIED_BY_TWO * v;
  const D3 x = D3(std::cos(static_cast<double>(phi))) * sinT;
  const D3 y = D3(std::sin(static_cast<double>(phi))) * sinT;
  const D3 z = cosT;
  const double aX = std::fabs(w[0].v), aY = std::fabs(w[1].v), aZ = std::fabs(w[2].v);
  const bool c0 = aX < aY && aX < aZ;
  const bool c1 = aY < aZ;
  const D3 vX = c0 ? D3(0.0) : c1 ? w[2] : w[1];
  const D3 vY = c0 ? w[2] : c1 ? D3(0.0) : -w[0];
  const D3 vZ = c0 ? -w[1] : c1 ? -w[0] : D3(0.0);
  const D3 vl = dsqrt(vX * vX + vY * vY + vZ * vZ);
  const D3 vnX = vX / vl, vnY = vY / vl, vnZ = vZ / vl;
  const D3 unX = vnY * w[2] - vnZ * w[1], unY = vnZ * w[0] - vnX * w[2], unZ = vnX * w[1] - vnY * w[0];
  const D3 rX = unX * x + vnX * y + w[0] * z;
  const D3 rY = unY * x + vnY * y + w[1] * z;
  const D3 rZ = unZ * x + vnZ * y + w[2] * z;
  const D3 rl = dsqrt(rX * rX + rY * rY + rZ * rZ);
  return (rX * D3(nX) + rY * D3(nY) + rZ * D3(nZ)) / rl;
}

// Differentiable mode: the surface records mapped to a parameter take it in
// place of their texture.  An albedo group replaces the albedo texture by one
// colour; a checkerboard group replaces the two colours of a checkerboard
// albedo texture, the tile rule staying the texture's.  Albedo slots number
// the colours of both: slot k < n_albedo is group k, slot n_albedo + 2 g + t
// is colour t (0: the texture's first colour, used on dark tiles) of
// checkerboard group g.
struct DiffMaps {
  int n_albedo, n_emission;
  const int32_t* surface_albedo;
  const int32_t* surface_emission;
  int n_surfaces;
  const float* albedo;
  const float* emission;
  int n_checker;
  const int32_t* surface_checker;  // null: no checkerboard groups
  const float* checker;            // n_checker x 2 x RGB
};

// The tangent of one path: its throughput in double, the derivative of the
// throughput with respect to the albedo slots the path has met so far (the
// only parameters it depends on), and the pixel's running sum of the
// derivatives of its samples, over all parameters.  Gradient rows: albedo
// groups x RGB, emissions, checkerboard groups x 2 x RGB, turbidity, sun
// elevation, sun azimuth.
struct PathTangent {
  static constexpr int kMaxSlots = 64;
  const DualSky* sky;
  int n_albedo, n_emission, i_emission, i_turbidity, i_elevation, i_azimuth;
  double m[3];
  int slots[kMaxSlots];
  double dm[kMaxSlots][3];
  int n_slots;
  double* dpix;  // (3 n_albedo + n_emission + 6 n_checker + 3) x 3

  // the gradient row of channel 0 of albedo slot `slot`
  int row(int slot) const { return 3 * slot + (slot >= n_albedo ? n_emission : 0); }
  void begin() {
    m[0] = m[1] = m[2] = 1.0;
    n_slots = 0;
  }
  int slot_row(int slot) {
    for (int k = 0; k < n_slots; ++k)
      if (slots[k] == slot) return k;
    slots[n_slots] = slot;
    dm[n_slots][0] = dm[n_slots][1] = dm[n_slots][2] = 0.0;
    return n_slots++;
  }
  // d(sample) += d(m) * g: the term m * g with g not depending on the albedos
  void through_throughput(const double g[3]) {
    for (int k = 0; k < n_slots; ++k)
      for (int c = 0; c < 3; ++c) dpix[(row(slots[k]) + c) * 3 + c] += dm[k][c] * g[c];
  }
  // m *= s, s free of parameters
  void scale(float s) {
    for (int c = 0; c < 3; ++c) m[c] *= s;
    for (int k = 0; k < n_slots; ++k)
      for (int c = 0; c < 3; ++c) dm[k][c] *= s;
  }
  // m *= a, a the albedo of slot `slot` (-1: a texture, no parameter)
  void albedo(int slot, const float a[3]) {
    for (int k = 0; k < n_slots; ++k)
      for (int c = 0; c < 3; ++c) dm[k][c] *= a[c];
    if (slot >= 0) {
      const int k = slot_row(slot);
      for (int c = 0; c < 3; ++c) dm[k][c] += m[c];  // (m a)' = m' a + m a'; m' a is applied above
    }
    for (int c = 0; c < 3; ++c) m[c] *= a[c];
  }
};

struct Counters {
  uint64_t rays = 0, samples = 0, box_tests = 0, tri_tests = 0, prim_tests = 0, sphere_tests = 0, terrain_steps = 0,
           sky_evals = 0;
};

struct Scene {
  const df_compiled_scene* a;
  const df_sky* sky;
  const df_params* p;
  int prim_count;
  const DiffMaps* diff = nullptr;  // differentiable mode, or null
};

struct Frame {
  int width, height;
  float* film;
  int32_t* film_samples;
  float* image_float;
  uint8_t* image_bgra;
  int32_t* primitive_offsets;
  int64_t* seeds;
  int clear;  // film flag FILM_FLAG_CLEAR for this pass
};

// ------------------------------------------------------------ work item

class Item {
 public:
  Item(const Scene& s, const Frame& f, int gid, Counters& c) : s_(s), f_(f), gid_(gid), cnt_(c) {
    std::memset(rec_, 0, sizeof rec_);
  }

  void run();  // GPURendererKernel.run (:252-308)
  // run() with the primary ray given instead of doCreatePrimaryRay's
  void run_ray(const float ray[6]);
  // doCreatePrimaryRay (:573-724): false for no ray (fisheye, outside the unit disk)
  bool primary_ray(float ray[6]) {
    const bool ok = create_primary_ray();
    std::memcpy(ray, ray_, sizeof ray_);
    return ok;
  }

  // exposed for unit tests
  float intersect(float ox, float oy, float oz, float dx, float dy, float dz, bool testing) {
    return do_intersect_primitives(ox, oy, oz, dx, dy, dz, testing);
  }
  const float* record() const { return rec_; }
  void sky_color(bool active, float dx, float dy, float dz, float out[3]) {
    do_sky(active, dx, dy, dz);
    std::memcpy(out, tmp_, sizeof tmp_);
  }
  float simplex3(float x, float y, float z) const { return simplex_noise_xyz(x, y, z); }
  float simplex2(float x, float y) const { return simplex_noise_xy(x, y); }
  float fbm3(float f, float g, float mn, float mx, int o, float x, float y, float z) const {
    return simplex_fbm_xyz(f, g, mn, mx, o, x, y, z);
  }
  int texture_color(int off) { return tex_color(off); }
  void set_record(const float* r) { std::memcpy(rec_, r, sizeof rec_); }
  // forward-mode derivatives of the path tracer's samples (null: none)
  void set_tangent(PathTangent* t) { tan_ = t; }

 private:
  const Scene& s_;
  const Frame& f_;
  int gid_;
  Counters& cnt_;
  float rec_[R_SIZE];
  float ray_[6];
  float tmp_[3];
  PathTangent* tan_ = nullptr;
  // the last sun sample: its random numbers, cosine, and whether it lit the point
  float sun_u_ = 0.0f, sun_v_ = 0.0f, sun_dp1_ = 0.0f;
  bool sun_lit_ = false;

  const float* cam() const { return s_.a->camera; }
  const float* p3() const { return s_.a->point3fs; }
  const float* v3() const { return s_.a->vector3fs; }

  // ---- RNG (AbstractKernel.java:438-440, 1101-1110)
  float next_float() {
    const uint64_t old = static_cast<uint64_t>(f_.seeds[gid_]);
    const uint64_t nw = (old * 0x5DEECE66Dull + 0xBull) & ((1ull << 48) - 1);
    f_.seeds[gid_] = static_cast<int64_t>(nw);
    const int bits = static_cast<int>(nw >> (48 - 24));
    return static_cast<float>(bits) * (1.0f / static_cast<float>(1 << 24));
  }

  // ---- film / image (AbstractImageKernel.java:333-472, 503-774)
  void film_add(float r, float g, float b) {
    cnt_.samples++;
    float* fd = f_.film + gid_ * 3;
    if (f_.clear) {
      fd[0] = r;
      fd[1] = g;
      fd[2] = b;
      f_.film_samples[gid_] = 1;
    } else {
      const int n = f_.film_samples[gid_] + 1;
      fd[0] = fd[0] + ((r - fd[0]) / static_cast<float>(n));
      fd[1] = fd[1] + ((g - fd[1]) / static_cast<float>(n));
      fd[2] = fd[2] + ((b - fd[2]) / static_cast<float>(n));
      f_.film_samples[gid_] = n;
    }
  }
  void film_set(float r, float g, float b) {
    cnt_.samples++;
    float* fd = f_.film + gid_ * 3;
    fd[0] = r;
    fd[1] = g;
    fd[2] = b;
    f_.film_samples[gid_] = 1;
  }
  void image_pass();

 public:
  void tonemap_only() { image_pass(); }

 private:
  // ---- helpers
  float modulo(float v) { return v - jfloor(v); }              // AbstractKernel.java:427-429
  static float remainder(float x, float y) {                    // AbstractKernel.java:533-537
    const int n = jf2i(x / y);
    return x - static_cast<float>(n) * y;
  }
  static float saturate(float v, float a, float b) {            // AbstractKernel.java:553-558
    const float mn = jminf(a, b), mx = jmaxf(a, b);
    return v < mn ? mn : v > mx ? mx : v;
  }
  static float lerp(float a, float b, float t) { return (1.0f - t) * a + t * b; }
  static float blerp(float v00, float v01, float v10, float v11, float tx, float ty) {
    return lerp(lerp(v00, v01, tx), lerp(v10, v11, tx), ty);
  }

  // ---- noise (AbstractKernel.java:572-917)
  float simplex_noise_xy(float x, float y) const;
  float simplex_noise_xyz(float x, float y, float z) const;
  float simplex_fbm_xyz(float frequency, float gain, float minimum, float maximum, int octaves, float x, float y,
                        float z) const {
    float amp = 1.0f, max_amp = 0.0f, freq = frequency, noise = 0.0f;
    for (int i = 0; i < octaves; ++i) {
      noise += simplex_noise_xyz(x * freq, y * freq, z * freq) * amp;
      max_amp += amp;
      amp *= gain;
      freq *= 2.0f;
    }
    noise /= max_amp;
    noise = noise * (maximum - minimum) / 2.0f + (maximum + minimum) / 2.0f;
    return noise;
  }
  float simplex_fractal_xy(float amplitude, float frequency, float gain, float lacunarity, int octaves, float x,
                           float y) const {
    float result = 0.0f, amp = amplitude, freq = frequency;
    for (int i = 0; i < octaves; ++i) {
      result += amp * simplex_noise_xy(x * freq, y * freq);
      amp *= gain;
      freq *= lacunarity;
    }
    return result;
  }

  // ---- kernel functions
  bool create_primary_ray();
  float do_intersect_primitives(float ox, float oy, float oz, float dx, float dy, float dz, bool testing);
  static float intersect_plane(float ox, float oy, float oz, float dx, float dy, float dz, float ax, float ay,
                               float az, float nx, float ny, float nz, float tmin, float tmax);
  static float intersect_sphere(float ox, float oy, float oz, float dx, float dy, float dz, float px, float py,
                                float pz, float r, float tmin, float tmax);
  static float intersect_terrain(float ox, float oy, float oz, float dx, float dy, float dz, uint64_t* steps);
  static float intersect_triangle(float ox, float oy, float oz, float dx, float dy, float dz, float ax, float ay,
                                  float az, float bx, float by, float bz, float cx, float cy, float cz, float tmin,
                                  float tmax);
  void props_plane(float ox, float oy, float oz, float dx, float dy, float dz, float t, int shape_off);
  void props_sphere(float ox, float oy, float oz, float dx, float dy, float dz, float t, int shape_off);
  void props_terrain(float ox, float oy, float oz, float dx, float dy, float dz, float t);
  void props_triangle(float ox, float oy, float oz, float dx, float dy, float dz, float t, int tri_off);
  void to_world(const float* o2w, const float* w2o);
  void normal_map_image(int prim_off);
  void normal_map_noise(int prim_off);
  int tex_color(int off);
  int tex_color2(int off);
  int tex_blend(int off);
  int tex_bullseye(int off);
  int tex_checker(int off);
  int checker_tile(int off);
  int tex_fbm(int off);
  int tex_image(int off);
  int tex_image_simple(int off);
  int tex_surface_normal();
  int tex_uv();
  int phong_shader0(bool check, float px, float py, float pz, float nx, float ny, float nz, float vx, float vy,
                    float vz, float ar, float ag, float ab, float kar, float kag, float kab, float kdr, float kdg,
                    float kdb, float ksr, float ksg, float ksb, float ns);
  void do_sky(bool active, float dx, float dy, float dz);
  void do_sun(bool active, float px, float py, float pz, float nx, float ny, float nz, float ar, float ag, float ab);

 public:
  void sun_direction(float pX, float pY, float pZ, float u, float v, float d[3]) const;

 private:
  void render_ao();
  void render_pt(bool sky_active, bool sun_active);

 public:
  void scatter(int mat, int aslot, float pX, float pY, float pZ, float sX, float sY, float sZ, float wX, float wY,
               float wZ, float dp2, float aR, float aG, float aB, float& ox, float& oy, float& oz, float& dx,
               float& dy, float& dz, float& mr, float& mg, float& mb);
  // the oriented normal w (s flipped to face the incoming direction d) and dp2 = 2 s.d (:3219-3225)
  static void orient(float sX, float sY, float sZ, float dx, float dy, float dz, float* w, float* dp2) {
    const float dp = sX * dx + sY * dy + sZ * dz;
    *dp2 = dp * 2.0f;
    const bool ok = dp < 0.0f;
    w[0] = ok ? sX : -sX;
    w[1] = ok ? sY : -sY;
    w[2] = ok ? sZ : -sZ;
  }

 private:
  void tan_sky(bool active, float dx, float dy, float dz);
  void tan_vertex(int aslot, int eslot, const float a[3], const float e[3], bool sun_term, float pX, float pY,
                  float pZ, float wX, float wY, float wZ);
  void render_rc(bool sky_active);
  void render_rm(bool sky_active);
  void render_rt(bool sky_active);
  void render_sn(bool sky_active);
  void render_wireframes();
  void shade_primary();
  float get_y(float x, float z) const {
    const df_params* p = s_.p;
    return simplex_fractal_xy(p->global_amplitude, p->global_frequency, p->global_gain, p->global_lacunarity,
                              p->global_octaves, x, z);
  }
};

// ------------------------------------------------------------ noise

float Item::simplex_noise_xy(float x, float y) const {  // AbstractKernel.java:762-806
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
  const int* P = kTables.perm;
  const int* P12 = kTables.perm12;
  const int gi0 = P12[ii + P[jj]];
  const int gi1 = P12[ii + i1 + P[jj + j1]];
  const int gi2 = P12[ii + 1 + P[jj + 1]];
  const float t0 = 0.5f - x0 * x0 - y0 * y0;
  const float n0 = t0 < 0.0f ? 0.0f : (t0 * t0) * (t0 * t0) * (kGrad3[gi0 * 3] * x0 + kGrad3[gi0 * 3 + 1] * y0);
  const float t1 = 0.5f - x1 * x1 - y1 * y1;
  const float n1 = t1 < 0.0f ? 0.0f : (t1 * t1) * (t1 * t1) * (kGrad3[gi1 * 3] * x1 + kGrad3[gi1 * 3 + 1] * y1);
  const float t2 = 0.5f - x2 * x2 - y2 * y2;
  const float n2 = t2 < 0.0f ? 0.0f : (t2 * t2) * (t2 * t2) * (kGrad3[gi2 * 3] * x2 + kGrad3[gi2 * 3 + 1] * y2);
  return 70.0f * (n0 + n1 + n2);
}

float Item::simplex_noise_xyz(float x, float y, float z) const {  // AbstractKernel.java:808-917
  const float F3 = 1.0f / 3.0f, G3 = 1.0f / 6.0f;
  const float s = (x + y + z) * F3;
  const int i = fast_floor(x + s);
  const int j = fast_floor(y + s);
  const int k = fast_floor(z + s);
  const float t = static_cast<float>(i + j + k) * G3;
  const float x0 = x - (static_cast<float>(i) - t);
  const float y0 = y - (static_cast<float>(j) - t);
  const float z0 = z - (static_cast<float>(k) - t);
  int i1, j1, k1, i2, j2, k2;
  if (x0 >= y0) {
    if (y0 >= z0) {
      i1 = 1; j1 = 0; k1 = 0; i2 = 1; j2 = 1; k2 = 0;
    } else if (x0 >= z0) {
      i1 = 1; j1 = 0; k1 = 0; i2 = 1; j2 = 0; k2 = 1;
    } else {
      i1 = 0; j1 = 0; k1 = 1; i2 = 1; j2 = 0; k2 = 1;
    }
  } else {
    if (y0 < z0) {
      i1 = 0; j1 = 0; k1 = 1; i2 = 0; j2 = 1; k2 = 1;
    } else if (x0 < z0) {
      i1 = 0; j1 = 1; k1 = 0; i2 = 0; j2 = 1; k2 = 1;
    } else {
      i1 = 0; j1 = 1; k1 = 0; i2 = 1; j2 = 1; k2 = 0;
    }
  }
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
  const int* P = kTables.perm;
  const int* P12 = kTables.perm12;
  const int gi0 = P12[ii + P[jj + P[kk]]];
  const int gi1 = P12[ii + i1 + P[jj + j1 + P[kk + k1]]];
  const int gi2 = P12[ii + i2 + P[jj + j2 + P[kk + k2]]];
  const int gi3 = P12[ii + 1 + P[jj + 1 + P[kk + 1]]];
  auto dot3 = [](int g, float a, float b, float c) {
    return kGrad3[g * 3] * a + kGrad3[g * 3 + 1] * b + kGrad3[g * 3 + 2] * c;
  };
  const float t0 = 0.6f - x0 * x0 - y0 * y0 - z0 * z0;
  const float n0 = t0 < 0.0f ? 0.0f : (t0 * t0) * (t0 * t0) * dot3(gi0, x0, y0, z0);
  const float t1 = 0.6f - x1 * x1 - y1 * y1 - z1 * z1;
  const float n1 = t1 < 0.0f ? 0.0f : (t1 * t1) * (t1 * t1) * dot3(gi1, x1, y1, z1);
  const float t2 = 0.6f - x2 * x2 - y2 * y2 - z2 * z2;
  const float n2 = t2 < 0.0f ? 0.0f : (t2 * t2) * (t2 * t2) * dot3(gi2, x2, y2, z2);
  const float t3 = 0.6f - x3 * x3 - y3 * y3 - z3 * z3;
  const float n3 = t3 < 0.0f ? 0.0f : (t3 * t3) * (t3 * t3) * dot3(gi3, x3, y3, z3);
  return 32.0f * (n0 + n1 + n2 + n3);
}

// ------------------------------------------------------------ primary ray

bool Item::create_primary_ray() {  // GPURendererKernel.java:573-724
  const float* c = cam();
  const int y = gid_ / f_.width;
  const int x = gid_ - y * f_.width;
  const float eyeX = c[1], eyeY = c[2], eyeZ = c[3];
  const float uX = c[7], uY = c[8], uZ = c[9];
  const float vX = c[10], vY = c[11], vZ = c[12];
  const float wX = c[13], wY = c[14], wZ = c[15];
  const float fovX1 = jtan(c[4] * 0.5f * PI_DIVIDED_BY_180);
  const float fovY1 = jtan(-c[5] * 0.5f * PI_DIVIDED_BY_180);
  const float hX = uX * fovX1, hY = uY * fovX1, hZ = uZ * fovX1;
  const float vtX = vX * fovY1, vtY = vY * fovY1, vtZ = vZ * fovY1;
  float sampleX = 0.5f, sampleY = 0.5f;
  if (s_.p->renderer_type == DF_RENDERER_PATH_TRACER) {
    sampleX = next_float();
    sampleY = next_float();
    sampleX = sampleX < 0.5f ? jsqrt(2.0f * sampleX) - 1.0f : 1.0f - jsqrt(2.0f - 2.0f * sampleX);
    sampleY = sampleY < 0.5f ? jsqrt(2.0f * sampleY) - 1.0f : 1.0f - jsqrt(2.0f - 2.0f * sampleY);
    if (planted(PLANT_CAMERA_BOX)) {
      sampleX = jsqrt(2.0f * sampleX + 2.0f) * 0.5f - 0.5f;  // any non-tent jitter in [-0.5, 0.5]
      sampleY = 0.0f;
    }
  }
  const float rx = planted(PLANT_CAMERA_RES) ? c[16] : c[16] - 1.0f;
  const float sx = (sampleX + static_cast<float>(x)) / rx;
  const float sy = (sampleY + static_cast<float>(y)) / (c[17] - 1.0f);
  const float sx0 = 2.0f * sx - 1.0f;
  const float sy0 = 2.0f * sy - 1.0f;
  float w = 1.0f;
  if (c[18] == 2.0f) {  // CAMERA_LENS_FISHEYE
    const float dp = sx0 * sx0 + sy0 * sy0;
    if (dp > 1.0f) return false;
    w = planted(PLANT_FISHEYE_W) ? 1.0f - dp : jsqrt(1.0f - dp);
  }
  const float mX = eyeX + wX * w, mY = eyeY + wY * w, mZ = eyeZ + wZ * w;
  const float ppX = mX + (hX * sx0) + (vtX * sy0);
  const float ppY = mY + (hY * sx0) + (vtY * sy0);
  const float ppZ = mZ + (hZ * sx0) + (vtZ * sy0);
  const float focal = c[6];
  const float piX = eyeX + (ppX - eyeX) * focal;
  const float piY = eyeY + (ppY - eyeY) * focal;
  const float piZ = eyeZ + (ppZ - eyeZ) * focal;
  const float aperture = c[0];
  float apX = eyeX, apY = eyeY, apZ = eyeZ;
  if (aperture > 0.00001f) {
    const float r1 = next_float();
    const float r2 = next_float();
    const float angle = PI_MULTIPLIED_BY_TWO * r1;
    const float dist = planted(PLANT_LENS_RADIUS) ? aperture * r2 : aperture * jsqrt(r2);
    const float ax = jcos(angle) * dist;
    const float ay = jsin(angle) * dist;
    apX = eyeX + uX * ax + vX * ay;
    apY = eyeY + uY * ax + vY * ay;
    apZ = eyeZ + uZ * ax + vZ * ay;
  }
  const float d0X = piX - apX, d0Y = piY - apY, d0Z = piZ - apZ;
  const float lr = jrsqrt(d0X * d0X + d0Y * d0Y + d0Z * d0Z);
  ray_[0] = apX;
  ray_[1] = apY;
  ray_[2] = apZ;
  ray_[3] = d0X * lr;
  ray_[4] = d0Y * lr;
  ray_[5] = d0Z * lr;
  return true;
}

// ------------------------------------------------------------ intersection

float Item::intersect_plane(float ox, float oy, float oz, float dx, float dy, float dz, float ax, float ay, float az,
                            float nx, float ny, float nz, float tmin, float tmax) {  // :1414-1431
  const float dp = nx * dx + ny * dy + nz * dz;
  if (dp < 0.0f || dp > 0.0f) {
    const float t = ((ax - ox) * nx + (ay - oy) * ny + (az - oz) * nz) / dp;
    if (t > tmin && t < tmax) return t;
  }
  return INFINITY_;
}

float Item::intersect_sphere(float ox, float oy, float oz, float dx, float dy, float dz, float px, float py,
                             float pz, float radius, float tmin, float tmax) {  // :1433-1467
  const float tox = ox - px, toy = oy - py, toz = oz - pz;
  const float r2 = radius * radius;
  const float a = dx * dx + dy * dy + dz * dz;
  const float b = 2.0f * (tox * dx + toy * dy + toz * dz);
  const float c = (tox * tox + toy * toy + toz * toz) - r2;
  const float disc2 = b * b - 4.0f * a * c;
  if (disc2 >= 0.0f) {
    const float disc = jsqrt(disc2);
    const float q = -0.5f * (b < 0.0f ? b - disc : b + disc);
    const float r0 = q / a;
    const float r1 = q == 0.0f ? r0 : c / q;
    const float t0 = jminf(r0, r1);
    const float t1 = jmaxf(r0, r1);
    if (t0 > tmin && t0 < tmax) return t0;
    if (t1 > tmin && t1 < tmax) return t1;
  }
  return INFINITY_;
}

float Item::intersect_terrain(float ox, float oy, float oz, float dx, float dy, float dz,
                             uint64_t* steps) {  // :1470-1506
  const float scale = 10.0f;
  const float sr = 1.0f / scale;
  float t = 0.0f;
  const float tmin = 0.001f, tmax = 10.0f;
  float td = 0.01f;
  float lh = 0.0f, ly = 0.0f;
  for (float tc = tmin; tc < tmax; tc += td) {
    ++*steps;
    const float px = ox * sr + dx * tc;
    const float py = oy * sr + dy * tc;
    const float pz = oz * sr + dz * tc;
    const float h = jsin(px) * jsin(pz);
    if (py < h) {
      t = tc - td + td * (lh - ly) / (py - ly - h + lh);
      tc = tmax;
    }
    td = 0.01f * tc;
    lh = h;
    ly = py;
  }
  t = t > EPSILON ? t : INFINITY_;
  return t;
}

float Item::intersect_triangle(float ox, float oy, float oz, float dx, float dy, float dz, float ax, float ay,
                               float az, float bx, float by, float bz, float cx, float cy, float cz, float tmin,
                               float tmax) {  // :1541-1594
  const float e0x = bx - ax, e0y = by - ay, e0z = bz - az;
  const float e1x = cx - ax, e1y = cy - ay, e1z = cz - az;
  const float v0x = dy * e1z - dz * e1y;
  const float v0y = dz * e1x - dx * e1z;
  const float v0z = dx * e1y - dy * e1x;
  const float det = e0x * v0x + e0y * v0y + e0z * v0z;
  float t = INFINITY_;
  if (det < -0.0001f || det > 0.0001f) {
    const float dr = 1.0f / det;
    const float v1x = ox - ax, v1y = oy - ay, v1z = oz - az;
    const float u = (v1x * v0x + v1y * v0y + v1z * v0z) * dr;
    if (u >= 0.0f && u <= 1.0f) {
      const float v2x = v1y * e0z - v1z * e0y;
      const float v2y = v1z * e0x - v1x * e0z;
      const float v2z = v1x * e0y - v1y * e0x;
      const float v = (dx * v2x + dy * v2y + dz * v2z) * dr;
      t = v >= 0.0f && u + v <= 1.0f ? (e1x * v2x + e1y * v2y + e1z * v2z) * dr : tmin;
      t = t > tmin && t < tmax ? t : INFINITY_;
    }
  }
  return t;
}

float Item::do_intersect_primitives(float ox, float oy, float oz, float dx, float dy, float dz,
                                    bool testing) {  // GPURendererKernel.java:1048-1411
  cnt_.rays++;
  const df_compiled_scene* A = s_.a;
  float closest = INFINITY_;
  int cm = -1, cp = -1, cst = -1, cso = -1;
  for (int i = 0; i < s_.prim_count; ++i) {
    float cur = INFINITY_;
    const int po = i * 3;
    const int mo = i * 16;
    const int st = A->primitives[po + 0];
    const int so = A->primitives[po + 1];
    const float* m = A->world_to_object + mo;
    cnt_.prim_tests++;
    const float oxo = m[0] * ox + m[1] * oy + m[2] * oz + m[3];
    const float oyo = m[4] * ox + m[5] * oy + m[6] * oz + m[7];
    const float ozo = m[8] * ox + m[9] * oy + m[10] * oz + m[11];
    const float dxo = m[0] * dx + m[1] * dy + m[2] * dz;
    const float dyo = m[4] * dx + m[5] * dy + m[6] * dz;
    const float dzo = m[8] * dx + m[9] * dy + m[10] * dz;
    const float rdx = 1.0f / dxo, rdy = 1.0f / dyo, rdz = 1.0f / dzo;
    if (st == 5) {  // TriangleMesh: skip-pointer BVH walk (:1097-1194)
      const int base = so;
      int rel = 0;
      while (rel != -1) {
        const int off = base + rel;
        const int* node = A->bvh + off;
        const float* mn = p3() + node[2];
        const float* mx = p3() + node[3];
        cnt_.box_tests++;
        const float t0x = (mn[0] - oxo) * rdx, t0y = (mn[1] - oyo) * rdy, t0z = (mn[2] - ozo) * rdz;
        const float t1x = (mx[0] - oxo) * rdx, t1y = (mx[1] - oyo) * rdy, t1z = (mx[2] - ozo) * rdz;
        const float tmax_ = jminf(jmaxf(t0x, t1x), jminf(jmaxf(t0y, t1y), jmaxf(t0z, t1z)));
        const float tmin_ = jmaxf(jminf(t0x, t1x), jmaxf(jminf(t0y, t1y), jminf(t0z, t1z)));
        const float t = tmin_ > tmax_ ? -1.0f
                        : tmin_ > 0.001f && tmin_ < closest ? tmin_
                        : tmax_ > 0.001f && tmax_ < closest ? tmax_
                                                             : -1.0f;
        if (t < 0.0f) {
          rel = node[1];
        } else if (node[0] == 1) {
          rel = node[4];
        } else {
          const int n = node[4];
          for (int j = 0; j < n; ++j) {
            const int to = node[5 + j];
            const float* a = p3() + A->triangles[to + 0];
            const float* b = p3() + A->triangles[to + 1];
            const float* c = p3() + A->triangles[to + 2];
            cnt_.tri_tests++;
            cur = intersect_triangle(oxo, oyo, ozo, dxo, dyo, dzo, a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1],
                                     c[2], 0.001f, closest);
            if (cur < closest) {
              closest = cur;
              cm = mo;
              cp = po;
              cst = 4;
              cso = to;
            }
            if (testing && cp != -1) return closest;
          }
          rel = node[1];
        }
      }
    } else if (st == 1) {
      const int* pl = A->planes + so;
      const float* a = p3() + pl[0];
      const float* n = v3() + pl[3];
      cur = intersect_plane(oxo, oyo, ozo, dxo, dyo, dzo, a[0], a[1], a[2], n[0], n[1], n[2], 0.001f, closest);
    } else if (st == 2) {
      cnt_.sphere_tests++;
      const float* pos = p3() + jf2i(A->spheres[so + 0]);
      cur = intersect_sphere(oxo, oyo, ozo, dxo, dyo, dzo, pos[0], pos[1], pos[2], A->spheres[so + 1], 0.001f,
                             closest);
    } else if (st == 3) {
      cur = intersect_terrain(oxo, oyo, ozo, dxo, dyo, dzo, &cnt_.terrain_steps);
    } else if (st == 4) {
      const float* a = p3() + A->triangles[so + 0];
      const float* b = p3() + A->triangles[so + 1];
      const float* c = p3() + A->triangles[so + 2];
      cnt_.tri_tests++;
      cur = intersect_triangle(oxo, oyo, ozo, dxo, dyo, dzo, a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2],
                               0.001f, closest);
    }
    if (cur < closest) {
      closest = cur;
      cm = mo;
      cp = po;
      cst = st;
      cso = so;
    }
    if (testing && cp != -1) return closest;
  }
  if (!testing) {
    if (cp != -1) {
      const float* w2o = A->world_to_object + cm;
      const float* o2w = A->object_to_world + cm;
      const float oxo = w2o[0] * ox + w2o[1] * oy + w2o[2] * oz + w2o[3];
      const float oyo = w2o[4] * ox + w2o[5] * oy + w2o[6] * oz + w2o[7];
      const float ozo = w2o[8] * ox + w2o[9] * oy + w2o[10] * oz + w2o[11];
      const float dxo = w2o[0] * dx + w2o[1] * dy + w2o[2] * dz;
      const float dyo = w2o[4] * dx + w2o[5] * dy + w2o[6] * dz;
      const float dzo = w2o[8] * dx + w2o[9] * dy + w2o[10] * dz;
      rec_[R_T] = closest;
      rec_[R_PRIM] = static_cast<float>(cp);
      rec_[R_SHAPE_TYPE] = static_cast<float>(cst);
      rec_[R_SHAPE_OFF] = static_cast<float>(cso);
      if (cst == 1) {
        props_plane(oxo, oyo, ozo, dxo, dyo, dzo, closest, cso);
      } else if (cst == 2) {
        props_sphere(oxo, oyo, ozo, dxo, dyo, dzo, closest, cso);
      } else if (cst == 3) {
        props_terrain(oxo, oyo, ozo, dxo, dyo, dzo, closest);
      } else if (cst == 4) {
        props_triangle(oxo, oyo, ozo, dxo, dyo, dzo, closest, cso);
      }
      to_world(o2w, w2o);
      normal_map_image(cp);
      normal_map_noise(cp);
    } else {
      rec_[R_T] = INFINITY_;
      rec_[R_PRIM] = -1.0f;
      rec_[R_SHAPE_TYPE] = -1.0f;
      rec_[R_SHAPE_OFF] = -1.0f;
    }
  }
  return closest;
}

// ------------------------------------------------------------ surface properties

void Item::props_plane(float ox, float oy, float oz, float dx, float dy, float dz, float t,
                       int so) {  // :2349-2450
  const int* pl = s_.a->planes + so;
  const float* A = p3() + pl[0];
  const float* B = p3() + pl[1];
  const float* C = p3() + pl[2];
  const float* N = v3() + pl[3];
  const float aX = A[0], aY = A[1], aZ = A[2], bX = B[0], bY = B[1], bZ = B[2], cX = C[0], cY = C[1], cZ = C[2];
  const float nX = N[0], nY = N[1], nZ = N[2];
  const float pX = ox + dx * t, pY = oy + dy * t, pZ = oz + dz * t;
  const float anX = jabs(nX), anY = jabs(nY), anZ = jabs(nZ);
  const bool isX = anX > anY && anX > anZ;
  const bool isY = anY > anZ;
  const float a1X = isX ? aY : isY ? aZ : aX;
  const float a1Y = isX ? aZ : isY ? aX : aY;
  const float b1X = isX ? cY - aX : isY ? cZ - aX : cX - aX;
  const float b1Y = isX ? cZ - aY : isY ? cX - aY : cY - aY;
  const float c1X = isX ? bY - aX : isY ? bZ - aX : bX - aX;
  const float c1Y = isX ? bZ - aY : isY ? bX - aY : bY - aY;
  const float det = b1X * c1Y - b1Y * c1X;
  const float dr = 1.0f / det;
  const float bNU = -b1Y * dr, bNV = b1X * dr, bND = (b1Y * a1X - b1X * a1Y) * dr;
  const float cNU = c1Y * dr, cNV = -c1X * dr, cND = (c1X * a1Y - c1Y * a1X) * dr;
  const float hU = isX ? pY : isY ? pZ : pX;
  const float hV = isX ? pZ : isY ? pX : pY;
  const float u = hU * bNU + hV * bNV + bND;
  const float v = hU * cNU + hV * cNV + cND;
  const float wX = nX, wY = nY, wZ = nZ;
  const bool c0 = anX < anY && anX < anZ;
  const bool c1 = anY < anZ;
  const float v0X = c0 ? 0.0f : c1 ? wZ : wY;
  const float v0Y = c0 ? wZ : c1 ? 0.0f : -nX;
  const float v0Z = c0 ? -wY : c1 ? -nX : 0.0f;
  const float vr = jrsqrt(v0X * v0X + v0Y * v0Y + v0Z * v0Z);
  const float v1X = v0X * vr, v1Y = v0Y * vr, v1Z = v0Z * vr;
  const float uX = v1Y * wZ - v1Z * wY, uY = v1Z * wX - v1X * wZ, uZ = v1X * wY - v1Y * wX;
  rec_[R_ONB_U] = uX; rec_[R_ONB_U + 1] = uY; rec_[R_ONB_U + 2] = uZ;
  rec_[R_ONB_V] = v1X; rec_[R_ONB_V + 1] = v1Y; rec_[R_ONB_V + 2] = v1Z;
  rec_[R_ONB_W] = wX; rec_[R_ONB_W + 1] = wY; rec_[R_ONB_W + 2] = wZ;
  rec_[R_P] = pX; rec_[R_P + 1] = pY; rec_[R_P + 2] = pZ;
  rec_[R_NG] = nX; rec_[R_NG + 1] = nY; rec_[R_NG + 2] = nZ;
  rec_[R_NS] = nX; rec_[R_NS + 1] = nY; rec_[R_NS + 2] = nZ;
  rec_[R_UV] = u; rec_[R_UV + 1] = v;
}

void Item::props_sphere(float ox, float oy, float oz, float dx, float dy, float dz, float t,
                        int so) {  // :2452-2492
  const float* pos = p3() + jf2i(s_.a->spheres[so]);
  const float pX = ox + dx * t, pY = oy + dy * t, pZ = oz + dz * t;
  const float n0X = pX - pos[0], n0Y = pY - pos[1], n0Z = pZ - pos[2];
  const float r = jrsqrt(n0X * n0X + n0Y * n0Y + n0Z * n0Z);
  const float nX = n0X * r, nY = n0Y * r, nZ = n0Z * r;
  const float u = 0.5f + (planted(PLANT_SPHERE_UV) ? jatan2(-nX, -nZ) : jatan2(-nZ, -nX)) * PI_MULTIPLIED_BY_TWO_RECIPROCAL;
  const float v = 0.5f - jasinpi(-nY);
  rec_[R_P] = pX; rec_[R_P + 1] = pY; rec_[R_P + 2] = pZ;
  rec_[R_NG] = nX; rec_[R_NG + 1] = nY; rec_[R_NG + 2] = nZ;
  rec_[R_NS] = nX; rec_[R_NS + 1] = nY; rec_[R_NS + 2] = nZ;
  rec_[R_UV] = u; rec_[R_UV + 1] = v;
}

void Item::props_terrain(float ox, float oy, float oz, float dx, float dy, float dz, float t) {  // :2495-2539
  const float sr = 1.0f / 10.0f;
  const float pX = ox * sr + dx * t, pY = oy * sr + dy * t, pZ = oz * sr + dz * t;
  const float e = 0.02f;
  const float nX = (jsin(pX - e) * jsin(pZ)) - (jsin(pX + e) * jsin(pZ));
  const float nY = 2.0f * e;
  const float nZ = (jsin(pX) * jsin(pZ - e)) - (jsin(pX) * jsin(pZ + e));
  const float r = jrsqrt(nX * nX + nY * nY + nZ * nZ);
  rec_[R_P] = pX; rec_[R_P + 1] = pY; rec_[R_P + 2] = pZ;
  rec_[R_UV] = pX; rec_[R_UV + 1] = pZ;
  rec_[R_NG] = nX * r; rec_[R_NG + 1] = nY * r; rec_[R_NG + 2] = nZ * r;
  rec_[R_NS] = nX * r; rec_[R_NS + 1] = nY * r; rec_[R_NS + 2] = nZ * r;
}

void Item::props_triangle(float ox, float oy, float oz, float dx, float dy, float dz, float t,
                          int to) {  // :2541-2650
  const int32_t* T = s_.a->triangles + to;
  const float* A = p3() + T[0];
  const float* B = p3() + T[1];
  const float* C = p3() + T[2];
  const float* nA = v3() + T[3];
  const float* nB = v3() + T[4];
  const float* nC = v3() + T[5];
  const float* uvA = s_.a->point2fs + T[6];
  const float* uvB = s_.a->point2fs + T[7];
  const float* uvC = s_.a->point2fs + T[8];
  const float* tA = v3() + T[9];
  const float* tB = v3() + T[10];
  const float* tC = v3() + T[11];
  const float pX = ox + dx * t, pY = oy + dy * t, pZ = oz + dz * t;
  const float abX = B[0] - A[0], abY = B[1] - A[1], abZ = B[2] - A[2];
  const float acX = C[0] - A[0], acY = C[1] - A[1], acZ = C[2] - A[2];
  const float v0X = dy * acZ - dz * acY, v0Y = dz * acX - dx * acZ, v0Z = dx * acY - dy * acX;
  const float v1X = ox - A[0], v1Y = oy - A[1], v1Z = oz - A[2];
  const float v2X = v1Y * abZ - v1Z * abY, v2Y = v1Z * abX - v1X * abZ, v2Z = v1X * abY - v1Y * abX;
  const float det = abX * v0X + abY * v0Y + abZ * v0Z;
  const float dr = 1.0f / det;
  float bu = (v1X * v0X + v1Y * v0Y + v1Z * v0Z) * dr;
  float bv = (dx * v2X + dy * v2Y + dz * v2Z) * dr;
  const float bw = 1.0f - bu - bv;
  if (planted(PLANT_BARYCENTRIC)) std::swap(bu, bv);
  const float u = bw * uvA[0] + bu * uvB[0] + bv * uvC[0];
  const float v = bw * uvA[1] + bu * uvB[1] + bv * uvC[1];
  rec_[R_T] = t;
  rec_[R_P] = pX; rec_[R_P + 1] = pY; rec_[R_P + 2] = pZ;
  rec_[R_UV] = u; rec_[R_UV + 1] = v;
  float n0X = 0.0f, n0Y = 0.0f, n0Z = 0.0f, t0X = 0.0f, t0Y = 0.0f, t0Z = 0.0f;
  if (s_.p->shader_type == DF_SHADER_FLAT) {
    n0X = abY * acZ - abZ * acY;
    n0Y = abZ * acX - abX * acZ;
    n0Z = abX * acY - abY * acX;
  } else if (s_.p->shader_type == DF_SHADER_GOURAUD) {
    n0X = nA[0] * bw + nB[0] * bu + nC[0] * bv;
    n0Y = nA[1] * bw + nB[1] * bu + nC[1] * bv;
    n0Z = nA[2] * bw + nB[2] * bu + nC[2] * bv;
    t0X = tA[0] * bw + tB[0] * bu + tC[0] * bv;
    t0Y = tA[1] * bw + tB[1] * bu + tC[1] * bv;
    t0Z = tA[2] * bw + tB[2] * bu + tC[2] * bv;
  }
  const float nr = jrsqrt(n0X * n0X + n0Y * n0Y + n0Z * n0Z);
  const float tr = jrsqrt(t0X * t0X + t0Y * t0Y + t0Z * t0Z);
  rec_[R_NG] = n0X * nr; rec_[R_NG + 1] = n0Y * nr; rec_[R_NG + 2] = n0Z * nr;
  rec_[R_NS] = n0X * nr; rec_[R_NS + 1] = n0Y * nr; rec_[R_NS + 2] = n0Z * nr;
  rec_[R_TAN] = t0X * tr; rec_[R_TAN + 1] = t0Y * tr; rec_[R_TAN + 2] = t0Z * tr;
}

void Item::to_world(const float* o, const float* w) {  // :3998-4073
  // normals/tangent/ONB by the transpose of worldToObject; point by objectToWorld
  auto nrm = [&](int off) {
    const float x = rec_[off], y = rec_[off + 1], z = rec_[off + 2];
    rec_[off] = w[0] * x + w[4] * y + w[8] * z;
    rec_[off + 1] = w[1] * x + w[5] * y + w[9] * z;
    rec_[off + 2] = w[2] * x + w[6] * y + w[10] * z;
  };
  nrm(R_ONB_U);
  nrm(R_ONB_V);
  nrm(R_ONB_W);
  const float x = rec_[R_P], y = rec_[R_P + 1], z = rec_[R_P + 2];
  rec_[R_P] = o[0] * x + o[1] * y + o[2] * z + o[3];
  rec_[R_P + 1] = o[4] * x + o[5] * y + o[6] * z + o[7];
  rec_[R_P + 2] = o[8] * x + o[9] * y + o[10] * z + o[11];
  nrm(R_NG);
  nrm(R_NS);
  nrm(R_TAN);
}

void Item::normal_map_noise(int po) {  // :2658-2712
  const int so = s_.a->primitives[po + 2];
  const float amount = s_.a->surfaces[so + 4];
  const float scale = s_.a->surfaces[so + 5];
  if (s_.p->normal_mapping == 1 && amount > 0.0f && scale > 0.0f) {
    const float x0 = rec_[R_P], y0 = rec_[R_P + 1], z0 = rec_[R_P + 2];
    const float sr = 1.0f / scale;
    const float x1 = x0 * sr, y1 = y0 * sr, z1 = z0 * sr;
    const float nx = simplex_fbm_xyz(scale, 0.5f, -0.26f, 0.26f, 16, x1, y1, z1);
    const float ny = planted(PLANT_NOISE_NORMAL) ? simplex_fbm_xyz(scale, 0.5f, -0.26f, 0.26f, 16, z1, x1, y1)
                                                 : simplex_fbm_xyz(scale, 0.5f, -0.26f, 0.26f, 16, y1, z1, x1);
    const float nz = simplex_fbm_xyz(scale, 0.5f, -0.26f, 0.26f, 16, z1, x1, y1);
    const float s1X = rec_[R_NS] + nx * amount;
    const float s1Y = rec_[R_NS + 1] + ny * amount;
    const float s1Z = rec_[R_NS + 2] + nz * amount;
    const float r = jrsqrt(s1X * s1X + s1Y * s1Y + s1Z * s1Z);
    rec_[R_NS] = s1X * r;
    rec_[R_NS + 1] = s1Y * r;
    rec_[R_NS + 2] = s1Z * r;
  }
}

void Item::normal_map_image(int po) {  // :2714-2846
  const int so = s_.a->primitives[po + 2];
  const int to = jf2i(s_.a->surfaces[so + 3]);
  const int tt = jf2i(s_.a->textures[to]);
  if (!(s_.p->normal_mapping == 1 && tt == 6)) return;
  const int rgb = tex_image_simple(to);
  const float r = ((((rgb >> 16) & 0xFF) * COLOR_RECIPROCAL) - 0.5f) * 2.0f;
  const float g = ((((rgb >> 8) & 0xFF) * COLOR_RECIPROCAL) - 0.5f) * 2.0f;
  const float b = ((((rgb >> 0) & 0xFF) * COLOR_RECIPROCAL) - (planted(PLANT_NORMAL_MAP_B) ? 0.5f : 0.0f)) * 2.0f;
  const float wX = rec_[R_NS], wY = rec_[R_NS + 1], wZ = rec_[R_NS + 2];
  const int type = s_.a->primitives[po];
  float sX, sY, sZ;
  if (type == 2) {
    const float v0X = -2.0f * PI * wY, v0Y = 2.0f * PI * wX, v0Z = 0.0f;
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
    const float uX = rec_[R_TAN], uY = rec_[R_TAN + 1], uZ = rec_[R_TAN + 2];
    const float v0X = wY * uZ - wZ * uY, v0Y = wZ * uX - wX * uZ, v0Z = wX * uY - wY * uX;
    const float vr = jrsqrt(v0X * v0X + v0Y * v0Y + v0Z * v0Z);
    const float v1X = v0X * vr, v1Y = v0Y * vr, v1Z = v0Z * vr;
    sX = r * uX + g * v1X + b * wX;
    sY = r * uY + g * v1Y + b * wY;
    sZ = r * uZ + g * v1Z + b * wZ;
  } else {
    const float aX = jabs(wX), aY = jabs(wY), aZ = jabs(wZ);
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
  rec_[R_NS] = sX * sr;
  rec_[R_NS + 1] = sY * sr;
  rec_[R_NS + 2] = sZ * sr;
}

// ------------------------------------------------------------ textures

static inline int pack_rgb(float r, float g, float b) {
  const int ri = jf2i(r), gi = jf2i(g), bi = jf2i(b);
  return ((ri & 0xFF) << 16) | ((gi & 0xFF) << 8) | (bi & 0xFF);
}

int Item::tex_color(int off) {  // :1596-1618
  const int type = jf2i(s_.a->textures[off]);
  switch (type) {
    case 1: return tex_blend(off);
    case 2: return tex_bullseye(off);
    case 3: return tex_checker(off);
    case 4: return jf2i(s_.a->textures[off + 2]);
    case 5: return tex_fbm(off);
    case 6: return tex_image(off);
    case 7: return tex_surface_normal();
    case 8: return tex_uv();
    default: return 0;
  }
}

int Item::tex_color2(int off) {  // :1620-1640 (no nested blend)
  const int type = jf2i(s_.a->textures[off]);
  switch (type) {
    case 2: return tex_bullseye(off);
    case 3: return tex_checker(off);
    case 4: return jf2i(s_.a->textures[off + 2]);
    case 5: return tex_fbm(off);
    case 6: return tex_image(off);
    case 7: return tex_surface_normal();
    case 8: return tex_uv();
    default: return 0;
  }
}

int Item::tex_blend(int off) {  // :1642-1666
  const float* T = s_.a->textures;
  const int ao = jf2i(T[off + 2]);
  const int bo = jf2i(T[off + 3]);
  const float f = T[off + 4];
  const int a = tex_color2(ao);
  const int b = tex_color2(bo);
  const float aR = ((a >> 16) & 0xFF) * COLOR_RECIPROCAL, aG = ((a >> 8) & 0xFF) * COLOR_RECIPROCAL,
              aB = (a & 0xFF) * COLOR_RECIPROCAL;
  const float bR = ((b >> 16) & 0xFF) * COLOR_RECIPROCAL, bG = ((b >> 8) & 0xFF) * COLOR_RECIPROCAL,
              bB = (b & 0xFF) * COLOR_RECIPROCAL;
  return pack_rgb(saturate((1.0f - f) * aR + f * bR, 0.0f, 1.0f) * 255.0f + 0.5f,
                  saturate((1.0f - f) * aG + f * bG, 0.0f, 1.0f) * 255.0f + 0.5f,
                  saturate((1.0f - f) * aB + f * bB, 0.0f, 1.0f) * 255.0f + 0.5f);
}

int Item::tex_bullseye(int off) {  // :1668-1712
  const float* T = s_.a->textures;
  const float x = rec_[R_P], y = rec_[R_P + 1], z = rec_[R_P + 2];
  const int a = jf2i(T[off + 2]);
  const int b = jf2i(T[off + 3]);
  const float aR = ((a >> 16) & 0xFF) * COLOR_RECIPROCAL, aG = ((a >> 8) & 0xFF) * COLOR_RECIPROCAL,
              aB = (a & 0xFF) * COLOR_RECIPROCAL;
  const float bR = ((b >> 16) & 0xFF) * COLOR_RECIPROCAL, bG = ((b >> 8) & 0xFF) * COLOR_RECIPROCAL,
              bB = (b & 0xFF) * COLOR_RECIPROCAL;
  const float length = jsqrt(x * x + y * y + z * z);
  const float value = remainder(length * (planted(PLANT_BULLSEYE) ? 0.5f : 0.25f), 1.0f);
  float cR = value > 0.5f ? aR : bR;
  float cG = value > 0.5f ? aG : bG;
  float cB = value > 0.5f ? aB : bB;
  const float m = value > 0.5f ? 0.8f : 1.2f;
  if (aR == bR && aG == bG && aB == bB) {
    cR *= m;
    cG *= m;
    cB *= m;
  }
  return pack_rgb(saturate(cR, 0.0f, 1.0f) * 255.0f + 0.5f, saturate(cG, 0.0f, 1.0f) * 255.0f + 0.5f,
                  saturate(cB, 0.0f, 1.0f) * 255.0f + 0.5f);
}

// The tile of the checkerboard texture at `off` under the hit record's uv
// (:1741-1747): 0 where the texture takes its first colour, 1 the second.
int Item::checker_tile(int off) {
  const float* T = s_.a->textures;
  const float u = rec_[R_UV], v = rec_[R_UV + 1];
  const float sU = T[off + 6], sV = T[off + 7];
  const float ca = T[off + 4], sa = T[off + 5];
  const float tu = modulo((u * ca - v * sa) * sU);
  const float tv = modulo((v * ca + u * sa) * sV);
  return (tu > 0.5f) != (tv > 0.5f) ? 0 : 1;
}

int Item::tex_checker(int off) {  // :1714-1767
  const float* T = s_.a->textures;
  const int c0 = jf2i(T[off + 2]);
  const int c1 = jf2i(T[off + 3]);
  const float c0R = ((c0 >> 16) & 0xFF) * COLOR_RECIPROCAL, c0G = ((c0 >> 8) & 0xFF) * COLOR_RECIPROCAL,
              c0B = (c0 & 0xFF) * COLOR_RECIPROCAL;
  const float c1R = ((c1 >> 16) & 0xFF) * COLOR_RECIPROCAL, c1G = ((c1 >> 8) & 0xFF) * COLOR_RECIPROCAL,
              c1B = (c1 & 0xFF) * COLOR_RECIPROCAL;
  const bool dark = checker_tile(off) == 0;
  const float m = dark ? 0.8f : 1.2f;
  float cR = dark ? c0R : c1R;
  float cG = dark ? c0G : c1G;
  float cB = dark ? c0B : c1B;
  if (c0R == c1R && c0G == c1G && c0B == c1B) {
    cR *= m;
    cG *= m;
    cB *= m;
  }
  return pack_rgb(saturate(cR, 0.0f, 1.0f) * 255.0f + 0.5f, saturate(cG, 0.0f, 1.0f) * 255.0f + 0.5f,
                  saturate(cB, 0.0f, 1.0f) * 255.0f + 0.5f);
}

int Item::tex_fbm(int off) {  // :1776-1815
  const float* T = s_.a->textures;
  const float x = rec_[R_P], y = rec_[R_P + 1], z = rec_[R_P + 2];
  const int ad = jf2i(T[off + 2]);
  const int mu = jf2i(T[off + 3]);
  const float aR = ((ad >> 16) & 0xFF) * COLOR_RECIPROCAL, aG = ((ad >> 8) & 0xFF) * COLOR_RECIPROCAL,
              aB = (ad & 0xFF) * COLOR_RECIPROCAL;
  const float mR = ((mu >> 16) & 0xFF) * COLOR_RECIPROCAL, mG = ((mu >> 8) & 0xFF) * COLOR_RECIPROCAL,
              mB = (mu & 0xFF) * COLOR_RECIPROCAL;
  const float freq = T[off + 4], gain = T[off + 5];
  const int oct = jf2i(T[off + 6]);
  const float n = planted(PLANT_FBM_TEXTURE) ? simplex_fbm_xyz(freq, gain, -1.0f, 1.0f, oct, x, y, z)
                                             : simplex_fbm_xyz(freq, gain, 0.0f, 1.0f, oct, x, y, z);
  return pack_rgb(saturate(n * mR + aR, 0.0f, 1.0f) * 255.0f + 0.5f, saturate(n * mG + aG, 0.0f, 1.0f) * 255.0f + 0.5f,
                  saturate(n * mB + aB, 0.0f, 1.0f) * 255.0f + 0.5f);
}

int Item::tex_image(int off) {  // :1817-1892
  const float* T = s_.a->textures;
  const float u = rec_[R_UV], v = rec_[R_UV + 1];
  const float width = T[off + 4], height = T[off + 5];
  const float sU = T[off + 6], sV = T[off + 7];
  const float ca = T[off + 2], sa = T[off + 3];
  const float u0 = (u * ca - v * sa);
  const float v0 = (v * ca + u * sa);
  const float u1 = remainder(u0 * sU * width, width);
  const float v1 = remainder(v0 * sV * height, height);
  const float u2 = u1 >= 0.0f ? u1 : width - jabs(u1);
  const float v2 = v1 >= 0.0f ? v1 : height - jabs(v1);
  const int x = jf2i(u2), y = jf2i(v2);
  const int w = jf2i(width);
  const int res = jf2i(width * height);
  const int i00 = y * w + x, i01 = y * w + x + 1, i10 = (y + 1) * w + x, i11 = (y + 1) * w + x + 1;
  auto fetch = [&](int i) { return i >= 0 && i < res ? jf2i(T[off + 8 + i]) : 0; };
  const int c00 = fetch(i00), c01 = fetch(i01), c10 = fetch(i10), c11 = fetch(i11);
  const float fx = u2 - static_cast<float>(x);
  const float fy = v2 - static_cast<float>(y);
  auto ch = [](int c, int sh) { return ((c >> sh) & 0xFF) * COLOR_RECIPROCAL; };
  const float r = blerp(ch(c00, 16), ch(c01, 16), ch(c10, 16), ch(c11, 16), fx, fy);
  const float g = blerp(ch(c00, 8), ch(c01, 8), ch(c10, 8), ch(c11, 8), fx, fy);
  const float b = blerp(ch(c00, 0), ch(c01, 0), ch(c10, 0), ch(c11, 0), fx, fy);
  return pack_rgb(saturate(r, 0.0f, 1.0f) * 255.0f + 0.5f, saturate(g, 0.0f, 1.0f) * 255.0f + 0.5f,
                  saturate(b, 0.0f, 1.0f) * 255.0f + 0.5f);
}

int Item::tex_image_simple(int off) {  // :1894-1929
  const float* T = s_.a->textures;
  const float u = rec_[R_UV], v = rec_[R_UV + 1];
  const float width = T[off + 4], height = T[off + 5];
  const float sU = T[off + 6], sV = T[off + 7];
  const float ca = T[off + 2], sa = T[off + 3];
  const float u0 = (u * ca - v * sa);
  const float v0 = (v * ca + u * sa);
  const float u1 = remainder(u0 * sU * width, width);
  const float v1 = remainder(v0 * sV * height, height);
  const float u2 = u1 >= 0.0f ? u1 : width - jabs(u1);
  const float v2 = v1 >= 0.0f ? v1 : height - jabs(v1);
  const int x = jf2i(u2), y = jf2i(v2);
  const int w = jf2i(width);
  const int res = jf2i(width * height);
  const int i = y * w + x;
  return i >= 0 && i < res ? jf2i(T[off + 8 + i]) : 0;
}

int Item::tex_surface_normal() {  // :1931-1945
  return pack_rgb(saturate((rec_[R_NS] + 1.0f) * 0.5f, 0.0f, 1.0f) * 255.0f + 0.5f,
                  saturate((rec_[R_NS + 1] + 1.0f) * 0.5f, 0.0f, 1.0f) * 255.0f + 0.5f,
                  saturate((rec_[R_NS + 2] + 1.0f) * 0.5f, 0.0f, 1.0f) * 255.0f + 0.5f);
}

int Item::tex_uv() {  // :1947-1960
  return pack_rgb(saturate(rec_[R_UV], 0.0f, 1.0f) * 255.0f + 0.5f, saturate(rec_[R_UV + 1], 0.0f, 1.0f) * 255.0f + 0.5f,
                  0.0f);
}

// ------------------------------------------------------------ lighting

int Item::phong_shader0(bool check, float pX, float pY, float pZ, float nX, float nY, float nZ, float vX, float vY,
                        float vZ, float aR, float aG, float aB, float kaR, float kaG, float kaB, float kdR, float kdG,
                        float kdB, float ksR, float ksG, float ksB, float ns) {  // :1962-2039
  float r = 0.0f, g = 0.0f, b = 0.0f;
  const float ia = 1.0f;
  r += aR * kaR * ia;
  g += aG * kaG * ia;
  b += aB * kaB * ia;
  const float lX = s_.sky->sun_origin[0] - pX, lY = s_.sky->sun_origin[1] - pY, lZ = s_.sky->sun_origin[2] - pZ;
  const float l2 = lX * lX + lY * lY + lZ * lZ;
  const float l = jsqrt(l2);
  const float lr = 1.0f / l;
  const float lnX = lX * lr, lnY = lY * lr, lnZ = lZ * lr;
  if (check) do_intersect_primitives(pX, pY, pZ, lnX, lnY, lnZ, false);
  const float t = check ? rec_[R_T] : INFINITY_;
  if (t >= INFINITY_ || t > l2) {
    const float ldn = lnX * nX + lnY * nY + lnZ * nZ;
    const float id = 1.0f;
    r += aR * kdR * jmaxf(ldn, 0.0f) * id;
    g += aG * kdG * jmaxf(ldn, 0.0f) * id;
    b += aB * kdB * jmaxf(ldn, 0.0f) * id;
    if (ldn > 0.0f) {
      const float rX = 2.0f * ldn * nX - lnX, rY = 2.0f * ldn * nY - lnY, rZ = 2.0f * ldn * nZ - lnZ;
      const float rr = jrsqrt(rX * rX + rY * rY + rZ * rZ);
      const float rdv = rX * rr * vX + rY * rr * vY + rZ * rr * vZ;
      const float sh = jpow(jmaxf(rdv, 0.0f), planted(PLANT_PHONG_SHADER) ? ns * 0.1f : ns);
      const float is = 1.0f;
      r += ksR * sh * is;
      g += ksG * sh * is;
      b += ksB * sh * is;
    }
  }
  r = saturate(r, 0.0f, 1.0f);
  g = saturate(g, 0.0f, 1.0f);
  b = saturate(b, 0.0f, 1.0f);
  return ((jf2i(r * 255.0f + 0.5f) & 0xFF) << 16) | ((jf2i(g * 255.0f + 0.5f) & 0xFF) << 8) |
         (jf2i(b * 255.0f + 0.5f) & 0xFF);
}

void Item::do_sky(bool active, float dx, float dy, float dz) {  // :2127-2236
  const df_sky* S = s_.sky;
  float d0X = dx * S->onb_u[0] + dy * S->onb_u[1] + dz * S->onb_u[2];
  float d0Y = dx * S->onb_v[0] + dy * S->onb_v[1] + dz * S->onb_v[2];
  float d0Z = dx * S->onb_w[0] + dy * S->onb_w[1] + dz * S->onb_w[2];
  float r = 0.01f, g = 0.01f, b = 0.01f;
  if (d0Z >= 0.0f && active) {
    cnt_.sky_evals++;
    d0Z = jmaxf(d0Z, 0.001f);
    const float lr = jrsqrt(d0X * d0X + d0Y * d0Y + d0Z * d0Z);
    const float d1X = d0X * lr, d1Y = d0Y * lr, d1Z = d0Z * lr;
    const float dp = d1X * S->sun_direction[0] + d1Y * S->sun_direction[1] + d1Z * S->sun_direction[2];
    const double theta0 = S->theta;
    const double theta1 = std::acos(jmaxd(jmind(static_cast<double>(d1Z), 1.0), -1.0));
    const double cosTheta0 = std::cos(theta0);
    const double cosTheta1 = std::cos(theta1);
    const double cosTheta1R = 1.0 / (cosTheta1 + 0.01);
    const double gamma = std::acos(jmaxd(jmind(static_cast<double>(dp), 1.0), -1.0));
    const double cosGamma = std::cos(gamma);
    const double* L = S->perez_relative_luminance;
    const double* X = S->perez_x;
    const double* Y = S->perez_y;
    const double L3 = planted(PLANT_PEREZ_SWAP) ? L[4] : L[3], L4 = planted(PLANT_PEREZ_SWAP) ? L[3] : L[4];
    const double lden = ((1.0 + L[0] * std::exp(L[1])) * (1.0 + L[2] * std::exp(L3 * theta0) + L4 * cosTheta0 * cosTheta0));
    const double lnum = ((1.0 + L[0] * std::exp(L[1] * cosTheta1R)) * (1.0 + L[2] * std::exp(L3 * gamma) + L4 * cosGamma * cosGamma));
    const double thetaD = planted(PLANT_PEREZ_THETA0) ? theta0 : theta1;  // the kernel's quirk: theta1
    const double lum = S->zenith_relative_luminance * lnum / lden * 1.0e-4;
    const double xden = ((1.0 + X[0] * std::exp(X[1])) * (1.0 + X[2] * std::exp(X[3] * thetaD) + X[4] * cosTheta0 * cosTheta0));
    const double xnum = ((1.0 + X[0] * std::exp(X[1] * cosTheta1R)) * (1.0 + X[2] * std::exp(X[3] * gamma) + X[4] * cosGamma * cosGamma));
    const double x = S->zenith_x * xnum / xden;
    const double yden = ((1.0 + Y[0] * std::exp(Y[1])) * (1.0 + Y[2] * std::exp(Y[3] * thetaD) + Y[4] * cosTheta0 * cosTheta0));
    const double ynum = ((1.0 + Y[0] * std::exp(Y[1] * cosTheta1R)) * (1.0 + Y[2] * std::exp(Y[3] * gamma) + Y[4] * cosGamma * cosGamma));
    const double y = S->zenith_y * ynum / yden;
    daylight_xyy_to_srgb(static_cast<float>(x), static_cast<float>(y), lum, &r, &g, &b);
  }
  tmp_[0] = r;
  tmp_[1] = g;
  tmp_[2] = b;
}

// The direction towards the sun that doCalculateColorForSun samples from the point p with the random numbers
// (u, v) (GPURendererKernel.java:2252-2320): uniform over the cone about the sun direction whose half-angle is
// that of a disk of radius 500 seen from p at the sun's origin.
void Item::sun_direction(float pX, float pY, float pZ, float u, float v, float d[3]) const {
  const df_sky* S = s_.sky;
  const float sdX = S->sun_direction_world[0], sdY = S->sun_direction_world[1], sdZ = S->sun_direction_world[2];
  const float sunR = 500.0f;
  const float sunR2 = sunR * sunR;
  const float toX = S->sun_origin[0] - pX, toY = S->sun_origin[1] - pY, toZ = S->sun_origin[2] - pZ;
  const float to2 = toX * toX + toY * toY + toZ * toZ;
  const float sinMax2 = planted(PLANT_SUN_CONE) ? sunR / to2 : sunR2 / to2;
  const float cosMax = jsqrt(jmaxf(0.0f, 1.0f - sinMax2));
  const float cosT = u * (cosMax - 1.0f) + 1.0f;
  const float sinT = jsqrt(jmaxf(0.0f, 1.0f - cosT * cosT));
  const float phi = PI_MULTIPLIED_BY_TWO * v;
  const float x = jcos(phi) * sinT, y = jsin(phi) * sinT, z = cosT;
  const float wX = sdX, wY = sdY, wZ = sdZ;
  const float aX = jabs(wX), aY = jabs(wY), aZ = jabs(wZ);
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
}

void Item::do_sun(bool active, float pX, float pY, float pZ, float nX, float nY, float nZ, float aR, float aG,
                  float aB) {  // :2238-2347
  float r = 0.0f, g = 0.0f, b = 0.0f;
  sun_lit_ = false;
  if (active) {
    const df_sky* S = s_.sky;
    const float sdX = S->sun_direction_world[0], sdY = S->sun_direction_world[1], sdZ = S->sun_direction_world[2];
    const float u = next_float();
    const float v = next_float();
    sun_u_ = u;
    sun_v_ = v;
    float d3[3];
    sun_direction(pX, pY, pZ, u, v, d3);
    const float dX = d3[0], dY = d3[1], dZ = d3[2];
    const float dp0 = dX * sdX + dY * sdY + dZ * sdZ;
    const float dp1 = dX * nX + dY * nY + dZ * nZ;
    if (dp0 > 0.0f && dp1 > 0.0f) {
      const float t = do_intersect_primitives(pX, pY, pZ, dX, dY, dZ, true);
      if (t >= INFINITY_ - 0.0001f) {
        sun_lit_ = true;
        sun_dp1_ = dp1;
        r = aR * 1.0f * dp1 * PI_RECIPROCAL;
        g = aG * 1.0f * dp1 * PI_RECIPROCAL;
        b = aB * 1.0f * dp1 * PI_RECIPROCAL;
      }
    }
  }
  tmp_[0] = r;
  tmp_[1] = g;
  tmp_[2] = b;
}

// ------------------------------------------------------------ renderers

void Item::render_ao() {  // :3029-3149
  float ox = ray_[0], oy = ray_[1], oz = ray_[2], dx = ray_[3], dy = ray_[4], dz = ray_[5];
  float pr = 0.0f, pg = 0.0f, pb = 0.0f;
  do_intersect_primitives(ox, oy, oz, dx, dy, dz, false);
  const float dist = rec_[R_T];
  const int po = jf2i(rec_[R_PRIM]);
  f_.primitive_offsets[gid_] = po;
  if (dist != INFINITY_ && po != -1) {
    const float pX = rec_[R_P], pY = rec_[R_P + 1], pZ = rec_[R_P + 2];
    const float sX = rec_[R_NS], sY = rec_[R_NS + 1], sZ = rec_[R_NS + 2];
    const float dp = sX * dx + sY * dy + sZ * dz;
    const bool ok = dp < 0.0f;
    const float w0X = ok ? sX : -sX, w0Y = ok ? sY : -sY, w0Z = ok ? sZ : -sZ;
    const float wr = jrsqrt(w0X * w0X + w0Y * w0Y + w0Z * w0Z);
    const float w1X = w0X * wr, w1Y = w0Y * wr, w1Z = w0Z * wr;
    const bool isY = jabs(w1X) > 0.1f;
    const float u0X = isY ? 0.0f : 1.0f;
    const float u0Y = isY ? 1.0f : 0.0f;
    const float u1X = u0Y * w1Z;
    const float u1Y = -(u0X * w1Z);
    const float u1Z = u0X * w1Y - u0Y * w1X;
    const float ur = jrsqrt(u1X * u1X + u1Y * u1Y + u1Z * u1Z);
    const float u2X = u1X * ur, u2Y = u1Y * ur, u2Z = u1Z * ur;
    const float v0X = w1Y * u2Z - w1Z * u2Y, v0Y = w1Z * u2X - w1X * u2Z, v0Z = w1X * u2Y - w1Y * u2X;
    const float xi = next_float();
    const float xj = next_float();
    const float phi = PI_MULTIPLIED_BY_TWO * xi;
    const float cp = jcos(phi), sp = jsin(phi);
    const float ct = planted(PLANT_AO_UNIFORM) ? 1.0f - xj : jsqrt(1.0f - xj);
    const float st = planted(PLANT_AO_UNIFORM) ? jsqrt(1.0f - ct * ct) : jsqrt(xj);
    const float d0X = cp * st, d0Y = sp * st, d0Z = ct;
    const float d1X = d0X * u2X + d0Y * v0X + d0Z * w1X;
    const float d1Y = d0X * u2Y + d0Y * v0Y + d0Z * w1Y;
    const float d1Z = d0X * u2Z + d0Y * v0Z + d0Z * w1Z;
    const float dr = jrsqrt(d1X * d1X + d1Y * d1Y + d1Z * d1Z);
    ox = pX + w1X * 0.01f;
    oy = pY + w1Y * 0.01f;
    oz = pZ + w1Z * 0.01f;
    dx = d1X * dr;
    dy = d1Y * dr;
    dz = d1Z * dr;
    const float t = do_intersect_primitives(ox, oy, oz, dx, dy, dz, false);
    const bool hit = t < s_.p->ao_maximum_distance;
    const float r = hit ? 1.0f : 0.0f, g = hit ? 1.0f : 0.0f, b = hit ? 1.0f : 0.0f;
    pr += (1.0f - r) * 1.0f + r * 0.0f;
    pg += (1.0f - g) * 1.0f + g * 0.0f;
    pb += (1.0f - b) * 1.0f + b * 0.0f;
  }
  film_add(pr, pg, pb);
}

void Item::render_pt(bool sky_active, bool sun_active) {  // :3151-3705
  const int depth_max = s_.p->pt_depth_maximum;
  const int depth_rr = s_.p->pt_depth_russian_roulette;
  int depth = 0;
  float ox = ray_[0], oy = ray_[1], oz = ray_[2], dx = ray_[3], dy = ray_[4], dz = ray_[5];
  float pr = 0.0f, pg = 0.0f, pb = 0.0f;
  float mr = 1.0f, mg = 1.0f, mb = 1.0f;
  const df_compiled_scene* A = s_.a;
  if (tan_) tan_->begin();
  do {
    do_intersect_primitives(ox, oy, oz, dx, dy, dz, false);
    const float dist = rec_[R_T];
    const int po = jf2i(rec_[R_PRIM]);
    if (depth == 0) f_.primitive_offsets[gid_] = po;
    if (dist == INFINITY_ || po == -1) {
      do_sky(sky_active, dx, dy, dz);
      if (tan_) tan_sky(sky_active, dx, dy, dz);
      pr += mr * tmp_[0] * PI_RECIPROCAL;
      pg += mg * tmp_[1] * PI_RECIPROCAL;
      pb += mb * tmp_[2] * PI_RECIPROCAL;
      film_add(pr, pg, pb);
      return;
    }
    const int so = A->primitives[po + 2];
    const float pX = rec_[R_P], pY = rec_[R_P + 1], pZ = rec_[R_P + 2];
    const float sX = rec_[R_NS], sY = rec_[R_NS + 1], sZ = rec_[R_NS + 2];
    float w3[3], dp2;
    orient(sX, sY, sZ, dx, dy, dz, w3, &dp2);
    const float wX = w3[0], wY = w3[1], wZ = w3[2];
    const int ta = jf2i(A->surfaces[so + 1]);
    const int te = jf2i(A->surfaces[so + 2]);
    const int alb = tex_color(ta);
    float aR = ((alb >> 16) & 0xFF) * COLOR_RECIPROCAL;
    float aG = ((alb >> 8) & 0xFF) * COLOR_RECIPROCAL;
    float aB = (alb & 0xFF) * COLOR_RECIPROCAL;
    const int emi = tex_color(te);
    float eR = ((emi >> 16) & 0xFF) * COLOR_RECIPROCAL;
    float eG = ((emi >> 8) & 0xFF) * COLOR_RECIPROCAL;
    float eB = (emi & 0xFF) * COLOR_RECIPROCAL;
    int aslot = -1, eslot = -1;
    if (s_.diff && so / 6 < s_.diff->n_surfaces) {  // differentiable mode: parameters replace the textures
      aslot = s_.diff->surface_albedo[so / 6];
      eslot = s_.diff->surface_emission[so / 6];
      if (aslot >= 0) {
        aR = s_.diff->albedo[aslot * 3 + 0];
        aG = s_.diff->albedo[aslot * 3 + 1];
        aB = s_.diff->albedo[aslot * 3 + 2];
      } else if (s_.diff->surface_checker && s_.diff->surface_checker[so / 6] >= 0) {
        const int k = 2 * s_.diff->surface_checker[so / 6] + (checker_tile(ta) ^ (planted(PLANT_CHECKER_TILE) ? 1 : 0));
        aslot = s_.diff->n_albedo + k;
        aR = s_.diff->checker[k * 3 + 0];
        aG = s_.diff->checker[k * 3 + 1];
        aB = s_.diff->checker[k * 3 + 2];
      }
      if (eslot >= 0) eR = eG = eB = s_.diff->emission[eslot];
    }
    const int mat = jf2i(A->surfaces[so]);
    if (mat == 1 || mat == 3 || mat == 4) {
      do_sun(sun_active, pX, pY, pZ, wX, wY, wZ, aR, aG, aB);
    } else {
      tmp_[0] = tmp_[1] = tmp_[2] = 0.0f;
    }
    if (tan_) {
      const float a3[3] = {aR, aG, aB}, e3[3] = {eR, eG, eB};
      tan_vertex(aslot, eslot, a3, e3, mat == 1 || mat == 3 || mat == 4, pX, pY, pZ, wX, wY, wZ);
    }
    pr += mr * (eR + tmp_[0]);
    pg += mg * (eG + tmp_[1]);
    pb += mb * (eB + tmp_[2]);
    depth++;
    if (depth >= depth_rr) {
      const float pdf = jmaxf(aR, jmaxf(aG, aB));
      const float rnd = next_float();
      if (rnd >= pdf) {
        const bool hit = do_intersect_primitives(ox, oy, oz, dx, dy, dz, true) < INFINITY_;
        if (!hit) {
          do_sky(sky_active, dx, dy, dz);
          pr += mr * tmp_[0] * PI_RECIPROCAL;
          pg += mg * tmp_[1] * PI_RECIPROCAL;
          pb += mb * tmp_[2] * PI_RECIPROCAL;
        }
        film_add(pr, pg, pb);
        return;
      }
      const float pdfr = 1.0f / pdf;
      aR *= pdfr;
      aG *= pdfr;
      aB *= pdfr;
    }
    scatter(mat, aslot, pX, pY, pZ, sX, sY, sZ, wX, wY, wZ, dp2, aR, aG, aB, ox, oy, oz, dx, dy, dz, mr, mg, mb);
  } while (depth < depth_max);
  const bool hit = do_intersect_primitives(ox, oy, oz, dx, dy, dz, true) < INFINITY_;
  if (!hit) {
    do_sky(sky_active, dx, dy, dz);
    if (tan_) tan_sky(sky_active, dx, dy, dz);
    pr += mr * tmp_[0] * PI_RECIPROCAL;
    pg += mg * tmp_[1] * PI_RECIPROCAL;
    pb += mb * tmp_[2] * PI_RECIPROCAL;
  }
  film_add(pr, pg, pb);
}

// The BSDF sample of material `mat` at a surface vertex (GPURendererKernel.java:3323-3686): replaces the ray
// (origin o, direction d, on entry the incoming direction) with the next one and multiplies the throughput m,
// given the hit point p, the shading normal s, the oriented normal w and dp2 = 2 s.d (orient()), and the
// albedo a (divided by the Russian-roulette probability when that applies).  `aslot` is the albedo's
// parameter group for the tangent (-1: none).  Draws 0 (Reflection, total internal reflection), 1 (Glass,
// ClearCoat's reflection) or 2 (Lambertian, Phong) further random numbers, 3 for ClearCoat's diffuse lobe.
void Item::scatter(int mat, int aslot, float pX, float pY, float pZ, float sX, float sY, float sZ, float wX,
                   float wY, float wZ, float dp2, float aR, float aG, float aB, float& ox, float& oy, float& oz,
                   float& dx, float& dy, float& dz, float& mr, float& mg, float& mb) {
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
      const float prob = planted(PLANT_FRESNEL_PROB) ? refl : 0.25f + 0.5f * refl;
      const float rp = refl / prob;
      const float tp = trans / (1.0f - prob);
      const bool isR = next_float() < prob;
      const float m = isR ? rp : tp;
      mr *= m;
      mg *= m;
      mb *= m;
      if (tan_) tan_->scale(m);
      if (isR) {
        const float off = planted(PLANT_CLEARCOAT_OFFSET) ? 0.01f : 0.02f;
        ox = pX + wX * off;
        oy = pY + wY * off;
        oz = pZ + wZ * off;
        dx = rX;
        dy = rY;
        dz = rZ;
      } else {
        const float u = next_float();
        const float v = next_float();
        const float phi = PI_MULTIPLIED_BY_TWO * u;
        const float ct = jsqrt(v), st = jsqrt(1.0f - v);
        const float x = jcos(phi) * st, y = jsin(phi) * st, z = ct;
        const bool isY = jabs(wX) > 0.1f;
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
        ox = pX + wX * 0.01f;
        oy = pY + wY * 0.01f;
        oz = pZ + wZ * 0.01f;
        dx = lX * lr;
        dy = lY * lr;
        dz = lZ * lr;
        mr *= aR;
        mg *= aG;
        mb *= aB;
        if (tan_) {
          const float a3[3] = {aR, aG, aB};
          tan_->albedo(aslot, a3);
        }
      }
    }
  } else if (mat == 3) {  // Lambertian :3465-3515
    const float u = next_float();
    const float v = next_float();
    const float phi = PI_MULTIPLIED_BY_TWO * u;
    const float ct = planted(PLANT_LAMBERT_UNIFORM) ? v : jsqrt(v);
    const float st = planted(PLANT_LAMBERT_UNIFORM) ? jsqrt(1.0f - v * v) : jsqrt(1.0f - v);
    const float x = jcos(phi) * st, y = jsin(phi) * st, z = ct;
    const bool isY = jabs(wX) > 0.1f;
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
    ox = pX + wX * 0.01f;
    oy = pY + wY * 0.01f;
    oz = pZ + wZ * 0.01f;
    dx = lX * lr;
    dy = lY * lr;
    dz = lZ * lr;
    mr *= aR;
    mg *= aG;
    mb *= aB;
    if (tan_) {
      const float a3[3] = {aR, aG, aB};
      tan_->albedo(aslot, a3);
    }
  } else if (mat == 4) {  // Phong :3516-3576
    const float exponent = planted(PLANT_PHONG_EXPONENT) ? 20.0f : 50.0f;
    const float u = next_float();
    const float v = next_float();
    const float phi = PI_MULTIPLIED_BY_TWO * u;
    const float ct = jpow(1.0f - v, 1.0f / (exponent + 1.0f));
    const float st = jsqrt(jmaxf(0.0f, 1.0f - ct * ct));
    const float x = jcos(phi) * st, y = jsin(phi) * st, z = ct;
    const float rwX = dx - sX * dp2, rwY = dy - sY * dp2, rwZ = dz - sZ * dp2;
    const float rwr = jrsqrt(rwX * rwX + rwY * rwY + rwZ * rwZ);
    const float rnX = rwX * rwr, rnY = rwY * rwr, rnZ = rwZ * rwr;
    const bool isY = jabs(rnX) > 0.1f;
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
    mg *= aG;
    mb *= aB;
    if (tan_) {
      const float a3[3] = {aR, aG, aB};
      tan_->albedo(aslot, a3);
    }
  } else if (mat == 2) {  // Glass :3577-3670
    const float a = 1.5f - 1.0f, b = 1.5f + 1.0f, c = 1.0f / 1.5f, d = 1.5f / 1.0f;
    const float e = (a * a) / (b * b);
    const bool in = sX * wX + sY * wY + sZ * wZ > 0.0f;
    const float nnt = planted(PLANT_SNELL_RATIO) ? (in ? d : c) : in ? c : d;
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
      const float st = jsqrt(tir);
      const float sc = in ? dw * nnt + st : -(dw * nnt + st);
      const float t0X = dx * nnt - sX * sc, t0Y = dy * nnt - sY * sc, t0Z = dz * nnt - sZ * sc;
      const float tr = jrsqrt(t0X * t0X + t0Y * t0Y + t0Z * t0Z);
      const float tX = t0X * tr, tY = t0Y * tr, tZ = t0Z * tr;
      const float a1 = (in ? -dw : tX * sX + tY * sY + tZ * sZ);
      const float a2 = 1.0f - a1;
      const float refl = e + (1.0f - e) * a2 * a2 * a2 * a2 * a2;
      const float trans = 1.0f - refl;
      const float prob = planted(PLANT_FRESNEL_PROB) ? refl : 0.25f + 0.5f * refl;
      const float rp = refl / prob;
      const float tp = trans / (1.0f - prob);
      const bool isR = next_float() < prob;
      const float m = isR ? rp : tp;
      mr *= m;
      mg *= m;
      mb *= m;
      if (tan_) tan_->scale(m);
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
    mg *= aG;
    mb *= aB;
    if (tan_) {
      const float a3[3] = {aR, aG, aB};
      tan_->albedo(aslot, a3);
    }
  }
}

// The sky term m * sky(d) / pi of a path that left the scene: through the
// albedos in m, and through turbidity, sun elevation and sun azimuth in the
// sky radiance.
void Item::tan_sky(bool active, float dx, float dy, float dz) {
  PathTangent& T = *tan_;
  const double g[3] = {static_cast<double>(tmp_[0]) * PI_RECIPROCAL, static_cast<double>(tmp_[1]) * PI_RECIPROCAL,
                       static_cast<double>(tmp_[2]) * PI_RECIPROCAL};
  T.through_throughput(g);
  D3 rgb[3];
  if (!active || !dual_sky_rgb(*T.sky, dx, dy, dz, rgb)) return;
  for (int c = 0; c < 3; ++c) {
    T.dpix[T.i_turbidity * 3 + c] += T.m[c] * rgb[c].t * PI_RECIPROCAL;
    T.dpix[T.i_elevation * 3 + c] += T.m[c] * rgb[c].e * PI_RECIPROCAL;
    T.dpix[T.i_azimuth * 3 + c] += T.m[c] * rgb[c].a * PI_RECIPROCAL;
  }
}

// The terms m * (e + sun) of a surface vertex: through the albedos in m, the
// light's own emission parameter, and the sun term a * dp1 / pi through the
// vertex's albedo and, at the sample's random numbers, the sun's elevation
// and azimuth.
void Item::tan_vertex(int aslot, int eslot, const float a[3], const float e[3], bool sun_term, float pX, float pY,
                      float pZ, float wX, float wY, float wZ) {
  PathTangent& T = *tan_;
  const double g[3] = {static_cast<double>(e[0]) + tmp_[0], static_cast<double>(e[1]) + tmp_[1],
                       static_cast<double>(e[2]) + tmp_[2]};
  T.through_throughput(g);
  if (eslot >= 0)
    for (int c = 0; c < 3; ++c) T.dpix[(T.i_emission + eslot) * 3 + c] += T.m[c];
  if (!sun_term || !sun_lit_) return;
  if (aslot >= 0)
    for (int c = 0; c < 3; ++c) T.dpix[(T.row(aslot) + c) * 3 + c] += T.m[c] * sun_dp1_ * PI_RECIPROCAL;
  const D3 dp1 = dual_sun_cosine(*T.sky, pX, pY, pZ, wX, wY, wZ, sun_u_, sun_v_);
  for (int c = 0; c < 3; ++c) {
    T.dpix[T.i_elevation * 3 + c] += T.m[c] * a[c] * dp1.e * PI_RECIPROCAL;
    T.dpix[T.i_azimuth * 3 + c] += T.m[c] * a[c] * dp1.a * PI_RECIPROCAL;
  }
}

void Item::render_rc(bool sky_active) {  // :3707-3800
  const float ox = ray_[0], oy = ray_[1], oz = ray_[2], dx = ray_[3], dy = ray_[4], dz = ray_[5];
  do_intersect_primitives(ox, oy, oz, dx, dy, dz, false);
  const float dist = rec_[R_T];
  const int po = jf2i(rec_[R_PRIM]);
  f_.primitive_offsets[gid_] = po;
  if (dist == INFINITY_ || po == -1) {
    do_sky(sky_active, dx, dy, dz);
    film_add(tmp_[0], tmp_[1], tmp_[2]);
    return;
  }
  const int so = s_.a->primitives[po + 2];
  const int alb = tex_color(jf2i(s_.a->surfaces[so + 1]));
  const float aR = ((alb >> 16) & 0xFF) * COLOR_RECIPROCAL;
  const float aG = ((alb >> 8) & 0xFF) * COLOR_RECIPROCAL;
  const float aB = (alb & 0xFF) * COLOR_RECIPROCAL;
  const float pX = rec_[R_P], pY = rec_[R_P + 1], pZ = rec_[R_P + 2];
  const float sX = rec_[R_NS], sY = rec_[R_NS + 1], sZ = rec_[R_NS + 2];
  const float dp = sX * dx + sY * dy + sZ * dz;
  const bool ok = dp < 0.0f;
  const float wX = ok ? sX : -sX, wY = ok ? sY : -sY, wZ = ok ? sZ : -sZ;
  const int c = phong_shader0(true, pX, pY, pZ, wX, wY, wZ, -dx, -dy, -dz, aR, aG, aB, 0.2f, 0.2f, 0.2f, 0.5f, 0.5f,
                              0.5f, 0.5f, 0.5f, 0.5f, 250.0f);
  film_add(((c >> 16) & 0xFF) * COLOR_RECIPROCAL, ((c >> 8) & 0xFF) * COLOR_RECIPROCAL, (c & 0xFF) * COLOR_RECIPROCAL);
}

void Item::render_rm(bool sky_active) {  // :3802-3864
  const float ox = ray_[0], oz = ray_[2];
  const float oy = get_y(ox, oz) + 0.1f;
  const float dx = ray_[3], dy = ray_[4], dz = ray_[5];
  const float dtm = 0.01f;
  float dt = dtm;
  const float mint = 0.001f;
  const float maxt = jmaxf(oy, 1.0f) * 20.0f;
  do_sky(sky_active, dx, dy, dz);
  float pr = tmp_[0], pg = tmp_[1], pb = tmp_[2];
  for (float t = mint; t < maxt; t += dt) {
    const float pX = ox + dx * t, pY = oy + dy * t, pZ = oz + dz * t;
    const float h = get_y(pX, pZ);
    if (pY < h) {
      const float nX = get_y(pX - EPSILON, pZ) - get_y(pX + EPSILON, pZ);
      const float nY = 2.0f * EPSILON;
      const float nZ = get_y(pX, pZ - EPSILON) - get_y(pX, pZ + EPSILON);
      const float nr = 1.0f / jsqrt(nX * nX + nY * nY + nZ * nZ);
      const float n = simplex_fbm_xyz(8.0f, 0.5f, 0.0f, 1.0f, 16, pX, pY, pZ);
      const float a = saturate(n * 1.0f + 0.1f, 0.0f, 1.0f);
      const int c = phong_shader0(false, pX, pY, pZ, nX * nr, nY * nr, nZ * nr, -dx, -dy, -dz, a, a, a, 0.2f, 0.2f,
                                  0.2f, 0.5f, 0.5f, 0.5f, 0.5f, 0.5f, 0.5f, 250.0f);
      pr = ((c >> 16) & 0xFF) * COLOR_RECIPROCAL;
      pg = ((c >> 8) & 0xFF) * COLOR_RECIPROCAL;
      pb = (c & 0xFF) * COLOR_RECIPROCAL;
      t = maxt;
    }
    dt = dtm * t;
  }
  film_add(pr, pg, pb);
}

void Item::render_rt(bool sky_active) {  // :3866-3996
  const int depth_max = 5;
  int depth = 0;
  float ox = ray_[0], oy = ray_[1], oz = ray_[2], dx = ray_[3], dy = ray_[4], dz = ray_[5];
  float pr = 0.0f, pg = 0.0f, pb = 0.0f;
  do {
    do_intersect_primitives(ox, oy, oz, dx, dy, dz, false);
    const float dist = rec_[R_T];
    const int po = jf2i(rec_[R_PRIM]);
    if (depth == 0) f_.primitive_offsets[gid_] = po;
    if (dist == INFINITY_ || po == -1) {
      do_sky(sky_active, dx, dy, dz);
      pr += tmp_[0];
      pg += tmp_[1];
      pb += tmp_[2];
      film_add(pr, pg, pb);
      return;
    }
    depth++;
    const int so = s_.a->primitives[po + 2];
    const int mat = jf2i(s_.a->surfaces[so]);
    const int alb = tex_color(jf2i(s_.a->surfaces[so + 1]));
    const float aR = ((alb >> 16) & 0xFF) / 255.0f;
    const float aG = ((alb >> 8) & 0xFF) / 255.0f;
    const float aB = (alb & 0xFF) / 255.0f;
    const float pX = rec_[R_P], pY = rec_[R_P + 1], pZ = rec_[R_P + 2];
    const float sX = rec_[R_NS], sY = rec_[R_NS + 1], sZ = rec_[R_NS + 2];
    const float dp = sX * dx + sY * dy + sZ * dz;
    const bool ok = dp < 0.0f;
    const float wX = ok ? sX : -sX, wY = ok ? sY : -sY, wZ = ok ? sZ : -sZ;
    const int c = phong_shader0(true, pX, pY, pZ, wX, wY, wZ, -dx, -dy, -dz, aR, aG, aB, 0.2f, 0.2f, 0.2f, 0.5f,
                                0.5f, 0.5f, 0.5f, 0.5f, 0.5f, 250.0f);
    pr += ((c >> 16) & 0xFF) * COLOR_RECIPROCAL;
    pg += ((c >> 8) & 0xFF) * COLOR_RECIPROCAL;
    pb += (c & 0xFF) * COLOR_RECIPROCAL;
    if (mat == 5) {
      const float dp2 = dp * 2.0f;
      ox = pX + sX * 0.000001f;
      oy = pY + sY * 0.000001f;
      oz = pZ + sZ * 0.000001f;
      dx = dx - sX * dp2;
      dy = dy - sY * dp2;
      dz = dz - sZ * dp2;
    } else {
      depth = depth_max;
    }
  } while (depth < depth_max);
  film_add(pr, pg, pb);
}

void Item::render_sn(bool sky_active) {  // :2848-2908
  const float ox = ray_[0], oy = ray_[1], oz = ray_[2], dx = ray_[3], dy = ray_[4], dz = ray_[5];
  do_intersect_primitives(ox, oy, oz, dx, dy, dz, false);
  const float dist = rec_[R_T];
  const int po = jf2i(rec_[R_PRIM]);
  f_.primitive_offsets[gid_] = po;
  if (dist == INFINITY_ || po == -1) {
    do_sky(sky_active, dx, dy, dz);
    film_add(tmp_[0], tmp_[1], tmp_[2]);
    return;
  }
  film_add((rec_[R_NS] + 1.0f) * 0.5f, (rec_[R_NS + 1] + 1.0f) * 0.5f, (rec_[R_NS + 2] + 1.0f) * 0.5f);
}

void Item::render_wireframes() {  // :2910-3027
  const float ox = ray_[0], oy = ray_[1], oz = ray_[2], dx = ray_[3], dy = ray_[4], dz = ray_[5];
  do_intersect_primitives(ox, oy, oz, dx, dy, dz, false);
  const float dist = rec_[R_T];
  const int po = jf2i(rec_[R_PRIM]);
  const float lw = PI * 0.5f / 4096.0f;
  const float lwc = jcos(lw);
  if (dist != INFINITY_ && po != -1) {
    const int st = jf2i(rec_[R_SHAPE_TYPE]);
    const int so = jf2i(rec_[R_SHAPE_OFF]);
    if (st == 4) {
      const float* c = cam();
      const float uX = c[7], uY = c[8], uZ = c[9], vX = c[10], vY = c[11], vZ = c[12], wX = c[13], wY = c[14],
                  wZ = c[15];
      const float pX = rec_[R_P], pY = rec_[R_P + 1], pZ = rec_[R_P + 2];
      const float pcX = pX * uX + pY * uY + pZ * uZ;
      const float pcY = pX * vX + pY * vY + pZ * vZ;
      const float pcZ = pX * wX + pY * wY + pZ * wZ;
      const float pcN = jrsqrt(pcX * pcX + pcY * pcY + pcZ * pcZ);
      const float* A = p3() + s_.a->triangles[so + 0];
      const float* B = p3() + s_.a->triangles[so + 1];
      const float* C = p3() + s_.a->triangles[so + 2];
      const float aX = A[0] * uX + A[1] * uY + A[2] * uZ, aY = A[0] * vX + A[1] * vY + A[2] * vZ,
                  aZ = A[0] * wX + A[1] * wY + A[2] * wZ;
      const float bX = B[0] * uX + B[1] * uY + B[2] * uZ, bY = B[0] * vX + B[1] * vY + B[2] * vZ,
                  bZ = B[0] * wX + B[1] * wY + B[2] * wZ;
      const float cX = C[0] * uX + C[1] * uY + C[2] * uZ, cY = C[0] * vX + C[1] * vY + C[2] * vZ,
                  cZ = C[0] * wX + C[1] * wY + C[2] * wZ;
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
      if (dpAC * prACN * pcN >= lwc || dpBA * prBAN * pcN >= lwc || dpCB * prCBN * pcN >= lwc) {
        film_set(0.0f, 0.0f, 0.0f);
      }
    }
  }
}

void Item::image_pass() {  // run() :282-307 and AbstractImageKernel.java:410-472, 503-774
  const int po = f_.primitive_offsets[gid_];
  const float tint_g = po > -1 && po == s_.p->selected_primitive_offset ? 1.0f : 0.0f;
  const float* fd = f_.film + gid_ * 3;
  float r = fd[0] + 0.0f, g = fd[1] + tint_g, b = fd[2] + 0.0f;
  const float ex = s_.p->tone_mapper_exposure;
  switch (s_.p->tone_mapper_type) {
    case DF_TONE_MAPPER_REINHARD: {
      const float rm = r * ex, gm = g * ex, bm = b * ex;
      r = rm / (1.0f + rm);
      g = gm / (1.0f + gm);
      b = bm / (1.0f + bm);
      break;
    }
    case DF_TONE_MAPPER_REINHARD_MODIFIED_1:
      r = 1.0f - jexp(-r * ex);
      g = 1.0f - jexp(-g * ex);
      b = 1.0f - jexp(-b * ex);
      break;
    case DF_TONE_MAPPER_REINHARD_MODIFIED_2: {
      const float lw2r = 1.0f / (4.0f * 4.0f);
      const float rm = r * ex, gm = g * ex, bm = b * ex;
      r = rm * (1.0f + rm * lw2r) / (1.0f + rm);
      g = gm * (1.0f + gm * lw2r) / (1.0f + gm);
      b = bm * (1.0f + bm * lw2r) / (1.0f + bm);
      break;
    }
    case DF_TONE_MAPPER_FILMIC_CURVE_ACES_MODIFIED: {
      const float A = 2.51f, B = 0.03f, C = 2.43f, D = 0.59f, E = 0.14f;
      auto f = [&](float x) {
        const float m = jmaxf(x * ex - 0.0f, -3.402823466e38f);
        return saturate((m * (A * m + B)) / (m * (C * m + D) + E), 0.0f, 1.0f);
      };
      r = f(r);
      g = f(g);
      b = f(b);
      break;
    }
    default:
      break;
  }
  // sRGB (image4j RGBColorSpace.SRGB: breakPoint 0.00304, gamma 2.4; slope,
  // slopeMatch and segmentOffset derived as in RGBColorSpace.java:148-158)
  const float bp = 0.00304f, gamma = 2.4f;
  const float slope = 1.0f / (gamma / jpow(bp, 1.0f / gamma - 1.0f) - gamma * bp + bp);
  const float slope_match = gamma * slope / jpow(bp, 1.0f / gamma - 1.0f);
  const float seg = slope_match * jpow(bp, 1.0f / gamma) - slope * bp;
  auto gc = [&](float v) {
    return v <= 0.0f ? 0.0f : v >= 1.0f ? 1.0f : v <= bp ? v * slope : slope_match * jpow(v, 1.0f / gamma) - seg;
  };
  r = gc(r);
  g = gc(g);
  b = gc(b);
  if (f_.image_float) {
    f_.image_float[gid_ * 3 + 0] = r;
    f_.image_float[gid_ * 3 + 1] = g;
    f_.image_float[gid_ * 3 + 2] = b;
  }
  const int R = jf2i(jmaxf(jminf(r * 255.0f + 0.5f, 255.0f), 0.0f));
  const int G = jf2i(jmaxf(jminf(g * 255.0f + 0.5f, 255.0f), 0.0f));
  const int B = jf2i(jmaxf(jminf(b * 255.0f + 0.5f, 255.0f), 0.0f));
  uint8_t* o = f_.image_bgra + gid_ * 4;
  o[0] = static_cast<uint8_t>(B);
  o[1] = static_cast<uint8_t>(G);
  o[2] = static_cast<uint8_t>(R);
  o[3] = 255;
}

void Item::run() {  // GPURendererKernel.run :252-308
  if (create_primary_ray()) {
    shade_primary();
  } else {
    film_set(0.0f, 0.0f, 0.0f);
  }
  image_pass();
}

void Item::shade_primary() {  // run() :258-300, the renderer dispatch and the wireframe overlay
  const bool sky_active = s_.sky->is_sky_active == 1;
  const bool sun_active = s_.sky->is_sun_active == 1;
  switch (s_.p->renderer_type) {
    case DF_RENDERER_AMBIENT_OCCLUSION: render_ao(); break;
    case DF_RENDERER_PATH_TRACER: render_pt(sky_active, sun_active); break;
    case DF_RENDERER_RAY_CASTER: render_rc(sky_active); break;
    case DF_RENDERER_RAY_MARCHER: render_rm(sky_active); break;
    case DF_RENDERER_RAY_TRACER: render_rt(sky_active); break;
    case DF_RENDERER_SURFACE_NORMALS: render_sn(sky_active); break;
    default: render_pt(sky_active, sun_active); break;
  }
  if (s_.p->wireframes == 1) render_wireframes();
}

void Item::run_ray(const float ray[6]) {
  std::memcpy(ray_, ray, sizeof ray_);
  shade_primary();
  image_pass();
}

}  // namespace

// ------------------------------------------------------------ C interface

extern "C" {

typedef struct oracle_frame {
  int32_t width, height;
  float* film;
  int32_t* film_samples;
  float* image_float;  // may be NULL
  uint8_t* image_bgra;
  int32_t* primitive_offsets;
  int64_t* seeds;
} oracle_frame;

typedef struct oracle_counters {
  uint64_t rays, samples, box_tests, tri_tests, prim_tests, sphere_tests, terrain_steps, sky_evals;
} oracle_counters;

// Runs `passes` passes of run() over pixels [pixel_begin, pixel_end) with
// `threads` workers over contiguous pixel chunks; the first pass honours
// `clear_first` (the film clear flag).  Returns 0 or DF_EINVAL.
int oracle_render(const df_compiled_scene* scene, const df_sky* sky, const df_params* params, oracle_frame* frame,
                  int passes, int clear_first, int threads, int64_t pixel_begin, int64_t pixel_end,
                  oracle_counters* counters) {
  if (!scene || !sky || !params || !frame || passes < 0) return DF_EINVAL;
  const int64_t total = static_cast<int64_t>(frame->width) * frame->height;
  if (pixel_begin < 0 || pixel_end > total || pixel_begin > pixel_end) return DF_EINVAL;
  Scene s{scene, sky, params, static_cast<int>(scene->primitives_len / 3)};
  if (threads < 1) threads = 1;
  std::vector<Counters> cs(threads);
  for (int pass = 0; pass < passes; ++pass) {
    Frame f{frame->width, frame->height, frame->film, frame->film_samples, frame->image_float, frame->image_bgra,
            frame->primitive_offsets, frame->seeds, (pass == 0 && clear_first) ? 1 : 0};
    const int64_t n = pixel_end - pixel_begin;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
      const int64_t b = pixel_begin + n * t / threads;
      const int64_t e = pixel_begin + n * (t + 1) / threads;
      pool.emplace_back([&, b, e, t]() {
        for (int64_t g = b; g < e; ++g) {
          Item it(s, f, static_cast<int>(g), cs[t]);
          it.run();
        }
      });
    }
    for (auto& th : pool) th.join();
  }
  if (counters) {
    std::memset(counters, 0, sizeof *counters);
    for (const Counters& c : cs) {
      counters->rays += c.rays;
      counters->samples += c.samples;
      counters->box_tests += c.box_tests;
      counters->tri_tests += c.tri_tests;
      counters->prim_tests += c.prim_tests;
      counters->sphere_tests += c.sphere_tests;
      counters->terrain_steps += c.terrain_steps;
      counters->sky_evals += c.sky_evals;
    }
  }
  return 0;
}

// Differentiable mode of the path tracer (no counterpart in the reference):
// the surface records mapped to continuous parameters (albedo group x RGB,
// white emission, checkerboard group x 2 colours x RGB; -1 keeps the
// texture), the parameter values, and the sky's parameters (the df_sky
// rendered with is Sky.set of them).
typedef struct oracle_diff {
  int32_t n_albedo, n_emission;
  const int32_t* surface_albedo;    // per surface record
  const int32_t* surface_emission;  // per surface record
  int64_t n_surfaces;
  const float* albedo;    // n_albedo x 3
  const float* emission;  // n_emission
  int32_t n_checker;
  const int32_t* surface_checker;  // per surface record, or null (no checkerboard groups)
  const float* checker;            // n_checker x 2 x 3: the first colour (dark tiles), the second
  double turbidity, elevation_deg, azimuth_deg;
} oracle_diff;

// L = sum over the crop [x0, x0 + crop_w) x [y0, y0 + crop_h) of
// (film - target)^2, the film being the mean of `passes` path-traced samples
// per pixel from the starting seeds given (width x height, not modified), and
// its gradient dL/d[albedo groups x RGB, emission, checkerboard groups x 2 x
// RGB, turbidity, sun elevation per degree, sun azimuth per degree] by
// forward-mode differentiation in double.  Needs the path tracer with Russian
// roulette off, no wireframes and depth <= 64.  A surface takes at most one
// albedo map, and a checkerboard group only a checkerboard albedo texture.
// Writes the crop's film (crop_w x crop_h x 3).  Returns 0 or DF_EINVAL.
int oracle_loss_gradient(const df_compiled_scene* scene, const df_sky* sky, const df_params* params,
                         const oracle_diff* diff, int32_t width, int32_t height, int32_t x0, int32_t y0,
                         int32_t crop_w, int32_t crop_h, const int64_t* seeds, const float* target, int passes,
                         int threads, float* film_out, double* loss_out, double* grad_out, int64_t n_grad) {
  if (!scene || !sky || !params || !diff || !seeds || !target || !film_out || !loss_out || !grad_out) return DF_EINVAL;
  if (params->renderer_type != DF_RENDERER_PATH_TRACER || params->wireframes == 1) return DF_EINVAL;
  if (params->pt_depth_russian_roulette <= params->pt_depth_maximum) return DF_EINVAL;
  if (params->pt_depth_maximum > PathTangent::kMaxSlots) return DF_EINVAL;
  if (width <= 0 || height <= 0 || x0 < 0 || y0 < 0 || crop_w <= 0 || crop_h <= 0 || x0 + crop_w > width ||
      y0 + crop_h > height || passes < 1)
    return DF_EINVAL;
  if (diff->n_albedo < 0 || diff->n_emission < 0 || diff->n_surfaces != static_cast<int64_t>(scene->surfaces_len / 6))
    return DF_EINVAL;
  const int n_checker = diff->surface_checker ? diff->n_checker : 0;
  if (n_checker < 0 || (n_checker > 0 && !diff->checker)) return DF_EINVAL;
  const int n = 3 * diff->n_albedo + diff->n_emission + 6 * n_checker + 3;
  if (n_grad != n) return DF_EINVAL;
  for (int64_t i = 0; i < diff->n_surfaces; ++i) {
    if (diff->surface_albedo[i] < -1 || diff->surface_albedo[i] >= diff->n_albedo) return DF_EINVAL;
    if (diff->surface_emission[i] < -1 || diff->surface_emission[i] >= diff->n_emission) return DF_EINVAL;
    if (!diff->surface_checker) continue;
    const int g = diff->surface_checker[i];
    if (g < -1 || g >= n_checker) return DF_EINVAL;
    if (g >= 0 && (diff->surface_albedo[i] >= 0 || jf2i(scene->textures[jf2i(scene->surfaces[i * 6 + 1])]) != 3))
      return DF_EINVAL;
  }
  const DualSky K = dual_sky(sky, diff->turbidity, diff->elevation_deg, diff->azimuth_deg);
  for (int i = 0; i < 3; ++i)  // the sky must be the one these parameters set up
    if (std::fabs(K.world[i].v - sky->sun_direction_world[i]) > 1e-5) return DF_EINVAL;
  if (std::fabs(diff->turbidity - sky->turbidity) > 1e-5 * diff->turbidity) return DF_EINVAL;

  const DiffMaps maps{diff->n_albedo,   diff->n_emission, diff->surface_albedo, diff->surface_emission,
                      static_cast<int>(diff->n_surfaces), diff->albedo,     diff->emission,       n_checker,
                      n_checker > 0 ? diff->surface_checker : nullptr,      diff->checker};
  Scene s{scene, sky, params, static_cast<int>(scene->primitives_len / 3), &maps};
  const size_t npix = static_cast<size_t>(width) * height;
  std::vector<float> film(npix * 3, 0.0f);
  std::vector<int32_t> film_samples(npix, 0), prim(npix, -1);
  std::vector<uint8_t> bgra(npix * 4, 0);
  std::vector<int64_t> sd(seeds, seeds + npix);
  if (threads < 1) threads = 1;
  if (threads > crop_h) threads = crop_h;
  std::vector<std::vector<double>> grads(threads, std::vector<double>(n, 0.0));
  std::vector<double> losses(threads, 0.0);
  std::vector<Counters> cs(threads);
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) {
    pool.emplace_back([&, t]() {
      std::vector<double> dpix(static_cast<size_t>(n) * 3);
      PathTangent pt;
      pt.sky = &K;
      pt.n_albedo = diff->n_albedo;
      pt.n_emission = diff->n_emission;
      pt.i_emission = 3 * diff->n_albedo;
      pt.i_turbidity = n - 3;
      pt.i_elevation = n - 2;
      pt.i_azimuth = n - 1;
      pt.dpix = dpix.data();
      std::vector<double>& g = grads[t];
      for (int y = y0 + crop_h * t / threads; y < y0 + crop_h * (t + 1) / threads; ++y) {
        for (int x = x0; x < x0 + crop_w; ++x) {
          const int gid = y * width + x;
          std::fill(dpix.begin(), dpix.end(), 0.0);
          for (int pass = 0; pass < passes; ++pass) {
            Frame f{width, height, film.data(), film_samples.data(), nullptr, bgra.data(), prim.data(), sd.data(),
                    pass == 0 ? 1 : 0};
            Item it(s, f, gid, cs[t]);
            it.set_tangent(&pt);
            it.run();
          }
          float* fo = film_out + (static_cast<size_t>(y - y0) * crop_w + (x - x0)) * 3;
          for (int c = 0; c < 3; ++c) {
            const double dl = static_cast<double>(film[gid * 3 + c]) - target[gid * 3 + c];
            fo[c] = film[gid * 3 + c];
            losses[t] += dl * dl;
            const double wgt = 2.0 * dl / passes;
            for (int i = 0; i < n; ++i) g[i] += wgt * dpix[i * 3 + c];
          }
        }
      }
    });
  }
  for (auto& th : pool) th.join();
  *loss_out = 0.0;
  for (int i = 0; i < n; ++i) grad_out[i] = 0.0;
  for (int t = 0; t < threads; ++t) {
    *loss_out += losses[t];
    for (int i = 0; i < n; ++i) grad_out[i] += grads[t][i];
  }
  return 0;
}

// The dual-number sky of the FP64 adjoint towards d: out[c] the radiance,
// out[3 + c] its derivative with respect to turbidity, out[6 + c] with
// respect to the sun elevation in degrees, out[9 + c] to the sun azimuth in
// degrees.  Returns 0 below the horizon.
int oracle_sky_dual(const df_sky* frame, double turbidity, double elevation_deg, double azimuth_deg, const float d[3],
                    double out[12]) {
  const DualSky K = dual_sky(frame, turbidity, elevation_deg, azimuth_deg);
  D3 rgb[3];
  if (!dual_sky_rgb(K, d[0], d[1], d[2], rgb)) return 0;
  for (int c = 0; c < 3; ++c) {
    out[c] = rgb[c].v;
    out[3 + c] = rgb[c].t;
    out[6 + c] = rgb[c].e;
    out[9 + c] = rgb[c].a;
  }
  return 1;
}

// The dual-number sun cosine of the FP64 adjoint: out[0] the cosine, out[1]
// its derivative with respect to the sun elevation in degrees, out[2] to the
// sun azimuth in degrees.
void oracle_sun_cosine_dual(const df_sky* frame, double turbidity, double elevation_deg, double azimuth_deg,
                            const float p[3], const float n[3], float u, float v, double out[3]) {
  const DualSky K = dual_sky(frame, turbidity, elevation_deg, azimuth_deg);
  const D3 c = dual_sun_cosine(K, p[0], p[1], p[2], n[0], n[1], n[2], u, v);
  out[0] = c.v;
  out[1] = c.e;
  out[2] = c.a;
}

// doCreatePrimaryRay (GPURendererKernel.java:573-724) for pixel (x, y) of a width x height frame with the
// camera array and renderer type given; advances *seed by the draws it makes.  Writes the ray (origin, unit
// direction) and returns 1, or returns 0 where the lens gives no ray.
int oracle_primary_ray(const float camera[19], int32_t renderer_type, int32_t width, int32_t height, int32_t x,
                       int32_t y, int64_t* seed, float ray_out[6]) {
  df_compiled_scene a{};
  a.camera = camera;
  a.camera_len = 19;
  df_params p{};
  p.renderer_type = renderer_type;
  df_sky sk{};
  Scene s{&a, &sk, &p, 0};
  std::vector<int64_t> seeds(static_cast<size_t>(width) * height, 0);
  const int gid = y * width + x;
  seeds[gid] = *seed;
  float film[3];
  int32_t n = 0, po = -1;
  uint8_t bgra[4];
  Frame f{width, height, film, &n, nullptr, bgra, &po, seeds.data(), 0};
  Counters c;
  Item it(s, f, gid, c);
  const bool ok = it.primary_ray(ray_out);
  *seed = seeds[gid];
  return ok ? 1 : 0;
}

// The path tracer's BSDF sample of material `material` (1 ClearCoat, 2 Glass, 3 Lambertian, 4 Phong,
// 5 Reflection; GPURendererKernel.java:3323-3686) for a ray arriving along the unit direction d at the point p
// with the shading normal s, the albedo given; advances *seed by the draws it makes.  out = next origin (3),
// next direction (3), the throughput factor (3) the sample multiplies the path's throughput by.
void oracle_scatter(int32_t material, const float p[3], const float s[3], const float d[3], const float albedo[3],
                    int64_t* seed, float out[9]) {
  df_compiled_scene a{};
  df_params pr{};
  df_sky sk{};
  Scene sc{&a, &sk, &pr, 0};
  float film[3];
  int32_t n = 0, po = -1;
  uint8_t bgra[4];
  Frame f{1, 1, film, &n, nullptr, bgra, &po, seed, 0};
  Counters c;
  Item it(sc, f, 0, c);
  float w[3], dp2;
  Item::orient(s[0], s[1], s[2], d[0], d[1], d[2], w, &dp2);
  float ox = 0.0f, oy = 0.0f, oz = 0.0f, dx = d[0], dy = d[1], dz = d[2], mr = 1.0f, mg = 1.0f, mb = 1.0f;
  it.scatter(material, -1, p[0], p[1], p[2], s[0], s[1], s[2], w[0], w[1], w[2], dp2, albedo[0], albedo[1],
             albedo[2], ox, oy, oz, dx, dy, dz, mr, mg, mb);
  const float o9[9] = {ox, oy, oz, dx, dy, dz, mr, mg, mb};
  std::memcpy(out, o9, sizeof o9);
}

// The sun direction doCalculateColorForSun samples from the point p with the random numbers (u, v)
// (GPURendererKernel.java:2252-2320).
void oracle_sun_direction(const df_sky* sky, const float p[3], float u, float v, float d_out[3]) {
  df_compiled_scene a{};
  df_params pr{};
  Scene sc{&a, sky, &pr, 0};
  int64_t seed = 0;
  float film[3];
  int32_t n = 0, po = -1;
  uint8_t bgra[4];
  Frame f{1, 1, film, &n, nullptr, bgra, &po, &seed, 0};
  Counters c;
  Item(sc, f, 0, c).sun_direction(p[0], p[1], p[2], u, v, d_out);
}

// run() for one pixel whose primary ray is given (origin, unit direction): the renderer of `params`, a cleared
// film, the seed *seed (advanced).  Writes the film sample, the primitive offset buffer and the BGRA bytes.
int oracle_render_ray(const df_compiled_scene* scene, const df_sky* sky, const df_params* params, const float ray[6],
                      int64_t* seed, float rgb_out[3], int32_t* primitive_offset, uint8_t bgra_out[4]) {
  if (!scene || !sky || !params || !ray || !seed || !rgb_out) return DF_EINVAL;
  Scene s{scene, sky, params, static_cast<int>(scene->primitives_len / 3)};
  int32_t n = 0, po = -1;
  uint8_t bgra[4];
  Frame f{1, 1, rgb_out, &n, nullptr, bgra, &po, seed, 1};
  Counters c;
  Item(s, f, 0, c).run_ray(ray);
  if (primitive_offset) *primitive_offset = po;
  if (bgra_out) std::memcpy(bgra_out, bgra, 4);
  return 0;
}

// Single ray query against the scene (doIntersectPrimitives); writes the
// 27-float hit record when testing == 0.  Returns t.
float oracle_intersect(const df_compiled_scene* scene, const df_sky* sky, const df_params* params, const float o[3],
                       const float d[3], int testing, float* record_out) {
  Scene s{scene, sky, params, static_cast<int>(scene->primitives_len / 3)};
  int64_t seed = 0;
  float film[3] = {0, 0, 0};
  int32_t n = 0, po = -1;
  uint8_t bgra[4];
  Frame f{1, 1, film, &n, nullptr, bgra, &po, &seed, 0};
  Counters c;
  Item it(s, f, 0, c);
  const float t = it.intersect(o[0], o[1], o[2], d[0], d[1], d[2], testing != 0);
  if (record_out) std::memcpy(record_out, it.record(), sizeof(float) * R_SIZE);
  return t;
}

// Sky radiance along a world direction (doCalculateColorForSky).
void oracle_sky(const df_sky* sky, const float d[3], int active, float rgb_out[3]) {
  df_compiled_scene empty{};
  df_params p{};
  Scene s{&empty, sky, &p, 0};
  int64_t seed = 0;
  float film[3];
  int32_t n = 0, po = -1;
  uint8_t bgra[4];
  Frame f{1, 1, film, &n, nullptr, bgra, &po, &seed, 0};
  Counters c;
  Item it(s, f, 0, c);
  it.sky_color(active != 0, d[0], d[1], d[2], rgb_out);
}

// Noise entry points (AbstractKernel simplexNoiseXY/XYZ and the fBm wrapper).
float oracle_simplex2(float x, float y) {
  df_compiled_scene empty{};
  df_params p{};
  df_sky sk{};
  Scene s{&empty, &sk, &p, 0};
  int64_t seed = 0;
  float film[3];
  int32_t n = 0, po = -1;
  uint8_t bgra[4];
  Frame f{1, 1, film, &n, nullptr, bgra, &po, &seed, 0};
  Counters c;
  return Item(s, f, 0, c).simplex2(x, y);
}

float oracle_simplex3(float x, float y, float z) {
  df_compiled_scene empty{};
  df_params p{};
  df_sky sk{};
  Scene s{&empty, &sk, &p, 0};
  int64_t seed = 0;
  float film[3];
  int32_t n = 0, po = -1;
  uint8_t bgra[4];
  Frame f{1, 1, film, &n, nullptr, bgra, &po, &seed, 0};
  Counters c;
  return Item(s, f, 0, c).simplex3(x, y, z);
}

float oracle_fbm3(float frequency, float gain, float minimum, float maximum, int octaves, float x, float y, float z) {
  df_compiled_scene empty{};
  df_params p{};
  df_sky sk{};
  Scene s{&empty, &sk, &p, 0};
  int64_t seed = 0;
  float film[3];
  int32_t n = 0, po = -1;
  uint8_t bgra[4];
  Frame f{1, 1, film, &n, nullptr, bgra, &po, &seed, 0};
  Counters c;
  return Item(s, f, 0, c).fbm3(frequency, gain, minimum, maximum, octaves, x, y, z);
}

// Texture lookup given a hit record (doGetTextureColor); returns packed RGB.
int32_t oracle_texture(const float* textures, size_t textures_len, int offset, const float record[27]) {
  df_compiled_scene a{};
  a.textures = textures;
  a.textures_len = textures_len;
  df_params p{};
  df_sky sk{};
  Scene s{&a, &sk, &p, 0};
  int64_t seed = 0;
  float film[3];
  int32_t n = 0, po = -1;
  uint8_t bgra[4];
  Frame f{1, 1, film, &n, nullptr, bgra, &po, &seed, 0};
  Counters c;
  Item it(s, f, 0, c);
  it.set_record(record);
  return it.texture_color(offset);
}

// The image stage of run() (selection tint, tone map, sRGB, BGRA) applied to
// one film colour; writes the gamma-corrected float colour and BGRA bytes.
void oracle_tonemap(const df_params* params, const float rgb[3], int32_t primitive_offset, float out_float[3],
                    uint8_t out_bgra[4]) {
  df_compiled_scene empty{};
  df_sky sk{};
  Scene s{&empty, &sk, params, 0};
  int64_t seed = 0;
  float film[3] = {rgb[0], rgb[1], rgb[2]};
  int32_t n = 1, po = primitive_offset;
  Frame f{1, 1, film, &n, out_float, out_bgra, &po, &seed, 0};
  Counters c;
  Item(s, f, 0, c).tonemap_only();
}

// Daylight chromaticity + luminance -> sRGB step of the sky (see above).
void oracle_daylight_xyy_to_srgb(float x, float y, double luminance, float rgb_out[3]) {
  daylight_xyy_to_srgb(x, y, luminance, &rgb_out[0], &rgb_out[1], &rgb_out[2]);
}

// Switches on the deliberate mistake `id` (enum Plant; 0 restores the faithful restatement) for the pin tests
// that show their checks fail on it.  Returns the previous value.
int oracle_plant(int id) {
  const int old = g_plant;
  g_plant = id;
  return old;
}

// The per-pixel RNG: advances *seed once and returns nextFloat().
float oracle_next_float(int64_t* seed) {
  const uint64_t old = static_cast<uint64_t>(*seed);
  const uint64_t nw = (old * 0x5DEECE66Dull + 0xBull) & ((1ull << 48) - 1);
  *seed = static_cast<int64_t>(nw);
  return static_cast<float>(static_cast<int>(nw >> 24)) * (1.0f / static_cast<float>(1 << 24));
}

}  // extern "C"
