// Forward-mode derivatives of the sky model for the adjoint: Sky.set
// (Sky.java:295-333) on the host and the sky and sun evaluations of the
// kernel (GPURendererKernel.java:2127-2236, 2238-2347) on the device, in
// double, each value carrying its derivatives with respect to the sky's
// three parameters (turbidity, sun elevation and sun azimuth, the angles in
// degrees).  The sun points towards (cos e sin a, sin e, cos e cos a) in the
// world (y up, azimuth from +z towards +x), the direction Sky.set receives.
#pragma once

#include <hip/hip_runtime.h>

#include "df_device_scene.h"

namespace dfk {

enum { kDualTurbidity = 0, kDualElevation = 1, kDualAzimuth = 2 };

__host__ __device__ inline Dual3 dconst(double v) { return Dual3{v, {0.0, 0.0, 0.0}}; }
__host__ __device__ inline Dual3 operator+(const Dual3& x, const Dual3& y) {
  return Dual3{x.v + y.v, {x.d[0] + y.d[0], x.d[1] + y.d[1], x.d[2] + y.d[2]}};
}
__host__ __device__ inline Dual3 operator-(const Dual3& x, const Dual3& y) {
  return Dual3{x.v - y.v, {x.d[0] - y.d[0], x.d[1] - y.d[1], x.d[2] - y.d[2]}};
}
__host__ __device__ inline Dual3 operator-(const Dual3& x) { return Dual3{-x.v, {-x.d[0], -x.d[1], -x.d[2]}}; }
__host__ __device__ inline Dual3 operator*(const Dual3& x, const Dual3& y) {
  return Dual3{x.v * y.v, {x.d[0] * y.v + x.v * y.d[0], x.d[1] * y.v + x.v * y.d[1], x.d[2] * y.v + x.v * y.d[2]}};
}
__host__ __device__ inline Dual3 operator*(const Dual3& x, double k) {
  return Dual3{x.v * k, {x.d[0] * k, x.d[1] * k, x.d[2] * k}};
}
__host__ __device__ inline Dual3 operator*(double k, const Dual3& x) { return x * k; }
__host__ __device__ inline Dual3 operator+(const Dual3& x, double k) { return Dual3{x.v + k, {x.d[0], x.d[1], x.d[2]}}; }
__host__ __device__ inline Dual3 operator+(double k, const Dual3& x) { return x + k; }
__host__ __device__ inline Dual3 operator-(double k, const Dual3& x) { return Dual3{k - x.v, {-x.d[0], -x.d[1], -x.d[2]}}; }
__host__ __device__ inline Dual3 operator/(const Dual3& x, const Dual3& y) {
  const double q = x.v / y.v, r = 1.0 / y.v;
  return Dual3{q, {(x.d[0] - q * y.d[0]) * r, (x.d[1] - q * y.d[1]) * r, (x.d[2] - q * y.d[2]) * r}};
}
// f(x) with f'(x) = df
__host__ __device__ inline Dual3 dapply(const Dual3& x, double f, double df) {
  return Dual3{f, {x.d[0] * df, x.d[1] * df, x.d[2] * df}};
}
__host__ __device__ inline Dual3 dexp(const Dual3& x) {
  const double f = ::exp(x.v);
  return dapply(x, f, f);
}
__host__ __device__ inline Dual3 dsin(const Dual3& x) { return dapply(x, ::sin(x.v), ::cos(x.v)); }
__host__ __device__ inline Dual3 dcos(const Dual3& x) { return dapply(x, ::cos(x.v), -::sin(x.v)); }
__host__ __device__ inline Dual3 dtan(const Dual3& x) {
  const double c = ::cos(x.v);
  return dapply(x, ::tan(x.v), 1.0 / (c * c));
}
// sqrt(max(x, 0)); zero with zero derivatives where x <= 0 (the centre of
// the sun's cone, where 1 - cos^2 stays zero for every parameter)
__host__ __device__ inline Dual3 dsqrt0(const Dual3& x) {
  if (!(x.v > 0.0)) return dconst(0.0);
  const double f = ::sqrt(x.v);
  return dapply(x, f, 0.5 / f);
}
// acos(clamp(x, -1, 1)): constant outside [-1, 1]
__host__ __device__ inline Dual3 dacos(const Dual3& x) {
  if (x.v >= 1.0) return dconst(0.0);
  if (x.v <= -1.0) return dconst(3.141592653589793);
  return dapply(x, ::acos(x.v), -1.0 / ::sqrt(1.0 - x.v * x.v));
}

// Sky.set for (turbidity, elevation, azimuth) in the sky frame (u, v, w).
inline DSkyDual sky_set_dual(const float onb_u[3], const float onb_v[3], const float onb_w[3], double turbidity,
                             double elevation_deg, double azimuth_deg) {
  DSkyDual S{};
  for (int i = 0; i < 3; ++i) {
    S.onb_u[i] = onb_u[i];
    S.onb_v[i] = onb_v[i];
    S.onb_w[i] = onb_w[i];
  }
  const double deg = 3.141592653589793 / 180.0;
  const Dual3 T{turbidity, {1.0, 0.0, 0.0}};
  const Dual3 e{elevation_deg * deg, {0.0, deg, 0.0}};
  const Dual3 a{azimuth_deg * deg, {0.0, 0.0, deg}};
  const Dual3 ce = dcos(e);
  S.world[0] = ce * dsin(a);
  S.world[1] = dsin(e);
  S.world[2] = ce * dcos(a);
  const float* frame[3] = {onb_u, onb_v, onb_w};
  for (int i = 0; i < 3; ++i)
    S.dir[i] = S.world[0] * static_cast<double>(frame[i][0]) + S.world[1] * static_cast<double>(frame[i][1]) +
               S.world[2] * static_cast<double>(frame[i][2]);
  const Dual3 len = dsqrt0(S.dir[0] * S.dir[0] + S.dir[1] * S.dir[1] + S.dir[2] * S.dir[2]);
  for (Dual3& c : S.dir) c = c / len;
  S.theta = dacos(S.dir[2]);
  const Dual3 th = S.theta, th2 = th * th, th3 = th2 * th, T2 = T * T;
  // zenith luminance (MathF.PI as the reference's float constant)
  const double pi_f = 3.1415927410125732;
  const Dual3 chi = (4.0 / 9.0 - T * (1.0 / 120.0)) * (pi_f - 2.0 * th);
  S.zl = ((4.0453 * T + -4.9710) * dtan(chi) - 0.2155 * T + 2.4192) * 1000.0;
  // zenith chromaticity: rows of (theta^3, theta^2, theta, 1) for T^2, T, 1
  const double zx[3][4] = {{0.00165, -0.00374, 0.00208, 0.0},
                           {-0.02902, 0.06377, -0.03202, 0.00394},
                           {0.11693, -0.21196, 0.06052, 0.25885}};
  const double zy[3][4] = {{0.00275, -0.00610, 0.00316, 0.0},
                           {-0.04212, 0.08970, -0.04153, 0.00515},
                           {0.15346, -0.26756, 0.06669, 0.26688}};
  const Dual3 powT[3] = {T2, T, dconst(1.0)};
  S.zx = dconst(0.0);
  S.zy = dconst(0.0);
  for (int r = 0; r < 3; ++r) {
    S.zx = S.zx + (zx[r][0] * th3 + zx[r][1] * th2 + zx[r][2] * th + zx[r][3]) * powT[r];
    S.zy = S.zy + (zy[r][0] * th3 + zy[r][1] * th2 + zy[r][2] * th + zy[r][3]) * powT[r];
  }
  // Perez coefficients A..E, each slope * T + intercept (Preetham et al. 1999)
  const double fits[3][5][2] = {
      {{0.17872, -1.46303}, {-0.35540, 0.42749}, {-0.02266, 5.32505}, {0.12064, -2.57705}, {-0.06696, 0.37027}},
      {{-0.01925, -0.25922}, {-0.06651, 0.00081}, {-0.00041, 0.21247}, {-0.06409, -0.89887}, {-0.00325, 0.04517}},
      {{-0.01669, -0.26078}, {-0.09495, 0.00921}, {-0.00792, 0.21023}, {-0.04405, -1.65369}, {-0.01092, 0.05291}}};
  for (int k = 0; k < 5; ++k) {
    S.pl[k] = fits[0][k][0] * T + fits[0][k][1];
    S.px[k] = fits[1][k][0] * T + fits[1][k][1];
    S.py[k] = fits[2][k][0] * T + fits[2][k][1];
  }
  return S;
}

// The sky radiance towards (dx, dy, dz), a direction that does not depend on
// the parameters, with its derivatives.  False below the horizon, where the
// sky is a constant.  The evaluation follows the kernel's, including its
// use of the view's zenith angle in the chromaticity denominators.
__device__ inline bool sky_rgb_dual(const DSkyDual& S, float dx, float dy, float dz, Dual3 out[3]) {
  const float d0X = dx * S.onb_u[0] + dy * S.onb_u[1] + dz * S.onb_u[2];
  const float d0Y = dx * S.onb_v[0] + dy * S.onb_v[1] + dz * S.onb_v[2];
  float d0Z = dx * S.onb_w[0] + dy * S.onb_w[1] + dz * S.onb_w[2];
  if (!(d0Z >= 0.0f)) return false;
  d0Z = d0Z > 0.001f ? d0Z : 0.001f;
  const double il = 1.0 / ::sqrt(static_cast<double>(d0X) * d0X + static_cast<double>(d0Y) * d0Y +
                                 static_cast<double>(d0Z) * d0Z);
  const double v[3] = {d0X * il, d0Y * il, d0Z * il};
  const double theta1 = ::acos(fmax(fmin(v[2], 1.0), -1.0));
  const double ct1r = 1.0 / (::cos(theta1) + 0.01);
  const Dual3 gamma = dacos(S.dir[0] * v[0] + S.dir[1] * v[1] + S.dir[2] * v[2]);
  const Dual3 cg = dcos(gamma);
  const Dual3 ct0 = dcos(S.theta);
  // Perez F(theta, gamma) = (1 + A exp(B / cos theta)) (1 + C exp(D gamma) + E cos^2 gamma)
  auto perez_num = [&](const Dual3* P) {
    return (1.0 + P[0] * dexp(P[1] * ct1r)) * (1.0 + P[2] * dexp(P[3] * gamma) + P[4] * cg * cg);
  };
  const Dual3 lden = (1.0 + S.pl[0] * dexp(S.pl[1])) * (1.0 + S.pl[2] * dexp(S.pl[3] * S.theta) + S.pl[4] * ct0 * ct0);
  const Dual3 lum = S.zl * perez_num(S.pl) / lden * 1.0e-4;
  const Dual3 xden = (1.0 + S.px[0] * dexp(S.px[1])) * (1.0 + S.px[2] * dexp(S.px[3] * theta1) + S.px[4] * ct0 * ct0);
  const Dual3 x = S.zx * perez_num(S.px) / xden;
  const Dual3 yden = (1.0 + S.py[0] * dexp(S.py[1])) * (1.0 + S.py[2] * dexp(S.py[3] * theta1) + S.py[4] * ct0 * ct0);
  const Dual3 y = S.zy * perez_num(S.py) / yden;
  // CIE daylight chromaticity (x, y) -> XYZ -> linear sRGB
  const Dual3 den = 0.0241 + 0.2562 * x - 0.7341 * y;
  const Dual3 m1 = (-1.3515 - 1.7703 * x + 5.9114 * y) / den;
  const Dual3 m2 = (0.03 - 31.4424 * x + 30.0717 * y) / den;
  const Dual3 cX = 10246.121 + 187.75537 * m1 + 213.14803 * m2;
  const Dual3 cY = 10676.695 + 192.59653 * m1 + 76.29494 * m2;
  const Dual3 cZ = 12372.504 + 3482.8765 * m1 + -235.71611 * m2;
  const Dual3 X = cX * lum / cY, Y = lum, Z = cZ * lum / cY;
  const Dual3 r = 3.2410042 * X + -1.5373994 * Y + -0.49861607 * Z;
  const Dual3 g = -0.9692241 * X + 1.8759298 * Y + 0.041554242 * Z;
  const Dual3 b = 0.05563942 * X + -0.20401107 * Y + 1.0571486 * Z;
  // the most negative channel lifted to zero, all three moved with it
  const Dual3* lo = r.v <= g.v ? &r : &g;
  if (b.v < lo->v) lo = &b;
  const Dual3 lift = lo->v < 0.0 ? -*lo : dconst(0.0);
  out[0] = r + lift;
  out[1] = g + lift;
  out[2] = b + lift;
  return true;
}

// The cosine between the normal n and the sun direction sampled with the
// draws (u, v) at p (the sun's origin 10000 units along its direction, its
// radius 500), with its derivatives: the sample's basis about the sun
// direction is the kernel's, chosen by the magnitudes of its components.
__device__ inline Dual3 sun_cosine_dual(const DSkyDual& S, float pX, float pY, float pZ, float nX, float nY, float nZ,
                                        float u, float v) {
  const Dual3* w = S.world;
  const Dual3 to[3] = {w[0] * 10000.0 + static_cast<double>(-pX), w[1] * 10000.0 + static_cast<double>(-pY),
                       w[2] * 10000.0 + static_cast<double>(-pZ)};
  const Dual3 to2 = to[0] * to[0] + to[1] * to[1] + to[2] * to[2];
  const Dual3 cos_max = dsqrt0(1.0 - dconst(250000.0) / to2);
  const Dual3 cos_t = static_cast<double>(u) * (cos_max + -1.0) + 1.0;
  const Dual3 sin_t = dsqrt0(1.0 - cos_t * cos_t);
  const double phi = static_cast<double>(6.2831854820251465f * v);
  const Dual3 lx = ::cos(phi) * sin_t, ly = ::sin(phi) * sin_t;
  const double aX = fabs(w[0].v), aY = fabs(w[1].v), aZ = fabs(w[2].v);
  Dual3 b[3];
  if (aX < aY && aX < aZ) {
    b[0] = dconst(0.0);
    b[1] = w[2];
    b[2] = -w[1];
  } else if (aY < aZ) {
    b[0] = w[2];
    b[1] = dconst(0.0);
    b[2] = -w[0];
  } else {
    b[0] = w[1];
    b[1] = -w[0];
    b[2] = dconst(0.0);
  }
  const Dual3 bl = dsqrt0(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  for (Dual3& c : b) c = c / bl;
  const Dual3 t[3] = {b[1] * w[2] - b[2] * w[1], b[2] * w[0] - b[0] * w[2], b[0] * w[1] - b[1] * w[0]};
  Dual3 r[3];
  for (int i = 0; i < 3; ++i) r[i] = t[i] * lx + b[i] * ly + w[i] * cos_t;
  const Dual3 rl = dsqrt0(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  return (r[0] * static_cast<double>(nX) + r[1] * static_cast<double>(nY) + r[2] * static_cast<double>(nZ)) / rl;
}

}  // namespace dfk
