// Adjoint (path-replay) kernels of the differentiable path tracer.
//
// Loss: L = sum over pixels and channels of (film - target)^2, film being the
// mean of `passes` path-traced samples.  Parameters: continuous albedo per
// surface group (RGB), emission per light (white radiance), the sky's
// turbidity and the sun's elevation.  Geometry is fixed, so visibility does
// not move; the gradient is the interior term.
//
// With Russian roulette off (its survival probability is the albedo's
// maximum, so it would make the path itself depend on the parameters), the
// directions and hits of a path do not depend on the parameters for a given
// RNG stream: a sample is a sum of terms beta_k * g_k, beta_k the product of
// the throughput factors of the vertices before k (s_j, times the albedo
// where the material multiplies it in) and g_k the emission, sun or sky
// term at k.  The adjoint kernel replays each pixel's paths from the seeds
// the forward pass started from, with the forward's own bounce code, and adds
// dL/dfilm / passes times the derivative of every term: of the albedo and
// emission it contains directly, of the earlier vertices' albedos through
// beta_k (a product over at most kMaxAdjointDepth factors, recomputed without
// division), and of turbidity and elevation through the sky radiance and the
// sun sample's cosine factor.  Those last two are smooth closed forms of the
// sky model (Sky.set on the host, the Perez evaluation here); their
// derivatives are central differences in double precision over skies set up
// at parameter offsets.  Partial gradients are summed per workgroup in LDS
// (double) and added to the output once per workgroup.
#include <hip/hip_runtime.h>

#include "df_kernels.h"
#include "df_pass_common.h"

namespace dfk {

// The sky radiance of doCalculateColorForSky (GPURendererKernel.java:2127-2236)
// for direction d, entirely in double precision (the forward rounds its last
// steps to float): the function whose parameter derivatives the adjoint
// takes.  Returns false below the horizon, where the sky is a constant.
__device__ bool sky_rgb_d(const DSky& K, float dx, float dy, float dz, double out[3]) {
  const float d0X = dx * K.onb_u[0] + dy * K.onb_u[1] + dz * K.onb_u[2];
  const float d0Y = dx * K.onb_v[0] + dy * K.onb_v[1] + dz * K.onb_v[2];
  float d0Z = dx * K.onb_w[0] + dy * K.onb_w[1] + dz * K.onb_w[2];
  if (!(d0Z >= 0.0f)) return false;
  d0Z = jmax(d0Z, 0.001f);
  const double l = 1.0 / ::sqrt(static_cast<double>(d0X) * d0X + static_cast<double>(d0Y) * d0Y +
                                static_cast<double>(d0Z) * d0Z);
  const double d1X = d0X * l, d1Y = d0Y * l, d1Z = d0Z * l;
  const double dp = d1X * K.sun_dir[0] + d1Y * K.sun_dir[1] + d1Z * K.sun_dir[2];
  const double theta1 = ::acos(fmax(fmin(d1Z, 1.0), -1.0));
  const double cosTheta1 = ::cos(theta1);
  const double cosTheta1R = 1.0 / (cosTheta1 + 0.01);
  const double gamma = ::acos(fmax(fmin(dp, 1.0), -1.0));
  const double cosGamma = ::cos(gamma);
  const double* L = K.pl;
  const double* X = K.px;
  const double* Y = K.py;
  const double lnum = (1.0 + L[0] * ::exp(L[1] * cosTheta1R)) * (1.0 + L[2] * ::exp(L[3] * gamma) + L[4] * cosGamma * cosGamma);
  const double lum = K.zenith_l * lnum / K.l_den * 1.0e-4;
  const double xden = K.x_den0 * (1.0 + X[2] * ::exp(X[3] * theta1) + K.x_c);
  const double xnum = (1.0 + X[0] * ::exp(X[1] * cosTheta1R)) * (1.0 + X[2] * ::exp(X[3] * gamma) + X[4] * cosGamma * cosGamma);
  const double x = K.zenith_x * xnum / xden;
  const double yden = K.y_den0 * (1.0 + Y[2] * ::exp(Y[3] * theta1) + K.y_c);
  const double ynum = (1.0 + Y[0] * ::exp(Y[1] * cosTheta1R)) * (1.0 + Y[2] * ::exp(Y[3] * gamma) + Y[4] * cosGamma * cosGamma);
  const double y = K.zenith_y * ynum / yden;
  const double cie0 = 1.0 / (0.0241 + 0.2562 * x - 0.7341 * y);
  const double cie1 = (-1.3515 - 1.7703 * x + 5.9114 * y) * cie0;
  const double cie2 = (0.03 - 31.4424 * x + 30.0717 * y) * cie0;
  const double cX = 10246.121 + cie1 * 187.75537 + cie2 * 213.14803;
  const double cY = 10676.695 + cie1 * 192.59653 + cie2 * 76.29494;
  const double cZ = 12372.504 + cie1 * 3482.8765 + cie2 * -235.71611;
  const double cR = cX * lum / cY, cG = lum, cB = cZ * lum / cY;
  double r = 3.2410042 * cR + -1.5373994 * cG + -0.49861607 * cB;
  double g = -0.9692241 * cR + 1.8759298 * cG + 0.041554242 * cB;
  double b = 0.05563942 * cR + -0.20401107 * cG + 1.0571486 * cB;
  const double w = fmax(0.0, -fmin(0.0, fmin(r, fmin(g, b))));
  out[0] = r + w;
  out[1] = g + w;
  out[2] = b + w;
  return true;
}

// dLoss/dfilm / passes per pixel channel, and the loss.
__global__ __launch_bounds__(256) void loss_kernel(const float* film, const float* target, float* weight, int n,
                                                   float scale, double* loss) {
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float dl = film[i] - target[i];
    weight[i] = 2.0f * dl * scale;
    acc += static_cast<double>(dl) * static_cast<double>(dl);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (__lane_id() == 0 && acc != 0.0) atomicAdd(loss, acc);
}

// Per-path bookkeeping of the replay: the throughput factors of the vertices
// so far.  Loops run over the compile-time maximum so that the arrays stay in
// registers.
struct PathFactors {
  float f[kMaxAdjointDepth][3];  // throughput factor of vertex j
  float s[kMaxAdjointDepth];     // its parameter-free part (f = s * albedo when slot >= 0)
  int slot[kMaxAdjointDepth];    // albedo group that multiplied the throughput, or -1
  int n;

  // d(beta)/d(albedo of vertex j's group), channel c: s_j * prod_{i != j} f_i.
  __device__ __forceinline__ float excl(int j, int c) const {
    float v = s[j];
#pragma unroll
    for (int i = 0; i < kMaxAdjointDepth; ++i)
      if (i < n && i != j) v *= f[i][c];
    return v;
  }
};

__device__ __forceinline__ void lds_add(double* g, int i, float v) {
  if (v != 0.0f) atomicAdd(&g[i], static_cast<double>(v));
}

// Gradient of a term beta * tail (tail per channel) through the albedos in beta.
__device__ __forceinline__ void through_beta(double* g, const PathFactors& pf, const float w[3], const float tail[3]) {
#pragma unroll
  for (int j = 0; j < kMaxAdjointDepth; ++j) {
    if (j < pf.n && pf.slot[j] >= 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) lds_add(g, pf.slot[j] * 3 + c, w[c] * tail[c] * pf.excl(j, c));
    }
  }
}

// The sky term beta * sky(d) / pi of a path that escaped: through the albedos
// in beta, and through turbidity and elevation (central differences).
template <int M>
__device__ __forceinline__ void sky_term(Tracer<M>& tr, double* g, const DAdjoint& A, const PathFactors& pf,
                                         const float w[3], const float beta[3], const float d[3], bool sky_active) {
  tr.sky(sky_active, d[0], d[1], d[2]);
  const float tail[3] = {tr.tmp[0] * DF_PI_RECIP, tr.tmp[1] * DF_PI_RECIP, tr.tmp[2] * DF_PI_RECIP};
  through_beta(g, pf, w, tail);
  if (!sky_active) return;
  double p[3], m[3];
  const int it = 3 * A.n_albedo + A.n_emission;
  if (sky_rgb_d(A.sky_off[0], d[0], d[1], d[2], p) && sky_rgb_d(A.sky_off[1], d[0], d[1], d[2], m)) {
    float acc = 0.0f;
    for (int c = 0; c < 3; ++c)
      acc += w[c] * beta[c] * static_cast<float>((p[c] - m[c]) / (2.0 * A.h_t)) * DF_PI_RECIP;
    lds_add(g, it, acc);
  }
  if (sky_rgb_d(A.sky_off[2], d[0], d[1], d[2], p) && sky_rgb_d(A.sky_off[3], d[0], d[1], d[2], m)) {
    float acc = 0.0f;
    for (int c = 0; c < 3; ++c)
      acc += w[c] * beta[c] * static_cast<float>((p[c] - m[c]) / (2.0 * A.h_e)) * DF_PI_RECIP;
    lds_add(g, it + 1, acc);
  }
}

// One path of render_pt (:3151-3705) replayed with the adjoint bookkeeping.
template <int M>
__device__ void replay_path(Tracer<M>& tr, double* g, const DAdjoint& A, const float ray[6], const float w[3],
                            bool sky_active, bool sun_active) {
  const int depth_max = tr.P.depth_max;
  int depth = 0;
  float o[3] = {ray[0], ray[1], ray[2]};
  float d[3] = {ray[3], ray[4], ray[5]};
  float m[3] = {1.0f, 1.0f, 1.0f};
  float p[3] = {0.0f, 0.0f, 0.0f};
  PathFactors pf;
  pf.n = 0;
  for (;;) {
    tr.template intersect_t<false>(o[0], o[1], o[2], d[0], d[1], d[2]);
    const float beta[3] = {m[0], m[1], m[2]};
    const float din[3] = {d[0], d[1], d[2]};
    Vtx v;
    v.sky_e[0] = &A.sky_off[2];
    v.sky_e[1] = &A.sky_off[3];
    v.h_e = A.h_e;
    v.miss = false;
    v.uses_albedo = false;
    v.aslot = v.eslot = -1;
    bool shadow;
    float sd[3], pv[3];
    const bool more = tr.pt_bounce(o, d, m, p, depth, sky_active, sun_active, shadow, sd, pv, &v);
    if (v.miss) {
      sky_term(tr, g, A, pf, w, beta, din, sky_active);
      return;
    }
    const bool lit = shadow && tr.sun_visible(tr.h.p[0], tr.h.p[1], tr.h.p[2], sd);
    // emission term beta * e: the light's own parameter, and the albedos in beta
    if (v.eslot >= 0) lds_add(g, 3 * A.n_albedo + v.eslot, w[0] * beta[0] + w[1] * beta[1] + w[2] * beta[2]);
    if (v.e[0] != 0.0f || v.e[1] != 0.0f || v.e[2] != 0.0f) through_beta(g, pf, w, v.e);
    // sun term beta * albedo * sun_w
    if (lit) {
      if (v.aslot >= 0)
        for (int c = 0; c < 3; ++c) lds_add(g, v.aslot * 3 + c, w[c] * beta[c] * v.sun_w);
      const float tail[3] = {v.a[0] * v.sun_w, v.a[1] * v.sun_w, v.a[2] * v.sun_w};
      through_beta(g, pf, w, tail);
      lds_add(g, 3 * A.n_albedo + A.n_emission + 1,
              (w[0] * beta[0] * v.a[0] + w[1] * beta[1] * v.a[1] + w[2] * beta[2] * v.a[2]) * v.dsun_w);
    }
    if (!more) return;
    if (pf.n < kMaxAdjointDepth) {
      const bool ua = v.uses_albedo && v.aslot >= 0;
#pragma unroll
      for (int j = 0; j < kMaxAdjointDepth; ++j) {
        if (j == pf.n) {
          pf.s[j] = v.s;
          pf.slot[j] = ua ? v.aslot : -1;
          for (int c = 0; c < 3; ++c) pf.f[j][c] = v.uses_albedo ? v.s * v.a[c] : v.s;
        }
      }
      ++pf.n;
    }
    if (depth >= depth_max) {  // the escape test after the last bounce (:3689-3700)
      if (!(tr.template intersect_t<true>(o[0], o[1], o[2], d[0], d[1], d[2]) < DF_INF)) {
        const float b2[3] = {m[0], m[1], m[2]};
        sky_term(tr, g, A, pf, w, b2, d, sky_active);
      }
      return;
    }
  }
}

// Occupancy hint: three waves per SIMD measured faster than the two an
// unconstrained allocation allows (and than four).
template <int M>
__global__ __launch_bounds__(256, 3) void adjoint_kernel(DScene S, DCamera C, DSky K, DParams P, DFrame F, DAdjoint A) {
  __shared__ int perm[512];
  __shared__ double g[kMaxDiffParams];
  const int tid = threadIdx.y * 16 + threadIdx.x;
  const int n_params = 3 * A.n_albedo + A.n_emission + 2;
  for (int i = tid; i < 512; i += 256) perm[i] = kPerm[i & 255];
  for (int i = tid; i < n_params; i += 256) g[i] = 0.0;
  __syncthreads();
  const int x = blockIdx.x * 16 + threadIdx.x;
  const int y = blockIdx.y * 16 + threadIdx.y;
  if (x < P.width && y < P.height) {
    const int gid = y * P.width + x;
    Tracer<M> tr(S, C, K, P, perm);
    tr.rng.s = F.seeds[gid];
    const float w[3] = {A.weight[gid * 3 + 0], A.weight[gid * 3 + 1], A.weight[gid * 3 + 2]};
    const bool sky_active = K.sky_active == 1;
    const bool sun_active = K.sun_active == 1;
    for (int pass = 0; pass < P.passes; ++pass) {
      float ray[6];
      if (tr.primary_ray(x, y, ray)) replay_path(tr, g, A, ray, w, sky_active, sun_active);
    }
    F.seeds[gid] = tr.rng.s;
  }
  __syncthreads();
  for (int i = tid; i < n_params; i += 256)
    if (g[i] != 0.0) atomicAdd(&A.grad[i], g[i]);
}

}  // namespace dfk

hipError_t dfk_launch_loss(const float* film, const float* target, float* weight, int n, float scale, double* loss,
                           hipStream_t stream) {
  const int blocks = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;
  hipLaunchKernelGGL(dfk::loss_kernel, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, stream, film, target, weight, n,
                     scale, loss);
  return hipGetLastError();
}

hipError_t dfk_launch_adjoint(const dfk::DScene& s, const dfk::DCamera& c, const dfk::DSky& k, const dfk::DParams& p,
                              const dfk::DFrame& f, const dfk::DAdjoint& a, int math_mode, hipStream_t stream) {
  const dim3 block(16, 16);
  const dim3 grid((p.width + 15) / 16, (p.height + 15) / 16);
  if (math_mode == 1)
    hipLaunchKernelGGL(dfk::adjoint_kernel<1>, grid, block, 0, stream, s, c, k, p, f, a);
  else
    hipLaunchKernelGGL(dfk::adjoint_kernel<0>, grid, block, 0, stream, s, c, k, p, f, a);
  return hipGetLastError();
}
