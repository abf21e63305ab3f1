// Adjoint (path-replay) kernels of the differentiable path tracer.
//
// Loss: L = sum over pixels and channels of (film - target)^2, film being the
// mean of `passes` path-traced samples.  Parameters: continuous albedo per
// surface group (RGB), emission per light (white radiance), the two colours
// of checkerboard albedo textures (RGB each), the sky's turbidity and the
// sun's elevation and azimuth.  Geometry is fixed, so visibility does not
// move; the gradient is the interior term.
//
// With Russian roulette off (its survival probability is the albedo's
// maximum, so it would make the path itself depend on the parameters), the
// directions and hits of a path do not depend on the parameters for a given
// RNG stream: a sample is a sum of terms beta_k * g_k, beta_k the product of
// the throughput factors of the vertices before k (s_j, times the albedo
// where the material multiplies it in) and g_k the emission, sun or sky
// term at k.  The adjoint kernel replays each pixel's paths from the seeds
// the forward pass started from, with the forward's own bounce code, and adds
// dL/dfilm / passes times the derivative of every term: of the albedo slot
// and emission it contains directly, of the earlier vertices' albedo slots
// through beta_k (a product over at most kMaxAdjointDepth factors,
// recomputed without division), and of the sky's parameters through the sky
// radiance and the sun sample's cosine factor.  Those last two are smooth
// closed forms of the sky model (Sky.set, the Perez evaluation, the sun cone
// sample); their derivatives come from dual numbers in double (Sky.set on the
// host, the evaluations here), with no finite differences.  Partial
// gradients are summed per workgroup in LDS (double) and added to the output
// once per workgroup.
#include <hip/hip_runtime.h>

#include "df_kernels.h"
#include "df_pass_common.h"

namespace dfk {

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

// The gradient entry of channel 0 of albedo slot `slot` (albedo groups come
// first, the checkerboard colours after the emissions).
__device__ __forceinline__ int albedo_row(const DAdjoint& A, int slot) {
  return 3 * slot + (slot >= A.n_albedo ? A.n_emission : 0);
}

// The entry of the sky's first parameter (turbidity), after every albedo
// slot and emission.
__device__ __forceinline__ int sky_row(const DAdjoint& A) { return 3 * A.n_albedo + A.n_emission + 6 * A.n_checker; }

// Per-path bookkeeping of the replay: the throughput factors of the vertices
// so far.  Loops run over the compile-time maximum so that the arrays stay in
// registers.
struct PathFactors {
  float f[kMaxAdjointDepth][3];  // throughput factor of vertex j
  float s[kMaxAdjointDepth];     // its parameter-free part (f = s * albedo when slot >= 0)
  int slot[kMaxAdjointDepth];    // albedo slot that multiplied the throughput, or -1
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
__device__ __forceinline__ void lds_add(double* g, int i, double v) {
  if (v != 0.0) atomicAdd(&g[i], v);
}

// Gradient of a term beta * tail (tail per channel) through the albedos in beta.
__device__ __forceinline__ void through_beta(double* g, const DAdjoint& A, const PathFactors& pf, const float w[3],
                                             const float tail[3]) {
#pragma unroll
  for (int j = 0; j < kMaxAdjointDepth; ++j) {
    if (j < pf.n && pf.slot[j] >= 0) {
      const int r = albedo_row(A, pf.slot[j]);
#pragma unroll
      for (int c = 0; c < 3; ++c) lds_add(g, r + c, w[c] * tail[c] * pf.excl(j, c));
    }
  }
}

// The sky term beta * sky(d) / pi of a path that escaped: through the albedos
// in beta, and through the sky's three parameters in the sky radiance.
template <int M>
__device__ __forceinline__ void sky_term(Tracer<M>& tr, double* g, const DAdjoint& A, const PathFactors& pf,
                                         const float w[3], const float beta[3], const float d[3], bool sky_active) {
  tr.sky(sky_active, d[0], d[1], d[2]);
  const float tail[3] = {tr.tmp[0] * DF_PI_RECIP, tr.tmp[1] * DF_PI_RECIP, tr.tmp[2] * DF_PI_RECIP};
  through_beta(g, A, pf, w, tail);
  if (!sky_active) return;
  Dual3 rgb[3];
  if (!sky_rgb_dual(*A.sky, d[0], d[1], d[2], rgb)) return;
  const int it = sky_row(A);
  for (int q = 0; q < 3; ++q) {
    double acc = 0.0;
    for (int c = 0; c < 3; ++c) acc += static_cast<double>(w[c] * beta[c]) * rgb[c].d[q];
    lds_add(g, it + q, acc * (1.0 / 3.141592653589793));
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
  // The sky term of a path whose ray escapes is its last contribution; it is
  // taken after the loop, with the escaping ray's throughput and direction,
  // where the wave's lanes are together again (the FP64 sky and its dual
  // numbers then run once, not once per bounce for the few lanes escaping
  // there).
  bool sky_due = false;
  float sky_beta[3] = {0.0f, 0.0f, 0.0f}, sky_d[3] = {0.0f, 0.0f, 0.0f};
  for (;;) {
    tr.template intersect_t<false>(o[0], o[1], o[2], d[0], d[1], d[2]);
    const float beta[3] = {m[0], m[1], m[2]};
    const float din[3] = {d[0], d[1], d[2]};
    Vtx v;
    v.sky = A.sky;
    v.miss = false;
    v.uses_albedo = false;
    v.aslot = v.eslot = -1;
    bool shadow;
    float sd[3], pv[3];
    const bool more = tr.pt_bounce(o, d, m, p, depth, sky_active, sun_active, shadow, sd, pv, &v);
    if (v.miss) {
      sky_due = true;
      for (int c = 0; c < 3; ++c) {
        sky_beta[c] = beta[c];
        sky_d[c] = din[c];
      }
      break;
    }
    const bool lit = shadow && tr.sun_visible(tr.h.p[0], tr.h.p[1], tr.h.p[2], sd);
    // emission term beta * e: the light's own parameter, and the albedos in beta
    if (v.eslot >= 0) lds_add(g, 3 * A.n_albedo + v.eslot, w[0] * beta[0] + w[1] * beta[1] + w[2] * beta[2]);
    if (v.e[0] != 0.0f || v.e[1] != 0.0f || v.e[2] != 0.0f) through_beta(g, A, pf, w, v.e);
    // sun term beta * albedo * sun_w
    if (lit) {
      if (v.aslot >= 0) {
        const int r = albedo_row(A, v.aslot);
        for (int c = 0; c < 3; ++c) lds_add(g, r + c, w[c] * beta[c] * v.sun_w);
      }
      const float tail[3] = {v.a[0] * v.sun_w, v.a[1] * v.sun_w, v.a[2] * v.sun_w};
      through_beta(g, A, pf, w, tail);
      const float wba = w[0] * beta[0] * v.a[0] + w[1] * beta[1] * v.a[1] + w[2] * beta[2] * v.a[2];
      lds_add(g, sky_row(A) + 1, wba * v.dsun_w[0]);
      lds_add(g, sky_row(A) + 2, wba * v.dsun_w[1]);
    }
    if (!more) break;
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
        sky_due = true;
        for (int c = 0; c < 3; ++c) {
          sky_beta[c] = m[c];
          sky_d[c] = d[c];
        }
      }
      break;
    }
  }
  if (sky_due) sky_term(tr, g, A, pf, w, sky_beta, sky_d, sky_active);
}

// Occupancy hint: three waves per SIMD measured faster than the two an
// unconstrained allocation allows (and than four).
template <int M>
__global__ __launch_bounds__(256, 3) void adjoint_kernel(DScene S, DCamera C, DSky K, DParams P, DFrame F, DAdjoint A) {
  __shared__ int perm[kPermLen];
  __shared__ double g[kMaxDiffParams];
  const int tid = threadIdx.y * 16 + threadIdx.x;
  const int n_params = sky_row(A) + 3;
  load_perm(perm, tid, 256);
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
