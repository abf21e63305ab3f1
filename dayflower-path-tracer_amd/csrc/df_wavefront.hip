// Wavefront path tracer for gfx950.
//
// The megakernel runs a whole path per work-item, so a wave holds lanes in
// every stage at once (BVH walk, terrain march, texture noise, sky) and the
// kernel's register footprint is that of the largest stage.  Here one pass
// of GPURendererKernel.run() (:3151-3705 for the path tracer) is split into
// kernels per stage, each over a queue of the paths still alive:
//
//   raygen           primary rays (camera, lens) for every pixel
//   per bounce b:    trace(b)     closest-hit queries of the live paths,
//                                 then any-hit queries of the previous
//                                 bounce's sun samples (resolving its
//                                 radiance)
//                    normals(b)   noise normal maps and albedo textures of
//                                 the paths whose hit surface has them
//                    shade(b)     hit record, emission, sun sample, Russian
//                                 roulette, scattered ray; appends the path
//                                 to the next bounce's queue and its sun
//                                 sample to the shadow queue
//   trace            the escape test (any-hit + sky) of the paths that took
//                    every bounce, then the last bounce's sun samples
//   film             film accumulation (adding an escaped path's sky term),
//                    wireframes and the image pass
//
// Every path draws from its pixel's LCG stream in the reference's order
// (raygen, then each shade), and the sun term is resolved from the two
// radiance values computed by shade, so a pass gives the megakernel's bits.
// The trace and shade kernels are persistent: each wave takes 64 queue
// entries at a time from a work head until the queue is drained, which
// balances the very uneven cost of incoherent rays.  Queues are striped over
// kQueueStripes lines and appended with one atomic per wave.  Path state is
// indexed by pixel and lives in DPaths.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "df_kernels.h"
#include "df_pass_common.h"

namespace dfk {

constexpr int32_t kFinish = 1 << 30;  // shadow-queue flag: the path ends once resolved
constexpr int kThreads = 256;

__device__ __forceinline__ void stage_perm(int* perm) {
  load_perm(perm, threadIdx.x + threadIdx.y * blockDim.x, blockDim.x * blockDim.y);
  __syncthreads();
}

// Global index of the calling wave (any block shape).
__device__ __forceinline__ int global_wave() {
  const int block = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const int per_block = (blockDim.x * blockDim.y * blockDim.z + 63) / 64;
  return block * per_block + (threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z)) / 64;
}

// Appends `value` for the lanes with `pred` set to the striped queue q (count
// group `counts`, stripe capacity `cap`): one atomic per wave, on the count of
// the wave's stripe, the lanes' entries in lane order.  Called by every active
// lane of the wave.
__device__ __forceinline__ void wave_append(int32_t* q, int64_t cap, int32_t* counts, bool pred, int32_t value) {
  const unsigned long long mask = __ballot(pred);
  if (mask == 0ull) return;
  const int lane = __lane_id();
  const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
  const int stripe = global_wave() % kQueueStripes;
  int base = 0;
  if (lane == leader) base = atomicAdd(counts + stripe * kLineInts, __popcll(mask));
  base = __shfl(base, leader, 64);
  if (pred) q[stripe * cap + base + __popcll(mask & ((1ull << lane) - 1ull))] = value;
}

// Appends `value` for the lanes with `pred` set to bin `key` of the binned
// queue q (bins of kQueueStripes stripes of `cap` entries; count groups of
// the bins kCtlGroup ints apart): per distinct key in the wave one ballot and
// one atomic, on the count of the wave's stripe of that bin.
__device__ __forceinline__ void wave_append_keyed(int32_t* q, int64_t cap, int32_t* counts, bool pred, int key,
                                                  int32_t value) {
  unsigned long long rest = __ballot(pred);
  const int lane = __lane_id();
  const int stripe = global_wave() % kQueueStripes;
  while (rest != 0ull) {
    const int leader = __ffsll(static_cast<unsigned long long>(rest)) - 1;
    const int k = __shfl(key, leader, 64);
    const unsigned long long mask = __ballot(pred && key == k);
    rest &= ~mask;
    int base = 0;
    if (lane == leader) base = atomicAdd(counts + static_cast<int64_t>(k) * kCtlGroup + stripe * kLineInts,
                                         __popcll(mask));
    base = __shfl(base, leader, 64);
    if (pred && key == k)
      q[(static_cast<int64_t>(k) * kQueueStripes + stripe) * cap + base + __popcll(mask & ((1ull << lane) - 1ull))] =
          value;
  }
}

// Runs body(valid, entry) over the striped queue q (count group `counts`,
// head group `heads`), 64 entries per wave at a time: a wave drains its own
// stripe first, then helps with the others (skipping, without an atomic, a
// stripe whose head has passed its end).  Every lane of a wave takes part in
// every call (valid is false past the end).
template <class Body>
__device__ __forceinline__ void drain(const int32_t* q, int64_t cap, const int32_t* counts, int32_t* heads,
                                      Body&& body) {
  const int first = global_wave() % kQueueStripes;
  for (int t = 0; t < kQueueStripes; ++t) {
    const int stripe = (first + t) % kQueueStripes;
    const int n = counts[stripe * kLineInts];
    const int chunks = (n + 63) / 64;
    int32_t* head = heads + stripe * kLineInts;
    if (t > 0 && __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= chunks) continue;
    const int32_t* qs = q + stripe * cap;
    for (;;) {
      int ch = 0;
      if (__lane_id() == 0) ch = atomicAdd(head, 1);
      ch = __shfl(ch, 0, 64);
      if (ch >= chunks) break;
      const int k = ch * 64 + __lane_id();
      const bool valid = k < n;
      body(valid, valid ? qs[k] : 0);
    }
  }
}

__device__ __forceinline__ float4 f4(const float v[3], float w) { return make_float4(v[0], v[1], v[2], w); }

__global__ void wf_reset(int32_t* ctl, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ctl[i] = 0;
}

template <int M>
__global__ __launch_bounds__(256) void wf_raygen(DScene S, DCamera C, DSky K, DParams P, DFrame F, DPaths W) {
  const int x = blockIdx.x * 16 + threadIdx.x;
  const int y = blockIdx.y * 16 + threadIdx.y;
  bool valid = false;
  int gid = 0;
  if (x < P.width && y < P.height) {
    gid = y * P.width + x;
    Tracer<M> tr(S, C, K, P, kPermTable.v);
    tr.rng.s = F.seeds[gid];
    float ray[6];
    valid = tr.primary_ray(x, y, ray);
    F.seeds[gid] = tr.rng.s;
    if (valid) {
      W.ray_o[gid] = make_float4(ray[0], ray[1], ray[2], 0.0f);
      W.ray_d[gid] = make_float4(ray[3], ray[4], ray[5], 0.0f);
      W.thru[gid] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
      W.rad[gid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (P.wireframes == 1) {
        W.ray0_o[gid] = W.ray_o[gid];
        W.ray0_d[gid] = W.ray_d[gid];
      }
    } else {
      W.sample[gid] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);  // filmSetColor(0)
    }
  }
  wave_append(W.queue[0], W.n, W.ctl, valid, gid);
}

// Occupancy hint: two waves per SIMD: shading waits on its gathers, and the
// registers let more of them be in flight per wave (shade stage 2.20 -> 2.14
// -> 2.13 ms per 1080p pass from four waves per SIMD to three to two, 2.42
// with five).  With the FP64 sky still inside it, four had measured faster
// than three.
//
// kBinned: the paths come in the trace kernel's shade bins (counts `count`
// and heads `head` are then the bins' groups, kCtlGroup ints apart), drained
// one bin after the other, so no wave mixes materials and the material switch
// branches the same way on every lane; else they come in the bounce's live
// queue q.  (Bodies specialised per bin at compile time measured slower: the
// seven copies made the kernel 93k instructions against 18k, past what the
// instruction cache holds.)
template <int M, bool kBinned>
__global__ __launch_bounds__(kThreads, 2) void wf_shade(DScene S, DCamera C, DSky K, DParams P, DFrame F, DPaths W,
                                                     const int32_t* q, const int32_t* count, int32_t* head,
                                                     int32_t* q_next, int32_t* count_next, int32_t* count_shadow,
                                                     int32_t* count_sky, int bounce) {
  __shared__ int perm[kPermLen];
  stage_perm(perm);
  Tracer<M> tr(S, C, K, P, perm);
  tr.noise_in = W.noise;  // computed by wf_normals
  const bool sky_active = K.sky_active == 1;
  const bool sun_active = K.sun_active == 1;
  auto shade = [&](int bin, bool valid, int32_t i) {
    bool more = false, shadow = false, sky = false;
    if (valid) {
      tr.noise_index = i;
      const float4 o4 = W.ray_o[i], d4 = W.ray_d[i], hv = W.hit[i], m4 = W.thru[i], p4 = W.rad[i];
      float o[3] = {o4.x, o4.y, o4.z}, d[3] = {d4.x, d4.y, d4.z};
      float m[3] = {m4.x, m4.y, m4.z}, p[3] = {p4.x, p4.y, p4.z};
      tr.fill_hit(o[0], o[1], o[2], d[0], d[1], d[2], hv.x, __float_as_int(hv.y), __float_as_int(hv.z),
                  __float_as_int(hv.w));
      tr.rng.s = F.seeds[i];
      int depth = bounce;
      float sd[3], pv[3];
      tr.sky_pending = false;
      more = tr.template pt_bounce<true>(o, d, m, p, depth, sky_active, sun_active, shadow, sd, pv, nullptr, bin);
      sky = tr.sky_pending;
      F.seeds[i] = tr.rng.s;
      W.rad[i] = f4(p, 0.0f);
      if (shadow) {
        W.sh_o[i] = f4(tr.h.p, 0.0f);
        W.sh_d[i] = f4(sd, 0.0f);
        W.rad_lit[i] = f4(pv, 0.0f);
      }
      if (more) {
        W.ray_o[i] = f4(o, 0.0f);
        W.ray_d[i] = f4(d, 0.0f);
        W.thru[i] = f4(m, 0.0f);
      } else if (!shadow && !sky) {
        W.sample[i] = f4(p, 0.0f);
      }
    }
    wave_append(q_next, W.n, count_next, valid && more, i);
    wave_append(W.shadow_queue, W.n, count_shadow, valid && shadow, more ? i : (i | kFinish));
    wave_append(W.sky_queue, W.n, count_sky, valid && sky, i);
  };
  if constexpr (kBinned) {
    for (int b = 0; b < kShadeBins; ++b) {
      // bin 6 (materials outside 1..5) reads the material at run time, as without bins
      const int as = b < 6 ? b : -1;
      drain(W.shade_bins + static_cast<int64_t>(b) * kQueueStripes * W.n, W.n, count + b * kCtlGroup,
            head + b * kCtlGroup, [&](bool valid, int32_t i) { shade(as, valid, i); });
    }
  } else {
    drain(q, W.n, count, head, [&](bool valid, int32_t i) { shade(-1, valid, i); });
  }
  unsigned long long cnt[kCounterCount] = {0, 0, 0, 0, 0, 0, 0, 0};
  take_counters(tr, cnt);
  flush_counters(cnt, F.counters, kStageShade);
}

// Noise normal maps (:2658-2712) of the paths the trace kernel queued,
// computed apart from shading: three 16-octave fBm evaluations at the hit
// point, the most expensive step of shading a hit, needed by the paths that
// hit such a surface.  Inside the shade kernel a wave ran them for those few
// lanes while the others waited (1.2 of its 3.2 ms per 1080p pass); here
// every lane has one.  The shade kernel takes the values from W.noise at
// the same step of its arithmetic, so its results are unchanged.
template <int M>
__global__ __launch_bounds__(kThreads) void wf_normals(DScene S, DCamera C, DSky K, DParams P, DPaths W,
                                                       const int32_t* count, int32_t* head) {
  __shared__ int perm[kPermLen];
  stage_perm(perm);
  Tracer<M> tr(S, C, K, P, perm);
  tr.noise_skip = true;  // fill_hit's own noise step: computed below instead
  drain(W.noise_queue, W.n, count, head, [&](bool valid, int32_t i) {
    if (!valid) return;
    const float4 o = W.ray_o[i], d = W.ray_d[i], hv = W.hit[i];
    const int ci = __float_as_int(hv.y);
    tr.fill_hit(o.x, o.y, o.z, d.x, d.y, d.z, hv.x, ci, __float_as_int(hv.z), __float_as_int(hv.w));
    const int so = S.prims[ci].surface_offset;
    float n[3];
    tr.noise_normal_values(S.surfaces[so + 5], n);
    // the hit record as the shade kernel's fill_hit leaves it, then the
    // albedo texture shade would evaluate on it
    tr.apply_noise_normal(S.surfaces[so + 4], n);
    const int alb = tr.tex_color(jf2i(S.surfaces[so + 1]));
    W.noise[i] = make_float4(n[0], n[1], n[2], __int_as_float(alb));
  });
}

// The ray queries of one bounce in one launch: first the closest-hit queries
// of the live paths - or, after the last bounce, their escape tests (any-hit
// + sky, :3689-3700) - then the sun's shadow queries of the previous bounce
// (any-hit; each resolves that bounce's radiance, and finishes the path if it
// ended there).  Both queues drain through the same persistent waves, so a
// short queue does not leave the GPU idle and a bounce pays one launch tail;
// the cheaper any-hit queries go last, where they fill that tail (measured
// 10.36 -> 10.07 ms per 1080p pass against the other order).  The two queues
// of one launch never touch the same path data: a shadow query writes the
// radiance (and the sample of a path that ended), a closest-hit query the hit
// record, an escape test the sample of a path that took every bounce.
// kWaves: the occupancy hint, waves per SIMD.  Seven (72 registers, some
// spilled) for launches whose queues fill the GPU: they are bound by VALU
// issue and the waves hide each other's latency (measured faster than five,
// six or eight).  Four (128 registers, none spilled) for the launches of the
// last bounces and the escape tests, whose queues leave most wave slots empty:
// the launch lasts as long as its slowest waves, and a wave without spill
// reloads in its walks finishes sooner (1080p pass: trace 7.27 -> 7.10 ms
// with four from bounce 4 or 5 on, 7.16 from bounce 3 or 6, three waves no
// faster; escape 0.281 -> 0.265 ms against five).
// kUnits: with the unit counters (Tracer).  kPairs: the mesh walks over the
// two-child records, each lane's stack (S.pair_stack entries of a key and a
// child) in the launch's dynamic LDS: the keys of the workgroup's lanes, then
// their children.
template <int M, bool kEscape, bool kUnits, bool kPairs, int kWaves>
__global__ __launch_bounds__(kThreads, kWaves) void wf_trace(DScene S, DCamera C, DSky K, DParams P, DFrame F, DPaths W,
                                                     const int32_t* shadow_count, int32_t* shadow_head,
                                                     const int32_t* q, const int32_t* count, int32_t* head,
                                                     int first_bounce, int32_t* noise_count, int32_t* bin_count,
                                                     int32_t* sky_count) {
  typedef Tracer<M, kUnits, kPairs> T;
  T tr(S, C, K, P, kPermTable.v);
  if constexpr (kPairs) {
    static_assert(T::kPairStride == kThreads, "one stack slot per lane");
    extern __shared__ int pair_stack[];
    tr.pstack_key = (typename T::lds_f*)(reinterpret_cast<float*>(pair_stack) + threadIdx.x);
    tr.pstack_ref = (typename T::lds_i*)(pair_stack + S.pair_stack * kThreads + threadIdx.x);
  }
  drain(q, W.n, count, head, [&](bool valid, int32_t i) {
    bool noisy = false, escaped = false;
    int bin = 0;
    if (valid) {
      const float4 o4 = W.ray_o[i], d4 = W.ray_d[i];
      if constexpr (kEscape) {
        // The escape test (:3689-3700) runs beside the last bounce's sun queries,
        // which may still change this path's radiance: a path whose ray escapes
        // goes to the sky pass, which stores its sky term m * sky / pi as the
        // sample (w = 2); a path whose ray hits something keeps its radiance
        // (w = 3); the film pass adds the final radiance, the same single float
        // addition the reference makes.
        escaped = !(tr.template intersect_t<true>(o4.x, o4.y, o4.z, d4.x, d4.y, d4.z) < DF_INF);
        if (!escaped) W.sample[i] = make_float4(0.0f, 0.0f, 0.0f, 3.0f);
      } else {
        int ci, cst, cso;
        const float t = tr.template query<false>(o4.x, o4.y, o4.z, d4.x, d4.y, d4.z, ci, cst, cso);
        W.hit[i] = make_float4(t, __int_as_float(ci), __int_as_float(cst), __int_as_float(cso));
        if (first_bounce) F.primitive_offsets[i] = ci == -1 ? -1 : ci * 3;
        noisy = tr.has_noise_normals(ci);
        if (P.shade_binned) bin = tr.shade_bin(ci);
      }
    }
    if constexpr (!kEscape) {
      wave_append(W.noise_queue, W.n, noise_count, noisy, i);
      if (P.shade_binned) wave_append_keyed(W.shade_bins, W.n, bin_count, valid, bin, i);
    } else {
      wave_append(W.sky_queue, W.n, sky_count, escaped, i);
    }
  });
  drain(W.shadow_queue, W.n, shadow_count, shadow_head, [&](bool valid, int32_t e) {
    if (!valid) return;
    const int32_t i = e & (kFinish - 1);
    const float4 o = W.sh_o[i], d = W.sh_d[i];
    const float sd[3] = {d.x, d.y, d.z};
    const bool lit = tr.sun_visible(o.x, o.y, o.z, sd);
    float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (lit) {
      p = W.rad_lit[i];
      W.rad[i] = p;
    }
    if (e & kFinish) {
      if (!lit) p = W.rad[i];
      W.sample[i] = make_float4(p.x, p.y, p.z, 0.0f);
    }
  });
  unsigned long long cnt[kCounterCount] = {0, 0, 0, 0, 0, 0, 0, 0};
  take_counters(tr, cnt);
  flush_counters(cnt, F.counters, kEscape ? kStageEscape : kStageTrace);
}

// The sky term of the pass's escaped paths (:2127-2236, the miss branch of
// the path loop and the escape test :3689-3700): m * sky(d) / pi of the ray
// that escaped, stored as the path's sample (w = 2) for the film pass to add
// to its radiance.  The shade and escape launches queue these paths instead
// of evaluating the FP64 sky for a few lanes of a wave each (and carrying its
// registers): here every lane has one.
template <int M>
__global__ __launch_bounds__(kThreads) void wf_sky(DScene S, DCamera C, DSky K, DParams P, DFrame F, DPaths W,
                                                   const int32_t* count, int32_t* head) {
  Tracer<M> tr(S, C, K, P, kPermTable.v);
  const bool sky_active = K.sky_active == 1;
  drain(W.sky_queue, W.n, count, head, [&](bool valid, int32_t i) {
    if (!valid) return;
    const float4 d4 = W.ray_d[i], m4 = W.thru[i];
    tr.sky(sky_active, d4.x, d4.y, d4.z);
    W.sample[i] = make_float4(m4.x * tr.tmp[0] * DF_PI_RECIP, m4.y * tr.tmp[1] * DF_PI_RECIP,
                              m4.z * tr.tmp[2] * DF_PI_RECIP, 2.0f);
  });
  unsigned long long cnt[kCounterCount] = {0, 0, 0, 0, 0, 0, 0, 0};
  take_counters(tr, cnt);
  flush_counters(cnt, F.counters, kStageShade);
}

// filmAddColor / filmSetColor, the wireframe overlay and, after the last pass,
// the image pass (run() :263-307).
template <int M>
__global__ __launch_bounds__(256) void wf_film(DScene S, DCamera C, DSky K, DParams P, DFrame F, DPaths W, int clear,
                                               int last) {
  __shared__ int perm[kPermLen];
  if (P.wireframes == 1) stage_perm(perm);
  const int x = blockIdx.x * 16 + threadIdx.x;
  const int y = blockIdx.y * 16 + threadIdx.y;
  unsigned long long cnt[kCounterCount] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (x < P.width && y < P.height) {
    const int gid = y * P.width + x;
    Tracer<M> tr(S, C, K, P, perm);
    float4 col = W.sample[gid];
    if (col.w >= 2.0f) {  // a path that took every bounce: its radiance, plus the sky term if it escaped
      const float4 r = W.rad[gid];
      col = col.w == 2.0f ? make_float4(r.x + col.x, r.y + col.y, r.z + col.z, 0.0f)
                          : make_float4(r.x, r.y, r.z, 0.0f);
    }
    const bool set = col.w != 0.0f;
    float film[3] = {F.film[gid * 3 + 0], F.film[gid * 3 + 1], F.film[gid * 3 + 2]};
    int nsamp = F.film_samples[gid];
    ++tr.samples;
    if (set || clear) {
      film[0] = col.x;
      film[1] = col.y;
      film[2] = col.z;
      nsamp = 1;
    } else {
      const int n = nsamp + 1;
      film[0] = film[0] + ((col.x - film[0]) / static_cast<float>(n));
      film[1] = film[1] + ((col.y - film[1]) / static_cast<float>(n));
      film[2] = film[2] + ((col.z - film[2]) / static_cast<float>(n));
      nsamp = n;
    }
    if (!set && P.wireframes == 1) {
      const float4 o = W.ray0_o[gid], d = W.ray0_d[gid];
      const float ray[6] = {o.x, o.y, o.z, d.x, d.y, d.z};
      if (tr.wireframe_edge(ray)) {
        ++tr.samples;
        film[0] = film[1] = film[2] = 0.0f;
        nsamp = 1;
      }
    }
    F.film[gid * 3 + 0] = film[0];
    F.film[gid * 3 + 1] = film[1];
    F.film[gid * 3 + 2] = film[2];
    F.film_samples[gid] = nsamp;
    if (last) image_pass<M>(P, film, F.primitive_offsets[gid], F.image_float + gid * 3, F.image_bgra + gid * 4);
    take_counters(tr, cnt);
  }
  flush_counters(cnt, F.counters, kStageFilm);
}

// Resident workgroups of a persistent kernel: occupancy x compute units.
template <class Kernel>
int resident_blocks(Kernel k, size_t dynamic_lds = 0) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1024;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1024;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kThreads, dynamic_lds) != hipSuccess || per_cu < 1)
    per_cu = 1;
  return cus * per_cu;
}

// Persistent grid sizes of the queue kernels, queried once per math mode (before
// any graph capture: the occupancy queries are not stream work).
struct Grids {
  int trace, shade, shade_binned, escape, normals;
  int trace_nu, escape_nu;      // the trace kernels without unit counters
  int trace_tail, trace_tail_nu;  // the last bounces' trace kernels (four waves per SIMD)
  int sky;
  // the two-child walk's trace kernels (with, without unit counters) for a
  // per-lane stack of s entries: [s]
  int trace_pairs[kPairStackMax + 1], trace_pairs_nu[kPairStackMax + 1];
};

// Bytes of dynamic LDS the two-child walk's trace kernels take for a per-lane
// stack of s entries.
inline size_t pair_stack_bytes(int s) { return static_cast<size_t>(s) * 2 * kThreads * sizeof(int32_t); }

template <int M>
const Grids& grids() {
  static const Grids g = [] {
    Grids r{resident_blocks(wf_trace<M, false, true, false, 7>),  resident_blocks(wf_shade<M, false>),
            resident_blocks(wf_shade<M, true>),                   resident_blocks(wf_trace<M, true, true, false, 4>),
            resident_blocks(wf_normals<M>),                       resident_blocks(wf_trace<M, false, false, false, 7>),
            resident_blocks(wf_trace<M, true, false, false, 4>),  resident_blocks(wf_trace<M, false, true, false, 4>),
            resident_blocks(wf_trace<M, false, false, false, 4>), resident_blocks(wf_sky<M>), {}, {}};
    for (int s = 1; s <= kPairStackMax; ++s) {
      r.trace_pairs[s] = resident_blocks(wf_trace<M, false, true, true, 7>, pair_stack_bytes(s));
      r.trace_pairs_nu[s] = resident_blocks(wf_trace<M, false, false, true, 7>, pair_stack_bytes(s));
    }
    return r;
  }();
  return g;
}

// One pass of the wavefront path tracer: every launch of one sample per pixel,
// in stream order.  `clear`: the film pass overwrites (the clear flag);
// `last`: the film pass ends with the image pass.
template <int M>
hipError_t launch_wavefront_pass(const DScene& s, const DCamera& c, const DSky& k, const DParams& p, const DFrame& f,
                                 const DPaths& w, hipStream_t stream, const dfk_marks* marks, int clear, int last) {
  const Grids& g = grids<M>();
  const bool units = p.count_units != 0;
  const bool pairs = p.bvh_pairs != 0 && s.pairs != nullptr && s.pair_stack >= 1 && s.pair_stack <= kPairStackMax;
  const size_t trace_lds = pairs ? pair_stack_bytes(s.pair_stack) : 0;
  const int g_trace = pairs ? (units ? g.trace_pairs[s.pair_stack] : g.trace_pairs_nu[s.pair_stack])
                            : (units ? g.trace : g.trace_nu);
  const int g_tail = units ? g.trace_tail : g.trace_tail_nu;
  const int g_shade = g.shade, g_shade_binned = g.shade_binned, g_escape = units ? g.escape : g.escape_nu,
            g_normals = g.normals;
  const auto trace_kernel = pairs ? (units ? wf_trace<M, false, true, true, 7> : wf_trace<M, false, false, true, 7>)
                                  : (units ? wf_trace<M, false, true, false, 7> : wf_trace<M, false, false, false, 7>);
  // the last bounces (from p.tail_bounce on; the two-child walk keeps its kernel)
  const auto tail_kernel = units ? wf_trace<M, false, true, false, 4> : wf_trace<M, false, false, false, 4>;
  const auto escape_kernel = units ? wf_trace<M, true, true, false, 4> : wf_trace<M, true, false, false, 4>;
  const int bounces = p.depth_max > 1 ? p.depth_max : 1;
  const int ctl_len = wavefront_ctl_len(p.depth_max);
  if (ctl_len > w.ctl_len || static_cast<int64_t>(p.width) * p.height > w.n) return hipErrorInvalidValue;
  // ctl: [0, B] live-path counts per bounce (B = escape queue), [B+1, 2B]
  // shadow counts (plus one that stays zero), then the work heads; each
  // count and head a group of kCtlGroup ints (one line per queue stripe).
  int32_t* ctl = w.ctl;
  auto group = [&](int g) { return ctl + static_cast<size_t>(g) * kCtlGroup; };
  auto live = [&](int b) { return group(b); };
  auto shadow_count = [&](int b) { return group(bounces + 1 + b); };
  auto head = [&](int h) { return group(2 * bounces + 2 + h); };
  auto noise_count = [&](int b) { return group(5 * bounces + 6 + 2 * b); };
  auto noise_head = [&](int b) { return group(5 * bounces + 7 + 2 * b); };
  // bounce b's shade bins: kShadeBins count groups, then kShadeBins head groups
  auto bin_count = [&](int b) { return group(7 * bounces + 6 + 2 * kShadeBins * b); };
  auto bin_head = [&](int b) { return group(7 * bounces + 6 + 2 * kShadeBins * b + kShadeBins); };
  int32_t* sky_count = group(7 * bounces + 6 + 2 * kShadeBins * bounces);
  int32_t* sky_head = group(7 * bounces + 7 + 2 * kShadeBins * bounces);
  const dim3 pix_block(16, 16);
  const dim3 pix_grid((p.width + 15) / 16, (p.height + 15) / 16);
  auto mark = [&](int stage, int begin) {
    if (marks) marks->mark(marks->user, stage, begin);
  };
  {
    hipLaunchKernelGGL(wf_reset, dim3((ctl_len + 255) / 256), dim3(256), 0, stream, ctl, ctl_len);
    mark(kStageRaygen, 1);
    hipLaunchKernelGGL(wf_raygen<M>, pix_grid, pix_block, 0, stream, s, c, k, p, f, w);
    mark(kStageRaygen, 0);
    for (int b = 0; b < bounces; ++b) {
      // shadow queries of bounce b - 1 (none before the first bounce: that count stays 0)
      int32_t* sc = b > 0 ? shadow_count(b - 1) : shadow_count(bounces);
      const bool tail = !pairs && b >= p.tail_bounce;
      mark(kStageTrace, 1);
      hipLaunchKernelGGL(tail ? tail_kernel : trace_kernel, dim3(tail ? g_tail : g_trace), dim3(kThreads),
                         tail ? 0 : trace_lds, stream, s, c, k, p, f, w, sc, head(3 * b), w.queue[b & 1], live(b),
                         head(3 * b + 1), b == 0 ? 1 : 0, noise_count(b), bin_count(b), nullptr);
      mark(kStageTrace, 0);
      mark(kStageShade, 1);  // the noise normal maps count as shading
      hipLaunchKernelGGL(wf_normals<M>, dim3(g_normals), dim3(kThreads), 0, stream, s, c, k, p, w, noise_count(b),
                         noise_head(b));
      if (p.shade_binned)
        hipLaunchKernelGGL((wf_shade<M, true>), dim3(g_shade_binned), dim3(kThreads), 0, stream, s, c, k, p, f, w,
                           w.queue[b & 1], bin_count(b), bin_head(b), w.queue[(b + 1) & 1], live(b + 1),
                           shadow_count(b), sky_count, b);
      else
        hipLaunchKernelGGL((wf_shade<M, false>), dim3(g_shade), dim3(kThreads), 0, stream, s, c, k, p, f, w,
                           w.queue[b & 1], live(b), head(3 * b + 2), w.queue[(b + 1) & 1], live(b + 1),
                           shadow_count(b), sky_count, b);
      mark(kStageShade, 0);
    }
    // the last bounce's sun queries and the escape tests of the paths that
    // took every bounce, in one launch (they write different path data; the
    // film pass adds the two)
    mark(kStageEscape, 1);
    hipLaunchKernelGGL(escape_kernel, dim3(g_escape), dim3(kThreads), 0, stream, s, c, k, p, f, w,
                       shadow_count(bounces - 1), head(3 * bounces), w.queue[bounces & 1], live(bounces),
                       head(3 * bounces + 1), 0, nullptr, nullptr, sky_count);
    mark(kStageEscape, 0);
    mark(kStageShade, 1);  // the sky terms count as shading
    hipLaunchKernelGGL(wf_sky<M>, dim3(g.sky), dim3(kThreads), 0, stream, s, c, k, p, f, w, sky_count, sky_head);
    mark(kStageShade, 0);
    mark(kStageFilm, 1);
    hipLaunchKernelGGL(wf_film<M>, pix_grid, pix_block, 0, stream, s, c, k, p, f, w, clear, last);
    mark(kStageFilm, 0);
  }
  return hipGetLastError();
}

template <int M>
hipError_t launch_wavefront(const DScene& s, const DCamera& c, const DSky& k, const DParams& p, const DFrame& f,
                            const DPaths& w, hipStream_t stream, const dfk_marks* marks) {
  for (int pass = 0; pass < p.passes; ++pass) {
    const hipError_t e = launch_wavefront_pass<M>(s, c, k, p, f, w, stream, marks, pass == 0 && p.clear_first ? 1 : 0,
                                                  pass == p.passes - 1 ? 1 : 0);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace dfk

hipError_t dfk_launch_pt_wavefront(const dfk::DScene& s, const dfk::DCamera& c, const dfk::DSky& k,
                                   const dfk::DParams& p, const dfk::DFrame& f, const dfk::DPaths& w, int math_mode,
                                   hipStream_t stream, const dfk_marks* marks) {
  return math_mode == 1 ? dfk::launch_wavefront<1>(s, c, k, p, f, w, stream, marks)
                        : dfk::launch_wavefront<0>(s, c, k, p, f, w, stream, marks);
}

void dfk_prepare_wavefront() {
  (void)dfk::grids<0>();
  (void)dfk::grids<1>();
}

hipError_t dfk_launch_pt_wavefront_pass(const dfk::DScene& s, const dfk::DCamera& c, const dfk::DSky& k,
                                        const dfk::DParams& p, const dfk::DFrame& f, const dfk::DPaths& w,
                                        int math_mode, hipStream_t stream, int clear, int last) {
  return math_mode == 1 ? dfk::launch_wavefront_pass<1>(s, c, k, p, f, w, stream, nullptr, clear, last)
                        : dfk::launch_wavefront_pass<0>(s, c, k, p, f, w, stream, nullptr, clear, last);
}
