// Host runtime of libdayflower_hip.so: the C ABI of include/dayflower_hip.h.
//
// Owns the device copies of the reference's flat arrays, validates every
// offset they contain before a kernel may dereference it, builds the
// device-private repack (DPrim / DNode / DTri, df_device_scene.h) and
// launches the kernels on one HIP stream per handle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/dayflower_hip.h"
#include "df_device_scene.h"
#include "df_dual.h"
#include "df_kernels.h"

using dfk::DBatchItem;
using dfk::DCamera;
using dfk::DCull;
using dfk::DFrame;
using dfk::DJob;
using dfk::DNode;
using dfk::DPair;
using dfk::kPairStackMax;
using dfk::DParams;
using dfk::DPrim;
using dfk::DScene;
using dfk::DSky;
using dfk::DTri;

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

// The path tracer's schedules the autotuner times for DF_PT_AUTO (every one
// gives the same bits): the megakernel with strip or square waves, the
// wavefront with one shade queue or with material bins.
constexpr int kTuneCandidates = 4;
constexpr df_schedule kCandidates[kTuneCandidates] = {
    {DF_PT_MEGAKERNEL, DF_TILE_ROWS, 0, 0, DF_TAIL_BOUNCE_DEFAULT},
    {DF_PT_MEGAKERNEL, DF_TILE_SQUARE, 0, 0, DF_TAIL_BOUNCE_DEFAULT},
    {DF_PT_WAVEFRONT, DF_TILE_ROWS, 0, 0, DF_TAIL_BOUNCE_DEFAULT},
    {DF_PT_WAVEFRONT, DF_TILE_ROWS, 1, 0, DF_TAIL_BOUNCE_DEFAULT},
};

// One launch sequence captured from a stream and instantiated, with the bytes
// of the by-value kernel arguments it was captured with: it is replayed while
// they are unchanged and captured again when they change.
struct GraphCache {
  hipGraphExec_t exec = nullptr;
  std::vector<unsigned char> key;
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    key.clear();
  }
};

// Host copy of the arrays (kept for re-validation and re-packing after
// transform / surface updates).
struct HostScene {
  std::vector<float> point2fs, point3fs, vector3fs, spheres, terrains, textures, surfaces, o2w, w2o;
  std::vector<int32_t> planes, triangles, bvh, primitives;
};

}  // namespace

struct df_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed = false;
  std::string err;

  HostScene hs;
  bool has_scene = false;
  std::vector<DevBuf> scene_bufs;
  DScene ds{};
  float camera[19] = {0};
  df_sky sky{};
  bool has_sky = false;
  df_params params{};

  int width = 0, height = 0;
  DevBuf film, film_samples, image_float, image_bgra, prim_offsets, seeds, counters;
  bool clear_pending = false;

  // wavefront path state (allocated on first use for the frame size); the
  // shade bins only once a binned schedule runs
  DevBuf path_f4[12], path_q[5], path_ctl, path_bins;
  dfk::DPaths paths{};

  // frame-loop graphs: one wavefront pass per (clear, last) flag pair, the
  // megakernel launch (df_set_graphs, default off: replaying measured 1-7 %
  // slower per pass than launching eagerly; off while stage timing is on)
  bool graphs = false;
  GraphCache g_wave[2][2], g_mega;

  // the wavefront trace kernels count units (df_set_unit_counters)
  bool count_units = true;

  // path-tracer schedule: fixed by df_set_schedule (sched_fixed), else for
  // DF_PT_AUTO the autotuner's choice once `tuned`, else the default
  // schedule of params.pt_engine (current_schedule)
  bool sched_fixed = false;
  bool tuned = false;
  df_schedule sched{};
  int tune_runs[kTuneCandidates] = {};     // passes timed per candidate
  float tune_ms[kTuneCandidates] = {};     // best time of one pass
  hipEvent_t tune_ev[2] = {nullptr, nullptr};

  // differentiable mode
  bool diff_on = false;
  int32_t diff_n_albedo = 0, diff_n_emission = 0, diff_n_checker = 0;
  double diff_sky_params[3] = {0.0, 0.0, 0.0};  // turbidity, sun elevation, sun azimuth
  std::vector<int32_t> diff_checker_host;       // per surface record: checkerboard group or -1
  DevBuf diff_surf_albedo, diff_surf_emission, diff_surf_checker, diff_albedo, diff_emission, diff_sky_dual,
      diff_weight, diff_grad, diff_loss, diff_seeds;
  size_t diff_n_surfaces = 0;

  // per-stage launch timing: event pairs recorded around each launch
  bool stage_timing = false;
  hipError_t mark_error = hipSuccess;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  struct Span {
    int stage;
    size_t begin, end;
  };
  std::vector<Span> spans;
  size_t open_begin = 0;
  double stage_ms[DF_STAGE_COUNT] = {};
  int64_t stage_launches[DF_STAGE_COUNT] = {};
};

namespace {

constexpr size_t kCounterBytes = sizeof(df_counters) * DF_STAGE_COUNT * dfk::kCounterSlots;


int fail(df_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

template <class T>
void key_add(std::vector<unsigned char>& k, const T& v) {
  const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
  k.insert(k.end(), p, p + sizeof(T));
}

// Enqueues `enqueue`'s launches on `stream` through the graph cache `g`:
// replays the instantiated graph when `key` equals the one it was captured
// with, else captures `enqueue` into a new graph and replays that.  No
// memset or allocation may be enqueued (the launch wrappers enqueue kernels
// only).  Returns the first HIP error.
template <class F>
hipError_t graph_launch(GraphCache& g, const std::vector<unsigned char>& key, hipStream_t stream, F&& enqueue) {
  if (!g.exec || g.key != key) {
    g.reset();
    hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    const hipError_t le = enqueue();
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(stream, &graph);
    if (le != hipSuccess || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return le != hipSuccess ? le : e;
    }
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      g.exec = nullptr;
      return e;
    }
    g.key = key;
  }
  return hipGraphLaunch(g.exec, stream);
}

int hip_fail(df_ctx* c, hipError_t e, const char* what) {
  return fail(c, DF_EHIP, "%s: %s", what, hipGetErrorString(e));
}

#define DF_HIP(ctx, call)                                   \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return hip_fail((ctx), e_, #call); \
  } while (0)

void free_buf(DevBuf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

int alloc_buf(df_ctx* c, DevBuf& b, size_t bytes) {
  free_buf(b);
  if (bytes == 0) bytes = 4;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    return fail(c, DF_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  b.bytes = bytes;
  return DF_OK;
}

template <typename T>
int upload(df_ctx* c, DevBuf& b, const std::vector<T>& v) {
  int rc = alloc_buf(c, b, v.size() * sizeof(T));
  if (rc) return rc;
  if (!v.empty()) DF_HIP(c, hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return DF_OK;
}

// Java (int) cast for offsets stored as floats.
inline int32_t jf2i(float v) {
  if (v != v) return 0;
  if (v >= 2147483648.0f) return INT32_MAX;
  if (v <= -2147483648.0f) return INT32_MIN;
  return static_cast<int32_t>(v);
}

// Whether surface record i's albedo texture is a checkerboard texture.
bool checker_albedo(const std::vector<float>& tex, const float* surfaces, size_t i) {
  const int32_t off = jf2i(surfaces[i * 6 + 1]);
  return off >= 0 && static_cast<size_t>(off) + 8 <= tex.size() && jf2i(tex[off]) == 3;
}

inline bool in_range(int64_t off, int64_t count, size_t len) { return off >= 0 && off + count <= static_cast<int64_t>(len); }

// ------------------------------------------------------------ validation

int check_texture(df_ctx* c, const HostScene& h, int32_t off, bool allow_blend) {
  if (!in_range(off, 2, h.textures.size())) return fail(c, DF_EINVAL, "texture offset %d out of range", off);
  const int type = jf2i(h.textures[off]);
  static const int kSizes[9] = {0, 5, 4, 8, 3, 7, 8, 3, 2};
  if (type < 1 || type > 8) return DF_OK;  // the kernel returns colour 0 for unknown types
  int64_t size = kSizes[type];
  if (type == 6) {
    if (!in_range(off, 8, h.textures.size())) return fail(c, DF_EINVAL, "image texture at %d truncated", off);
    const float w = h.textures[off + 4], hh = h.textures[off + 5];
    if (!(w >= 1.0f && hh >= 1.0f && w <= 65536.0f && hh <= 65536.0f))
      return fail(c, DF_EINVAL, "image texture at %d has bad size", off);
    size = 8 + static_cast<int64_t>(jf2i(w * hh));
  }
  if (!in_range(off, size, h.textures.size())) return fail(c, DF_EINVAL, "texture at %d truncated", off);
  if (type == 1) {
    if (!allow_blend) return DF_OK;  // nested blends are colour 0 in the reference
    int rc = check_texture(c, h, jf2i(h.textures[off + 2]), false);
    if (rc) return rc;
    rc = check_texture(c, h, jf2i(h.textures[off + 3]), false);
    if (rc) return rc;
  }
  return DF_OK;
}

int check_surface(df_ctx* c, const HostScene& h, int32_t so) {
  if (!in_range(so, 6, h.surfaces.size())) return fail(c, DF_EINVAL, "surface offset %d out of range", so);
  for (int k = 1; k <= 3; ++k) {
    int rc = check_texture(c, h, jf2i(h.surfaces[so + k]), true);
    if (rc) return rc;
  }
  return DF_OK;
}

int check_triangle(df_ctx* c, const HostScene& h, int32_t to) {
  if (!in_range(to, 12, h.triangles.size())) return fail(c, DF_EINVAL, "triangle offset %d out of range", to);
  const int32_t* t = h.triangles.data() + to;
  for (int k = 0; k < 3; ++k)
    if (!in_range(t[k], 3, h.point3fs.size())) return fail(c, DF_EINVAL, "triangle %d position offset bad", to);
  for (int k = 3; k < 6; ++k)
    if (!in_range(t[k], 3, h.vector3fs.size())) return fail(c, DF_EINVAL, "triangle %d normal offset bad", to);
  for (int k = 6; k < 9; ++k)
    if (!in_range(t[k], 2, h.point2fs.size())) return fail(c, DF_EINVAL, "triangle %d uv offset bad", to);
  for (int k = 9; k < 12; ++k)
    if (!in_range(t[k], 3, h.vector3fs.size())) return fail(c, DF_EINVAL, "triangle %d tangent offset bad", to);
  return DF_OK;
}

DTri pack_triangle(const HostScene& h, int32_t to, bool last) {
  DTri d{};
  const int32_t* t = h.triangles.data() + to;
  const float* A = h.point3fs.data() + t[0];
  const float* B = h.point3fs.data() + t[1];
  const float* C = h.point3fs.data() + t[2];
  for (int k = 0; k < 3; ++k) {
    d.a[k] = A[k];
    d.e0[k] = B[k] - A[k];  // the kernel's edge0 (GPURendererKernel.java:1543-1545)
    d.e1[k] = C[k] - A[k];  // edge1 (:1548-1550)
  }
  d.tri_offset = to;
  d.last = last ? 1 : 0;
  return d;
}

// Re-lays one mesh's skip-pointer BVH segment (SceneCompiler.java:892-995)
// as DNodes in the same pre-order, validating every field.
int repack_mesh(df_ctx* c, const HostScene& h, int32_t base, std::vector<DNode>& nodes, std::vector<DTri>& tris,
                int32_t* root_out) {
  const size_t len = h.bvh.size();
  std::unordered_map<int32_t, int32_t> index_of;  // relative offset -> node index
  struct Pending {
    int32_t node, rel_next, rel_left;
  };
  std::vector<Pending> pending;
  const int32_t first = static_cast<int32_t>(nodes.size());
  int32_t rel = 0;
  int64_t open = 1;  // subtrees still to be read (pre-order completeness)
  while (open > 0) {
    const int64_t off = static_cast<int64_t>(base) + rel;
    if (!in_range(off, 5, len)) return fail(c, DF_EINVAL, "BVH node at %lld out of range", static_cast<long long>(off));
    const int32_t* n = h.bvh.data() + off;
    const int type = n[0];
    if (type != 1 && type != 2) return fail(c, DF_EINVAL, "BVH node at %lld has type %d", static_cast<long long>(off), type);
    if (!in_range(n[2], 3, h.point3fs.size()) || !in_range(n[3], 3, h.point3fs.size()))
      return fail(c, DF_EINVAL, "BVH node at %lld has bad bound offsets", static_cast<long long>(off));
    DNode d{};
    for (int k = 0; k < 3; ++k) {
      d.mn[k] = h.point3fs[n[2] + k];
      d.mx[k] = h.point3fs[n[3] + k];
    }
    const int32_t idx = static_cast<int32_t>(nodes.size());
    index_of[rel] = idx;
    int32_t size = 5;
    if (type == 1) {
      d.tri = -1;
      pending.push_back({idx, n[1], n[4]});
      open += 1;
    } else {
      const int32_t count = n[4];
      if (count < 0 || !in_range(off, 5 + static_cast<int64_t>(count), len))
        return fail(c, DF_EINVAL, "BVH leaf at %lld has bad triangle count", static_cast<long long>(off));
      if (count == 0) {
        d.tri = -1;
        pending.push_back({idx, n[1], n[1]});  // descends straight to its skip target
      } else {
        d.tri = static_cast<int32_t>(tris.size());
        for (int32_t k = 0; k < count; ++k) {
          const int rc = check_triangle(c, h, n[5 + k]);
          if (rc) return rc;
          tris.push_back(pack_triangle(h, n[5 + k], k == count - 1));
        }
        pending.push_back({idx, n[1], -2});
      }
      size = 5 + count;
      open -= 1;
    }
    nodes.push_back(d);
    rel += size;
  }
  for (const Pending& p : pending) {
    auto resolve = [&](int32_t r, int32_t* out) -> bool {
      if (r == -1) {
        *out = -1;
        return true;
      }
      auto it = index_of.find(r);
      if (it == index_of.end()) return false;
      *out = it->second;
      return true;
    };
    int32_t next;
    if (!resolve(p.rel_next, &next)) return fail(c, DF_EINVAL, "BVH skip pointer %d does not name a node", p.rel_next);
    nodes[p.node].next = next;
    if (p.rel_left != -2) {
      int32_t left;
      if (!resolve(p.rel_left, &left)) return fail(c, DF_EINVAL, "BVH left pointer %d does not name a node", p.rel_left);
      nodes[p.node].tri = -left - 2;
    }
  }
  // A traversal must terminate: every pointer moves strictly forward in
  // pre-order except -1, which ends the walk.
  for (int32_t i = first; i < static_cast<int32_t>(nodes.size()); ++i) {
    const DNode& d = nodes[i];
    if (d.next != -1 && d.next <= i) return fail(c, DF_EINVAL, "BVH skip pointer does not advance");
    if (d.tri < 0) {
      const int32_t left = -d.tri - 2;
      if (left != -1 && left <= i) return fail(c, DF_EINVAL, "BVH left pointer does not advance");
    }
  }
  *root_out = first;
  return DF_OK;
}

// World-space bounds of a primitive, for the ray query's bounds cull.  The
// kernel tests the world ray against them before the reference's per-shape
// test: the ray parameter t is the same in world and object space (the
// object ray is the world ray under the affine world->object map), so the
// box's t-interval in world space contains the object-space interval of the
// shape's box whenever the box maps onto the object box.  The object box is
// mapped with the inverse of world_to_object computed here in double (the
// kernel transforms rays with world_to_object), and enlarged so that the
// float rounding of both tests stays inside it.
// The primitive's world-space bounding sphere (DCull): the object-space
// sphere of a sphere primitive, or the sphere around a mesh's BVH root box
// (a ray that misses the root box reaches no triangle), mapped to the world
// by the inverse of world_to_object with the radius scaled by the map's
// largest stretch.  Primitives under a strongly anisotropic map (stretch
// ratio above 4) are never skipped: the enlargement argument assumes a
// near-conformal map.
DCull cull_bounds(const HostScene& h, const std::vector<DNode>& nodes, const DPrim& p) {
  (void)h;
  DCull out{};
  double c[3], r;
  if (p.type == 2) {
    for (int k = 0; k < 3; ++k) c[k] = p.geo[k];
    r = std::fabs(static_cast<double>(p.geo[3]));
  } else if (p.type == 5) {
    const DNode& n = nodes[p.node_root];
    double d2 = 0.0;
    for (int k = 0; k < 3; ++k) {
      c[k] = 0.5 * (static_cast<double>(n.mn[k]) + n.mx[k]);
      const double e = static_cast<double>(n.mx[k]) - n.mn[k];
      d2 += e * e;
    }
    r = 0.5 * std::sqrt(d2);
  } else {
    return out;
  }
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(c[k])) return out;
  if (!std::isfinite(r) || !(r > 0.0)) return out;
  const float* m = p.w2o;
  const double a[3][3] = {{m[0], m[1], m[2]}, {m[4], m[5], m[6]}, {m[8], m[9], m[10]}};
  const double b[3] = {m[3], m[7], m[11]};
  const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                     a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                     a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  if (!std::isfinite(det) || std::fabs(det) < 1e-12) return out;
  double inv[3][3];
  inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
  inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
  inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
  inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
  inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
  inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
  inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
  inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
  inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
  // extreme eigenvalues of M = inv^T inv (closed form for symmetric 3x3)
  double M[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M[i][j] = inv[0][i] * inv[0][j] + inv[1][i] * inv[1][j] + inv[2][i] * inv[2][j];
  const double p1 = M[0][1] * M[0][1] + M[0][2] * M[0][2] + M[1][2] * M[1][2];
  const double q = (M[0][0] + M[1][1] + M[2][2]) / 3.0;
  double e_max, e_min;
  if (p1 < 1e-30 * q * q) {
    e_max = std::max({M[0][0], M[1][1], M[2][2]});
    e_min = std::min({M[0][0], M[1][1], M[2][2]});
  } else {
    const double p2 = (M[0][0] - q) * (M[0][0] - q) + (M[1][1] - q) * (M[1][1] - q) +
                      (M[2][2] - q) * (M[2][2] - q) + 2.0 * p1;
    const double pp = std::sqrt(p2 / 6.0);
    double B[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) B[i][j] = (M[i][j] - (i == j ? q : 0.0)) / pp;
    const double detB = B[0][0] * (B[1][1] * B[2][2] - B[1][2] * B[2][1]) -
                        B[0][1] * (B[1][0] * B[2][2] - B[1][2] * B[2][0]) +
                        B[0][2] * (B[1][0] * B[2][1] - B[1][1] * B[2][0]);
    const double phi = std::acos(std::max(-1.0, std::min(1.0, detB / 2.0))) / 3.0;
    e_max = q + 2.0 * pp * std::cos(phi);
    e_min = q + 2.0 * pp * std::cos(phi + 2.0 * M_PI / 3.0);
  }
  if (!(e_min > 0.0) || e_max > 16.0 * e_min) return out;  // stretch ratio above 4
  const double stretch = std::sqrt(e_max) * (1.0 + 1e-6);
  double cw[3];
  for (int k = 0; k < 3; ++k)
    cw[k] = inv[k][0] * (c[0] - b[0]) + inv[k][1] * (c[1] - b[1]) + inv[k][2] * (c[2] - b[2]);
  const double rw = r * stretch * 1.0001;
  for (int k = 0; k < 3; ++k) {
    if (!std::isfinite(cw[k])) return out;
    out.c[k] = static_cast<float>(cw[k]);
  }
  // the float centre is off by at most half an ulp of each coordinate: fold that into the radius
  double ulp = 0.0;
  for (int k = 0; k < 3; ++k) ulp += std::fabs(cw[k] - out.c[k]);
  out.r = std::nextafter(static_cast<float>(rw + ulp), INFINITY);
  // the sphere test's float error in world units is below 1e-7 (e_max / e_min) |c - o|^2 / rw
  const double k = 1e-6 * (e_max / e_min) / rw;
  // The kernel enlarges the radius to rw + 1e-4 |c - o| + k |c - o|^2 (ten times the float error of the
  // box and sphere tests); with 1e-4 |c - o| <= 5e-5 (|c - o|^2 / s + s) for any s > 0 it evaluates
  // a + b |c - o|^2 >= that, a = rw + 5e-5 s and b = k + 5e-5 / s, with s the distance of the sphere from
  // the origin (equality at |c - o| = s).
  const double s = std::sqrt(cw[0] * cw[0] + cw[1] * cw[1] + cw[2] * cw[2]) + rw + 1.0;
  out.r = std::nextafter(static_cast<float>(static_cast<double>(out.r) + 5e-5 * s), INFINITY);
  out.k = std::nextafter(static_cast<float>(k + 5e-5 / s), INFINITY);
  if (!std::isfinite(out.r) || !std::isfinite(out.k)) return DCull{};
  out.on = 1;
  return out;
}

int build_prims(df_ctx* c, const HostScene& h, std::vector<DPrim>& prims, std::vector<DNode>& nodes,
                std::vector<DTri>& tris, std::vector<DCull>& cull) {
  const size_t np = h.primitives.size() / 3;
  if (h.primitives.size() % 3 != 0 || np == 0) return fail(c, DF_EINVAL, "primitives length %zu", h.primitives.size());
  if (h.o2w.size() < np * 16 || h.w2o.size() < np * 16)
    return fail(c, DF_EINVAL, "transform arrays hold fewer than 16 floats per primitive");
  prims.assign(np, DPrim{});
  cull.assign((np + 7) / 8 * 8, DCull{});  // padded: the skip pass reads the spheres eight at a time
  for (size_t i = 0; i < np; ++i) {
    DPrim& p = prims[i];
    p.type = h.primitives[i * 3 + 0];
    p.shape_offset = h.primitives[i * 3 + 1];
    p.surface_offset = h.primitives[i * 3 + 2];
    p.node_root = -1;
    int rc = check_surface(c, h, p.surface_offset);
    if (rc) return rc;
    for (int k = 0; k < 12; ++k) {
      p.w2o[k] = h.w2o[i * 16 + k];
      p.o2w[k] = h.o2w[i * 16 + k];
    }
    const int32_t so = p.shape_offset;
    switch (p.type) {
      case 1: {  // plane
        if (!in_range(so, 4, h.planes.size())) return fail(c, DF_EINVAL, "plane offset %d out of range", so);
        const int32_t* pl = h.planes.data() + so;
        for (int k = 0; k < 3; ++k)
          if (!in_range(pl[k], 3, h.point3fs.size())) return fail(c, DF_EINVAL, "plane %d point offset bad", so);
        if (!in_range(pl[3], 3, h.vector3fs.size())) return fail(c, DF_EINVAL, "plane %d normal offset bad", so);
        for (int k = 0; k < 3; ++k) {
          p.geo[k] = h.point3fs[pl[0] + k];
          p.geo[3 + k] = h.vector3fs[pl[3] + k];
        }
        break;
      }
      case 2: {  // sphere
        if (!in_range(so, 2, h.spheres.size())) return fail(c, DF_EINVAL, "sphere offset %d out of range", so);
        const int32_t po = jf2i(h.spheres[so]);
        if (!in_range(po, 3, h.point3fs.size())) return fail(c, DF_EINVAL, "sphere %d position offset bad", so);
        for (int k = 0; k < 3; ++k) p.geo[k] = h.point3fs[po + k];
        p.geo[3] = h.spheres[so + 1];
        break;
      }
      case 3:  // terrain: its parameters are not read by the march
        if (!in_range(so, 5, h.terrains.size())) return fail(c, DF_EINVAL, "terrain offset %d out of range", so);
        break;
      case 4: {  // triangle
        rc = check_triangle(c, h, so);
        if (rc) return rc;
        p.node_root = static_cast<int32_t>(tris.size());
        tris.push_back(pack_triangle(h, so, true));
        break;
      }
      case 5: {  // triangle mesh
        rc = repack_mesh(c, h, so, nodes, tris, &p.node_root);
        if (rc) return rc;
        break;
      }
      default:
        break;  // unknown shapes are never hit (the reference skips them)
    }
    cull[i] = cull_bounds(h, nodes, p);
  }
  // transform classes: primitives whose world -> object rows are equal bit
  // for bit map a ray to the same object-space ray
  {
    std::unordered_map<std::string, int32_t> first_of;
    for (size_t i = 0; i < np; ++i) {
      const std::string key(reinterpret_cast<const char*>(prims[i].w2o), sizeof prims[i].w2o);
      prims[i].xform = first_of.emplace(key, static_cast<int32_t>(i)).first->second;
    }
  }
  for (size_t i = np; i-- > 0;)  // runs of consecutive meshes
    prims[i].run_end = prims[i].type == 5 && i + 1 < np && prims[i + 1].type == 5 ? prims[i + 1].run_end
                                                                                  : static_cast<int32_t>(i + 1);
  return DF_OK;
}

template <typename T>
std::vector<T> copy_in(const T* p, size_t n) {
  return (p && n) ? std::vector<T>(p, p + n) : std::vector<T>();
}

// Per-frame camera constants: tan(fovX * 0.5 * PI/180) and tan(-fovY * ...)
// rounded once to float, exactly as run() evaluates them per work-item
// (GPURendererKernel.java:602-605).
DCamera make_camera(const float cam[19]) {
  DCamera d{};
  std::memcpy(d.c, cam, sizeof d.c);
  const float pi_180 = static_cast<float>(M_PI) / 180.0f;
  d.fov_x_tan = static_cast<float>(std::tan(static_cast<double>(cam[4] * 0.5f * pi_180)));
  d.fov_y_tan = static_cast<float>(std::tan(static_cast<double>(-cam[5] * 0.5f * pi_180)));
  return d;
}

DSky make_sky(const df_sky& s) {
  DSky d{};
  for (int k = 0; k < 3; ++k) {
    d.onb_u[k] = s.onb_u[k];
    d.onb_v[k] = s.onb_v[k];
    d.onb_w[k] = s.onb_w[k];
    d.sun_dir_world[k] = s.sun_direction_world[k];
    d.sun_dir[k] = s.sun_direction[k];
    d.sun_origin[k] = s.sun_origin[k];
  }
  d.theta = s.theta;
  d.sky_active = s.is_sky_active;
  d.sun_active = s.is_sun_active;
  d.zenith_l = s.zenith_relative_luminance;
  d.zenith_x = s.zenith_x;
  d.zenith_y = s.zenith_y;
  for (int k = 0; k < 5; ++k) {
    d.pl[k] = s.perez_relative_luminance[k];
    d.px[k] = s.perez_x[k];
    d.py[k] = s.perez_y[k];
  }
  // GPURendererKernel.java:2150-2203, the parts that do not depend on the ray.
  const double theta0 = static_cast<double>(s.theta);
  const double ct0 = std::cos(theta0);
  const double* L = d.pl;
  d.l_den = ((1.0 + L[0] * std::exp(L[1])) * (1.0 + L[2] * std::exp(L[3] * theta0) + L[4] * ct0 * ct0));
  d.x_den0 = (1.0 + d.px[0] * std::exp(d.px[1]));
  d.x_c = d.px[4] * ct0 * ct0;
  d.y_den0 = (1.0 + d.py[0] * std::exp(d.py[1]));
  d.y_c = d.py[4] * ct0 * ct0;
  return d;
}

inline float jpowf(float a, float b) { return static_cast<float>(std::pow(static_cast<double>(a), static_cast<double>(b))); }

DParams make_params(const df_params& p, int w, int h, int passes, bool clear) {
  DParams d{};
  d.renderer_type = p.renderer_type;
  d.tone_mapper_type = p.tone_mapper_type;
  d.exposure = p.tone_mapper_exposure;
  d.depth_max = p.pt_depth_maximum;
  d.depth_rr = p.pt_depth_russian_roulette;
  d.ao_max = p.ao_maximum_distance;
  d.normal_mapping = p.normal_mapping;
  d.wireframes = p.wireframes;
  d.shader_type = p.shader_type;
  d.selected = p.selected_primitive_offset;
  d.ga = p.global_amplitude;
  d.gf = p.global_frequency;
  d.gg = p.global_gain;
  d.gl = p.global_lacunarity;
  d.go = p.global_octaves;
  d.width = w;
  d.height = h;
  d.passes = passes;
  d.clear_first = clear ? 1 : 0;
  d.count_units = 1;
  d.tail_bounce = DF_TAIL_BOUNCE_DEFAULT;
  // image4j RGBColorSpace.SRGB (breakPoint 0.00304, gamma 2.4) and its
  // derived slope / slopeMatch / segmentOffset (RGBColorSpace.java:148-158).
  const float bp = 0.00304f, gamma = 2.4f;
  const float slope = 1.0f / (gamma / jpowf(bp, 1.0f / gamma - 1.0f) - gamma * bp + bp);
  const float slope_match = gamma * slope / jpowf(bp, 1.0f / gamma - 1.0f);
  d.srgb_break = bp;
  d.srgb_gamma_inv = 1.0f / gamma;
  d.srgb_slope = slope;
  d.srgb_slope_match = slope_match;
  d.srgb_segment = slope_match * jpowf(bp, 1.0f / gamma) - slope * bp;
  return d;
}

int validate_params(df_ctx* c, const df_params& p) {
  if (p.renderer_type < 1 || p.renderer_type > 6) return fail(c, DF_EINVAL, "renderer_type %d", p.renderer_type);
  if (p.tone_mapper_type < 1 || p.tone_mapper_type > 4) return fail(c, DF_EINVAL, "tone_mapper_type %d", p.tone_mapper_type);
  if (p.shader_type < 1 || p.shader_type > 2) return fail(c, DF_EINVAL, "shader_type %d", p.shader_type);
  // Any depth is accepted as in the reference (a value below 1 still traces
  // one bounce); the upper bound keeps a path with unit albedo everywhere
  // from looping for billions of bounces.
  if (p.pt_depth_maximum > 256) return fail(c, DF_EINVAL, "pt_depth_maximum %d > 256", p.pt_depth_maximum);
  if (p.global_octaves < 0 || p.global_octaves > 64) return fail(c, DF_EINVAL, "global_octaves %d", p.global_octaves);
  if (p.math_mode != DF_MATH_FAST && p.math_mode != DF_MATH_EXACT) return fail(c, DF_EINVAL, "math_mode %d", p.math_mode);
  if (p.pt_engine < DF_PT_AUTO || p.pt_engine > DF_PT_WAVEFRONT) return fail(c, DF_EINVAL, "pt_engine %d", p.pt_engine);
  return DF_OK;
}

df_params reference_defaults() {
  df_params p{};
  p.renderer_type = DF_RENDERER_PATH_TRACER;
  p.tone_mapper_type = DF_TONE_MAPPER_FILMIC_CURVE_ACES_MODIFIED;
  p.tone_mapper_exposure = 1.0f;
  p.pt_depth_maximum = 5;
  p.pt_depth_russian_roulette = 5;
  p.ao_maximum_distance = 200.0f;
  p.normal_mapping = 1;
  p.wireframes = 0;
  p.shader_type = DF_SHADER_GOURAUD;
  p.selected_primitive_offset = -1;
  p.global_amplitude = 0.5f;
  p.global_frequency = 0.2f;
  p.global_gain = 0.5f;
  p.global_lacunarity = 2.0f;
  p.global_octaves = 2;
  p.math_mode = DF_MATH_FAST;
  p.pt_engine = DF_PT_AUTO;
  return p;
}

// The positions of the terrain height march (GPURendererKernel.java:1484-1501:
// tc starts at 0.001 and grows by td = 0.01 * tc while tc < 10), in the same
// float arithmetic as the device loop; they do not depend on the ray.
std::vector<float> terrain_march_positions() {
  std::vector<float> v;
  float td = 0.01f;
  for (float tc = 0.001f; tc < 10.0f; tc += td) {
    v.push_back(tc);
    td = 0.01f * tc;
  }
  return v;
}

// The two-child records of every mesh's tree (DPair, indexed like the
// nodes) and the deepest per-lane stack their walk needs: one entry per
// first child on the path from the root to a tree node, plus one.  Returns
// false (no records) when some tree node has other than two children or the
// stack would be deeper than kPairStackMax.
bool build_pairs(const std::vector<DNode>& nodes, const std::vector<DPrim>& prims, std::vector<DPair>& pairs,
                 int32_t* stack_depth) {
  pairs.assign(nodes.size(), DPair{});
  *stack_depth = 0;
  const int32_t n = static_cast<int32_t>(nodes.size());
  auto valid = [&](int32_t i) { return i >= 0 && i < n; };
  auto is_tree = [&](int32_t i) { return nodes[i].tri < 0 && -nodes[i].tri - 2 != nodes[i].next; };
  auto ref_of = [&](int32_t i) -> int32_t {
    if (nodes[i].tri >= 0) return -nodes[i].tri - 2;  // leaf
    return is_tree(i) ? i : -1;                       // tree node, or a leaf without triangles
  };
  for (const DPrim& p : prims) {
    if (p.type != 5 || !valid(p.node_root)) continue;
    std::vector<std::pair<int32_t, int32_t>> todo;  // (tree node, first children on its path)
    if (is_tree(p.node_root)) todo.push_back({p.node_root, 0});
    while (!todo.empty()) {
      const auto [i, firsts] = todo.back();
      todo.pop_back();
      const int32_t c0 = -nodes[i].tri - 2;
      if (!valid(c0)) return false;
      const int32_t c1 = nodes[c0].next;
      if (!valid(c1) || c1 == nodes[i].next || nodes[c1].next != nodes[i].next) return false;
      if (firsts + 1 > kPairStackMax) return false;
      *stack_depth = std::max(*stack_depth, firsts + 1);
      DPair& r = pairs[i];
      const int32_t ch[2] = {c0, c1};
      for (int k = 0; k < 2; ++k) {
        for (int a = 0; a < 3; ++a) {
          r.mn[a][k] = nodes[ch[k]].mn[a];
          r.mx[a][k] = nodes[ch[k]].mx[a];
        }
        r.ref[k] = ref_of(ch[k]);
      }
      if (is_tree(c1)) todo.push_back({c1, firsts});
      if (is_tree(c0)) todo.push_back({c0, firsts + 1});
    }
  }
  return true;
}

int rebuild_device_scene(df_ctx* c) {
  std::vector<DPrim> prims;
  std::vector<DNode> nodes;
  std::vector<DTri> tris;
  std::vector<DCull> cull;
  int rc = build_prims(c, c->hs, prims, nodes, tris, cull);
  if (rc) return rc;
  // buffers: 0 point2fs 1 point3fs 2 vector3fs 3 planes 4 spheres 5 terrains
  // 6 triangles 7 textures 8 surfaces 9 primitives 10 prims 11 nodes 12 tris
  // 13 terrain march positions 14 cull spheres 15 two-child records
  c->scene_bufs.resize(16);
  std::vector<DPair> pairs;
  int32_t pair_stack = 0;
  if (!build_pairs(nodes, prims, pairs, &pair_stack)) {
    pairs.clear();
    pair_stack = 0;
  }
  const std::vector<float> march = terrain_march_positions();
  const HostScene& h = c->hs;
  if ((rc = upload(c, c->scene_bufs[0], h.point2fs)) || (rc = upload(c, c->scene_bufs[1], h.point3fs)) ||
      (rc = upload(c, c->scene_bufs[2], h.vector3fs)) || (rc = upload(c, c->scene_bufs[3], h.planes)) ||
      (rc = upload(c, c->scene_bufs[4], h.spheres)) || (rc = upload(c, c->scene_bufs[5], h.terrains)) ||
      (rc = upload(c, c->scene_bufs[6], h.triangles)) || (rc = upload(c, c->scene_bufs[7], h.textures)) ||
      (rc = upload(c, c->scene_bufs[8], h.surfaces)) || (rc = upload(c, c->scene_bufs[9], h.primitives)) ||
      (rc = upload(c, c->scene_bufs[10], prims)) || (rc = upload(c, c->scene_bufs[11], nodes)) ||
      (rc = upload(c, c->scene_bufs[12], tris)) || (rc = upload(c, c->scene_bufs[13], march)) ||
      (rc = upload(c, c->scene_bufs[14], cull)) || (rc = upload(c, c->scene_bufs[15], pairs)))
    return rc;
  DScene& d = c->ds;
  d.point2fs = static_cast<const float*>(c->scene_bufs[0].p);
  d.point3fs = static_cast<const float*>(c->scene_bufs[1].p);
  d.vector3fs = static_cast<const float*>(c->scene_bufs[2].p);
  d.planes = static_cast<const int32_t*>(c->scene_bufs[3].p);
  d.spheres = static_cast<const float*>(c->scene_bufs[4].p);
  d.terrains = static_cast<const float*>(c->scene_bufs[5].p);
  d.triangles = static_cast<const int32_t*>(c->scene_bufs[6].p);
  d.textures = static_cast<const float*>(c->scene_bufs[7].p);
  d.surfaces = static_cast<const float*>(c->scene_bufs[8].p);
  d.primitives = static_cast<const int32_t*>(c->scene_bufs[9].p);
  d.prims = static_cast<const DPrim*>(c->scene_bufs[10].p);
  d.nodes = static_cast<const DNode*>(c->scene_bufs[11].p);
  d.tris = static_cast<const DTri*>(c->scene_bufs[12].p);
  d.march_tc = static_cast<const float*>(c->scene_bufs[13].p);
  d.cull = static_cast<const DCull*>(c->scene_bufs[14].p);
  d.pairs = pairs.empty() ? nullptr : static_cast<const DPair*>(c->scene_bufs[15].p);
  d.pair_stack = pair_stack;
  d.prim_count = static_cast<int32_t>(prims.size());
  d.march_n = static_cast<int32_t>(march.size());
  return DF_OK;
}

int need_frame(df_ctx* c) {
  if (!c) return DF_EINVAL;
  if (c->width <= 0 || !c->film.p) return fail(c, DF_ESTATE, "df_update has not been called");
  return DF_OK;
}

size_t pixels(const df_ctx* c) { return static_cast<size_t>(c->width) * static_cast<size_t>(c->height); }

void free_paths(df_ctx* c) {
  for (DevBuf& b : c->path_f4) free_buf(b);
  for (DevBuf& b : c->path_q) free_buf(b);
  free_buf(c->path_ctl);
  free_buf(c->path_bins);
  c->paths = dfk::DPaths{};
}

// Path state of the wavefront path tracer for the current frame size and a
// maximum depth: per pixel 12 float4 records and an entry in every stripe of
// 4 queues (448 bytes).
int ensure_paths(df_ctx* c, int depth_max, bool bins) {
  const size_t n = pixels(c);
  const int ctl_len = dfk::wavefront_ctl_len(depth_max);
  int rc;
  if (c->paths.n != static_cast<int64_t>(n) || c->paths.ctl_len < ctl_len) {
    DF_HIP(c, hipStreamSynchronize(c->stream));  // earlier launches may still read the old buffers
    free_paths(c);
    for (DevBuf& b : c->path_f4)
      if ((rc = alloc_buf(c, b, n * sizeof(float4)))) return rc;
    for (DevBuf& b : c->path_q)
      if ((rc = alloc_buf(c, b, n * dfk::kQueueStripes * sizeof(int32_t)))) return rc;
    const int ctl_alloc = ctl_len > 64 * dfk::kCtlGroup ? ctl_len : 64 * dfk::kCtlGroup;
    if ((rc = alloc_buf(c, c->path_ctl, static_cast<size_t>(ctl_alloc) * sizeof(int32_t)))) return rc;
    dfk::DPaths& w = c->paths;
    float4** f4[12] = {&w.ray_o, &w.ray_d, &w.thru, &w.rad, &w.hit, &w.sh_o, &w.sh_d, &w.rad_lit, &w.sample,
                       &w.ray0_o, &w.ray0_d, &w.noise};
    for (int i = 0; i < 12; ++i) *f4[i] = static_cast<float4*>(c->path_f4[i].p);
    w.queue[0] = static_cast<int32_t*>(c->path_q[0].p);
    w.queue[1] = static_cast<int32_t*>(c->path_q[1].p);
    w.shadow_queue = static_cast<int32_t*>(c->path_q[2].p);
    w.noise_queue = static_cast<int32_t*>(c->path_q[3].p);
    w.sky_queue = static_cast<int32_t*>(c->path_q[4].p);
    w.ctl = static_cast<int32_t*>(c->path_ctl.p);
    w.n = static_cast<int64_t>(n);
    w.ctl_len = ctl_alloc;
  }
  if (bins && !c->path_bins.p) {
    DF_HIP(c, hipStreamSynchronize(c->stream));
    if ((rc = alloc_buf(c, c->path_bins, n * dfk::kShadeBins * dfk::kQueueStripes * sizeof(int32_t)))) return rc;
    c->paths.shade_bins = static_cast<int32_t*>(c->path_bins.p);
  }
  return DF_OK;
}

// DF_PT_AUTO chooses by measurement: the first passes of a frame time each
// schedule of kCandidates in turn, one pass at a time, until each has run
// kTunePasses times, and the one with the shortest best pass is kept (every
// schedule gives the same bits, so passes split between them change nothing
// in the result).  Changing the scene, the resolution or the parameters
// starts over.
constexpr int kTunePasses = 2;

bool tuning(const df_ctx* c) {
  return c->params.renderer_type == DF_RENDERER_PATH_TRACER && c->params.pt_engine == DF_PT_AUTO &&
         !c->sched_fixed && !c->tuned;
}

void reset_tuning(df_ctx* c) {
  c->tuned = false;
  for (int k = 0; k < kTuneCandidates; ++k) {
    c->tune_runs[k] = 0;
    c->tune_ms[k] = 0.0f;
  }
}

// The schedule of the next passes once tuning is over: the fixed one, the
// tuned one, or the default of the engine params.pt_engine names (the other
// renderers run as the megakernel).
df_schedule current_schedule(const df_ctx* c) {
  if (c->sched_fixed || (c->tuned && c->params.pt_engine == DF_PT_AUTO)) return c->sched;
  df_schedule d{DF_PT_MEGAKERNEL, DF_TILE_ROWS, 0, 0, DF_TAIL_BOUNCE_DEFAULT};
  if (c->params.renderer_type == DF_RENDERER_PATH_TRACER && c->params.pt_engine == DF_PT_WAVEFRONT)
    d.engine = DF_PT_WAVEFRONT;
  return d;
}

// The path tracer's schedule for this frame.
bool use_wavefront(const df_ctx* c) {
  return c->params.renderer_type == DF_RENDERER_PATH_TRACER && current_schedule(c).engine == DF_PT_WAVEFRONT;
}

hipEvent_t take_event(df_ctx* c) {
  if (c->ev_used == c->ev_pool.size()) {
    hipEvent_t e = nullptr;
    const hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) {
      c->mark_error = r;
      return nullptr;
    }
    c->ev_pool.push_back(e);
  }
  return c->ev_pool[c->ev_used++];
}

// dfk_marks hook: an event before and after each launch.
void stage_mark(void* user, int stage, int begin) {
  df_ctx* c = static_cast<df_ctx*>(user);
  hipEvent_t e = take_event(c);
  if (!e) return;
  const hipError_t r = hipEventRecord(e, c->stream);
  if (r != hipSuccess) c->mark_error = r;
  if (begin) {
    c->open_begin = c->ev_used - 1;
  } else {
    c->spans.push_back({stage, c->open_begin, c->ev_used - 1});
  }
}

// Folds the recorded spans into the per-stage totals (after a stream sync).
int collect_spans(df_ctx* c) {
  for (const df_ctx::Span& sp : c->spans) {
    float ms = 0.0f;
    DF_HIP(c, hipEventElapsedTime(&ms, c->ev_pool[sp.begin], c->ev_pool[sp.end]));
    c->stage_ms[sp.stage] += ms;
    c->stage_launches[sp.stage] += 1;
  }
  c->spans.clear();
  c->ev_used = 0;
  return DF_OK;
}

int read_back(df_ctx* c, void* dst, const DevBuf& b, size_t bytes) {
  DF_HIP(c, hipStreamSynchronize(c->stream));
  DF_HIP(c, hipMemcpy(dst, b.p, bytes, hipMemcpyDeviceToHost));
  return DF_OK;
}

// The differentiable-mode fields of the kernel parameters.
void apply_diff(const df_ctx* c, DParams& p) {
  p.diff_on = c->diff_on ? 1 : 0;
  if (!c->diff_on) return;
  p.n_surfaces = static_cast<int32_t>(c->diff_n_surfaces);
  p.surface_albedo = static_cast<const int32_t*>(c->diff_surf_albedo.p);
  p.surface_emission = static_cast<const int32_t*>(c->diff_surf_emission.p);
  p.surface_checker = static_cast<const int32_t*>(c->diff_surf_checker.p);
  p.diff_albedo = static_cast<const float*>(c->diff_albedo.p);
  p.diff_emission = static_cast<const float*>(c->diff_emission.p);
  p.checker_base = c->diff_n_albedo;
}

DFrame frame_of(df_ctx* c) {
  DFrame f{};
  f.film = static_cast<float*>(c->film.p);
  f.film_samples = static_cast<int32_t*>(c->film_samples.p);
  f.image_float = static_cast<float*>(c->image_float.p);
  f.image_bgra = static_cast<uint8_t*>(c->image_bgra.p);
  f.primitive_offsets = static_cast<int32_t*>(c->prim_offsets.p);
  f.seeds = static_cast<uint64_t*>(c->seeds.p);
  f.counters = static_cast<unsigned long long*>(c->counters.p);
  return f;
}

}  // namespace

struct df_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed = false;
  std::string err;
  std::vector<df_ctx*> jobs;
  std::vector<int32_t> passes;
  DevBuf djobs, ditems, dheads;
  // the group launches of the last plan as a graph (df_batch_set_graphs,
  // default off like the renderer's)
  bool graphs = false;
  GraphCache g_launch;
};

namespace {

int bfail(df_batch* b, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (b) b->err = buf;
  return code;
}

#define DF_BHIP(b, call)                                                             \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) return bfail((b), DF_EHIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

// The batch's work tables for one df_batch_execute: the job table, the
// (job, 16x16 tile) items grouped by (math mode, renderer) launch, and each
// group's range of items.
struct BatchPlan {
  std::vector<DJob> jobs;
  std::vector<DBatchItem> items;
  std::vector<int> group_key;       // math_mode * 8 + renderer_type, one entry per group
  std::vector<size_t> group_begin;  // items of group g: [group_begin[g], group_begin[g + 1])
  std::vector<size_t> wave_jobs;    // jobs rendered on their own wavefront schedule instead
};

// ---- allocation-free batch execution (df_batch_execute_noalloc) ----------
// The work tables live in one static device array per GPU: the group work
// heads (zeroed by a fill kernel on every call), the job table, the items.
// Every call on one GPU shares that GPU's array, so calls that overlap on one
// GPU would race: the entry point serialises the calls of this process per
// device and synchronises before it returns.
constexpr int32_t kNoAllocMaxJobs = 1024;
constexpr int32_t kNoAllocMaxItems = 262144;
constexpr int32_t kNoAllocMaxGroups = 64;  // at most 2 math modes x 6 renderers occur
constexpr size_t kScratchHeadsBytes = 256;
constexpr size_t kScratchJobsBytes = (kNoAllocMaxJobs * sizeof(DJob) + 255) / 256 * 256;
constexpr size_t kScratchBytes = kScratchHeadsBytes + kScratchJobsBytes + kNoAllocMaxItems * sizeof(DBatchItem);
static_assert(kNoAllocMaxGroups * sizeof(int) <= kScratchHeadsBytes, "heads region");
constexpr int kMaxScratchDevices = 64;

std::mutex g_scratch_setup_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};
std::mutex g_scratch_call_mu[kMaxScratchDevices];

__global__ void zero_ints(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Makes d hold at least `bytes` (kept when it already does, so that the
// tables of a repeated batch stay at the addresses its graph was captured with).
int balloc(df_batch* b, DevBuf& d, size_t bytes) {
  if (bytes == 0) bytes = 4;
  if (d.p && d.bytes >= bytes) return DF_OK;
  free_buf(d);
  const hipError_t e = hipMalloc(&d.p, bytes);
  if (e != hipSuccess) {
    d.p = nullptr;
    return bfail(b, DF_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  d.bytes = bytes;
  return DF_OK;
}

}  // namespace

// The scratch array of df_batch_execute_noalloc: kScratchBytes (about 4.7 MiB)
// per GPU, held until the process ends.  Host code never uses this name as an
// address: scratch_address() looks up each device's copy.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// The device address of `scratch` on `device` (which must be current), looked
// up once per device with hipGetSymbolAddress.
int scratch_address(int device, unsigned char** out) {
  if (device < 0 || device >= kMaxScratchDevices) return DF_ENODEVICE;
  std::lock_guard<std::mutex> lock(g_scratch_setup_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return DF_EHIP;
    g_scratch_ptr[device] = ptr;
  }
  *out = static_cast<unsigned char*>(g_scratch_ptr[device]);
  return DF_OK;
}

// Builds the batch's work tables: per job its scene, camera, sky, parameters
// and frame; per group (math mode, renderer) its (job, tile) items, most
// expensive first.  With `allow_wave`, a path-tracer job whose own schedule is
// the wavefront (fixed with df_set_schedule, given by its pt_engine, or chosen
// by its autotuner: large mesh scenes) is left out of the items and rendered
// on that schedule, its path state allocated here.  Waits for every job's own
// stream.
int batch_plan(df_batch* b, BatchPlan& plan, bool allow_wave) {
  const size_t n = b->jobs.size();
  struct Item {
    double cost;
    int group;
    DBatchItem it;
  };
  plan.jobs.assign(n, DJob{});
  std::vector<Item> items;
  plan.group_key.clear();
  plan.wave_jobs.clear();
  for (size_t j = 0; j < n; ++j) {
    df_ctx* c = b->jobs[j];
    DF_BHIP(b, hipStreamSynchronize(c->stream));
    DJob& hj = plan.jobs[j];
    hj.S = c->ds;
    hj.C = make_camera(c->camera);
    hj.K = make_sky(c->sky);
    hj.P = make_params(c->params, c->width, c->height, b->passes[j], c->clear_pending);
    apply_diff(c, hj.P);
    hj.F = frame_of(c);
    if (b->passes[j] == 0) continue;
    if (allow_wave && use_wavefront(c)) {
      if (const int rc = ensure_paths(c, c->params.pt_depth_maximum, current_schedule(c).shade_binned != 0))
        return bfail(b, rc, "job %zu: %s", j, df_last_error(c));
      plan.wave_jobs.push_back(j);
      continue;
    }
    const int key = c->params.math_mode * 8 + c->params.renderer_type;
    int g = 0;
    while (g < static_cast<int>(plan.group_key.size()) && plan.group_key[g] != key) ++g;
    if (g == static_cast<int>(plan.group_key.size())) plan.group_key.push_back(key);
    // work estimate of a tile: passes x (primitive loop + BVH depth)
    const double cost = static_cast<double>(b->passes[j]) *
                        (16.0 + c->ds.prim_count + 8.0 * std::log2(1.0 + static_cast<double>(c->hs.bvh.size())));
    const int tw = (c->width + 15) / 16, th = (c->height + 15) / 16;
    for (int ty = 0; ty < th; ++ty)
      for (int tx = 0; tx < tw; ++tx) items.push_back({cost, g, {static_cast<int32_t>(j), tx, ty, 0}});
  }
  // per group, the most expensive items first (longest-processing-time order)
  std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& c) {
    return a.group != c.group ? a.group < c.group : a.cost > c.cost;
  });
  plan.items.resize(items.size());
  for (size_t i = 0; i < items.size(); ++i) plan.items[i] = items[i].it;
  plan.group_begin.assign(plan.group_key.size() + 1, items.size());
  for (size_t i = items.size(); i-- > 0;) plan.group_begin[items[i].group] = i;
  return DF_OK;
}

// One persistent launch per group over the tables already on the device
// (heads zeroed), timed with the batch's events; waits for the batch.
int batch_launch(df_batch* b, const BatchPlan& plan, const DJob* djobs, const DBatchItem* ditems, int* dheads) {
  auto enqueue = [&]() -> hipError_t {
    for (size_t g = 0; g < plan.group_key.size(); ++g) {
      const size_t begin = plan.group_begin[g], end = plan.group_begin[g + 1];
      const hipError_t e = dfk_launch_batch(djobs, ditems + begin, static_cast<int>(end - begin), dheads + g,
                                            plan.group_key[g] / 8, plan.group_key[g] % 8, b->stream);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  };
  DF_BHIP(b, hipEventRecord(b->ev0, b->stream));
  if (b->graphs) {
    // the persistent grid sizes are queried (n_items 0: no launch) before any capture
    for (size_t g = 0; g < plan.group_key.size(); ++g)
      DF_BHIP(b, dfk_launch_batch(djobs, ditems, 0, dheads, plan.group_key[g] / 8, plan.group_key[g] % 8, b->stream));
    std::vector<unsigned char> key;
    key_add(key, djobs);
    key_add(key, ditems);
    key_add(key, dheads);
    for (size_t g = 0; g < plan.group_key.size(); ++g) {
      key_add(key, plan.group_key[g]);
      key_add(key, plan.group_begin[g]);
    }
    key_add(key, plan.items.size());
    DF_BHIP(b, graph_launch(b->g_launch, key, b->stream, enqueue));
  } else {
    DF_BHIP(b, enqueue());
  }
  for (const size_t j : plan.wave_jobs) {  // on their own (wavefront) schedule, on the batch's stream
    df_ctx* c = b->jobs[j];
    DParams par = plan.jobs[j].P;
    par.shade_binned = current_schedule(c).shade_binned;
    par.bvh_pairs = current_schedule(c).bvh_pairs;
    par.tail_bounce = current_schedule(c).tail_bounce;
    par.count_units = c->count_units ? 1 : 0;
    DF_BHIP(b, dfk_launch_pt_wavefront(c->ds, plan.jobs[j].C, plan.jobs[j].K, par, plan.jobs[j].F, c->paths,
                                       c->params.math_mode, b->stream, nullptr));
  }
  DF_BHIP(b, hipEventRecord(b->ev1, b->stream));
  DF_BHIP(b, hipStreamSynchronize(b->stream));
  b->timed = true;
  for (df_ctx* c : b->jobs) c->clear_pending = false;
  return DF_OK;
}

}  // namespace

extern "C" {

int df_device_count(int32_t* count) {
  if (!count) return DF_EINVAL;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return DF_OK;
}

int df_validate_scene(const df_compiled_scene* s, char* msg, size_t cap) {
  if (!s) return DF_EINVAL;
  df_ctx scratch;  // error holder only; no HIP call is made
  int rc = DF_OK;
  if (!s->camera || s->camera_len != 19) {
    rc = fail(&scratch, DF_EINVAL, "camera must hold 19 floats");
  } else if (!s->primitives || s->primitives_len < 3) {
    rc = fail(&scratch, DF_EINVAL, "scene has no primitives");
  } else {
    HostScene h;
    h.point2fs = copy_in(s->point2fs, s->point2fs_len);
    h.point3fs = copy_in(s->point3fs, s->point3fs_len);
    h.vector3fs = copy_in(s->vector3fs, s->vector3fs_len);
    h.planes = copy_in(s->planes, s->planes_len);
    h.spheres = copy_in(s->spheres, s->spheres_len);
    h.terrains = copy_in(s->terrains, s->terrains_len);
    h.triangles = copy_in(s->triangles, s->triangles_len);
    h.bvh = copy_in(s->bvh, s->bvh_len);
    h.textures = copy_in(s->textures, s->textures_len);
    h.surfaces = copy_in(s->surfaces, s->surfaces_len);
    h.primitives = copy_in(s->primitives, s->primitives_len);
    h.o2w = copy_in(s->object_to_world, s->object_to_world_len);
    h.w2o = copy_in(s->world_to_object, s->world_to_object_len);
    std::vector<DPrim> prims;
    std::vector<DNode> nodes;
    std::vector<DTri> tris;
    std::vector<DCull> cull;
    rc = build_prims(&scratch, h, prims, nodes, tris, cull);
  }
  if (msg && cap) {
    std::snprintf(msg, cap, "%s", rc ? scratch.err.c_str() : "ok");
  }
  return rc;
}

int df_create(df_ctx** out, int32_t device) {
  if (!out) return DF_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return DF_ENODEVICE;
  if (device < 0 || device >= n) return DF_ENODEVICE;
  df_ctx* c = new (std::nothrow) df_ctx;
  if (!c) return DF_ENOMEM;
  c->device = device;
  c->params = reference_defaults();
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    df_destroy(c);
    return DF_EHIP;
  }
  if (alloc_buf(c, c->counters, kCounterBytes) != DF_OK || hipMemset(c->counters.p, 0, kCounterBytes) != hipSuccess) {
    df_destroy(c);
    return DF_ENOMEM;
  }
  *out = c;
  return DF_OK;
}

void df_destroy(df_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (DevBuf& b : c->scene_bufs) free_buf(b);
  free_buf(c->film);
  free_buf(c->film_samples);
  free_buf(c->image_float);
  free_buf(c->image_bgra);
  free_buf(c->prim_offsets);
  free_buf(c->seeds);
  free_buf(c->counters);
  free_paths(c);
  for (DevBuf* b : {&c->diff_surf_albedo, &c->diff_surf_emission, &c->diff_surf_checker, &c->diff_albedo,
                    &c->diff_emission, &c->diff_sky_dual, &c->diff_weight, &c->diff_grad, &c->diff_loss,
                    &c->diff_seeds})
    free_buf(*b);
  for (auto& row : c->g_wave)
    for (GraphCache& g : row) g.reset();
  c->g_mega.reset();
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->tune_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int df_load_scene(df_ctx* c, const df_compiled_scene* s) {
  if (!c || !s) return DF_EINVAL;
  if (!s->camera || s->camera_len != 19) return fail(c, DF_EINVAL, "camera must hold 19 floats");
  if (!s->primitives || s->primitives_len < 3) return fail(c, DF_EINVAL, "scene has no primitives");
  DF_HIP(c, hipSetDevice(c->device));
  DF_HIP(c, hipStreamSynchronize(c->stream));
  HostScene h;
  h.point2fs = copy_in(s->point2fs, s->point2fs_len);
  h.point3fs = copy_in(s->point3fs, s->point3fs_len);
  h.vector3fs = copy_in(s->vector3fs, s->vector3fs_len);
  h.planes = copy_in(s->planes, s->planes_len);
  h.spheres = copy_in(s->spheres, s->spheres_len);
  h.terrains = copy_in(s->terrains, s->terrains_len);
  h.triangles = copy_in(s->triangles, s->triangles_len);
  h.bvh = copy_in(s->bvh, s->bvh_len);
  h.textures = copy_in(s->textures, s->textures_len);
  h.surfaces = copy_in(s->surfaces, s->surfaces_len);
  h.primitives = copy_in(s->primitives, s->primitives_len);
  h.o2w = copy_in(s->object_to_world, s->object_to_world_len);
  h.w2o = copy_in(s->world_to_object, s->world_to_object_len);
  std::swap(c->hs, h);
  const int rc = rebuild_device_scene(c);
  if (rc) {
    std::swap(c->hs, h);
    c->has_scene = false;
    return rc;
  }
  std::memcpy(c->camera, s->camera, sizeof c->camera);
  c->has_scene = true;
  c->diff_on = false;  // the parameter maps index the previous scene's surface records
  reset_tuning(c);
  return DF_OK;
}

int df_update(df_ctx* c, int32_t w, int32_t h, int32_t local_size_hint) {
  (void)local_size_hint;
  if (!c) return DF_EINVAL;
  if (w <= 0 || h <= 0 || static_cast<int64_t>(w) * h > (1ll << 28)) return fail(c, DF_EINVAL, "bad resolution %dx%d", w, h);
  DF_HIP(c, hipSetDevice(c->device));
  DF_HIP(c, hipStreamSynchronize(c->stream));
  const size_t n = static_cast<size_t>(w) * h;
  free_paths(c);
  reset_tuning(c);
  // no frame until every buffer has the new size: a failed allocation below
  // leaves the handle without one (DF_ESTATE) rather than with mixed sizes
  c->width = 0;
  c->height = 0;
  int rc;
  if ((rc = alloc_buf(c, c->film, n * 3 * sizeof(float))) || (rc = alloc_buf(c, c->film_samples, n * sizeof(int32_t))) ||
      (rc = alloc_buf(c, c->image_float, n * 3 * sizeof(float))) || (rc = alloc_buf(c, c->image_bgra, n * 4)) ||
      (rc = alloc_buf(c, c->prim_offsets, n * sizeof(int32_t))) || (rc = alloc_buf(c, c->seeds, n * sizeof(uint64_t))))
    return rc;
  DF_HIP(c, hipMemset(c->film.p, 0, n * 3 * sizeof(float)));
  DF_HIP(c, hipMemset(c->film_samples.p, 0, n * sizeof(int32_t)));
  DF_HIP(c, hipMemset(c->image_float.p, 0, n * 3 * sizeof(float)));
  DF_HIP(c, hipMemset(c->image_bgra.p, 0, n * 4));
  DF_HIP(c, hipMemset(c->prim_offsets.p, 0xFF, n * sizeof(int32_t)));  // Arrays.fill(primitiveOffsets, -1)
  c->width = w;
  c->height = h;
  c->clear_pending = false;
  DF_HIP(c, dfk_launch_seed(static_cast<uint64_t*>(c->seeds.p), n, 0, c->stream));
  DF_HIP(c, hipStreamSynchronize(c->stream));
  return DF_OK;
}

int df_seed(df_ctx* c, uint64_t master) {
  int rc = need_frame(c);
  if (rc) return rc;
  DF_HIP(c, hipSetDevice(c->device));
  DF_HIP(c, dfk_launch_seed(static_cast<uint64_t*>(c->seeds.p), pixels(c), master, c->stream));
  DF_HIP(c, hipStreamSynchronize(c->stream));
  return DF_OK;
}

int df_set_seeds(df_ctx* c, const int64_t* seeds, size_t n) {
  int rc = need_frame(c);
  if (rc) return rc;
  if (!seeds || n != pixels(c)) return fail(c, DF_EINVAL, "seed count %zu != %zu", n, pixels(c));
  DF_HIP(c, hipStreamSynchronize(c->stream));
  DF_HIP(c, hipMemcpy(c->seeds.p, seeds, n * sizeof(int64_t), hipMemcpyHostToDevice));
  return DF_OK;
}

int df_get_seeds(df_ctx* c, int64_t* seeds, size_t n) {
  int rc = need_frame(c);
  if (rc) return rc;
  if (!seeds || n != pixels(c)) return fail(c, DF_EINVAL, "seed count %zu != %zu", n, pixels(c));
  return read_back(c, seeds, c->seeds, n * sizeof(int64_t));
}

int df_upload_camera(df_ctx* c, const float* cam, size_t n) {
  if (!c || !cam) return DF_EINVAL;
  if (n != 19) return fail(c, DF_EINVAL, "camera must hold 19 floats");
  std::memcpy(c->camera, cam, sizeof c->camera);
  return DF_OK;
}

int df_upload_transforms(df_ctx* c, const float* o2w, const float* w2o, int32_t n_prim) {
  if (!c || !o2w || !w2o) return DF_EINVAL;
  if (!c->has_scene) return fail(c, DF_ESTATE, "no scene loaded");
  if (n_prim < 0 || static_cast<size_t>(n_prim) * 3 != c->hs.primitives.size())
    return fail(c, DF_EINVAL, "transform count %d != primitive count %zu", n_prim, c->hs.primitives.size() / 3);
  DF_HIP(c, hipSetDevice(c->device));
  DF_HIP(c, hipStreamSynchronize(c->stream));
  c->hs.o2w.assign(o2w, o2w + static_cast<size_t>(n_prim) * 16);
  c->hs.w2o.assign(w2o, w2o + static_cast<size_t>(n_prim) * 16);
  return rebuild_device_scene(c);
}

int df_upload_surfaces(df_ctx* c, const float* surfaces, size_t n) {
  if (!c || !surfaces) return DF_EINVAL;
  if (!c->has_scene) return fail(c, DF_ESTATE, "no scene loaded");
  if (c->diff_on) {  // a checkerboard group's surface must keep a checkerboard albedo texture
    for (size_t i = 0; i < c->diff_checker_host.size() && i * 6 + 1 < n; ++i)
      if (c->diff_checker_host[i] >= 0 && !checker_albedo(c->hs.textures, surfaces, i))
        return fail(c, DF_ESTATE, "surface %zu is in a checkerboard group: its albedo texture must stay one", i);
  }
  DF_HIP(c, hipSetDevice(c->device));
  DF_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<float> old = c->hs.surfaces;
  c->hs.surfaces.assign(surfaces, surfaces + n);
  const int rc = rebuild_device_scene(c);
  if (rc) {
    c->hs.surfaces.swap(old);
    (void)rebuild_device_scene(c);
  }
  return rc;
}

int df_upload_sky(df_ctx* c, const df_sky* sky) {
  if (!c || !sky) return DF_EINVAL;
  c->sky = *sky;
  c->has_sky = true;
  return DF_OK;
}

int df_toggle_sky(df_ctx* c) {
  if (!c) return DF_EINVAL;
  c->sky.is_sky_active = c->sky.is_sky_active == 0 ? 1 : 0;
  return DF_OK;
}

int df_toggle_sun(df_ctx* c) {
  if (!c) return DF_EINVAL;
  c->sky.is_sun_active = c->sky.is_sun_active == 0 ? 1 : 0;
  return DF_OK;
}

int df_set_params(df_ctx* c, const df_params* p) {
  if (!c || !p) return DF_EINVAL;
  const int rc = validate_params(c, *p);
  if (rc) return rc;
  if (std::memcmp(&c->params, p, sizeof *p) != 0) reset_tuning(c);
  c->params = *p;
  return DF_OK;
}

int df_get_params(df_ctx* c, df_params* p) {
  if (!c || !p) return DF_EINVAL;
  *p = c->params;
  return DF_OK;
}

int df_toggle_primitive_selection(df_ctx* c, int32_t x, int32_t y) {
  int rc = need_frame(c);
  if (rc) return rc;
  const int64_t index = static_cast<int64_t>(y) * c->width + x;
  if (index < 0 || index >= static_cast<int64_t>(pixels(c))) return DF_OK;  // outside: no change (:358)
  int32_t po = -1;
  DF_HIP(c, hipStreamSynchronize(c->stream));
  DF_HIP(c, hipMemcpy(&po, static_cast<int32_t*>(c->prim_offsets.p) + index, sizeof po, hipMemcpyDeviceToHost));
  c->params.selected_primitive_offset = po == c->params.selected_primitive_offset ? -1 : po;
  return DF_OK;
}

int df_toggle_primitive_material(df_ctx* c) {
  if (!c) return DF_EINVAL;
  const int32_t sel = c->params.selected_primitive_offset;
  if (sel == -1 || !c->has_scene) return DF_OK;
  if (!in_range(sel, 3, c->hs.primitives.size())) return fail(c, DF_EINVAL, "selected primitive %d out of range", sel);
  const int32_t so = c->hs.primitives[sel + 2];
  const int old_type = jf2i(c->hs.surfaces[so]);
  const int ord = old_type - 1;
  if (ord < 0 || ord >= 5) return DF_OK;
  std::vector<float> s = c->hs.surfaces;
  s[so] = static_cast<float>((ord + 1) % 5 + 1);  // ClearCoat, Glass, Lambertian, Phong, Reflection
  return df_upload_surfaces(c, s.data(), s.size());
}

int df_clear(df_ctx* c) {
  if (!c) return DF_EINVAL;
  c->clear_pending = true;
  return DF_OK;
}

int df_execute(df_ctx* c, int32_t passes) {
  int rc = need_frame(c);
  if (rc) return rc;
  if (!c->has_scene) return fail(c, DF_ESTATE, "no scene loaded");
  if (!c->has_sky) return fail(c, DF_ESTATE, "no sky uploaded");
  if (passes < 0) return fail(c, DF_EINVAL, "passes %d", passes);
  if (passes == 0) return DF_OK;
  DF_HIP(c, hipSetDevice(c->device));
  const DCamera cam = make_camera(c->camera);
  const DSky sky = make_sky(c->sky);
  const DFrame f = frame_of(c);
  const bool may_wave = use_wavefront(c) || tuning(c);
  const bool may_bin = tuning(c) || (use_wavefront(c) && current_schedule(c).shade_binned);
  if (may_wave && (rc = ensure_paths(c, c->params.pt_depth_maximum, may_bin))) return rc;
  const dfk_marks marks{stage_mark, c};
  const dfk_marks* mk = c->stage_timing ? &marks : nullptr;
  c->mark_error = hipSuccess;
  // `n` passes with one schedule; the film is cleared before the first if `clear`
  // through the frame-loop graphs unless the launches are timed one by one
  const bool graph = c->graphs && !mk;
  auto run = [&](const df_schedule& sc, int n, bool clear) -> int {
    const bool wave = c->params.renderer_type == DF_RENDERER_PATH_TRACER && sc.engine == DF_PT_WAVEFRONT;
    DParams par = make_params(c->params, c->width, c->height, n, clear);
    apply_diff(c, par);
    par.shade_binned = wave && sc.shade_binned ? 1 : 0;
    par.wave_tile = sc.wave_tile;
    par.count_units = c->count_units ? 1 : 0;
    par.bvh_pairs = wave && sc.bvh_pairs ? 1 : 0;
    par.tail_bounce = sc.tail_bounce;
    const int math = c->params.math_mode;
    if (wave && graph) {
      // one graph per pass; the wavefront kernels read neither passes nor clear_first
      dfk_prepare_wavefront();
      DParams p1 = par;
      p1.passes = 1;
      for (int pass = 0; pass < n; ++pass) {
        const int cl = pass == 0 && clear ? 1 : 0, last = pass == n - 1 ? 1 : 0;
        p1.clear_first = cl;
        std::vector<unsigned char> key;
        key_add(key, c->ds);
        key_add(key, cam);
        key_add(key, sky);
        key_add(key, p1);
        key_add(key, f);
        key_add(key, c->paths);
        key_add(key, math);
        DF_HIP(c, graph_launch(c->g_wave[cl][last], key, c->stream, [&] {
          return dfk_launch_pt_wavefront_pass(c->ds, cam, sky, p1, f, c->paths, math, c->stream, cl, last);
        }));
      }
    } else if (wave) {
      DF_HIP(c, dfk_launch_pt_wavefront(c->ds, cam, sky, par, f, c->paths, math, c->stream, mk));
    } else if (graph) {
      std::vector<unsigned char> key;
      key_add(key, c->ds);
      key_add(key, cam);
      key_add(key, sky);
      key_add(key, par);
      key_add(key, f);
      key_add(key, math);
      DF_HIP(c, graph_launch(c->g_mega, key, c->stream,
                             [&] { return dfk_launch_render(c->ds, cam, sky, par, f, math, c->stream, nullptr); }));
    } else {
      DF_HIP(c, dfk_launch_render(c->ds, cam, sky, par, f, math, c->stream, mk));
    }
    return DF_OK;
  };
  DF_HIP(c, hipEventRecord(c->ev0, c->stream));
  int done = 0;
  while (done < passes && tuning(c)) {
    for (hipEvent_t& e : c->tune_ev)
      if (!e) DF_HIP(c, hipEventCreate(&e));
    int w = 0;  // the candidate timed least often so far (the first of them)
    for (int k = 1; k < kTuneCandidates; ++k)
      if (c->tune_runs[k] < c->tune_runs[w]) w = k;
    DF_HIP(c, hipEventRecord(c->tune_ev[0], c->stream));
    if ((rc = run(kCandidates[w], 1, c->clear_pending && done == 0))) return rc;
    DF_HIP(c, hipEventRecord(c->tune_ev[1], c->stream));
    DF_HIP(c, hipEventSynchronize(c->tune_ev[1]));
    float ms = 0.0f;
    DF_HIP(c, hipEventElapsedTime(&ms, c->tune_ev[0], c->tune_ev[1]));
    c->tune_ms[w] = c->tune_runs[w] == 0 ? ms : std::min(c->tune_ms[w], ms);
    ++c->tune_runs[w];
    if (c->tune_runs[w] >= kTunePasses) {
      bool all = true;
      int best = 0;
      for (int k = 0; k < kTuneCandidates; ++k) {
        all = all && c->tune_runs[k] >= kTunePasses;
        if (c->tune_ms[k] < c->tune_ms[best]) best = k;
      }
      if (all) {
        c->sched = kCandidates[best];
        c->tuned = true;
      }
    }
    ++done;
  }
  if (done < passes && (rc = run(current_schedule(c), passes - done, c->clear_pending && done == 0))) return rc;
  DF_HIP(c, hipEventRecord(c->ev1, c->stream));
  if (c->mark_error != hipSuccess) return hip_fail(c, c->mark_error, "stage timing event");
  c->timed = true;
  c->clear_pending = false;  // clearFilmFlags() after the pass
  return DF_OK;
}

int df_get_engine(df_ctx* c, int32_t* engine) {
  if (!c || !engine) return DF_EINVAL;
  *engine = tuning(c) ? DF_PT_AUTO : current_schedule(c).engine;
  return DF_OK;
}

int df_get_schedule(df_ctx* c, df_schedule* out) {
  if (!c || !out) return DF_EINVAL;
  *out = current_schedule(c);
  if (tuning(c)) out->engine = DF_PT_AUTO;
  return DF_OK;
}

int df_get_bvh_pair_stack(df_ctx* c, int32_t* depth) {
  if (!c || !depth) return DF_EINVAL;
  if (!c->has_scene) return fail(c, DF_ESTATE, "no scene loaded");
  *depth = c->ds.pairs ? c->ds.pair_stack : 0;
  return DF_OK;
}

int df_set_schedule(df_ctx* c, const df_schedule* s) {
  if (!c) return DF_EINVAL;
  if (!s) {
    c->sched_fixed = false;
    reset_tuning(c);
    return DF_OK;
  }
  if ((s->engine != DF_PT_MEGAKERNEL && s->engine != DF_PT_WAVEFRONT) ||
      (s->wave_tile != DF_TILE_ROWS && s->wave_tile != DF_TILE_SQUARE) ||
      (s->shade_binned != 0 && s->shade_binned != 1) || (s->bvh_pairs != 0 && s->bvh_pairs != 1) ||
      s->tail_bounce < 0)
    return fail(c, DF_EINVAL, "schedule (engine %d, wave_tile %d, shade_binned %d, bvh_pairs %d, tail_bounce %d)",
                s->engine, s->wave_tile, s->shade_binned, s->bvh_pairs, s->tail_bounce);
  c->sched = *s;
  c->sched_fixed = true;
  return DF_OK;
}

int df_synchronize(df_ctx* c) {
  if (!c) return DF_EINVAL;
  DF_HIP(c, hipStreamSynchronize(c->stream));
  return DF_OK;
}

int df_last_execute_ms(df_ctx* c, float* ms) {
  if (!c || !ms) return DF_EINVAL;
  if (!c->timed) return fail(c, DF_ESTATE, "nothing executed yet");
  DF_HIP(c, hipEventSynchronize(c->ev1));
  DF_HIP(c, hipEventElapsedTime(ms, c->ev0, c->ev1));
  return DF_OK;
}

int df_get_image_bgra(df_ctx* c, uint8_t* out, size_t n) {
  int rc = need_frame(c);
  if (rc) return rc;
  if (!out || n != pixels(c) * 4)
    return fail(c, DF_EINVAL, "resolutionX * resolutionY * 4 != length: %d, %d, %zu", c->width, c->height, n);
  return read_back(c, out, c->image_bgra, n);
}

int df_get_image_float(df_ctx* c, float* out, size_t n) {
  int rc = need_frame(c);
  if (rc) return rc;
  if (!out || n != pixels(c) * 3) return fail(c, DF_EINVAL, "image float length %zu != %zu", n, pixels(c) * 3);
  return read_back(c, out, c->image_float, n * sizeof(float));
}

int df_get_primitive_offsets(df_ctx* c, int32_t* out, size_t n) {
  int rc = need_frame(c);
  if (rc) return rc;
  if (!out || n != pixels(c)) return fail(c, DF_EINVAL, "primitive offsets length %zu != %zu", n, pixels(c));
  return read_back(c, out, c->prim_offsets, n * sizeof(int32_t));
}

int df_get_film(df_ctx* c, float* rgb, size_t rgb_n, int32_t* samples, size_t samples_n) {
  int rc = need_frame(c);
  if (rc) return rc;
  if (!rgb || rgb_n != pixels(c) * 3 || !samples || samples_n != pixels(c))
    return fail(c, DF_EINVAL, "film lengths %zu / %zu do not match %zu pixels", rgb_n, samples_n, pixels(c));
  rc = read_back(c, rgb, c->film, rgb_n * sizeof(float));
  if (rc) return rc;
  return read_back(c, samples, c->film_samples, samples_n * sizeof(int32_t));
}

int df_set_film(df_ctx* c, const float* rgb, size_t rgb_n, const int32_t* samples, size_t samples_n) {
  int rc = need_frame(c);
  if (rc) return rc;
  if (!rgb || rgb_n != pixels(c) * 3 || !samples || samples_n != pixels(c))
    return fail(c, DF_EINVAL, "film lengths %zu / %zu do not match %zu pixels", rgb_n, samples_n, pixels(c));
  DF_HIP(c, hipStreamSynchronize(c->stream));
  DF_HIP(c, hipMemcpy(c->film.p, rgb, rgb_n * sizeof(float), hipMemcpyHostToDevice));
  DF_HIP(c, hipMemcpy(c->film_samples.p, samples, samples_n * sizeof(int32_t), hipMemcpyHostToDevice));
  return DF_OK;
}

// Sums the per-wave copies of the device counters: out[stage][counter].
static int read_counters(df_ctx* c, uint64_t out[DF_STAGE_COUNT][8]) {
  std::vector<uint64_t> raw(kCounterBytes / sizeof(uint64_t));
  DF_HIP(c, hipMemcpyAsync(raw.data(), c->counters.p, kCounterBytes, hipMemcpyDeviceToHost, c->stream));
  DF_HIP(c, hipStreamSynchronize(c->stream));
  for (int st = 0; st < DF_STAGE_COUNT; ++st)
    for (int i = 0; i < 8; ++i) {
      uint64_t v = 0;
      for (int s = 0; s < dfk::kCounterSlots; ++s) v += raw[(static_cast<size_t>(st) * dfk::kCounterSlots + s) * 8 + i];
      out[st][i] = v;
    }
  return DF_OK;
}

int df_get_counters(df_ctx* c, df_counters* out) {
  if (!c || !out) return DF_EINVAL;
  DF_HIP(c, hipStreamSynchronize(c->stream));
  uint64_t all[DF_STAGE_COUNT][8];
  if (const int rc = read_counters(c, all)) return rc;
  uint64_t sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int st = 0; st < DF_STAGE_COUNT; ++st)
    for (int i = 0; i < 8; ++i) sum[i] += all[st][i];
  std::memcpy(out, sum, sizeof sum);
  return DF_OK;
}

int df_reset_counters(df_ctx* c) {
  if (!c) return DF_EINVAL;
  DF_HIP(c, hipStreamSynchronize(c->stream));
  DF_HIP(c, hipMemset(c->counters.p, 0, kCounterBytes));
  c->spans.clear();
  c->ev_used = 0;
  for (int st = 0; st < DF_STAGE_COUNT; ++st) {
    c->stage_ms[st] = 0.0;
    c->stage_launches[st] = 0;
  }
  return DF_OK;
}

int df_set_graphs(df_ctx* c, int32_t enabled) {
  if (!c) return DF_EINVAL;
  DF_HIP(c, hipSetDevice(c->device));
  DF_HIP(c, hipStreamSynchronize(c->stream));
  c->graphs = enabled != 0;
  if (!c->graphs) {
    for (auto& row : c->g_wave)
      for (GraphCache& g : row) g.reset();
    c->g_mega.reset();
  }
  return DF_OK;
}

int df_check_terrain_march(int32_t device, int32_t math_mode, const float* rays, size_t n, float* t_skip,
                           float* t_plain, uint32_t* evals) {
  if (!rays || !t_skip || !t_plain || !evals) return DF_EINVAL;
  if (math_mode != DF_MATH_FAST && math_mode != DF_MATH_EXACT) return DF_EINVAL;
  if (n == 0) return DF_OK;
  if (n > (size_t{1} << 26)) return DF_EINVAL;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return DF_ENODEVICE;
  if (device < 0 || device >= count) return DF_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return DF_EHIP;
  const std::vector<float> tab = terrain_march_positions();
  float *d_rays = nullptr, *d_tab = nullptr, *d_skip = nullptr, *d_plain = nullptr;
  uint32_t* d_evals = nullptr;
  int rc = DF_OK;
  if (hipMalloc(&d_rays, n * 6 * sizeof(float)) != hipSuccess || hipMalloc(&d_tab, tab.size() * sizeof(float)) != hipSuccess ||
      hipMalloc(&d_skip, n * sizeof(float)) != hipSuccess || hipMalloc(&d_plain, n * sizeof(float)) != hipSuccess ||
      hipMalloc(&d_evals, n * sizeof(uint32_t)) != hipSuccess) {
    rc = DF_ENOMEM;
  } else if (hipMemcpy(d_rays, rays, n * 6 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(d_tab, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
             dfk_launch_terrain_check(d_rays, static_cast<int64_t>(n), d_tab, static_cast<int>(tab.size()), d_skip,
                                      d_plain, d_evals, math_mode, nullptr) != hipSuccess ||
             hipMemcpy(t_skip, d_skip, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
             hipMemcpy(t_plain, d_plain, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
             hipMemcpy(evals, d_evals, n * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess) {
    rc = DF_EHIP;
  }
  for (void* p : {static_cast<void*>(d_rays), static_cast<void*>(d_tab), static_cast<void*>(d_skip),
                  static_cast<void*>(d_plain), static_cast<void*>(d_evals)})
    if (p) (void)hipFree(p);
  return rc;
}

int df_check_reciprocal(int32_t device, uint64_t* tested, uint64_t* mismatches, uint32_t* first_mismatch) {
  if (!tested || !mismatches || !first_mismatch) return DF_EINVAL;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return DF_ENODEVICE;
  if (device < 0 || device >= count) return DF_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return DF_EHIP;
  // 2^-14 < |x| < 2^126: bit patterns (0x38800000, 0x7E800000) without the sign
  const uint32_t lo = 0x38800001u, hi = 0x7E800000u;
  unsigned long long* d = nullptr;
  if (hipMalloc(&d, 3 * sizeof(unsigned long long)) != hipSuccess) return DF_ENOMEM;
  const unsigned long long init[3] = {0ull, 0ull, 0xFFFFFFFFull};
  int rc = DF_OK;
  unsigned long long out[3] = {0ull, 0ull, 0ull};
  if (hipMemcpy(d, init, sizeof init, hipMemcpyHostToDevice) != hipSuccess ||
      dfk_launch_reciprocal_check(lo, hi, d, d + 1, reinterpret_cast<uint32_t*>(d + 2), nullptr) != hipSuccess ||
      hipMemcpy(out, d, sizeof out, hipMemcpyDeviceToHost) != hipSuccess)
    rc = DF_EHIP;
  (void)hipFree(d);
  *tested = out[0];
  *mismatches = out[1];
  *first_mismatch = static_cast<uint32_t>(out[2]);
  return rc;
}

int df_set_unit_counters(df_ctx* c, int32_t enabled) {
  if (!c) return DF_EINVAL;
  c->count_units = enabled != 0;
  return DF_OK;
}

int df_set_stage_timing(df_ctx* c, int32_t enabled) {
  if (!c) return DF_EINVAL;
  DF_HIP(c, hipStreamSynchronize(c->stream));
  int rc = collect_spans(c);
  if (rc) return rc;
  c->stage_timing = enabled != 0;
  return DF_OK;
}

int df_get_stage_stats(df_ctx* c, int32_t stage, df_counters* counters, double* ms, int64_t* launches) {
  if (!c) return DF_EINVAL;
  if (stage < 0 || stage >= DF_STAGE_COUNT) return fail(c, DF_EINVAL, "stage %d", stage);
  DF_HIP(c, hipStreamSynchronize(c->stream));
  int rc = collect_spans(c);
  if (rc) return rc;
  if (counters) {
    uint64_t all[DF_STAGE_COUNT][8];
    if (const int r = read_counters(c, all)) return r;
    std::memcpy(counters, all[stage], sizeof(df_counters));
  }
  if (ms) *ms = c->stage_ms[stage];
  if (launches) *launches = c->stage_launches[stage];
  return DF_OK;
}

const char* df_last_error(df_ctx* c) {
  if (!c) return "null handle";
  return c->err.c_str();
}

// ------------------------------------------------------------ differentiable path tracer

// Whether the sky the renderer holds is Sky.set of (turbidity, elevation,
// azimuth): its sun direction within 1e-5 of the parameters' and its
// turbidity within 1e-5 relative.
bool sky_matches(const df_sky& k, const double q[3]) {
  const double deg = M_PI / 180.0;
  const double e = q[1] * deg, a = q[2] * deg;
  const double w[3] = {std::cos(e) * std::sin(a), std::sin(e), std::cos(e) * std::cos(a)};
  for (int i = 0; i < 3; ++i)
    if (!(std::fabs(w[i] - k.sun_direction_world[i]) <= 1e-5)) return false;
  return std::fabs(q[0] - k.turbidity) <= 1e-5 * std::fabs(q[0]);
}

int df_set_diff(df_ctx* c, const df_diff* d) {
  if (!c || !d) return DF_EINVAL;
  if (!c->has_scene) return fail(c, DF_ESTATE, "no scene loaded");
  const size_t n_surf = c->hs.surfaces.size() / 6;
  if (d->n_surfaces != n_surf || !d->surface_albedo || !d->surface_emission)
    return fail(c, DF_EINVAL, "surface maps must hold one entry per surface record (%zu)", n_surf);
  const int32_t nc = d->n_checker;
  if (d->n_albedo < 0 || d->n_emission < 0 || nc < 0 ||
      3 * d->n_albedo + d->n_emission + 6 * nc + 3 > dfk::kMaxDiffParams)
    return fail(c, DF_EINVAL, "too many parameters (%d albedo groups, %d emissions, %d checkerboard groups)",
                d->n_albedo, d->n_emission, nc);
  if ((d->n_albedo > 0 && !d->albedo) || (d->n_emission > 0 && !d->emission) ||
      (nc > 0 && (!d->checker || !d->surface_checker)))
    return fail(c, DF_EINVAL, "parameter arrays missing");
  const std::vector<float>& tex = c->hs.textures;
  const std::vector<float>& surf = c->hs.surfaces;
  for (size_t i = 0; i < n_surf; ++i) {
    if (d->surface_albedo[i] < -1 || d->surface_albedo[i] >= d->n_albedo)
      return fail(c, DF_EINVAL, "surface %zu: albedo group %d", i, d->surface_albedo[i]);
    if (d->surface_emission[i] < -1 || d->surface_emission[i] >= d->n_emission)
      return fail(c, DF_EINVAL, "surface %zu: emission index %d", i, d->surface_emission[i]);
    const int32_t g = nc > 0 ? d->surface_checker[i] : -1;
    if (g < -1 || g >= nc) return fail(c, DF_EINVAL, "surface %zu: checkerboard group %d", i, g);
    if (g >= 0) {
      if (d->surface_albedo[i] >= 0)
        return fail(c, DF_EINVAL, "surface %zu: both an albedo group and a checkerboard group", i);
      if (!checker_albedo(tex, surf.data(), i))
        return fail(c, DF_EINVAL, "surface %zu: a checkerboard group needs a checkerboard albedo texture", i);
    }
  }
  const double q[3] = {d->turbidity, d->elevation_deg, d->azimuth_deg};
  if (!sky_matches(c->sky, q))
    return fail(c, DF_EINVAL, "the current sky is not Sky.set of turbidity %g, elevation %g, azimuth %g", q[0], q[1],
                q[2]);
  DF_HIP(c, hipSetDevice(c->device));
  DF_HIP(c, hipStreamSynchronize(c->stream));
  const std::vector<int32_t> sa(d->surface_albedo, d->surface_albedo + n_surf);
  const std::vector<int32_t> se(d->surface_emission, d->surface_emission + n_surf);
  std::vector<int32_t> sc(n_surf, -1);
  if (nc > 0) sc.assign(d->surface_checker, d->surface_checker + n_surf);
  // albedo slots: the groups, then each checkerboard group's two colours
  std::vector<float> al(d->albedo, d->albedo + 3 * d->n_albedo);
  if (nc > 0) al.insert(al.end(), d->checker, d->checker + 6 * nc);
  if (al.empty()) al.assign(3, 0.0f);
  std::vector<float> em(d->emission, d->emission + d->n_emission);
  if (em.empty()) em.assign(1, 0.0f);
  const std::vector<dfk::DSkyDual> dual = {
      dfk::sky_set_dual(c->sky.onb_u, c->sky.onb_v, c->sky.onb_w, q[0], q[1], q[2])};
  int rc;
  if ((rc = upload(c, c->diff_surf_albedo, sa)) || (rc = upload(c, c->diff_surf_emission, se)) ||
      (rc = upload(c, c->diff_surf_checker, sc)) || (rc = upload(c, c->diff_albedo, al)) ||
      (rc = upload(c, c->diff_emission, em)) || (rc = upload(c, c->diff_sky_dual, dual)))
    return rc;
  c->diff_n_surfaces = n_surf;
  c->diff_n_albedo = d->n_albedo;
  c->diff_n_emission = d->n_emission;
  c->diff_n_checker = nc;
  c->diff_checker_host = sc;
  std::copy(q, q + 3, c->diff_sky_params);
  c->diff_on = true;
  return DF_OK;
}

int df_clear_diff(df_ctx* c) {
  if (!c) return DF_EINVAL;
  c->diff_on = false;
  return DF_OK;
}

int df_loss_gradient(df_ctx* c, const float* target, size_t n, int32_t passes, double* loss, double* grad,
                     size_t n_grad, float* forward_ms, float* adjoint_ms) {
  int rc = need_frame(c);
  if (rc) return rc;
  if (!c->diff_on) return fail(c, DF_ESTATE, "df_set_diff has not been called");
  if (!target || n != pixels(c) * 3) return fail(c, DF_EINVAL, "target length %zu != %zu", n, pixels(c) * 3);
  if (passes < 1) return fail(c, DF_EINVAL, "passes %d", passes);
  const df_params& p = c->params;
  if (p.renderer_type != DF_RENDERER_PATH_TRACER) return fail(c, DF_EINVAL, "the adjoint needs the path tracer");
  if (grad) {
    const size_t need = static_cast<size_t>(3 * c->diff_n_albedo + c->diff_n_emission + 6 * c->diff_n_checker + 3);
    if (n_grad != need) return fail(c, DF_EINVAL, "gradient length %zu != %zu", n_grad, need);
    if (!sky_matches(c->sky, c->diff_sky_params))
      return fail(c, DF_ESTATE, "the sky changed after df_set_diff: its parameters are no longer the diff's");
    if (p.pt_depth_maximum > dfk::kMaxAdjointDepth)
      return fail(c, DF_EINVAL, "the adjoint supports pt_depth_maximum <= %d", dfk::kMaxAdjointDepth);
    if (p.pt_depth_russian_roulette <= p.pt_depth_maximum)
      return fail(c, DF_EINVAL, "the adjoint needs Russian roulette off (pt_depth_russian_roulette > pt_depth_maximum)");
  }
  DF_HIP(c, hipSetDevice(c->device));
  DF_HIP(c, hipStreamSynchronize(c->stream));  // nothing queued still uses the buffers below
  const size_t np = pixels(c);
  if (c->diff_weight.bytes != np * 3 * sizeof(float) &&
      (rc = alloc_buf(c, c->diff_weight, np * 3 * sizeof(float))))
    return rc;
  if (!c->diff_loss.p && (rc = alloc_buf(c, c->diff_loss, sizeof(double)))) return rc;
  if (c->diff_seeds.bytes != np * sizeof(uint64_t) && (rc = alloc_buf(c, c->diff_seeds, np * sizeof(uint64_t))))
    return rc;
  // the forward pass starts from these seeds; the replay starts from them again
  DF_HIP(c, hipMemcpyAsync(c->diff_seeds.p, c->seeds.p, np * sizeof(uint64_t), hipMemcpyDeviceToDevice, c->stream));
  c->clear_pending = true;
  if ((rc = df_execute(c, passes))) return rc;
  float fms = 0.0f;
  DF_HIP(c, hipEventSynchronize(c->ev1));
  DF_HIP(c, hipEventElapsedTime(&fms, c->ev0, c->ev1));
  if (forward_ms) *forward_ms = fms;
  // every copy below is ordered on the renderer's stream (a plain hipMemcpy
  // would not wait for it: the stream is non-blocking)
  DF_HIP(c, hipMemcpyAsync(c->diff_weight.p, target, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
  DF_HIP(c, hipMemsetAsync(c->diff_loss.p, 0, sizeof(double), c->stream));
  // weights in place: dLoss/dfilm / passes
  DF_HIP(c, dfk_launch_loss(static_cast<const float*>(c->film.p), static_cast<const float*>(c->diff_weight.p),
                            static_cast<float*>(c->diff_weight.p), static_cast<int>(n), 1.0f / static_cast<float>(passes),
                            static_cast<double*>(c->diff_loss.p), c->stream));
  if (grad) {
    const size_t ng = n_grad;
    if (c->diff_grad.bytes != ng * sizeof(double) && (rc = alloc_buf(c, c->diff_grad, ng * sizeof(double)))) return rc;
    DF_HIP(c, hipMemsetAsync(c->diff_grad.p, 0, ng * sizeof(double), c->stream));
    // replay from the forward's starting seeds; the replay leaves the seeds as the forward did
    DF_HIP(c, hipMemcpyAsync(c->seeds.p, c->diff_seeds.p, np * sizeof(uint64_t), hipMemcpyDeviceToDevice, c->stream));
    dfk::DAdjoint a{};
    a.sky = static_cast<const dfk::DSkyDual*>(c->diff_sky_dual.p);
    a.weight = static_cast<const float*>(c->diff_weight.p);
    a.grad = static_cast<double*>(c->diff_grad.p);
    a.n_albedo = c->diff_n_albedo;
    a.n_emission = c->diff_n_emission;
    a.n_checker = c->diff_n_checker;
    const DCamera cam = make_camera(c->camera);
    const DSky sky = make_sky(c->sky);
    DParams par = make_params(c->params, c->width, c->height, passes, false);
    apply_diff(c, par);
    DF_HIP(c, hipEventRecord(c->ev0, c->stream));
    DF_HIP(c, dfk_launch_adjoint(c->ds, cam, sky, par, frame_of(c), a, p.math_mode, c->stream));
    DF_HIP(c, hipEventRecord(c->ev1, c->stream));
    DF_HIP(c, hipEventSynchronize(c->ev1));
    float ams = 0.0f;
    DF_HIP(c, hipEventElapsedTime(&ams, c->ev0, c->ev1));
    if (adjoint_ms) *adjoint_ms = ams;
    DF_HIP(c, hipMemcpyAsync(grad, c->diff_grad.p, ng * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  double l = 0.0;
  DF_HIP(c, hipMemcpyAsync(&l, c->diff_loss.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  DF_HIP(c, hipStreamSynchronize(c->stream));
  if (loss) *loss = l;
  return DF_OK;
}

// ------------------------------------------------------------ batches of scenes

int df_batch_create(df_batch** out, int32_t device) {
  if (!out) return DF_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return DF_ENODEVICE;
  if (device < 0 || device >= n) return DF_ENODEVICE;
  df_batch* b = new (std::nothrow) df_batch;
  if (!b) return DF_ENOMEM;
  b->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
    df_batch_destroy(b);
    return DF_EHIP;
  }
  // set-up of the allocation-free entry point: this device's scratch address
  unsigned char* sp = nullptr;
  if (scratch_address(device, &sp) != DF_OK) {
    df_batch_destroy(b);
    return DF_EHIP;
  }
  *out = b;
  return DF_OK;
}

void df_batch_destroy(df_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (df_ctx* c : b->jobs) df_destroy(c);
  b->g_launch.reset();
  free_buf(b->djobs);
  free_buf(b->ditems);
  free_buf(b->dheads);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int df_batch_add(df_batch* b, const df_compiled_scene* scene, const df_sky* sky, const df_params* params,
                 int32_t width, int32_t height, int32_t passes, uint64_t seed, int32_t* index) {
  if (!b || !scene || !sky || !params) return DF_EINVAL;
  if (passes < 0) return bfail(b, DF_EINVAL, "passes %d", passes);
  df_ctx* c = nullptr;
  int rc = df_create(&c, b->device);
  if (rc) return bfail(b, rc, "df_create failed (%d)", rc);
  if ((rc = df_load_scene(c, scene)) || (rc = df_upload_sky(c, sky)) || (rc = df_set_params(c, params)) ||
      (rc = df_update(c, width, height, 0)) || (rc = df_seed(c, seed))) {
    bfail(b, rc, "job %zu: %s", b->jobs.size(), df_last_error(c));
    df_destroy(c);
    return rc;
  }
  b->jobs.push_back(c);
  b->passes.push_back(passes);
  if (index) *index = static_cast<int32_t>(b->jobs.size() - 1);
  return DF_OK;
}

int df_batch_count(df_batch* b, int32_t* count) {
  if (!b || !count) return DF_EINVAL;
  *count = static_cast<int32_t>(b->jobs.size());
  return DF_OK;
}

int df_batch_job(df_batch* b, int32_t index, df_ctx** ctx) {
  if (!b || !ctx) return DF_EINVAL;
  if (index < 0 || index >= static_cast<int32_t>(b->jobs.size())) return bfail(b, DF_EINVAL, "job %d", index);
  DF_BHIP(b, hipStreamSynchronize(b->stream));
  *ctx = b->jobs[index];
  return DF_OK;
}

int df_batch_execute(df_batch* b) {
  if (!b) return DF_EINVAL;
  const size_t n = b->jobs.size();
  if (n == 0) return DF_OK;
  DF_BHIP(b, hipSetDevice(b->device));
  BatchPlan plan;
  int rc;
  if ((rc = batch_plan(b, plan, true))) return rc;
  if ((rc = balloc(b, b->djobs, n * sizeof(DJob))) ||
      (rc = balloc(b, b->ditems, plan.items.size() * sizeof(DBatchItem))) ||
      (rc = balloc(b, b->dheads, (plan.group_key.size() + 1) * sizeof(int))))
    return rc;
  DF_BHIP(b, hipMemcpy(b->djobs.p, plan.jobs.data(), n * sizeof(DJob), hipMemcpyHostToDevice));
  if (!plan.items.empty())
    DF_BHIP(b, hipMemcpy(b->ditems.p, plan.items.data(), plan.items.size() * sizeof(DBatchItem),
                         hipMemcpyHostToDevice));
  DF_BHIP(b, hipMemset(b->dheads.p, 0, (plan.group_key.size() + 1) * sizeof(int)));
  return batch_launch(b, plan, static_cast<const DJob*>(b->djobs.p), static_cast<const DBatchItem*>(b->ditems.p),
                      static_cast<int*>(b->dheads.p));
}

int df_batch_execute_noalloc(df_batch* b) {
  if (!b) return DF_EINVAL;
  const size_t n = b->jobs.size();
  if (n == 0) return DF_OK;
  if (n > static_cast<size_t>(kNoAllocMaxJobs))
    return bfail(b, DF_EINVAL, "df_batch_execute_noalloc: %zu jobs exceed the maximum of %d", n, kNoAllocMaxJobs);
  DF_BHIP(b, hipSetDevice(b->device));
  unsigned char* sp = nullptr;
  if (scratch_address(b->device, &sp) != DF_OK)
    return bfail(b, DF_EHIP, "df_batch_execute_noalloc: no scratch address on device %d", b->device);
  BatchPlan plan;
  int rc;
  if ((rc = batch_plan(b, plan, false))) return rc;  // the persistent batch kernel only: no path state
  if (plan.items.size() > static_cast<size_t>(kNoAllocMaxItems))
    return bfail(b, DF_EINVAL, "df_batch_execute_noalloc: %zu tiles exceed the maximum of %d", plan.items.size(),
                 kNoAllocMaxItems);
  if (plan.group_key.size() > static_cast<size_t>(kNoAllocMaxGroups))
    return bfail(b, DF_EINVAL, "df_batch_execute_noalloc: %zu launch groups exceed the maximum of %d",
                 plan.group_key.size(), kNoAllocMaxGroups);
  std::lock_guard<std::mutex> lock(g_scratch_call_mu[b->device]);
  int* heads = reinterpret_cast<int*>(sp);
  DJob* jobs = reinterpret_cast<DJob*>(sp + kScratchHeadsBytes);
  DBatchItem* items = reinterpret_cast<DBatchItem*>(sp + kScratchHeadsBytes + kScratchJobsBytes);
  DF_BHIP(b, hipMemcpyAsync(jobs, plan.jobs.data(), n * sizeof(DJob), hipMemcpyHostToDevice, b->stream));
  if (!plan.items.empty())
    DF_BHIP(b, hipMemcpyAsync(items, plan.items.data(), plan.items.size() * sizeof(DBatchItem),
                              hipMemcpyHostToDevice, b->stream));
  hipLaunchKernelGGL(zero_ints, dim3(1), dim3(kNoAllocMaxGroups), 0, b->stream, heads, kNoAllocMaxGroups);
  DF_BHIP(b, hipGetLastError());
  return batch_launch(b, plan, jobs, items, heads);
}

int df_batch_set_graphs(df_batch* b, int32_t enabled) {
  if (!b) return DF_EINVAL;
  DF_BHIP(b, hipSetDevice(b->device));
  DF_BHIP(b, hipStreamSynchronize(b->stream));
  b->graphs = enabled != 0;
  if (!b->graphs) b->g_launch.reset();
  return DF_OK;
}

int df_batch_last_execute_ms(df_batch* b, float* ms) {
  if (!b || !ms) return DF_EINVAL;
  if (!b->timed) return bfail(b, DF_ESTATE, "nothing executed yet");
  DF_BHIP(b, hipEventElapsedTime(ms, b->ev0, b->ev1));
  return DF_OK;
}

const char* df_batch_last_error(df_batch* b) {
  if (!b) return "null handle";
  return b->err.c_str();
}

}  // extern "C"
