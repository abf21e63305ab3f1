// Launch wrappers exported by df_kernels.hip and df_wavefront.hip to the host runtime.
#pragma once

#include <hip/hip_runtime.h>

#include "df_device_scene.h"

// Optional hook around each kernel launch (stage = DF_STAGE_*): begin = 1 is
// called just before the launch is enqueued, begin = 0 just after.
struct dfk_marks {
  void (*mark)(void* user, int stage, int begin);
  void* user;
};

hipError_t dfk_launch_render(const dfk::DScene& s, const dfk::DCamera& c, const dfk::DSky& k, const dfk::DParams& p,
                             const dfk::DFrame& f, int math_mode, hipStream_t stream,
                             const dfk_marks* marks = nullptr);

// Persistent launch over a batch's (job, tile) items of one math mode and
// renderer type; head must be zero.
hipError_t dfk_launch_batch(const dfk::DJob* jobs, const dfk::DBatchItem* items, int n_items, int* head,
                            int math_mode, int renderer_type, hipStream_t stream);

hipError_t dfk_launch_seed(uint64_t* seeds, uint64_t n, uint64_t master, hipStream_t stream);

// recip_rn against the division for every float with bit pattern in [lo, hi)
// (both signs): counts into tested / mismatches, smallest mismatch into first
// (df_check_reciprocal).
hipError_t dfk_launch_reciprocal_check(uint32_t lo, uint32_t hi, unsigned long long* tested,
                                       unsigned long long* mismatches, uint32_t* first, hipStream_t stream);

// n rays (6 floats each) through the terrain march of the kernels and
// through the reference's march over every position (df_check_terrain_march).
hipError_t dfk_launch_terrain_check(const float* rays, int64_t n, const float* tc_tab, int tn, float* t_skip,
                                    float* t_plain, uint32_t* evals, int math_mode, hipStream_t stream);

// One or more passes of the path tracer as wavefront kernels (df_wavefront.hip);
// w must hold at least width * height slots and wavefront_ctl_len(depth) counters.
hipError_t dfk_launch_pt_wavefront(const dfk::DScene& s, const dfk::DCamera& c, const dfk::DSky& k,
                                   const dfk::DParams& p, const dfk::DFrame& f, const dfk::DPaths& w, int math_mode,
                                   hipStream_t stream, const dfk_marks* marks = nullptr);

// Queries the wavefront kernels' persistent grid sizes (done once; call before
// capturing a pass into a graph).
void dfk_prepare_wavefront();

// One pass of the same (the launch sequence a frame loop captures in a graph):
// `clear` overwrites the film, `last` ends with the image pass.
hipError_t dfk_launch_pt_wavefront_pass(const dfk::DScene& s, const dfk::DCamera& c, const dfk::DSky& k,
                                        const dfk::DParams& p, const dfk::DFrame& f, const dfk::DPaths& w,
                                        int math_mode, hipStream_t stream, int clear, int last);

// Differentiable path tracer (df_adjoint.hip): dLoss/dfilm weights and the
// loss of a film against a target, then the path-replay adjoint.
hipError_t dfk_launch_loss(const float* film, const float* target, float* weight, int n, float scale, double* loss,
                           hipStream_t stream);
hipError_t dfk_launch_adjoint(const dfk::DScene& s, const dfk::DCamera& c, const dfk::DSky& k, const dfk::DParams& p,
                              const dfk::DFrame& f, const dfk::DAdjoint& a, int math_mode, hipStream_t stream);
