// Launch wrappers exported by df_kernels.hip to the host runtime.
#pragma once

#include <hip/hip_runtime.h>

#include "df_device_scene.h"

hipError_t dfk_launch_render(const dfk::DScene& s, const dfk::DCamera& c, const dfk::DSky& k, const dfk::DParams& p,
                             const dfk::DFrame& f, int math_mode, hipStream_t stream);

hipError_t dfk_launch_seed(uint64_t* seeds, uint64_t n, uint64_t master, hipStream_t stream);
