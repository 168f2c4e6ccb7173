// pip_adj_launch.h -- interface between the C-ABI (pipgpu.hip) and the
// adjoint kernels, which are compiled in their own translation unit
// (pipgpu_adjoint.hip) so that each unit can take its own code-generation
// flags.  Internal to libpipgpu; not part of include/pipgpu.h.
#pragma once

#include <hip/hip_runtime.h>

#include "pip_adj_items.h"

namespace pip {

// One reverse SSP-RK stage on `stream`: the primitives of the stage input
// when a.make_V (and then the stage-1 ghost zeroing), the face VJPs of the
// plasma and of the neutrals over the x and y faces, and the gather over the
// interior (whose extra blocks zero the stage-1 ghost ring when the
// primitives were saved).  N = ix * jx.
hipError_t adj_stage_launch(const AdjArgs& a, hipStream_t stream);

// The adjoint seed: fold the ghost entries of lam onto the cells they image
// (the boundary map's transpose), then zero the ghost ring.
hipError_t adj_seed_launch(double* lam, int nvt, int ix, int jx, long long N, const Bnd& bnd, hipStream_t stream);

}  // namespace pip
