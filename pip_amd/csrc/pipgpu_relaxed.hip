// pipgpu_relaxed.hip -- the fused SSP-RK stage kernel (k_stage of
// pip_kernels.h) built a second time in relaxed arithmetic, for
// pip_set_arith(ctx, PIP_ARITH_RELAXED).
//
// This file is compiled on its own with FMA contraction (-ffp-contract=fast)
// and PIP_RELAXED = 1, which makes the device quotients products with the
// refined reciprocal (no residual correction, no fixup), the divisions and
// square roots of the get_Pr / set_IR / cfl terms the same sequences, and
// min / max the hardware instructions.  Operations, their order, tiles,
// halos and boundary maps are the exact kernel's; the headers are compiled
// in namespace pipr so that both builds live in one library.
#include <hip/hip_runtime.h>
#include <string.h>

#define PIP_NS pipr
#define PIP_RELAXED 1
#include "pip_kernels.h"
#include "pip_relaxed.h"

namespace pipr {

template <int NDIM, int NV, bool IR, int TX, int TY, int NT>
static int launch(const StageArgs& a, hipStream_t stream, const char** err) {
  constexpr size_t lds = sizeof(double) * (size_t)TileGeom<NDIM, NV, TX, TY>::LDS_DOUBLES;
  auto kern = k_stage<NDIM, NV, IR, false, TX, TY, NT, false>;
  static bool attr_done = false;
  if (!attr_done) {
    const hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) {
      *err = hipGetErrorString(e);
      return -3;
    }
    attr_done = true;
  }
  const int ntx = (a.nx + TX - 1) / TX;
  const int nty = (NDIM == 2) ? (a.ny + TY - 1) / TY : 1;
  hipLaunchKernelGGL(kern, dim3(ntx * nty), dim3(NT), lds, stream, a);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    *err = hipGetErrorString(e);
    return -3;
  }
  return 0;
}

template <int NV, bool IR>
static int launch_2d(const StageArgs& a, int tx, int ty, hipStream_t s, const char** err) {
  if (tx == 16 && ty == 15) return launch<2, NV, IR, 16, 15, 256>(a, s, err);
  if (tx == 32 && ty == 7) return launch<2, NV, IR, 32, 7, 256>(a, s, err);
  if (tx == 16 && ty == 7) return launch<2, NV, IR, 16, 7, 128>(a, s, err);
  return -1;
}

}  // namespace pipr

int pip_relaxed_covers(int ndim, int nv, int ir, int tx, int ty) {
  if (ndim == 1) return nv == 8 && !ir;
  const bool tile = (tx == 16 && ty == 15) || (tx == 32 && ty == 7) || (tx == 16 && ty == 7);
  return ndim == 2 && nv == 9 && tile;
}

int pip_relaxed_stage(const void* args, size_t args_bytes, int ndim, int nv, int ir, int tx, int ty,
                      void* stream, const char** err) {
  if (args_bytes != sizeof(pipr::StageArgs)) return -2;
  if (!pip_relaxed_covers(ndim, nv, ir, tx, ty)) return -1;
  pipr::StageArgs a;
  memcpy(&a, args, sizeof(a));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (ndim == 1) return pipr::launch<1, 8, false, 255, 1, 256>(a, s, err);
  return ir ? pipr::launch_2d<9, true>(a, tx, ty, s, err) : pipr::launch_2d<9, false>(a, tx, ty, s, err);
}
