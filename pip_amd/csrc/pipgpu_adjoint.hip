// pipgpu_adjoint.hip -- the adjoint kernels of pip_adjoint.h and their
// launches (pip_adj_launch.h).  Compiled on its own, with the exact build's
// contraction flag (-ffp-contract=off; the VJP bodies enable contraction
// themselves) and the compiler's default instruction scheduling.
#include <hip/hip_runtime.h>

#include "pip_adj_launch.h"
#include "pip_adjoint.h"

namespace pip {

hipError_t adj_stage_launch(const AdjArgs& a, hipStream_t stream) {
  const dim3 cells((unsigned)((a.N + 255) / 256));
  if (a.make_V) hipLaunchKernelGGL(k_adj_cell, cells, dim3(256), 0, stream, a);
  const dim3 gf((unsigned)(adj_plasma_face_groups(a.nx, a.ny) + adj_edge_face_blocks(a.nx, a.ny)));
  hipLaunchKernelGGL(k_adj_plasma_faces, gf, dim3(ADJ_PBS), 0, stream, a);
  const int ghost = (a.istep == 1 && !a.make_V) ? adj_ghost_blocks(a.ix, a.jx) : 0;
  hipLaunchKernelGGL(k_adj_gather, dim3((unsigned)(adj_gather_tiles(a.nx, a.ny) + ghost)), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t adj_seed_launch(double* lam, int nvt, int ix, int jx, long long N, const Bnd& bnd, hipStream_t stream) {
  const dim3 cells((unsigned)((N + 255) / 256));
  hipLaunchKernelGGL(k_adj_fold_seed, cells, dim3(256), 0, stream, lam, nvt, ix, jx, N, bnd);
  hipLaunchKernelGGL(k_zero_ghosts, cells, dim3(256), 0, stream, lam, nvt, ix, jx, N);
  return hipGetLastError();
}

}  // namespace pip
