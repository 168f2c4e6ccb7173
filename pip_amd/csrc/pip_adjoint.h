// pip_adjoint.h -- HIP kernels of the discrete adjoint stage (one work item
// of pip_adj_items.h per thread) and the step bookkeeping of recorded and
// replayed runs.
#pragma once

#include "pip_adj_items.h"
#include "pip_kernels.h"

namespace pip {

__global__ void __launch_bounds__(256) k_adj_cell(AdjArgs a) {
  adj_cell_item(a, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

template <int FL>
__global__ void __launch_bounds__(256) k_adj_faces(AdjArgs a, int dir) {
  adj_face_item<FL>(a, dir, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void __launch_bounds__(256) k_adj_gather(AdjArgs a) {
  adj_gather_item(a, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void __launch_bounds__(256) k_adj_fold_seed(double* lam, int nvt, int ix, int jx, long long N,
                                                       Bnd bnd) {
  adj_fold_seed_item(lam, nvt, ix, jx, N, bnd, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void __launch_bounds__(256) k_zero_ghosts(double* x, int nvt, int ix, int jx, long long N) {
  zero_ghosts_item(x, nvt, ix, jx, N, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}


// Step bookkeeping for recorded / replayed runs: hist[2k] = dt, hist[2k+1]
// = cmax of step k.
__global__ void k_record_dt(const DevScalars* sc, double* hist, int k) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  hist[2 * k] = sc->active ? sc->dt : 0.0;
  hist[2 * k + 1] = sc->cmax;
}
__global__ void k_set_dt(DevScalars* sc, const double* hist, int k) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  sc->dt = hist[2 * k];
  sc->cmax = hist[2 * k + 1];
  sc->active = 1;
  sc->last = 0;
}

}  // namespace pip
