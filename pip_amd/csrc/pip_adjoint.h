// pip_adjoint.h -- HIP kernels of the discrete adjoint stage (one work item
// of pip_adj_items.h per thread).  Compiled in pipgpu_adjoint.hip only.
#pragma once

#include <hip/hip_runtime.h>

#include "pip_adj_items.h"

namespace pip {

__global__ void __launch_bounds__(256) k_adj_cell(AdjArgs a) {
  adj_cell_item(a, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// The gather covers the interior cells with 64x4 tiles, one per 256-thread
// workgroup.  Workgroups are dealt round-robin to the 8 XCDs; the tile order
// gives each XCD a contiguous range of tiles (and of face blocks, below), so
// that the rows a tile shares with its neighbours (y stencil) meet in one L2.
constexpr int ADJ_TW = 64, ADJ_TH = 4;

__device__ __forceinline__ int xcd_order(int bid, int nb) {
  const int q = nb / 8, rr = nb % 8, x = bid % 8, k = bid / 8;
  return x * q + (x < rr ? x : rr) + k;
}

__host__ __device__ inline int adj_tiles(int w, int h) {
  return ((w + ADJ_TW - 1) / ADJ_TW) * ((h + ADJ_TH - 1) / ADJ_TH);
}

// tile t of a w x h index space -> this thread's item (x, y); false outside
__device__ __forceinline__ bool adj_tile_item(int t, int w, int h, int& x, int& y) {
  const int ntx = (w + ADJ_TW - 1) / ADJ_TW;
  x = (t % ntx) * ADJ_TW + (int)(threadIdx.x % ADJ_TW);
  y = (t / ntx) * ADJ_TH + (int)(threadIdx.x / ADJ_TW);
  return x < w && y < h;
}

// Faces of both directions of one fluid in one launch, 256 consecutive
// faces (row-major) per workgroup: the first adj_face_blocks(nx+1, ny)
// workgroups take the x faces, the rest the y faces.  Rows of nx+1 x faces
// are not a multiple of 64, so rectangular tiles would leave waves with a
// single live lane running the whole VJP.
__host__ __device__ inline int adj_face_blocks(int w, int h) {
  return (int)(((long long)w * h + 255) / 256);
}

template <int FL>
__global__ void __launch_bounds__(256, FL ? 2 : 1) k_adj_faces(AdjArgs a) {
  const int nbx = adj_face_blocks(a.nx + 1, a.ny);
  int b = xcd_order(blockIdx.x, gridDim.x);
  const int dir = (b < nbx) ? 0 : 1;
  if (dir) b -= nbx;
  const int w = dir ? a.nx : a.nx + 1, h = dir ? a.ny + 1 : a.ny;
  const long long t = (long long)b * 256 + threadIdx.x;
  if (t >= (long long)w * h) return;
  adj_face_item<FL>(a, dir, t);
}

// Interior tiles; in stage 1 (lam_in then receives the step's result) the
// blocks past them zero the ghost ring of lam_in, adj_ghost_blocks of them.
__host__ __device__ inline int adj_ghost_blocks(int ix, int jx) {
  return (2 * ix + 2 * jx - 8 + 255) / 128;  // 4 ix + 4 (jx - 4) ring cells, 256 per block
}

__global__ void __launch_bounds__(256) k_adj_gather(AdjArgs a) {
  const int nt = adj_tiles(a.nx, a.ny);
  if ((int)blockIdx.x >= nt) {
    const int g = ((int)blockIdx.x - nt) * 256 + (int)threadIdx.x;
    int i, j;
    if (g < 4 * a.ix) {
      const int r = g / a.ix;
      j = r < 2 ? r : a.jx - 4 + r;
      i = g % a.ix;
    } else {
      const int h = g - 4 * a.ix, c = h % 4;
      j = 2 + h / 4;
      i = c < 2 ? c : a.ix - 4 + c;
      if (j >= a.jx - 2) return;
    }
    const long long k = (long long)j * a.ix + i;
#pragma unroll
    for (int n = 0; n < ADJ_NVT; n++) a.lam_in[n * a.N + k] = 0.0;
    return;
  }
  int x, y;
  if (!adj_tile_item(xcd_order(blockIdx.x, nt), a.nx, a.ny, x, y)) return;
  adj_gather_item(a, (long long)(y + 2) * a.ix + (x + 2));
}

__global__ void __launch_bounds__(256) k_adj_fold_seed(double* lam, int nvt, int ix, int jx, long long N,
                                                       Bnd bnd) {
  adj_fold_seed_item(lam, nvt, ix, jx, N, bnd, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void __launch_bounds__(256) k_zero_ghosts(double* x, int nvt, int ix, int jx, long long N) {
  zero_ghosts_item(x, nvt, ix, jx, N, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}


}  // namespace pip
