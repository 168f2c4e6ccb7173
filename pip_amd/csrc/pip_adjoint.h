// pip_adjoint.h -- HIP kernels of the discrete adjoint stage (the work
// items of pip_adj_items.h).  Compiled in pipgpu_adjoint.hip only.
//
// A stage is two launches (three when the primitives of the stage input were
// not saved by the replay):
//   k_adj_plasma_faces  the plasma's face VJPs (HLLD with GLM) over every
//                       face, written as (left, right) adjoint pairs, plus
//                       the neutrals' VJPs at the faces next to the boundary;
//   k_adj_gather        per tile of interior cells: the neutrals' face VJPs
//                       of every face whose stencil reaches the tile, into
//                       shared memory, then the gather of both fluids' face
//                       contributions, the pointwise part and u2v^T.
#pragma once

#include <hip/hip_runtime.h>

#include "pip_adj_items.h"

namespace pip {

__global__ void __launch_bounds__(256) k_adj_cell(AdjArgs a) {
  adj_cell_item(a, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Workgroups are dealt round-robin to the 8 XCDs; xcd_order gives each XCD a
// contiguous range of tiles and face blocks, so that the rows neighbours
// share (y stencil) meet in one L2.
__device__ __forceinline__ int xcd_order(int bid, int nb) {
  const int q = nb / 8, rr = nb % 8, x = bid % 8, k = bid / 8;
  return x * q + (x < rr ? x : rr) + k;
}

// The plasma's faces of both directions in one launch, in batches of
// ADJ_PBS consecutive faces (row-major): adj_face_blocks(nx+1, ny) batches
// of x faces, then adj_face_blocks(nx, ny+1) of y faces, ADJ_PF batches per
// workgroup of ADJ_PBS threads (two waves: at one wave per SIMD the launch
// ends on finer units than with four-wave workgroups; 64 threads measured
// slower).  Rows of nx+1 x faces are not a multiple of 64, so rectangular
// tiles would leave waves with a single live lane running the whole VJP.
// The first adj_edge_face_blocks workgroups compute the neutrals' boundary
// faces that the gather reads at ghost images (adj_edge_face_item):
// dispatched first, they share the first round with plasma workgroups
// instead of running alone after the last one (7.7 us of the launch at
// 512^2).
constexpr int ADJ_PBS = 128;
__host__ __device__ inline int adj_face_blocks(int w, int h) {
  return (int)(((long long)w * h + ADJ_PBS - 1) / ADJ_PBS);
}
__host__ __device__ inline int adj_plasma_face_blocks(int nx, int ny) {
  return adj_face_blocks(nx + 1, ny) + adj_face_blocks(nx, ny + 1);
}
__host__ __device__ inline int adj_edge_face_blocks(int nx, int ny) {
  return (int)((adj_edge_faces(nx, ny) + ADJ_PBS - 1) / ADJ_PBS);
}

// Batch b of ADJ_PBS plasma faces (b < adj_plasma_face_blocks): this thread's
// face, as direction, + cell and its offset; false past the last face.
__device__ __forceinline__ bool adj_plasma_face_of(const AdjArgs& a, int b, int& dir, int& fi, int& fj,
                                                   long long& oR) {
  const int nbx = adj_face_blocks(a.nx + 1, a.ny);
  if (b >= adj_plasma_face_blocks(a.nx, a.ny)) return false;
  dir = (b < nbx) ? 0 : 1;
  if (dir) b -= nbx;
  const int w = dir ? a.nx : a.nx + 1, h = dir ? a.ny + 1 : a.ny;
  const long long t = (long long)b * ADJ_PBS + threadIdx.x;
  if (t >= (long long)w * h) return false;
  fi = (int)(t % w) + 2;
  fj = (int)(t / w) + 2;
  oR = (long long)fj * a.ix + fi;
  return true;
}

// Each plasma workgroup takes two consecutive batches: a thread issues both
// faces' stencil loads before the first face's VJP, so that the second
// face's load latency passes under it (one wave per SIMD leaves no other
// wave to hide it; three faces per thread fit in 475 registers but left a
// longer last round and measured slower).
constexpr int ADJ_PF = 2;
__host__ __device__ inline int adj_plasma_face_groups(int nx, int ny) {
  return (adj_plasma_face_blocks(nx, ny) + ADJ_PF - 1) / ADJ_PF;
}

__global__ void __launch_bounds__(ADJ_PBS, 1) k_adj_plasma_faces(AdjArgs a) {
  const int nbe = adj_edge_face_blocks(a.nx, a.ny);
  if ((int)blockIdx.x < nbe) {
    adj_edge_face_item(a, (long long)blockIdx.x * ADJ_PBS + threadIdx.x);
    return;
  }
  const int g = xcd_order((int)blockIdx.x - nbe, adj_plasma_face_groups(a.nx, a.ny));
  int dir[ADJ_PF], fi[ADJ_PF], fj[ADJ_PF];
  long long oR[ADJ_PF];
  bool ok[ADJ_PF];
  FaceRaw<0> raw[ADJ_PF];
#pragma unroll
  for (int q = 0; q < ADJ_PF; q++) {
    dir[q] = fi[q] = fj[q] = 0;
    oR[q] = 0;
    ok[q] = adj_plasma_face_of(a, ADJ_PF * g + q, dir[q], fi[q], fj[q], oR[q]);
    if (ok[q]) adj_face_raw<0>(a, dir[q], fi[q], fj[q], raw[q]);
  }
  FaceIn<0> in[ADJ_PF];
#pragma unroll
  for (int q = 0; q < ADJ_PF; q++)
    if (ok[q]) adj_face_from_raw<0>(a, dir[q], raw[q], in[q]);
#pragma unroll
  for (int q = 0; q < ADJ_PF; q++) {
    if (!ok[q]) continue;
    double lvl[ADJ_NV], lvr[ADJ_NV];
    adj_face_compute<0>(a, in[q], lvl, lvr);
    adj_face_store<0>(a, dir[q], oR[q], in[q].codes, lvl, lvr);
  }
}

// ---------------------------------------------------------------- gather
// Tiles of GT_X x GT_Y interior cells, one per 256-thread workgroup (the
// last 48 threads only take part in the face phase).  The faces whose
// stencils reach the tile: x faces with + cell c-1 .. c+GT_X+1 on its rows
// (GT_X + 3 per row), y faces likewise in its columns; 16 x 13 makes them
// 247 + 256 = 503, two passes of the workgroup, 1.21 face VJPs per face of
// the grid.  Their neutral adjoint pairs (5 x 16 B) and packed routing codes
// (20 bits) take 42 KB of shared memory, so three workgroups share a CU at
// 168 VGPRs.
constexpr int GT_X = 16, GT_Y = 13;
constexpr int GT_NFX = (GT_X + 3) * GT_Y;
constexpr int GT_NFY = GT_X * (GT_Y + 3);
constexpr int GT_NF = GT_NFX + GT_NFY;
static_assert(GT_X * GT_Y <= 256 && GT_NF <= 2 * 256, "gather tile");

__host__ __device__ inline int adj_gather_tiles(int nx, int ny) {
  return ((nx + GT_X - 1) / GT_X) * ((ny + GT_Y - 1) / GT_Y);
}

// Stage 1 (lam_in then receives the step's result): the blocks past the
// tiles zero the ghost ring of lam_in, adj_ghost_blocks of them.
__host__ __device__ inline int adj_ghost_blocks(int ix, int jx) {
  return (2 * ix + 2 * jx - 8 + 255) / 128;  // 4 ix + 4 (jx - 4) ring cells, 256 per block
}

// The neutral contributions of one direction at a tile cell from the face
// table: pairs sp[q * GT_NF + f] of plane NV + q and codes sc[f] at the
// faces f2, f1, f0, fm with + cells c+2, c+1, c, c-1 (faces outside the grid
// hold zero pairs).  The same products and sums as adj_gather_fluid<5>.
__device__ __forceinline__ void adj_gather_neutral_lds(const FacePair* sp, const unsigned* sc, int f2, int f1,
                                                       int f0, int fm, double* lv) {
  const unsigned k2 = sc[f2], k1 = sc[f1], k0 = sc[f0], km = sc[fm];
#pragma unroll
  for (int q = 0; q < 5; q++) {
    const int sh = 4 * q;
    const FacePair p2 = sp[q * GT_NF + f2], p1 = sp[q * GT_NF + f1], p0 = sp[q * GT_NF + f0],
                   pm = sp[q * GT_NF + fm];
    const double acc = (p2.l * wl((k2 >> sh) & 3u, 0)) +
                       (p1.l * wl((k1 >> sh) & 3u, 1) + p1.r * wr((k1 >> (sh + 2)) & 3u, 0)) +
                       (p0.l * wl((k0 >> sh) & 3u, 2) + p0.r * wr((k0 >> (sh + 2)) & 3u, 1)) +
                       (pm.r * wr((km >> (sh + 2)) & 3u, 2));
    lv[ADJ_NV + q] += acc;
  }
}

__global__ void __launch_bounds__(256, 3) k_adj_gather(AdjArgs a) {
  __shared__ FacePair sp[5 * GT_NF];
  __shared__ unsigned sc[GT_NF];
  const int nt = adj_gather_tiles(a.nx, a.ny);
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
  const int t = xcd_order(blockIdx.x, nt);
  const int ntx = (a.nx + GT_X - 1) / GT_X;
  const int i0 = (t % ntx) * GT_X + 2, j0 = (t / ntx) * GT_Y + 2;  // first cell of the tile
  // ---- the neutrals' face VJPs of the tile's faces
#pragma unroll 1
  for (int f = threadIdx.x; f < GT_NF; f += 256) {
    int dir, fi, fj;
    if (f < GT_NFX) {
      dir = 0;
      fi = i0 - 1 + f % (GT_X + 3);
      fj = j0 + f / (GT_X + 3);
    } else {
      const int g = f - GT_NFX;
      dir = 1;
      fi = i0 + g % GT_X;
      fj = j0 - 1 + g / GT_X;
    }
    const bool ok = (dir == 0) ? (fi >= 2 && fi <= a.nx + 2 && fj <= a.ny + 1)
                               : (fi <= a.nx + 1 && fj >= 2 && fj <= a.ny + 2);
    double lvl[5], lvr[5];
    unsigned code = 0u;
    if (ok) {
      adj_face_vjp<1>(a, dir, fi, fj, lvl, lvr, [&](unsigned long long c) { code = (unsigned)c; });
    } else {
#pragma unroll
      for (int q = 0; q < 5; q++) {
        lvl[q] = 0.0;
        lvr[q] = 0.0;
      }
    }
    // variable q in the sweep's order is plane adj_perm(NV + q, dir)
#pragma unroll
    for (int q = 0; q < 5; q++) sp[(adj_perm(ADJ_NV + q, dir) - ADJ_NV) * GT_NF + f] = FacePair{lvl[q], lvr[q]};
    sc[f] = code;
  }
  __syncthreads();
  // ---- the gather at the tile's cells
  if ((int)threadIdx.x >= GT_X * GT_Y) return;
  const int lx = (int)threadIdx.x % GT_X, ly = (int)threadIdx.x / GT_X;
  const int ci = i0 + lx, cj = j0 + ly;
  if (ci > a.nx + 1 || cj > a.ny + 1) return;
  double lv[ADJ_NVT];
#pragma unroll
  for (int n = 0; n < ADJ_NVT; n++) lv[n] = 0.0;
  {
    const int b = ly * (GT_X + 3) + lx;  // x face with + cell ci-1
    adj_gather_neutral_lds(sp, sc, b + 3, b + 2, b + 1, b, lv);
  }
  {
    const int b = GT_NFX + ly * GT_X + lx;  // y face with + cell cj-1
    adj_gather_neutral_lds(sp, sc, b + 3 * GT_X, b + 2 * GT_X, b + GT_X, b, lv);
  }
  // The sums are formed here: left to the compiler, the shared-memory loads
  // and their weights stay live into the plasma gather (~260 registers).
#pragma unroll
  for (int n = ADJ_NV; n < ADJ_NVT; n++) asm volatile("" : "+v"(lv[n]));
  adj_gather_at<false>(a, ci, cj, 0u, lv);
#pragma unroll
  for (int n = 0; n < ADJ_NV; n++) asm volatile("" : "+v"(lv[n]));
  for_each_image(ci, cj, a.ix, a.jx, a.bnd,
                 [&](int xi, int yj, unsigned m) { adj_gather_at(a, xi, yj, m, lv); });
  adj_cell_finish(a, (long long)cj * a.ix + ci, lv);
}

__global__ void __launch_bounds__(256) k_adj_fold_seed(double* lam, int nvt, int ix, int jx, long long N,
                                                       Bnd bnd) {
  adj_fold_seed_item(lam, nvt, ix, jx, N, bnd, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void __launch_bounds__(256) k_zero_ghosts(double* x, int nvt, int ix, int jx, long long N) {
  zero_ghosts_item(x, nvt, ix, jx, N, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

}  // namespace pip
