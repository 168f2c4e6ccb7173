// pip_relaxed.h -- interface between the C-ABI (pipgpu.hip) and the second
// build of the stage kernels in relaxed arithmetic (pipgpu_relaxed.hip).
// Internal to libpipgpu; not part of include/pipgpu.h.
#pragma once

#include <stddef.h>

// Launch one SSP-RK stage of the relaxed build on `stream`.  `args` points to
// the exact build's StageArgs (same definition, same layout), `args_bytes`
// is its size; ndim / nv / ir select the instantiation and tx / ty the tile.
// Returns 0, or -1 for a combination the relaxed build does not cover, -2
// for a mismatched argument block and -3 for a HIP error (*err names it).
int pip_relaxed_stage(const void* args, size_t args_bytes, int ndim, int nv, int ir, int tx, int ty,
                      void* stream, const char** err);

// 1 when the relaxed build covers (ndim, nv, ir, tx, ty).
int pip_relaxed_covers(int ndim, int nv, int ir, int tx, int ty);
