// Pooled problem creation for the serving engine (t3x_serve.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "../../include/t3x.h"

namespace t3x {

// Device memory, stream and pinned host buffer lent to one problem: its
// device arrays are carved out of [base, base + size) (at most the
// t3x_footprint of the model), its work runs on `stream`, and destroying the
// problem releases none of them.
struct PoolSlot {
  void* base = nullptr;
  size_t size = 0;
  hipStream_t stream = nullptr;
  double* hostBuf = nullptr;  // 256 doubles, pinned
  // coarsest solve in one launch per recurrence where the process's share of
  // the device's work-group slots allows it
  bool oneLaunchCoarse = true;
};

// t3x_create with the built-in or explicit one-case boundary conditions, on
// lent resources.
int create_pooled(const uint8_t* voxels, int nely, int nelx, int nelz, int nFix, const int* fixNodes,
                  const uint8_t* fixState, int nLoad, const int* loadNodes, const double* loadVec,
                  const t3x_params* params, const PoolSlot& slot, t3x_problem** out);

}  // namespace t3x
