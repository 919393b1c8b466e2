// Dataflow D-ILU sweeps (one persistent launch per triangular solve).
#pragma once
#include <hip/hip_runtime.h>

#include "p3d_internal.hpp"

namespace p3d {

struct SweepPlan {
  int ready = 0;
};

// Returns 0 when the dataflow sweep can run on this grid.
int build_sweep_plan(SweepPlan& plan, const Grid3& g, int num_cus);
void free_sweep_plan(SweepPlan& plan);

// z = (LU)^-1 r with the matrix-free D-ILU factors (dinv = 1/pivots).
template <typename T>
void launch_dilu_sweep(const SweepPlan& plan, const Metric& m, const Grid3& g, const float* dinv,
                       const T* r, float* yv, T* z, const CgState* guard, hipStream_t s);

}  // namespace p3d
