// Dataflow D-ILU sweeps.
#include "p3d_sweep.hpp"

namespace p3d {

int build_sweep_plan(SweepPlan& plan, const Grid3&, int) {
  plan.ready = 0;
  return 1;
}

void free_sweep_plan(SweepPlan& plan) { plan.ready = 0; }

template <typename T>
void launch_dilu_sweep(const SweepPlan&, const Metric&, const Grid3&, const float*, const T*,
                       float*, T*, const CgState*, hipStream_t) {}

template void launch_dilu_sweep<double>(const SweepPlan&, const Metric&, const Grid3&,
                                        const float*, const double*, float*, double*,
                                        const CgState*, hipStream_t);
template void launch_dilu_sweep<float>(const SweepPlan&, const Metric&, const Grid3&,
                                       const float*, const float*, float*, float*,
                                       const CgState*, hipStream_t);

}  // namespace p3d
