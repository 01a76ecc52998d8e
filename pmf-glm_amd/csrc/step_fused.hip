// Fused tiled GLM step (placeholder until the tiled kernel lands).
#include "launch.hpp"

namespace pmf {

bool fused_supported(const Grid&, int, int, int) { return false; }

template <class T>
bool launch_step_fused(const T*, T*, const Grid&, const StepParams<T>&, double*, double*, hipStream_t) {
  return false;
}
template bool launch_step_fused<float>(const float*, float*, const Grid&, const StepParams<float>&, double*, double*,
                                       hipStream_t);
template bool launch_step_fused<double>(const double*, double*, const Grid&, const StepParams<double>&, double*,
                                        double*, hipStream_t);

}  // namespace pmf
