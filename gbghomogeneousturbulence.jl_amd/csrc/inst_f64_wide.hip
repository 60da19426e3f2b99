// Plain-FP64 instantiation of the HouLi and PhaseShift pipelines.
#include "pipeline.hpp"

namespace hdns {
std::unique_ptr<Pipeline> make_pipeline_f64_wide(int n, int dealias) {
  if (dealias == DEALIAS_HOU_LI) return make_pipeline_wide_t<double, DEALIAS_HOU_LI>(n);
  if (dealias == DEALIAS_PHASE_SHIFT) return make_pipeline_wide_t<double, DEALIAS_PHASE_SHIFT>(n);
  return nullptr;
}
}  // namespace hdns
