// FP32-bulk instantiation of the HouLi and PhaseShift pipelines.
#include "pipeline.hpp"

namespace hdns {
std::unique_ptr<Pipeline> make_pipeline_f32_wide(int n, int dealias) {
  if (dealias == DEALIAS_HOU_LI) return make_pipeline_wide_t<float, DEALIAS_HOU_LI>(n);
  if (dealias == DEALIAS_PHASE_SHIFT) return make_pipeline_wide_t<float, DEALIAS_PHASE_SHIFT>(n);
  return nullptr;
}
}  // namespace hdns
