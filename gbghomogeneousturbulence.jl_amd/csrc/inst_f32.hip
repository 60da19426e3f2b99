// FP32-bulk instantiation of the stage pipeline for every supported grid size.
#include "pipeline.hpp"

namespace hdns {
std::unique_ptr<Pipeline> make_pipeline_f32(int n) { return make_pipeline_t<float>(n); }
}  // namespace hdns
