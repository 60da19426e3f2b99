// Plain-FP64 instantiation of the stage pipeline for every supported grid size.
#include "pipeline.hpp"

namespace hdns {
std::unique_ptr<Pipeline> make_pipeline_f64(int n) { return make_pipeline_t<double>(n); }
}  // namespace hdns
