// FP32 instantiations of the plan building and execution templates.
#include "jt_impl.hpp"

namespace jt_api {
template void build<float>(jt_plan* p);
template void apply_flags<float>(Impl<float>& I, uint32_t flags);
template int exec_t<float>(jt_plan* p, int op, float* a, bool scale, cpx<float>* ws, hipStream_t st);
template std::unique_ptr<RaggedPlan<float>> make_ragged_tuned<float>(const int64_t* lengths, int64_t count, bool tune,
                                                                std::string* desc);
}  // namespace jt_api
