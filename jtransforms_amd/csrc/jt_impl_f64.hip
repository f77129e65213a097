// FP64 instantiations of the plan building and execution templates.
#include "jt_impl.hpp"

namespace jt_api {
template void build<double>(jt_plan* p);
template void apply_flags<double>(Impl<double>& I, uint32_t flags);
template int exec_t<double>(jt_plan* p, int op, double* a, bool scale, cpx<double>* ws, hipStream_t st);
template std::unique_ptr<RaggedPlan<double>> make_ragged_tuned<double>(const int64_t* lengths, int64_t count, bool tune,
                                                                std::string* desc);
}  // namespace jt_api
