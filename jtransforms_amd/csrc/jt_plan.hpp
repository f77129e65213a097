// The plan object shared by the C ABI (jt_api.hip) and the per-precision
// implementation units (jt_impl_f64.hip, jt_impl_f32.hip, which instantiate
// the kernels; split so that they compile in parallel).
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/jtransforms_amd.h"
#include "jt_multid.hpp"
#include "jt_r2r.hpp"
#include "jt_ragged.hpp"

namespace jt_api {
using namespace jt;

enum Family { FAM_FFT = 0, FAM_DCT = 1, FAM_DST = 2, FAM_DHT = 3, FAM_RAGGED = 4 };

constexpr int NUM_OPS = 8;  // JT_COMPLEX_FORWARD .. JT_INVERSE

template <typename T>
struct Impl {
    // FFT: one complex plan per distinct axis length (shared when two axes
    // have the same length) and a real plan along the last axis
    std::shared_ptr<Fft1D<T>> ax[3];
    std::unique_ptr<Real1D<T>> real;
    std::shared_ptr<R2R1D<T>> r2r[3];
    std::unique_ptr<RaggedPlan<T>> ragged;
    // One device buffer, sized at creation: an op uses its first op_ws[op]
    // complex elements as workspace and the next ceil(elements(op)/2) as the
    // staging area of host-array calls and of device calls whose offset
    // misaligns the (re, im) pairs.  Execution never allocates.
    DevBuf ws;
};

}  // namespace jt_api

struct jt_plan {
    std::string desc;  // tuning record (ragged batches)
    int family = jt_api::FAM_FFT;
    int precision = JT_F64;
    int ndim = 1;
    int64_t dims[3] = {1, 1, 1};
    int64_t batch = 1;
    uint32_t flags = 0;
    bool pow2_all = false;
    // workspace (complex elements) each op touches; -1: op not defined
    int64_t op_ws[jt_api::NUM_OPS] = {-1, -1, -1, -1, -1, -1, -1, -1};
    // Calls on one plan are serialised: `mu` guards the host-side state (the
    // graph cache of ragged batches, flags) during a call, and the device work
    // of calls that touch the plan's buffers (workspace, staging, the ragged
    // queue) is ordered across streams by `chain`, recorded after each such
    // call on its stream and waited for by the next one on another stream.
    std::mutex mu;
    hipEvent_t chain = nullptr;
    hipStream_t chain_st = nullptr;
    bool chain_set = false;
    std::unique_ptr<jt_api::Impl<double>> d;
    std::unique_ptr<jt_api::Impl<float>> f;
    ~jt_plan() {
        if (chain) (void)hipEventDestroy(chain);
    }
};

namespace jt_api {

// records the message for jt_last_error() and returns the code (jt_api.hip)
int set_err(int code, const std::string& msg);

inline int64_t total(const jt_plan* p) {
    int64_t t = 1;
    for (int i = 0; i < p->ndim; i++) t *= p->dims[i];
    return t;
}

// Number of array elements an op reads and writes (jt_plan_elements); -1 if
// the op is not defined for the plan's family.
inline int64_t op_elements(const jt_plan* p, int op) {
    if (op < 0 || op >= NUM_OPS) return -1;
    if (p->family == FAM_RAGGED) {
        const int64_t t = p->precision == JT_F64 ? p->d->ragged->total : p->f->ragged->total;
        return (op == JT_COMPLEX_FORWARD || op == JT_COMPLEX_INVERSE) ? 2 * t : -1;
    }
    const int64_t tot = total(p) * p->batch;
    if (p->family != FAM_FFT) return (op == JT_FORWARD || op == JT_INVERSE) ? tot : -1;
    switch (op) {
        case JT_COMPLEX_FORWARD:
        case JT_COMPLEX_INVERSE:
        case JT_REAL_FORWARD_FULL:
        case JT_REAL_INVERSE_FULL:
            return 2 * tot;
        case JT_REAL_FORWARD:
        case JT_REAL_INVERSE:
            return tot;
        default:
            return -1;
    }
}

template <typename T>
Impl<T>& impl(jt_plan* p);
template <>
inline Impl<double>& impl<double>(jt_plan* p) { return *p->d; }
template <>
inline Impl<float>& impl<float>(jt_plan* p) { return *p->f; }

// defined in jt_impl.hpp, instantiated in jt_impl_f64.hip / jt_impl_f32.hip
template <typename T>
void build(jt_plan* p);
template <typename T>
void apply_flags(Impl<T>& I, uint32_t flags);
// runs `op` on pair-aligned device data `a`; `ws` is the op's workspace
// (op_ws[op] complex elements)
template <typename T>
int exec_t(jt_plan* p, int op, T* a, bool scale, cpx<T>* ws, hipStream_t st);
template <typename T>
std::unique_ptr<RaggedPlan<T>> make_ragged_tuned(const int64_t* lengths, int64_t count, bool tune,
                                                 std::string* desc);

}  // namespace jt_api
