// The plan object shared by the C ABI (jt_api.hip) and the per-precision
// implementation units (jt_impl_f64.hip, jt_impl_f32.hip, which instantiate
// the kernels; split so that they compile in parallel).
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/jtransforms_amd.h"
#include "jt_multid.hpp"
#include "jt_r2r.hpp"
#include "jt_ragged.hpp"

namespace jt_api {
using namespace jt;

enum Family { FAM_FFT = 0, FAM_DCT = 1, FAM_DST = 2, FAM_DHT = 3, FAM_RAGGED = 4 };

template <typename T>
struct Impl {
    // FFT: one complex / real plan per distinct axis length
    std::unique_ptr<Fft1D<T>> ax[3];
    std::unique_ptr<Real1D<T>> real;  // along the last axis
    std::unique_ptr<R2R1D<T>> r2r[3];
    std::unique_ptr<RaggedPlan<T>> ragged;
    DevBuf ws;       // workspace
    DevBuf staging;  // host-array execution
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
    std::unique_ptr<jt_api::Impl<double>> d;
    std::unique_ptr<jt_api::Impl<float>> f;
};

namespace jt_api {

// records the message for jt_last_error() and returns the code (jt_api.hip)
int set_err(int code, const std::string& msg);

inline int64_t total(const jt_plan* p) {
    int64_t t = 1;
    for (int i = 0; i < p->ndim; i++) t *= p->dims[i];
    return t;
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
template <typename T>
int exec_t(jt_plan* p, int op, T* a, bool scale, hipStream_t st);
template <typename T>
std::unique_ptr<RaggedPlan<T>> make_ragged_tuned(const int64_t* lengths, int64_t count, bool tune,
                                                 std::string* desc);

}  // namespace jt_api
