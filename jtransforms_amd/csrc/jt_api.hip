// C ABI of the transform library (declared in include/jtransforms_amd.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/jtransforms_amd.h"
#include "jt_multid.hpp"
#include "jt_r2r.hpp"
#include "jt_ragged.hpp"

using namespace jt;

namespace {

thread_local std::string g_err;

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

}  // namespace

struct jt_plan {
    std::string desc;  // tuning record (ragged batches)
    int family = FAM_FFT;
    int precision = JT_F64;
    int ndim = 1;
    int64_t dims[3] = {1, 1, 1};
    int64_t batch = 1;
    uint32_t flags = 0;
    bool pow2_all = false;
    std::unique_ptr<Impl<double>> d;
    std::unique_ptr<Impl<float>> f;
};

namespace {

int set_err(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

template <typename F>
int guarded(F&& fn) {
    try {
        g_err.clear();
        // The HIP runtime may be shared with the caller (e.g. PyTorch); drop
        // any error state its calls left behind so the checks after this
        // library's own launches report only those launches.
        (void)hipGetLastError();
        return fn();
    } catch (const JtError& e) {
        return set_err(e.code, e.what());
    } catch (const std::bad_alloc&) {
        return set_err(JT_OOM, "host allocation failed");
    } catch (const std::exception& e) {
        return set_err(JT_HIP_ERROR, e.what());
    }
}

int64_t total(const jt_plan* p) {
    int64_t t = 1;
    for (int i = 0; i < p->ndim; i++) t *= p->dims[i];
    return t;
}

template <typename T>
Impl<T>& impl(jt_plan* p);
template <>
Impl<double>& impl<double>(jt_plan* p) { return *p->d; }
template <>
Impl<float>& impl<float>(jt_plan* p) { return *p->f; }

// workspace requirement (complex elements)
template <typename T>
int64_t ws_need(jt_plan* p) {
    Impl<T>& I = impl<T>(p);
    int64_t w = 0;
    const int64_t tot = total(p) * p->batch;
    for (int a = 0; a < p->ndim; a++) {
        if (!I.ax[a]) continue;
        // (real-input multi-D passes run the column axes on half-width data,
        // which needs no more than the complex passes)
        int64_t lines = tot / p->dims[a];
        w = std::max(w, axis_ws(*I.ax[a], lines));
    }
    if (I.real) {
        int64_t lines = tot / p->dims[p->ndim - 1];
        w = std::max(w, I.real->ws_elems(lines));
        w = std::max(w, 2 * tot);  // *Full ops expand to complex
    }
    for (int a = 0; a < 3; a++)
        if (I.r2r[a]) w = std::max(w, I.r2r[a]->ws_elems(tot / p->dims[a]) + tot);
    return w;
}

// split-FP16 MFMA radix-16 stages on every compile-time complex axis plan
template <typename T>
void apply_flags(Impl<T>& I, uint32_t flags) {
    const bool mf = sizeof(T) == 4 && (flags & JT_FLAG_SPLIT_FP16);
    for (int a = 0; a < 3; a++) {
        if (!I.ax[a]) continue;
        Fft1D<T>& p = *I.ax[a];
        std::vector<int> r;
        p.mf = mf && p.kind == K_TPL && mf_radices(p.n, r);
        if (p.mf && !p.tw_mf.p) upload(p.tw_mf, stage_twiddles_mf<T>(r));
    }
}

template <typename T>
void build(jt_plan* p) {
    auto I = std::make_unique<Impl<T>>();
    if (p->family == FAM_FFT) {
        for (int a = 0; a < p->ndim; a++) {
            bool dup = false;
            for (int b = 0; b < a; b++)
                if (p->dims[b] == p->dims[a]) dup = true;
            (void)dup;
            I->ax[a] = make_fft1d<T>(p->dims[a]);
        }
        // packed real transforms run the last axis as a real plan; the other
        // axes of a multi-D real transform are complex passes of full length
        I->real = make_real1d<T>(p->dims[p->ndim - 1]);
    } else {
        for (int a = 0; a < p->ndim; a++) I->r2r[a] = make_r2r1d<T>(p->family, p->dims[a]);
    }
    apply_flags(*I, p->flags);
    if constexpr (sizeof(T) == 8) p->d = std::move(I);
    else p->f = std::move(I);
    int64_t w = ws_need<T>(p);
    impl<T>(p).ws.alloc((size_t)std::max<int64_t>(w, 1) * sizeof(cpx<T>));
}

int create(jt_plan** out, int family, int precision, int ndim, const int64_t* dims, int64_t batch,
           uint32_t flags) {
    return guarded([&]() -> int {
        if (!out) return set_err(JT_BAD_ARG, "null plan pointer");
        *out = nullptr;
        if (precision != JT_F64 && precision != JT_F32) return set_err(JT_BAD_ARG, "unknown precision");
        if (ndim < 1 || ndim > 3 || !dims) return set_err(JT_BAD_ARG, "ndim must be 1, 2 or 3");
        if (ndim == 1) {
            if (dims[0] < 1) return set_err(JT_BAD_SIZE, "n must be greater than 0");
        } else if (ndim == 2) {
            if (dims[0] <= 1 || dims[1] <= 1)
                return set_err(JT_BAD_SIZE, "rows and columns must be greater than 1");
        } else {
            if (dims[0] <= 1 || dims[1] <= 1 || dims[2] <= 1)
                return set_err(JT_BAD_SIZE, "slices, rows and columns must be greater than 1");
        }
        if (batch < 1) return set_err(JT_BAD_SIZE, "batch must be greater than 0");
        int64_t tot = batch;
        for (int i = 0; i < ndim; i++) {
            if (dims[i] > (int64_t(1) << 40) / tot) return set_err(JT_TOO_BIG, "data array is too big");
            tot *= dims[i];
        }
        if (family == FAM_DCT && ndim == 1 && dims[0] > (int64_t(1) << 28))
            return set_err(JT_TOO_BIG, "n must be smaller or equal to 268435456");
        auto p = std::make_unique<jt_plan>();
        p->family = family;
        p->precision = precision;
        p->ndim = ndim;
        p->batch = batch;
        p->flags = flags;
        p->pow2_all = true;
        for (int i = 0; i < ndim; i++) {
            p->dims[i] = dims[i];
            if (!is_pow2(dims[i])) p->pow2_all = false;
        }
        if (precision == JT_F64) build<double>(p.get());
        else build<float>(p.get());
        *out = p.release();
        return JT_OK;
    });
}

// ---------------------------------------------------------------------------
// execution
// ---------------------------------------------------------------------------
template <typename T, int DIR>
void run_complex_nd(jt_plan* p, cpx<T>* a, bool scale, hipStream_t st) {
    Impl<T>& I = impl<T>(p);
    cpx<T>* ws = I.ws.template as<cpx<T>>();
    const int nd = p->ndim;
    const int64_t B = p->batch;
    const T s = (DIR > 0 && scale) ? T(1.0 / (double)total(p)) : T(1);
    if (nd == 1) {
        c2c_contig<T, DIR>(*I.ax[0], a, a, B, s, ws, st);
        return;
    }
    // (no streaming loads or stores here: the passes after the first read
    // what the previous pass left in the caches, and streaming the first
    // pass's loads measured slower overall: 4096^2 FP64 0.42 -> 0.48 ms)
    if (nd == 2) {
        const int64_t r = p->dims[0], c = p->dims[1];
        axis_c2c<T, DIR>(*I.ax[1], a, LineMap{B * r, B * r, c, 0, 1}, T(1), ws, st);
        axis_c2c<T, DIR>(*I.ax[0], a, LineMap{B * c, c, 1, r * c, c}, s, ws, st);
        return;
    }
    const int64_t S = p->dims[0], r = p->dims[1], c = p->dims[2];
    axis_c2c<T, DIR>(*I.ax[2], a, LineMap{B * S * r, B * S * r, c, 0, 1}, T(1), ws, st);
    axis_c2c<T, DIR>(*I.ax[1], a, LineMap{B * S * c, c, 1, r * c, c}, T(1), ws, st);
    axis_c2c<T, DIR>(*I.ax[0], a, LineMap{B * r * c, r * c, 1, S * r * c, r * c}, s, ws, st);
}

// packed real 2D/3D (power-of-two sizes)
template <typename T>
void run_real_forward_nd(jt_plan* p, T* a, hipStream_t st) {
    Impl<T>& I = impl<T>(p);
    cpx<T>* ws = I.ws.template as<cpx<T>>();
    cpx<T>* z = reinterpret_cast<cpx<T>*>(a);
    if (p->ndim == 2) {
        const int64_t r = p->dims[0], c = p->dims[1], ch = c / 2;
        real_forward<T>(*I.real, a, r, ws, st);
        axis_c2c<T, -1>(*I.ax[0], z, LineMap{ch, ch, 1, 0, ch}, T(1), ws, st);
        { JT_PROF(st, "rdft2d_fix");
        hipLaunchKernelGGL((rdft2d_fix<T>), ew_grid(r / 2), dim3(256), 0, st, a, r, c, 1);
        }
        JT_HIP(hipGetLastError());
        return;
    }
    const int64_t S = p->dims[0], r = p->dims[1], c = p->dims[2], ch = c / 2;
    real_forward<T>(*I.real, a, S * r, ws, st);
    axis_c2c<T, -1>(*I.ax[1], z, LineMap{S * ch, ch, 1, r * ch, ch}, T(1), ws, st);
    axis_c2c<T, -1>(*I.ax[0], z, LineMap{r * ch, r * ch, 1, 0, r * ch}, T(1), ws, st);
    { JT_PROF(st, "rdft3d_fix");
    hipLaunchKernelGGL((rdft3d_fix<T>), ew_grid(S * (r / 2 + 1)), dim3(256), 0, st, a, S, r, c, 1);
    }
    JT_HIP(hipGetLastError());
}

template <typename T>
void run_real_inverse_nd(jt_plan* p, T* a, bool scale, hipStream_t st) {
    Impl<T>& I = impl<T>(p);
    cpx<T>* ws = I.ws.template as<cpx<T>>();
    cpx<T>* z = reinterpret_cast<cpx<T>*>(a);
    if (p->ndim == 2) {
        const int64_t r = p->dims[0], c = p->dims[1], ch = c / 2;
        { JT_PROF(st, "rdft2d_fix");
        hipLaunchKernelGGL((rdft2d_fix<T>), ew_grid(r / 2), dim3(256), 0, st, a, r, c, -1);
        }
        JT_HIP(hipGetLastError());
        axis_c2c<T, 1>(*I.ax[0], z, LineMap{ch, ch, 1, 0, ch}, scale ? T(1.0 / (double)r) : T(1), ws, st);
        real_inverse<T>(*I.real, a, r, scale, ws, st);
        return;
    }
    const int64_t S = p->dims[0], r = p->dims[1], c = p->dims[2], ch = c / 2;
    { JT_PROF(st, "rdft3d_fix");
    hipLaunchKernelGGL((rdft3d_fix<T>), ew_grid(S * (r / 2 + 1)), dim3(256), 0, st, a, S, r, c, -1);
    }
    JT_HIP(hipGetLastError());
    axis_c2c<T, 1>(*I.ax[0], z, LineMap{r * ch, r * ch, 1, 0, r * ch}, scale ? T(1.0 / (double)S) : T(1),
                   ws, st);
    axis_c2c<T, 1>(*I.ax[1], z, LineMap{S * ch, ch, 1, r * ch, ch}, scale ? T(1.0 / (double)r) : T(1), ws,
                   st);
    real_inverse<T>(*I.real, a, S * r, scale, ws, st);
}

// realForwardFull / realInverseFull for 2D/3D: complex transform of the real
// input (first prod(dims) elements), written as interleaved complex.
template <typename T, int DIR>
void run_real_full_nd(jt_plan* p, T* a, bool scale, hipStream_t st) {
    Impl<T>& I = impl<T>(p);
    const int64_t tot = total(p);
    // expand in place from the back: reals -> complex (use workspace copy)
    cpx<T>* ws = I.ws.template as<cpx<T>>();
    { JT_PROF(st, "real_to_complex");
    hipLaunchKernelGGL((real_to_complex<T>), ew_grid(tot), dim3(256), 0, st, a, ws, tot, (int64_t)1);
    }
    JT_HIP(hipGetLastError());
    { JT_PROF(st, "copy_complex");
    hipLaunchKernelGGL((copy_complex<T>), ew_grid(tot), dim3(256), 0, st, ws, reinterpret_cast<cpx<T>*>(a),
                       tot);
    }
    JT_HIP(hipGetLastError());
    run_complex_nd<T, DIR>(p, reinterpret_cast<cpx<T>*>(a), scale, st);
}

template <typename T>
int exec_t(jt_plan* p, int op, T* a, bool scale, hipStream_t st) {
    Impl<T>& I = impl<T>(p);
    cpx<T>* ws = I.ws.template as<cpx<T>>();
    if (p->family == FAM_RAGGED) {
        if (op != JT_COMPLEX_FORWARD && op != JT_COMPLEX_INVERSE)
            return set_err(JT_BAD_ARG, "a ragged batch supports complexForward / complexInverse");
        ragged_execute<T>(*I.ragged, reinterpret_cast<cpx<T>*>(a), op == JT_COMPLEX_FORWARD ? -1 : 1, scale, st,
                          (p->flags & JT_FLAG_GRAPH) != 0);
        return JT_OK;
    }
    if (p->family != FAM_FFT) {
        if (op != JT_FORWARD && op != JT_INVERSE) return set_err(JT_BAD_ARG, "op not defined for this plan");
        run_r2r_nd<T>(p->family, p->ndim, p->dims, I.r2r, a, op == JT_FORWARD, scale, ws, st);
        return JT_OK;
    }
    const int nd = p->ndim;
    const int64_t n = p->dims[0];
    switch (op) {
        case JT_COMPLEX_FORWARD:
            if (nd == 1 && n == 1) return JT_OK;
            run_complex_nd<T, -1>(p, reinterpret_cast<cpx<T>*>(a), false, st);
            return JT_OK;
        case JT_COMPLEX_INVERSE:
            if (nd == 1 && n == 1) return JT_OK;
            run_complex_nd<T, 1>(p, reinterpret_cast<cpx<T>*>(a), scale, st);
            return JT_OK;
        case JT_REAL_FORWARD:
            if (nd == 1) {
                real_forward<T>(*I.real, a, p->batch, ws, st);
                return JT_OK;
            }
            if (!p->pow2_all)
                return set_err(JT_NOT_POW2, nd == 2 ? "rows and columns must be power of two numbers"
                                                    : "slices, rows and columns must be power of two numbers");
            run_real_forward_nd<T>(p, a, st);
            return JT_OK;
        case JT_REAL_INVERSE:
            if (nd == 1) {
                real_inverse<T>(*I.real, a, p->batch, scale, ws, st);
                return JT_OK;
            }
            if (!p->pow2_all)
                return set_err(JT_NOT_POW2, nd == 2 ? "rows and columns must be power of two numbers"
                                                    : "slices, rows and columns must be power of two numbers");
            run_real_inverse_nd<T>(p, a, scale, st);
            return JT_OK;
        case JT_REAL_FORWARD_FULL:
            if (nd == 1) {
                if (p->batch != 1) return set_err(JT_UNSUPPORTED, "realForwardFull on a batch plan");
                real_full<T, -1>(*I.real, a, 1, false, ws, st);
                return JT_OK;
            }
            run_real_full_nd<T, -1>(p, a, false, st);
            return JT_OK;
        case JT_REAL_INVERSE_FULL:
            if (nd == 1) {
                if (p->batch != 1) return set_err(JT_UNSUPPORTED, "realInverseFull on a batch plan");
                real_full<T, 1>(*I.real, a, 1, scale, ws, st);
                return JT_OK;
            }
            run_real_full_nd<T, 1>(p, a, scale, st);
            return JT_OK;
        default:
            return set_err(JT_BAD_ARG, "op not defined for this plan");
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// exported entry points
// ---------------------------------------------------------------------------
extern "C" {

int jt_plan_create_fft1d(jt_plan** out, int precision, int64_t n) {
    return create(out, FAM_FFT, precision, 1, &n, 1, 0);
}

int jt_plan_create_fft2d(jt_plan** out, int precision, int64_t rows, int64_t columns) {
    int64_t d[2] = {rows, columns};
    return create(out, FAM_FFT, precision, 2, d, 1, 0);
}

int jt_plan_create_fft3d(jt_plan** out, int precision, int64_t slices, int64_t rows, int64_t columns) {
    int64_t d[3] = {slices, rows, columns};
    return create(out, FAM_FFT, precision, 3, d, 1, 0);
}

int jt_plan_create_dct(jt_plan** out, int precision, int ndim, const int64_t* dims) {
    return create(out, FAM_DCT, precision, ndim, dims, 1, 0);
}

int jt_plan_create_dst(jt_plan** out, int precision, int ndim, const int64_t* dims) {
    return create(out, FAM_DST, precision, ndim, dims, 1, 0);
}

int jt_plan_create_dht(jt_plan** out, int precision, int ndim, const int64_t* dims) {
    return create(out, FAM_DHT, precision, ndim, dims, 1, 0);
}

int jt_plan_create_fft1d_batch(jt_plan** out, int precision, int64_t n, int64_t batch, uint32_t flags) {
    if (flags & ~(uint32_t)JT_FLAG_SPLIT_FP16) return set_err(JT_BAD_ARG, "unknown flags");
    if ((flags & JT_FLAG_SPLIT_FP16) && precision != JT_F32)
        return set_err(JT_BAD_ARG, "split-FP16 form is defined for FP32 plans only");
    return create(out, FAM_FFT, precision, 1, &n, batch, flags);
}

}  // extern "C"

namespace {

// Tuned schedules by (precision, length list), for the life of the process.
std::mutex g_tune_mu;
std::map<std::pair<int, std::vector<int64_t>>, RTune> g_tuned;

template <typename T>
std::unique_ptr<RaggedPlan<T>> make_ragged_tuned(const int64_t* lengths, int64_t count, bool tune,
                                                 std::string* desc) {
    std::unique_ptr<RaggedPlan<T>> P;
    if (!tune) {
        P = make_ragged<T>(lengths, count);
    } else {
        auto key = std::make_pair((int)sizeof(T), std::vector<int64_t>(lengths, lengths + count));
        RTune known;
        bool have = false;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tuned.find(key);
            if (it != g_tuned.end()) {
                known = it->second;
                have = true;
            }
        }
        if (have) {
            P = make_ragged<T>(lengths, count, known);
            *desc = "autotuned (remembered)";
        } else {
            std::string log;
            P = ragged_autotune<T>(lengths, count, &log);
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tuned[key] = P->tune;
            for (auto& c : log)
                if (c == '\n') c = ';';
            *desc = "autotuned: " + log;
        }
    }
    *desc = std::string("ragged ") + (sizeof(T) == 8 ? "f64" : "f32") + " transforms=" + std::to_string(P->ntr) +
            " points=" + std::to_string(P->total) + " schedule " + P->tune.str() + (desc->empty() ? "" : " ") + *desc;
    return P;
}

}  // namespace

extern "C" {

int jt_plan_create_ragged(jt_plan** out, int precision, int64_t count, const int64_t* lengths, uint32_t flags) {
    return guarded([&]() -> int {
        if (!out) return set_err(JT_BAD_ARG, "null plan pointer");
        *out = nullptr;
        if (precision != JT_F64 && precision != JT_F32) return set_err(JT_BAD_ARG, "unknown precision");
        if (count < 1 || !lengths) return set_err(JT_BAD_SIZE, "count must be greater than 0");
        if (flags & ~(uint32_t)(JT_FLAG_GRAPH | JT_FLAG_AUTOTUNE)) return set_err(JT_BAD_ARG, "unknown flags");
        for (int64_t i = 0; i < count; i++)
            if (lengths[i] < 1) return set_err(JT_BAD_SIZE, "n must be greater than 0");
        auto p = std::make_unique<jt_plan>();
        p->family = FAM_RAGGED;
        p->precision = precision;
        p->ndim = 1;
        p->dims[0] = count;
        p->flags = flags;
        const bool tune = (flags & JT_FLAG_AUTOTUNE) != 0;
        if (precision == JT_F64) {
            p->d = std::make_unique<Impl<double>>();
            p->d->ragged = make_ragged_tuned<double>(lengths, count, tune, &p->desc);
        } else {
            p->f = std::make_unique<Impl<float>>();
            p->f->ragged = make_ragged_tuned<float>(lengths, count, tune, &p->desc);
        }
        *out = p.release();
        return JT_OK;
    });
}

int64_t jt_plan_describe(const jt_plan* p, char* buf, int64_t len) {
    if (!p) return -1;
    std::string d = p->desc;
    if (d.empty()) {
        d = std::string(p->family == FAM_FFT ? "fft" : p->family == FAM_DCT ? "dct" : p->family == FAM_DST ? "dst" : "dht") +
            (p->precision == JT_F64 ? " f64" : " f32") + " dims=";
        for (int i = 0; i < p->ndim; i++) d += (i ? "x" : "") + std::to_string(p->dims[i]);
        if (p->batch > 1) d += " batch=" + std::to_string(p->batch);
        if (p->flags & JT_FLAG_SPLIT_FP16) d += " split-fp16";
    }
    if (buf && len > 0) {
        size_t n = std::min<size_t>((size_t)len - 1, d.size());
        std::memcpy(buf, d.data(), n);
        buf[n] = 0;
    }
    return (int64_t)d.size();
}

int jt_plan_sync_status(jt_plan* p) {
    return guarded([&]() -> int {
        if (!p) return set_err(JT_BAD_ARG, "null plan");
        JT_HIP(hipDeviceSynchronize());
        if (p->family == FAM_RAGGED) {
            int e = 0;
            DevBuf& d = p->precision == JT_F64 ? p->d->ragged->d_err : p->f->ragged->d_err;
            JT_HIP(hipMemcpy(&e, d.p, sizeof(int), hipMemcpyDeviceToHost));
            if (e) return set_err(JT_HIP_ERROR, "a ragged-batch wait timed out; results are incomplete");
        }
        return JT_OK;
    });
}

int jt_plan_set_flags(jt_plan* p, uint32_t flags) {
    return guarded([&]() -> int {
        if (!p) return set_err(JT_BAD_ARG, "null plan");
        if (flags & ~(uint32_t)JT_FLAG_SPLIT_FP16) return set_err(JT_BAD_ARG, "unknown flags");
        if ((flags & JT_FLAG_SPLIT_FP16) && p->precision != JT_F32)
            return set_err(JT_BAD_ARG, "split-FP16 form is defined for FP32 plans only");
        if (p->family != FAM_FFT) return set_err(JT_BAD_ARG, "flags apply to FFT plans");
        if (!p->f) return set_err(JT_BAD_ARG, "split-FP16 form is defined for FP32 plans only");
        p->flags = flags;
        apply_flags(*p->f, flags);
        return JT_OK;
    });
}

int jt_plan_destroy(jt_plan* plan) {
    return guarded([&]() -> int {
        delete plan;
        return JT_OK;
    });
}

int64_t jt_plan_elements(const jt_plan* p, int op) {
    if (!p) return -1;
    if (p->family == FAM_RAGGED) {
        int64_t t = p->precision == JT_F64 ? p->d->ragged->total : p->f->ragged->total;
        return (op == JT_COMPLEX_FORWARD || op == JT_COMPLEX_INVERSE) ? 2 * t : -1;
    }
    const int64_t tot = total(p) * p->batch;
    if (p->family != FAM_FFT) return tot;
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

int jt_execute(jt_plan* p, int op, void* data, int64_t offset, int scale, void* stream) {
    return guarded([&]() -> int {
        if (!p || !data) return set_err(JT_BAD_ARG, "null argument");
        if (offset < 0) return set_err(JT_BAD_ARG, "negative offset");
        if (jt_plan_elements(p, op) < 0) return set_err(JT_BAD_ARG, "op not defined for this plan");
        hipStream_t st = reinterpret_cast<hipStream_t>(stream);
        if (p->precision == JT_F64) return exec_t<double>(p, op, static_cast<double*>(data) + offset, scale != 0, st);
        return exec_t<float>(p, op, static_cast<float*>(data) + offset, scale != 0, st);
    });
}

int jt_execute_host(jt_plan* p, int op, void* data, int64_t offset, int scale) {
    return guarded([&]() -> int {
        if (!p || !data) return set_err(JT_BAD_ARG, "null argument");
        if (offset < 0) return set_err(JT_BAD_ARG, "negative offset");
        int64_t ne = jt_plan_elements(p, op);
        if (ne < 0) return set_err(JT_BAD_ARG, "op not defined for this plan");
        size_t es = p->precision == JT_F64 ? 8 : 4;
        size_t bytes = (size_t)ne * es;
        DevBuf& stg = p->precision == JT_F64 ? p->d->staging : p->f->staging;
        if (stg.bytes < bytes) stg.alloc(bytes);
        char* h = static_cast<char*>(data) + offset * es;
        JT_HIP(hipMemcpy(stg.p, h, bytes, hipMemcpyHostToDevice));
        int rc = jt_execute(p, op, stg.p, 0, scale, nullptr);
        if (rc != JT_OK) return rc;
        rc = jt_plan_sync_status(p);
        if (rc != JT_OK) return rc;
        JT_HIP(hipMemcpy(h, stg.p, bytes, hipMemcpyDeviceToHost));
        return JT_OK;
    });
}

const char* jt_status_string(int s) {
    switch (s) {
        case JT_OK: return "ok";
        case JT_BAD_SIZE: return "bad size";
        case JT_NOT_POW2: return "sizes must be power of two numbers";
        case JT_TOO_BIG: return "data array is too big";
        case JT_HIP_ERROR: return "HIP runtime error";
        case JT_OOM: return "out of device memory";
        case JT_BAD_ARG: return "bad argument";
        case JT_UNSUPPORTED: return "unsupported";
        default: return "unknown status";
    }
}

const char* jt_last_error(void) { return g_err.c_str(); }

int jt_profile_begin(void) {
    return guarded([&]() -> int {
        prof_clear();
        prof_on().store(true);
        return JT_OK;
    });
}

int64_t jt_profile_end(char* buf, int64_t len) {
    // the summary of the last recording stays readable until the next begin
    static std::string last;
    if (prof_on().exchange(false)) {
        try {
            last = prof_summary();
            prof_clear();
        } catch (...) {
            return -1;
        }
    }
    if (buf && len > 0) {
        int64_t k = std::min<int64_t>(len - 1, (int64_t)last.size());
        std::memcpy(buf, last.data(), (size_t)k);
        buf[k] = 0;
    }
    return (int64_t)last.size();
}

const char* jt_version(void) { return "jtransforms-amd 0.1 (gfx950)"; }

}  // extern "C"
