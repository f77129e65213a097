// C ABI of the transform library (declared in include/jtransforms_amd.h).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "jt_plan.hpp"

using namespace jt;

namespace jt_api {

thread_local std::string g_err;

int set_err(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

}  // namespace jt_api

using namespace jt_api;

// Device scratch of jt_execute_noalloc: one array per GPU, loaded with the
// code object and held until the process ends.  Host code never uses the
// array's name as an address (on the host it is a stand-in symbol): the
// device address comes from hipGetSymbolAddress, once per device.
constexpr size_t JT_NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[JT_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// This device's copy of `scratch`, looked up on its first use.
char* scratch_on_current_device() {
    int dev = 0;
    JT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) throw JtError(ST_UNSUPPORTED, "device ordinal beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        JT_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw JtError(ST_HIP, "hipGetSymbolAddress(scratch) returned no address");
        g_scratch_ptr[dev] = ptr;
    }
    return static_cast<char*>(g_scratch_ptr[dev]);
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
        // DoubleDCT_1D (dct/DoubleDCT_1D.java:93-100), and DoubleDST_1D through
        // the DoubleDCT_1D it builds (dst/DoubleDST_1D.java:59-65), reject
        // 2^28 < n unless they run on large arrays, which they do for lengths
        // beyond a 32-bit Java array; this library has no array-size limit
        // of its own and accepts those
        if ((family == FAM_DCT || family == FAM_DST) && ndim == 1 && dims[0] > (int64_t(1) << 28) &&
            dims[0] <= INT32_MAX)
            return set_err(JT_TOO_BIG,
                           "n must be smaller or equal to 268435456 when useLargeArrays argument is set to false");
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
        JT_HIP(hipEventCreateWithFlags(&p->chain, hipEventDisableTiming));
        *out = p.release();
        return JT_OK;
    });
}

// Does `op` touch the plan's device buffers (beyond the tables, which are
// read-only)?  Then calls on different streams are ordered (see jt_plan).
bool touches_plan_buffers(const jt_plan* p, int op, bool staged) {
    return staged || p->family == FAM_RAGGED || p->op_ws[op] > 0;
}

// The plan's workspace of `op`, and its staging area after it.
char* ws_base(jt_plan* p) { return static_cast<char*>(p->precision == JT_F64 ? p->d->ws.p : p->f->ws.p); }
char* staging_area(jt_plan* p, int op) {
    const size_t cs = p->precision == JT_F64 ? 16 : 8;
    return ws_base(p) + (size_t)p->op_ws[op] * cs;
}

// Enqueues `op` on device memory; the caller holds p->mu.
int execute_locked(jt_plan* p, int op, char* ptr, int scale, hipStream_t st) {
    const int64_t ne = op_elements(p, op);
    if (ne < 0 || p->op_ws[op] < 0) return set_err(JT_BAD_ARG, "op not defined for this plan");
    const size_t es = p->precision == JT_F64 ? 8 : 4;
    // The kernels access (re, im) pairs of 2*es bytes: an offset that
    // misaligns them (the reference accepts any offa) runs on an aligned copy
    // of the range in the plan's staging area, on the same stream.
    const bool staged = reinterpret_cast<uintptr_t>(ptr) % (2 * es) != 0;
    bool chained = touches_plan_buffers(p, op, staged);
    if (chained) {
        // under the caller's stream capture the captured graph orders the
        // work; an event recorded outside it cannot be waited for there
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        JT_HIP(hipStreamIsCapturing(st, &cs));
        if (cs != hipStreamCaptureStatusNone) chained = false;
    }
    if (chained && p->chain_set && p->chain_st != st) JT_HIP(hipStreamWaitEvent(st, p->chain, 0));
    const size_t bytes = (size_t)ne * es;
    char* d = ptr;
    if (staged) {
        d = staging_area(p, op);
        JT_HIP(hipMemcpyAsync(d, ptr, bytes, hipMemcpyDeviceToDevice, st));
    }
    int rc;
    if (p->precision == JT_F64)
        rc = exec_t<double>(p, op, reinterpret_cast<double*>(d), scale != 0,
                            reinterpret_cast<cpx<double>*>(ws_base(p)), st);
    else
        rc = exec_t<float>(p, op, reinterpret_cast<float*>(d), scale != 0,
                           reinterpret_cast<cpx<float>*>(ws_base(p)), st);
    if (rc == JT_OK && staged) JT_HIP(hipMemcpyAsync(ptr, d, bytes, hipMemcpyDeviceToDevice, st));
    if (chained) {
        JT_HIP(hipEventRecord(p->chain, st));
        p->chain_st = st;
        p->chain_set = true;
    }
    return rc;
}

// Reads and clears the ragged batch's failure flag (set by a dependency wait
// that timed out in any launch since the last check).
int take_device_status(jt_plan* p) {
    if (p->family != FAM_RAGGED) return JT_OK;
    DevBuf& d = p->precision == JT_F64 ? p->d->ragged->d_err : p->f->ragged->d_err;
    int e = 0;
    JT_HIP(hipMemcpy(&e, d.p, sizeof(int), hipMemcpyDeviceToHost));
    if (!e) return JT_OK;
    const int zero = 0;
    JT_HIP(hipMemcpy(d.p, &zero, sizeof(int), hipMemcpyHostToDevice));
    return set_err(JT_HIP_ERROR, "a ragged-batch wait timed out; results are incomplete");
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
        int64_t pts = 0;
        if (precision == JT_F64) {
            p->d = std::make_unique<Impl<double>>();
            p->d->ragged = make_ragged_tuned<double>(lengths, count, tune, &p->desc);
            pts = p->d->ragged->total;
            p->d->ws.alloc((size_t)std::max<int64_t>(pts, 1) * sizeof(cpx<double>));  // staging
        } else {
            p->f = std::make_unique<Impl<float>>();
            p->f->ragged = make_ragged_tuned<float>(lengths, count, tune, &p->desc);
            pts = p->f->ragged->total;
            p->f->ws.alloc((size_t)std::max<int64_t>(pts, 1) * sizeof(cpx<float>));  // staging
        }
        p->op_ws[JT_COMPLEX_FORWARD] = p->op_ws[JT_COMPLEX_INVERSE] = 0;
        JT_HIP(hipEventCreateWithFlags(&p->chain, hipEventDisableTiming));
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
        std::lock_guard<std::mutex> lk(p->mu);
        return take_device_status(p);
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
        std::lock_guard<std::mutex> lk(p->mu);
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
    return op_elements(p, op);
}

int jt_execute(jt_plan* p, int op, void* data, int64_t offset, int scale, void* stream) {
    return guarded([&]() -> int {
        if (!p || !data) return set_err(JT_BAD_ARG, "null argument");
        if (offset < 0) return set_err(JT_BAD_ARG, "negative offset");
        if (op_elements(p, op) < 0) return set_err(JT_BAD_ARG, "op not defined for this plan");
        const size_t es = p->precision == JT_F64 ? 8 : 4;
        std::lock_guard<std::mutex> lk(p->mu);
        return execute_locked(p, op, static_cast<char*>(data) + offset * (int64_t)es, scale,
                              reinterpret_cast<hipStream_t>(stream));
    });
}

int jt_execute_noalloc(jt_plan* p, int op, void* data, int64_t offset, int scale, void* stream) {
    return guarded([&]() -> int {
        if (!p || !data) return set_err(JT_BAD_ARG, "null argument");
        if (offset < 0) return set_err(JT_BAD_ARG, "negative offset");
        const int64_t ne = op_elements(p, op);
        if (ne < 0 || p->op_ws[op] < 0) return set_err(JT_BAD_ARG, "op not defined for this plan");
        if (p->family == FAM_RAGGED)
            return set_err(JT_UNSUPPORTED, "ragged batches keep their work queue in the plan; use jt_execute");
        const size_t es = p->precision == JT_F64 ? 8 : 4;
        const size_t cs = 2 * es;
        char* ptr = static_cast<char*>(data) + offset * (int64_t)es;
        const bool staged = reinterpret_cast<uintptr_t>(ptr) % cs != 0;
        const size_t bytes = (size_t)ne * es;
        const size_t need = (size_t)p->op_ws[op] * cs + (staged ? (bytes + cs - 1) / cs * cs : 0);
        if (need > JT_NOALLOC_SCRATCH_BYTES)
            return set_err(JT_TOO_BIG, "jt_execute_noalloc: this op needs " + std::to_string(need) +
                                           " bytes of scratch, more than its " +
                                           std::to_string(JT_NOALLOC_SCRATCH_BYTES) + "-byte device array");
        char* ws = scratch_on_current_device();
        hipStream_t st = reinterpret_cast<hipStream_t>(stream);
        std::lock_guard<std::mutex> lk(p->mu);  // host-side plan state only
        char* d = ptr;
        if (staged) {
            d = ws + (size_t)p->op_ws[op] * cs;
            JT_HIP(hipMemcpyAsync(d, ptr, bytes, hipMemcpyDeviceToDevice, st));
        }
        int rc;
        if (p->precision == JT_F64)
            rc = exec_t<double>(p, op, reinterpret_cast<double*>(d), scale != 0, reinterpret_cast<cpx<double>*>(ws),
                                st);
        else
            rc = exec_t<float>(p, op, reinterpret_cast<float*>(d), scale != 0, reinterpret_cast<cpx<float>*>(ws), st);
        if (rc == JT_OK && staged) JT_HIP(hipMemcpyAsync(ptr, d, bytes, hipMemcpyDeviceToDevice, st));
        return rc;
    });
}

int jt_execute_host(jt_plan* p, int op, void* data, int64_t offset, int scale) {
    return guarded([&]() -> int {
        if (!p || !data) return set_err(JT_BAD_ARG, "null argument");
        if (offset < 0) return set_err(JT_BAD_ARG, "negative offset");
        const int64_t ne = op_elements(p, op);
        if (ne < 0 || p->op_ws[op] < 0) return set_err(JT_BAD_ARG, "op not defined for this plan");
        const size_t es = p->precision == JT_F64 ? 8 : 4;
        const size_t bytes = (size_t)ne * es;
        char* h = static_cast<char*>(data) + offset * (int64_t)es;
        // The whole call holds the plan: its staging area is the plan's.
        std::lock_guard<std::mutex> lk(p->mu);
        if (p->chain_set) JT_HIP(hipEventSynchronize(p->chain));  // earlier device calls are done
        char* d = staging_area(p, op);
        hipStream_t st = nullptr;
        JT_HIP(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
        int rc = execute_locked(p, op, d, scale, st);
        if (rc != JT_OK) return rc;
        JT_HIP(hipStreamSynchronize(st));
        rc = take_device_status(p);
        if (rc != JT_OK) return rc;
        JT_HIP(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost));
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
