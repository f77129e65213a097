// Plan building and execution, templated on the precision; instantiated by
// jt_impl_f64.hip and jt_impl_f32.hip.
#pragma once
#include <map>
#include <mutex>

#include "jt_plan.hpp"

namespace jt_api {

// Workspace (complex elements) that op touches; -1 if the op is not defined.
// Mirrors the buffer uses of the execution paths below.
template <typename T>
int64_t op_ws_need(jt_plan* p, int op) {
    Impl<T>& I = impl<T>(p);
    if (op_elements(p, op) < 0) return -1;
    if (p->family == FAM_RAGGED) return 0;  // the batch's buffers live in its RaggedPlan
    const int64_t tot = total(p) * p->batch;
    if (p->family != FAM_FFT) {
        int64_t w = 0;
        for (int a = 0; a < p->ndim; a++) w = std::max(w, I.r2r[a]->ws_elems(tot / p->dims[a]));
        return w;
    }
    // complex passes over every axis (run_complex_nd)
    int64_t wc = 0;
    if (p->ndim == 1) wc = I.ax[0]->workspace(p->batch);
    else
        for (int a = 0; a < p->ndim; a++) wc = std::max(wc, axis_ws(*I.ax[a], tot / p->dims[a]));
    switch (op) {
        case JT_COMPLEX_FORWARD:
        case JT_COMPLEX_INVERSE:
            return wc;
        case JT_REAL_FORWARD:
        case JT_REAL_INVERSE: {
            const Real1D<T>& r = *I.real;
            const int64_t lines = tot / p->dims[p->ndim - 1];
            // even n: half-length complex plan; odd n: full complex copy
            int64_t w = r.half ? r.half->workspace(lines) : lines * r.n + r.full->workspace(lines);
            // multi-D: column passes over half-width planes (within wc)
            if (p->ndim > 1) w = std::max(w, wc);
            return w;
        }
        default:  // *Full: reals expanded to complex in the workspace
            if (p->ndim == 1)  // (one line: batch plans reject the *Full ops)
                return p->batch > 1 ? 0 : I.real->n + I.real->full->workspace(1);
            return std::max(tot, wc);
    }
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
            for (int b = 0; b < a && !I->ax[a]; b++)
                if (p->dims[b] == p->dims[a]) I->ax[a] = I->ax[b];
            if (!I->ax[a]) I->ax[a] = make_fft1d<T>(p->dims[a]);
        }
        // packed real transforms run the last axis as a real plan; the other
        // axes of a multi-D real transform are complex passes of full length
        I->real = make_real1d<T>(p->dims[p->ndim - 1]);
    } else {
        for (int a = 0; a < p->ndim; a++) {
            for (int b = 0; b < a && !I->r2r[a]; b++)
                if (p->dims[b] == p->dims[a]) I->r2r[a] = I->r2r[b];
            if (!I->r2r[a]) I->r2r[a] = make_r2r1d<T>(p->family, p->dims[a]);
        }
    }
    apply_flags(*I, p->flags);
    if constexpr (sizeof(T) == 8) p->d = std::move(I);
    else p->f = std::move(I);
    // one buffer: each op's workspace followed by its staging area
    int64_t w = 1;
    for (int op = 0; op < NUM_OPS; op++) {
        p->op_ws[op] = op_ws_need<T>(p, op);
        if (p->op_ws[op] < 0) continue;
        w = std::max(w, p->op_ws[op] + (op_elements(p, op) + 1) / 2);
    }
    impl<T>(p).ws.alloc((size_t)w * sizeof(cpx<T>));
}

// ---------------------------------------------------------------------------
// execution
// ---------------------------------------------------------------------------
template <typename T, int DIR>
void run_complex_nd(jt_plan* p, cpx<T>* a, bool scale, cpx<T>* ws, hipStream_t st) {
    Impl<T>& I = impl<T>(p);
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
void run_real_forward_nd(jt_plan* p, T* a, cpx<T>* ws, hipStream_t st) {
    Impl<T>& I = impl<T>(p);
    cpx<T>* z = reinterpret_cast<cpx<T>*>(a);
    if (p->ndim == 2) {
        const int64_t r = p->dims[0], c = p->dims[1], ch = c / 2;
        real_forward<T>(*I.real, a, r, ws, st);
        if (col_fft_fix2d<T, -1, 1>(*I.ax[0], z, r, ch, T(1), st)) return;
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
void run_real_inverse_nd(jt_plan* p, T* a, bool scale, cpx<T>* ws, hipStream_t st) {
    Impl<T>& I = impl<T>(p);
    cpx<T>* z = reinterpret_cast<cpx<T>*>(a);
    if (p->ndim == 2) {
        const int64_t r = p->dims[0], c = p->dims[1], ch = c / 2;
        const T cs = scale ? T(1.0 / (double)r) : T(1);
        if (!col_fft_fix2d<T, 1, -1>(*I.ax[0], z, r, ch, cs, st)) {
            { JT_PROF(st, "rdft2d_fix");
            hipLaunchKernelGGL((rdft2d_fix<T>), ew_grid(r / 2), dim3(256), 0, st, a, r, c, -1);
            }
            JT_HIP(hipGetLastError());
            axis_c2c<T, 1>(*I.ax[0], z, LineMap{ch, ch, 1, 0, ch}, cs, ws, st);
        }
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
void run_real_full_nd(jt_plan* p, T* a, bool scale, cpx<T>* ws, hipStream_t st) {
    const int64_t tot = total(p);
    // reals -> complex through the workspace, then the complex transform
    { JT_PROF(st, "real_to_complex");
    hipLaunchKernelGGL((real_to_complex<T>), ew_grid(tot), dim3(256), 0, st, a, ws, tot, (int64_t)1);
    }
    JT_HIP(hipGetLastError());
    { JT_PROF(st, "copy_complex");
    hipLaunchKernelGGL((copy_complex<T>), ew_grid(tot), dim3(256), 0, st, ws, reinterpret_cast<cpx<T>*>(a),
                       tot);
    }
    JT_HIP(hipGetLastError());
    run_complex_nd<T, DIR>(p, reinterpret_cast<cpx<T>*>(a), scale, ws, st);
}

template <typename T>
int exec_t(jt_plan* p, int op, T* a, bool scale, cpx<T>* ws, hipStream_t st) {
    Impl<T>& I = impl<T>(p);
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
            run_complex_nd<T, -1>(p, reinterpret_cast<cpx<T>*>(a), false, ws, st);
            return JT_OK;
        case JT_COMPLEX_INVERSE:
            if (nd == 1 && n == 1) return JT_OK;
            run_complex_nd<T, 1>(p, reinterpret_cast<cpx<T>*>(a), scale, ws, st);
            return JT_OK;
        case JT_REAL_FORWARD:
            if (nd == 1) {
                real_forward<T>(*I.real, a, p->batch, ws, st);
                return JT_OK;
            }
            if (!p->pow2_all)
                return set_err(JT_NOT_POW2, nd == 2 ? "rows and columns must be power of two numbers"
                                                    : "slices, rows and columns must be power of two numbers");
            run_real_forward_nd<T>(p, a, ws, st);
            return JT_OK;
        case JT_REAL_INVERSE:
            if (nd == 1) {
                real_inverse<T>(*I.real, a, p->batch, scale, ws, st);
                return JT_OK;
            }
            if (!p->pow2_all)
                return set_err(JT_NOT_POW2, nd == 2 ? "rows and columns must be power of two numbers"
                                                    : "slices, rows and columns must be power of two numbers");
            run_real_inverse_nd<T>(p, a, scale, ws, st);
            return JT_OK;
        case JT_REAL_FORWARD_FULL:
            if (nd == 1) {
                if (p->batch != 1) return set_err(JT_UNSUPPORTED, "realForwardFull on a batch plan");
                real_full<T, -1>(*I.real, a, 1, false, ws, st);
                return JT_OK;
            }
            run_real_full_nd<T, -1>(p, a, false, ws, st);
            return JT_OK;
        case JT_REAL_INVERSE_FULL:
            if (nd == 1) {
                if (p->batch != 1) return set_err(JT_UNSUPPORTED, "realInverseFull on a batch plan");
                real_full<T, 1>(*I.real, a, 1, scale, ws, st);
                return JT_OK;
            }
            run_real_full_nd<T, 1>(p, a, scale, ws, st);
            return JT_OK;
        default:
            return set_err(JT_BAD_ARG, "op not defined for this plan");
    }
}

// Tuned schedules by (precision, length list), for the life of the process.
inline std::mutex g_tune_mu;
inline std::map<std::pair<int, std::vector<int64_t>>, RTune> g_tuned;

template <typename T>
std::unique_ptr<RaggedPlan<T>> make_ragged_tuned(const int64_t* lengths, int64_t count, bool tune,
                                                 std::string* desc) {
    std::unique_ptr<RaggedPlan<T>> P;
    if (!tune) {
        // JT_RAGGED_SCHEDULE="wgs=2,nt=512,tplx=0,chunk=8,order=0,lag=16" (any
        // subset; the format of jt_plan_describe) fixes the schedule of
        // untuned plans, for tests and measurements of one schedule
        RTune t;
        if (const char* e = std::getenv("JT_RAGGED_SCHEDULE")) {
            std::string sch(e);
            size_t pos = 0;
            while (pos < sch.size()) {
                size_t end = sch.find(',', pos);
                if (end == std::string::npos) end = sch.size();
                const std::string kv = sch.substr(pos, end - pos);
                const size_t eq = kv.find('=');
                if (eq != std::string::npos) {
                    const std::string k = kv.substr(0, eq);
                    const int v = std::atoi(kv.c_str() + eq + 1);
                    if (k == "wgs") t.wgs = v;
                    else if (k == "nt") t.nt = v;
                    else if (k == "tplx") t.tplx = v;
                    else if (k == "chunk") t.chunk = v;
                    else if (k == "order") t.order = v;
                    else if (k == "lag") t.lag = v;
                    else throw JtError(ST_BAD_ARG, "JT_RAGGED_SCHEDULE: unknown key " + k);
                }
                pos = end + 1;
            }
        }
        P = make_ragged<T>(lengths, count, t);
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

}  // namespace jt_api
