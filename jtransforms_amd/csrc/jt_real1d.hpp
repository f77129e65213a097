// 1D real-data transforms on top of the complex engine: realForward /
// realInverse in the reference's packed layout (DoubleFFT_1D.java:436-561,
// :946-989) and the full-spectrum variants (:678-755, :1112-1195).
#pragma once
#include "jt_fft1d.hpp"

namespace jt {

// ---- elementwise helpers ----------------------------------------------------

// packed post pass in place over lines of N complex (= 2N reals), pairs (k, N-k)
template <typename T>
__global__ void r2c_post_inplace(cpx<T>* __restrict__ z, const cpx<T>* __restrict__ rtw, int64_t loff, int64_t N,
                                 int64_t lines, T scale) {
    int64_t half = N / 2 + 1;
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i < lines * half; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t b = i / half, k = i % half;
        cpx<T>* a = z + b * N;
        if (k == 0) {
            cpx<T> z0 = a[0];
            a[0] = cpx<T>{(z0.x + z0.y) * scale, (z0.x - z0.y) * scale};
            continue;
        }
        int64_t kk = N - k;
        cpx<T> A = a[k];
        cpx<T> Zc = a[kk];
        cpx<T> B = {Zc.x, -Zc.y};
        cpx<T> fe = cadd(A, B), d = csub(A, B);
        cpx<T> fo = {d.y, -d.x};
        cpx<T> wf = cmul(cmul(rtw[k >> 8], rtw[loff + (k & 255)]), fo);
        T h = T(0.5) * scale;
        a[k] = cpx<T>{(fe.x + wf.x) * h, (fe.y + wf.y) * h};
        if (kk != k) a[kk] = cpx<T>{(fe.x - wf.x) * h, -(fe.y - wf.y) * h};
    }
}

// packed pre pass in place (inverse of r2c_post up to the factor 2N)
template <typename T>
__global__ void c2r_pre_inplace(cpx<T>* __restrict__ z, const cpx<T>* __restrict__ rtw, int64_t loff, int64_t N,
                                int64_t lines) {
    int64_t half = N / 2 + 1;
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i < lines * half; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t b = i / half, k = i % half;
        cpx<T>* a = z + b * N;
        if (k == 0) {
            cpx<T> x = a[0];
            a[0] = cpx<T>{x.x + x.y, x.x - x.y};
            continue;
        }
        int64_t kk = N - k;
        cpx<T> Xk = a[k];
        cpx<T> Xn = a[kk];
        cpx<T> Xc = {Xn.x, -Xn.y};
        cpx<T> E = cadd(Xk, Xc), D = csub(Xk, Xc);
        const cpx<T> wk = cmul(rtw[k >> 8], rtw[loff + (k & 255)]);
        cpx<T> w = {wk.x, -wk.y};
        cpx<T> O = cmul(D, w);
        a[k] = cpx<T>{E.x - O.y, E.y + O.x};
        if (kk != k) a[kk] = cpx<T>{E.x + O.y, O.x - E.y};
    }
}

// reals (line stride n) -> complex with zero imaginary parts (line stride n)
template <typename T>
__global__ void real_to_complex(const T* __restrict__ x, cpx<T>* __restrict__ y, int64_t n,
                                int64_t lines) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i < lines * n; i += (int64_t)gridDim.x * blockDim.x) y[i] = cpx<T>{x[i], T(0)};
}

// full spectrum (complex n) -> packed reals (n), both line stride n
template <typename T>
__global__ void pack_real(const cpx<T>* __restrict__ X, T* __restrict__ a, int64_t n, int64_t lines) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i < lines * n; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t b = i / n, j = i % n;
        const cpx<T>* Xb = X + b * n;
        T v;
        if (j == 0) v = Xb[0].x;
        else if (j == 1) v = (n % 2 == 0) ? Xb[n / 2].x : Xb[(n - 1) / 2].y;
        else if (j % 2 == 0) v = Xb[j / 2].x;
        else v = Xb[j / 2].y;
        a[i] = v;
    }
}

// packed reals -> full Hermitian spectrum (complex n)
template <typename T>
__global__ void unpack_real(const T* __restrict__ a, cpx<T>* __restrict__ X, int64_t n, int64_t lines) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i < lines * n; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t b = i / n, k = i % n;
        const T* ab = a + b * n;
        int64_t kk = k <= n / 2 ? k : n - k;
        bool cj = k > n / 2;
        T re, im;
        if (kk == 0) {
            re = ab[0];
            im = 0;
        } else if (n % 2 == 0 && kk == n / 2) {
            re = ab[1];
            im = 0;
        } else if (n % 2 == 1 && kk == (n - 1) / 2) {
            re = ab[2 * kk];
            im = ab[1];
        } else {
            re = ab[2 * kk];
            im = ab[2 * kk + 1];
        }
        X[i] = cpx<T>{re, cj ? -im : im};
    }
}

// real parts of complex lines -> reals, times s
template <typename T>
__global__ void complex_real_part(const cpx<T>* __restrict__ X, T* __restrict__ a, int64_t count, T s) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i < count; i += (int64_t)gridDim.x * blockDim.x) a[i] = X[i].x * s;
}

template <typename T>
__global__ void copy_complex(const cpx<T>* __restrict__ s, cpx<T>* __restrict__ d, int64_t count) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i < count; i += (int64_t)gridDim.x * blockDim.x) d[i] = s[i];
}

// ---- real plan ----------------------------------------------------------------

template <typename T>
struct Real1D {
    int64_t n = 0;
    std::unique_ptr<Fft1D<T>> full;  // length n (odd n, *Full ops)
    std::unique_ptr<Fft1D<T>> half;  // length n/2 (even n)
    // w^k = exp(-2 pi i k/n): with a one-pass half plan, w^k for k in
    // [0, n/4 + 1] (the IO hooks of its kernel); otherwise (large n) as two levels,
    // w^k = rtw[k >> 8] * rtw[rlo + (k & 255)] for k in [0, n/4 + 1], so the
    // table stays O(n / 256) long
    DevBuf rtw;
    int64_t rlo = 0;
    int64_t ws_elems(int64_t lines) const {
        int64_t w = 0;
        if (half) w = std::max<int64_t>(w, half->workspace(lines));
        if (full) w = std::max<int64_t>(w, lines * n + full->workspace(lines));
        return w;
    }
};

template <typename T>
std::unique_ptr<Real1D<T>> make_real1d(int64_t n) {
    auto r = std::make_unique<Real1D<T>>();
    r->n = n;
    r->full = make_fft1d<T>(n);
    if (n % 2 == 0 && n >= 2) {
        r->half = make_fft1d<T>(n / 2);
        std::vector<cpx<T>> h;
        if (r->half->single()) {
            h.resize(n / 4 + 2);
            for (int64_t k = 0; k < (int64_t)h.size(); k++) h[k] = expm<T>(k, n);
        } else {
            const int64_t nhi = ((n / 4 + 1) >> 8) + 1;
            for (int64_t i = 0; i < nhi; i++) h.push_back(expm<T>(i << 8, n));
            r->rlo = nhi;
            for (int64_t i = 0; i < 256; i++) h.push_back(expm<T>(i, n));
        }
        upload(r->rtw, h);
    }
    return r;
}

// realForward on `lines` lines of n reals at stride n.
template <typename T>
void real_forward(const Real1D<T>& r, T* a, int64_t lines, cpx<T>* ws, hipStream_t st) {
    const int64_t n = r.n;
    if (n == 1 || lines <= 0) return;
    if (r.half) {
        const int64_t N = n / 2;
        cpx<T>* z = reinterpret_cast<cpx<T>*>(a);
        if (r.half->single()) {
            IOR2C<T> io{z, a, r.rtw.template as<cpx<T>>(), LineMap{lines, lines, N, 0, 1}, lines, (int)N,
                        T(1), false};
            run_single<T, -1, false>(*r.half, io, st);
        } else {
            c2c_contig<T, -1>(*r.half, z, z, lines, T(1), ws, st);
            { JT_PROF(st, "r2c_post_inplace");
            hipLaunchKernelGGL((r2c_post_inplace<T>), ew_grid(lines * (N / 2 + 1)), dim3(256), 0, st, z,
                               r.rtw.template as<cpx<T>>(), r.rlo, N, lines, T(1));
            }
            JT_HIP(hipGetLastError());
        }
        return;
    }
    // odd n: full complex transform of the real input, then pack
    cpx<T>* X = ws;
    cpx<T>* ws2 = ws + lines * n;
    { JT_PROF(st, "real_to_complex");
    hipLaunchKernelGGL((real_to_complex<T>), ew_grid(lines * n), dim3(256), 0, st, a, X, n, lines);
    }
    JT_HIP(hipGetLastError());
    c2c_contig<T, -1>(*r.full, X, X, lines, T(1), ws2, st);
    { JT_PROF(st, "pack_real");
    hipLaunchKernelGGL((pack_real<T>), ew_grid(lines * n), dim3(256), 0, st, X, a, n, lines);
    }
    JT_HIP(hipGetLastError());
}

// Unscaled realInverse returns (n/2) x for powers of two and n x otherwise
// (DoubleFFT_1D.java:965 vs :977/:983).
template <typename T>
inline T real_inverse_factor(int64_t n, bool scale) {
    if (scale) return T(1.0 / (double)n);
    return (is_pow2(n) && n > 1) ? T(0.5) : T(1);
}

template <typename T>
void real_inverse(const Real1D<T>& r, T* a, int64_t lines, bool scale, cpx<T>* ws, hipStream_t st) {
    const int64_t n = r.n;
    if (n == 1 || lines <= 0) return;
    const T f = real_inverse_factor<T>(n, scale);
    if (r.half) {
        const int64_t N = n / 2;
        cpx<T>* z = reinterpret_cast<cpx<T>*>(a);
        if (r.half->single()) {
            IOC2R<T> io{a, z, r.rtw.template as<cpx<T>>(), LineMap{lines, lines, N, 0, 1}, lines, (int)N, f,
                        false};
            run_single<T, 1, false>(*r.half, io, st);
        } else {
            { JT_PROF(st, "c2r_pre_inplace");
            hipLaunchKernelGGL((c2r_pre_inplace<T>), ew_grid(lines * (N / 2 + 1)), dim3(256), 0, st, z,
                               r.rtw.template as<cpx<T>>(), r.rlo, N, lines);
            }
            JT_HIP(hipGetLastError());
            c2c_contig<T, 1>(*r.half, z, z, lines, f, ws, st);
        }
        return;
    }
    cpx<T>* X = ws;
    cpx<T>* ws2 = ws + lines * n;
    { JT_PROF(st, "unpack_real");
    hipLaunchKernelGGL((unpack_real<T>), ew_grid(lines * n), dim3(256), 0, st, a, X, n, lines);
    }
    JT_HIP(hipGetLastError());
    c2c_contig<T, 1>(*r.full, X, X, lines, T(1), ws2, st);
    // the backward sum is n x; unscaled non-power-of-two result is n x
    { JT_PROF(st, "complex_real_part");
    hipLaunchKernelGGL((complex_real_part<T>), ew_grid(lines * n), dim3(256), 0, st, X, a, lines * n,
                       scale ? T(1.0 / (double)n) : T(1));
    }
    JT_HIP(hipGetLastError());
}

// realForwardFull / realInverseFull: complex transform of the real input
// a[0..n) written as 2n values (dir -1 forward, +1 inverse with scale).
template <typename T, int DIR>
void real_full(const Real1D<T>& r, T* a, int64_t lines, bool scale, cpx<T>* ws, hipStream_t st) {
    const int64_t n = r.n;
    if (lines <= 0) return;
    // lines of the *input* are n reals apart in the reference's batched use
    // (one line per call there), outputs 2n reals apart
    cpx<T>* X = ws;
    cpx<T>* ws2 = ws + lines * n;
    { JT_PROF(st, "real_to_complex");
    hipLaunchKernelGGL((real_to_complex<T>), ew_grid(lines * n), dim3(256), 0, st, a, X, n, lines);
    }
    JT_HIP(hipGetLastError());
    T s = (DIR > 0 && scale) ? T(1.0 / (double)n) : T(1);
    c2c_contig<T, DIR>(*r.full, X, X, lines, s, ws2, st);
    { JT_PROF(st, "copy_complex");
    hipLaunchKernelGGL((copy_complex<T>), ew_grid(lines * n), dim3(256), 0, st, X,
                       reinterpret_cast<cpx<T>*>(a), lines * n);
    }
    JT_HIP(hipGetLastError());
}

}  // namespace jt
