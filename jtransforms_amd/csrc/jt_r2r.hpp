// DCT-II/III, DST-II/III and DHT (1D/2D/3D) on the complex engine.
//
// Reference semantics (dct/DoubleDCT_1D.java:169-434, dst/DoubleDST_1D.java:
// 96-318, dht/DoubleDHT_1D.java:94-268, DHT_2D/3D yTransform):
//   DCT forward  C_k = sum_j x_j cos(pi (2j+1) k / 2n); returned as C_k for
//                power-of-two n, 2 C_k otherwise; scale -> orthonormal.
//   DCT inverse  power of two: a_0 + sum_{k>=1} a_k cos(pi k (2m+1)/2n);
//                otherwise (1/n)(a_0/2 + sum_{k>=1} ...); scale -> orthonormal.
//   DST          negate odd inputs, DCT-II, reverse (inverse mirrors it).
//   DHT          H_k = Re X_k - Im X_k; inverse = forward (x 1/n if scale).
//
// Even n with a single-kernel n/2 plan runs fused in one Stockham launch per
// axis: the pre hook gathers the line with Makhoul's even/odd permutation
// (v_j = x_2j, v_{n-1-j} = x_{2j+1}) into n/2 complex points, the post hook
// applies the real-FFT split and the exp(-i pi k/2n) twiddle and stores the
// n results (with the DST's sign/reverse maps folded into the index maps).
// Other lengths use a mirrored 2n-point complex transform (the reference's
// own non-power-of-two DCT route, DoubleDCT_1D.java:195-242).
#pragma once
#include "jt_multid.hpp"

namespace jt {

enum { R2R_DCT = 1, R2R_DST = 2, R2R_DHT = 3 };

// ---------------------------------------------------------------------------
// scale conventions
// ---------------------------------------------------------------------------
struct R2RScale {
    double f0;    // extra factor on element 0 (forward: applied to output 0;
                  // inverse: applied to input 0)
    double g;     // overall factor
};

// Forward DCT-II from the unnormalised C_k.
inline R2RScale dct2_scale(int64_t n, bool scale) {
    if (scale) return {1.0 / std::sqrt(2.0), std::sqrt(2.0 / (double)n)};
    return {1.0, is_pow2(n) ? 1.0 : 2.0};
}
// Inverse DCT-III as g * y(a') with y(a) = a_0/2 + sum_{k>=1} a_k cos(.) and
// a'_0 = f0 a_0.
inline R2RScale dct3_scale(int64_t n, bool scale) {
    if (scale) return {std::sqrt(2.0), std::sqrt(2.0 / (double)n)};
    if (is_pow2(n)) return {2.0, 1.0};
    return {1.0, 1.0 / (double)n};
}

// ---------------------------------------------------------------------------
// fused IO policy.  Real data: element k of line l at data[base(l) + k*es].
// N = n/2 complex points per line.  q[k] = exp(-i pi k / 2n), k in [0, n];
// rtw[k] = exp(-2 pi i k / n), k in [0, n/4 + 1].
// ---------------------------------------------------------------------------
template <typename T, int KIND, bool FWD>
struct IOR2R {
    static constexpr bool pre_lds = true, post_lds = true;
    // lines of reals: a column group's row segment is W * sizeof(T) bytes
    static constexpr bool real_lines = true;
    static const char* tag() {
        return KIND == R2R_DHT ? "dht" : (KIND == R2R_DCT ? (FWD ? "dct2" : "dct3") : (FWD ? "dst2" : "dst3"));
    }
    std::string desc() const { return ",es=" + std::to_string(m.es); }
    T* data;
    const cpx<T>* q;
    const cpx<T>* rtw;
    LineMap m;  // real units
    int64_t nlines;
    int n, N;
    T f0, g;
    bool xcd;
    __device__ __forceinline__ int64_t base(int64_t l) const { return m.base(l); }
    __device__ cpx<T> load(int64_t, int) const { return cpx<T>{T(0), T(0)}; }
    __device__ void store(int64_t, int, cpx<T>) const {}

    __device__ __forceinline__ T rd(int64_t b, int k) const { return data[b + (int64_t)k * m.es]; }
    __device__ __forceinline__ void wr(int64_t b, int k, T v) const { data[b + (int64_t)k * m.es] = v; }

    // forward input x_j with the DST sign map
    __device__ __forceinline__ T xin(int64_t b, int j) const {
        T v = rd(b, j);
        if (KIND == R2R_DST && (j & 1)) v = -v;
        return v;
    }
    // inverse input a'_k (DST reads reversed), 0 for k == n
    __device__ __forceinline__ T ain(int64_t b, int k) const {
        if (k >= n) return T(0);
        T v = KIND == R2R_DST ? rd(b, n - 1 - k) : rd(b, k);
        return k == 0 ? v * f0 : v;
    }

    // forward: element i of the line -> its slot in the N complex inputs
    __device__ __forceinline__ void put_fwd(int i, T v, T* re, T* im) const {
        int j;
        if constexpr (KIND == R2R_DHT) j = i;  // natural pairs z_m = x_2m + i x_2m+1
        else j = (i & 1) ? n - 1 - (i >> 1) : (i >> 1);  // Makhoul: v_j = x_2j, v_{n-1-j} = x_2j+1
        int mm = j >> 1;
        if (j & 1) im[lpos(mm)] = v;
        else re[lpos(mm)] = v;
    }
    __device__ __forceinline__ T fwd_in(int64_t b, int i) const {
        if constexpr (KIND == R2R_DHT) return rd(b, i);
        else return xin(b, i);
    }
    // inverse: Z_k, Z_{N-k} from a'_k, a'_{n-k}, a'_{N-k}, a'_{N+k}
    __device__ __forceinline__ void inv_pair(int k, T ak, T ank, T akk, T ankk, T* re, T* im) const {
        int kk = N - k;
        cpx<T> Vk = cmulc(cpx<T>{ak, -ank}, q[k]);
        cpx<T> Vc = cmulc(cpx<T>{akk, -ankk}, q[kk]);
        cpx<T> Xc = {Vc.x, -Vc.y};  // conj V_{N-k}
        cpx<T> E = cadd(Vk, Xc), D = csub(Vk, Xc);
        cpx<T> w = {rtw[k].x, -rtw[k].y};
        cpx<T> O = cmul(D, w);
        re[lpos(k)] = E.x - O.y;
        im[lpos(k)] = E.y + O.x;
        if (kk != k) {
            re[lpos(kk)] = E.x + O.y;
            im[lpos(kk)] = O.x - E.y;
        }
    }

    template <int CN, int CTL>
    __device__ void pre(int64_t, int64_t b, T* re, T* im, int t, int nt) const {
        if constexpr (FWD || KIND == R2R_DHT) {
            if constexpr (CN > 0) {
                // 2N reals per line: issue every load before the LDS writes
                constexpr int IT = (2 * CN) / CTL;
                T vals[IT];
#pragma unroll
                for (int q2 = 0; q2 < IT; q2++) vals[q2] = fwd_in(b, t + q2 * CTL);
#pragma unroll
                for (int q2 = 0; q2 < IT; q2++) put_fwd(t + q2 * CTL, vals[q2], re, im);
            } else {
                for (int i = t; i < n; i += nt) put_fwd(i, fwd_in(b, i), re, im);
            }
        } else {
            // DCT-III / DST-III: V_k = conj(q_k)(a'_k - i a'_{n-k}), then the
            // real-inverse split into Z (N points)
            if (t == 0) {
                T v0 = ain(b, 0);
                T vN = T(1.41421356237309504880168872420970) * ain(b, N);
                re[lpos(0)] = v0 + vN;
                im[lpos(0)] = v0 - vN;
            }
            if constexpr (CN > 0) {
                constexpr int IT = (CN / 2 + CTL - 1) / CTL;
                T a0[IT], a1[IT], a2[IT], a3[IT];
#pragma unroll
                for (int q2 = 0; q2 < IT; q2++) {
                    int k = 1 + t + q2 * CTL;
                    if (k <= CN / 2) {
                        a0[q2] = ain(b, k);
                        a1[q2] = ain(b, n - k);
                        a2[q2] = ain(b, N - k);
                        a3[q2] = ain(b, N + k);
                    }
                }
#pragma unroll
                for (int q2 = 0; q2 < IT; q2++) {
                    int k = 1 + t + q2 * CTL;
                    if (k <= CN / 2) inv_pair(k, a0[q2], a1[q2], a2[q2], a3[q2], re, im);
                }
            } else {
                for (int k = 1 + t; k <= N / 2; k += nt)
                    inv_pair(k, ain(b, k), ain(b, n - k), ain(b, N - k), ain(b, N + k), re, im);
            }
        }
    }

    // output index map (DST reverses the DCT-II output)
    __device__ __forceinline__ void out_fwd(int64_t b, int k, T v) const {
        T s = g;
        if (KIND != R2R_DHT && k == 0) s *= f0;
        if (KIND == R2R_DST) wr(b, n - 1 - k, v * s);
        else wr(b, k, v * s);
    }

    template <int CN, int CTL>
    __device__ void post(int64_t, int64_t b, const T* re, const T* im, int t, int nt) const {
        if constexpr (FWD || KIND == R2R_DHT) {
            // real-FFT split V_k / V_{N-k} from Z, then the transform outputs
            for (int k = t; k <= N / 2; k += nt) {
                if (k == 0) {
                    T zr = re[lpos(0)], zi = im[lpos(0)];
                    T V0 = zr + zi, VN = zr - zi;
                    if constexpr (KIND == R2R_DHT) {
                        out_fwd(b, 0, V0);
                        out_fwd(b, N, VN);
                    } else {
                        out_fwd(b, 0, V0);
                        out_fwd(b, N, T(0.70710678118654752440084436210485) * VN);
                    }
                    continue;
                }
                int kk = N - k;
                cpx<T> A = {re[lpos(k)], im[lpos(k)]};
                cpx<T> B = {re[lpos(kk)], -im[lpos(kk)]};
                cpx<T> fe = cadd(A, B), d = csub(A, B);
                cpx<T> fo = {d.y, -d.x};
                cpx<T> wf = cmul(rtw[k], fo);
                cpx<T> Vk = {(fe.x + wf.x) * T(0.5), (fe.y + wf.y) * T(0.5)};
                cpx<T> Vc = {(fe.x - wf.x) * T(0.5), -(fe.y - wf.y) * T(0.5)};  // V_{N-k}
                if constexpr (KIND == R2R_DHT) {
                    out_fwd(b, k, Vk.x - Vk.y);
                    out_fwd(b, n - k, Vk.x + Vk.y);
                    if (kk != k) {
                        out_fwd(b, kk, Vc.x - Vc.y);
                        out_fwd(b, N + k, Vc.x + Vc.y);
                    }
                } else {
                    cpx<T> u = cmul(q[k], Vk);
                    out_fwd(b, k, u.x);
                    out_fwd(b, n - k, -u.y);
                    if (kk != k) {
                        cpx<T> uc = cmul(q[kk], Vc);
                        out_fwd(b, kk, uc.x);
                        out_fwd(b, N + k, -uc.y);
                    }
                }
            }
        } else {
            // inverse: LDS holds n * (v_2m + i v_2m+1); x_2j = v_j, x_2j+1 = v_{n-1-j}
            for (int i = t; i < n; i += nt) {
                int j = (i & 1) ? n - 1 - (i >> 1) : (i >> 1);
                int mm = j >> 1;
                T v = (j & 1) ? im[lpos(mm)] : re[lpos(mm)];
                v *= g;
                if (KIND == R2R_DST && (i & 1)) v = -v;
                wr(b, i, v);
            }
        }
    }
};

// ---------------------------------------------------------------------------
// fallback kernels (any n): through a complex transform of length 2n (DCT,
// DST) or n (DHT) on gathered lines
// ---------------------------------------------------------------------------
template <typename T, int KIND, bool FWD>
__global__ void r2r_fb_pre(const T* __restrict__ x, cpx<T>* __restrict__ y, LineMap m, int64_t n,
                           int64_t L, const cpx<T>* __restrict__ q, T f0) {
    // L = complex length per line (2n or n)
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i < m.nlines * L; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t l = i / L, j = i % L;
        int64_t b = m.base(l);
        cpx<T> v{T(0), T(0)};
        if (KIND == R2R_DHT) {
            v.x = x[b + j * m.es];
        } else if (FWD) {
            int64_t src = j < n ? j : 2 * n - 1 - j;  // mirrored extension
            T xv = x[b + src * m.es];
            if (KIND == R2R_DST && (src & 1)) xv = -xv;
            v.x = xv;
        } else {
            // u_k = c_k a'_k exp(+i pi k/2n), k < n; zero above
            if (j < n) {
                int64_t kk = KIND == R2R_DST ? n - 1 - j : j;
                T a = x[b + kk * m.es];
                T c = j == 0 ? T(0.5) * f0 : T(1);
                cpx<T> w = {q[j].x, -q[j].y};
                v = cpx<T>{a * c * w.x, a * c * w.y};
            }
        }
        y[i] = v;
    }
}

template <typename T, int KIND, bool FWD>
__global__ void r2r_fb_post(const cpx<T>* __restrict__ y, T* __restrict__ x, LineMap m, int64_t n,
                            int64_t L, const cpx<T>* __restrict__ q, T f0, T g) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i < m.nlines * n; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t l = i / n, k = i % n;
        int64_t b = m.base(l);
        const cpx<T>* yl = y + l * L;
        T v;
        if (KIND == R2R_DHT) {
            v = (yl[k].x - yl[k].y) * g;
            x[b + k * m.es] = v;
        } else if (FWD) {
            // DFT_2n(mirror) = 2 exp(i pi k/2n) C_k
            cpx<T> u = cmul(q[k], yl[k]);
            v = T(0.5) * u.x * g;
            if (k == 0) v *= f0;
            int64_t dst = KIND == R2R_DST ? n - 1 - k : k;
            x[b + dst * m.es] = v;
        } else {
            v = yl[k].x * g * T(2);
            if (KIND == R2R_DST && (k & 1)) v = -v;
            x[b + k * m.es] = v;
        }
    }
}

// The multi-D DHT: separable 1D transforms along each axis give
// T(k) = sum x cas(a1) cas(a2) (...); the reference then turns T into the
// transform with cas(a1 + a2 (+ a3)) (its yTransform, DoubleDHT_2D.java:
// 1288-1309, DoubleDHT_3D.java:2314-2356).  2D: cas(x+y) = (cas(x)cas(y) +
// cas(-x)cas(y) + cas(x)cas(-y) - cas(-x)cas(-y)) / 2, so over the group
// {r, -r} x {c, -c} with T values A, B, C, D at (r,c), (-r,c), (r,-c),
// (-r,-c) each member moves by E = (A + D - B - C) / 2 (A - E, B + E, ...).
template <typename T>
__global__ void ytransform2d(T* __restrict__ a, int64_t rows, int64_t cols) {
    int64_t nr = rows / 2 + 1, nc = cols / 2 + 1;
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i < nr * nc; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = i / nc, c = i % nc;
        int64_t mr = (rows - r) % rows, mc = (cols - c) % cols;
        // each group {r,mr} x {c,mc} once: skip the duplicate visits that
        // the reference's loop makes for r == rows/2 or c == cols/2 (they are
        // self-mirrored, so the update is the identity there)
        T A = a[r * cols + c], B = a[mr * cols + c], C = a[r * cols + mc], D = a[mr * cols + mc];
        T E = ((A + D) - (B + C)) / 2;
        if (r == mr || c == mc) continue;  // E is zero for self-mirrored groups
        a[r * cols + c] = A - E;
        a[mr * cols + c] = B + E;
        a[r * cols + mc] = C + E;
        a[mr * cols + mc] = D - E;
    }
}

// 3D: from cas(x+y+z) = (cas(-x)cas(y)cas(z) + cas(x)cas(-y)cas(z) +
// cas(x)cas(y)cas(-z) - cas(-x)cas(-y)cas(-z)) / 2 the true transform at k is
//   H(k) = (T(-k1,k2,k3) + T(k1,-k2,k3) + T(k1,k2,-k3) - T(-k1,-k2,-k3)) / 2
// of the separable one T.  One thread per orbit of (s, r, c) under negation
// of each index: member b (bit 0 slice, bit 1 row, bit 2 column negated)
// gets (T[b^1] + T[b^2] + T[b^4] - T[b^7]) / 2; self-mirrored indices give
// coinciding members that receive equal values.
template <typename T>
__global__ void ytransform3d(T* __restrict__ a, int64_t sl, int64_t rows, int64_t cols) {
    const int64_t ns = sl / 2 + 1, nr = rows / 2 + 1, nc = cols / 2 + 1;
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i < ns * nr * nc; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = i / (nr * nc), r = (i / nc) % nr, c = i % nc;
        const int64_t sp[2] = {s * rows * cols, ((sl - s) % sl) * rows * cols};
        const int64_t rp[2] = {r * cols, ((rows - r) % rows) * cols};
        const int64_t cp[2] = {c, (cols - c) % cols};
        int64_t at[8];
        T t[8];
#pragma unroll
        for (int m = 0; m < 8; m++) {
            at[m] = sp[m & 1] + rp[(m >> 1) & 1] + cp[(m >> 2) & 1];
            t[m] = a[at[m]];
        }
#pragma unroll
        for (int m = 0; m < 8; m++) a[at[m]] = (t[m ^ 1] + t[m ^ 2] + t[m ^ 4] - t[m ^ 7]) / 2;
    }
}

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
template <typename T>
struct R2R1D {
    int kind = R2R_DCT;
    int64_t n = 0;
    std::unique_ptr<Fft1D<T>> half;  // n/2, fused path
    std::unique_ptr<Fft1D<T>> big;   // 2n (DCT/DST) or n (DHT), fallback
    DevBuf q, rtw;
    bool fused() const { return half && half->single(); }
    int64_t L() const { return kind == R2R_DHT ? n : 2 * n; }
    int64_t ws_elems(int64_t lines) const {
        if (fused() || n == 1) return 0;
        return lines * L() + big->workspace(lines);
    }
};

template <typename T>
std::unique_ptr<R2R1D<T>> make_r2r1d(int family, int64_t n) {
    auto r = std::make_unique<R2R1D<T>>();
    r->kind = family == 1 ? R2R_DCT : (family == 2 ? R2R_DST : R2R_DHT);
    r->n = n;
    if (n == 1) return r;
    std::vector<cpx<T>> q(n + 1);
    for (int64_t k = 0; k <= n; k++) q[k] = expm<T>(k, 4 * n);
    upload(r->q, q);
    if (n % 2 == 0 && n >= 4) {
        auto h = make_fft1d<T>(n / 2);
        if (h->single()) {
            r->half = std::move(h);
            std::vector<cpx<T>> w(n / 4 + 2);
            for (int64_t k = 0; k < (int64_t)w.size(); k++) w[k] = expm<T>(k, n);
            upload(r->rtw, w);
            return r;
        }
    }
    r->big = make_fft1d<T>(r->L());
    return r;
}

template <typename T, int KIND, bool FWD>
void r2r_axis_k(const R2R1D<T>& r, T* data, LineMap lm, bool scale, cpx<T>* ws, hipStream_t st) {
    const int64_t n = r.n;
    // DHT: `scale` is only passed for the inverse (forward(a) has no scale)
    R2RScale sc = (KIND == R2R_DHT) ? R2RScale{1.0, scale ? 1.0 / (double)n : 1.0}
                                    : (FWD ? dct2_scale(n, scale) : dct3_scale(n, scale));
    if (r.fused()) {
        const int64_t N = n / 2;
        // inverse: LDS holds n * v; y(a') = (n/2) v  -> store factor g*(n/2)/n = g/2
        T g = FWD || KIND == R2R_DHT ? T(sc.g) : T(sc.g * 0.5);
        IOR2R<T, KIND, FWD> io{data, r.q.template as<cpx<T>>(), r.rtw.template as<cpx<T>>(), lm, lm.nlines,
                               (int)n, (int)N, T(sc.f0), g, lm.es != 1};
        if (lm.es == 1) run_single<T, (FWD || KIND == R2R_DHT) ? -1 : 1, false>(*r.half, io, st);
        else run_single<T, (FWD || KIND == R2R_DHT) ? -1 : 1, true>(*r.half, io, st);
        return;
    }
    const int64_t L = r.L();
    cpx<T>* y = ws;
    cpx<T>* ws2 = ws + lm.nlines * L;
    { JT_PROF(st, "r2r_fb_pre");
    hipLaunchKernelGGL((r2r_fb_pre<T, KIND, FWD>), ew_grid(lm.nlines * L), dim3(256), 0, st, data, y, lm, n, L,
                       r.q.template as<cpx<T>>(), T(sc.f0));
    }
    JT_HIP(hipGetLastError());
    if (KIND == R2R_DHT || FWD) c2c_contig<T, -1>(*r.big, y, y, lm.nlines, T(1), ws2, st);
    else c2c_contig<T, 1>(*r.big, y, y, lm.nlines, T(1), ws2, st);
    // inverse: backward 2n sum of u gives y(a') directly (the 1/2 on a_0 is
    // in u_0): fb_post multiplies by 2*g*0.5 -> use g/2 here
    T gpost = (KIND == R2R_DHT || FWD) ? T(sc.g) : T(sc.g * 0.5);
    { JT_PROF(st, "r2r_fb_post");
    hipLaunchKernelGGL((r2r_fb_post<T, KIND, FWD>), ew_grid(lm.nlines * n), dim3(256), 0, st, y, data, lm, n, L,
                       r.q.template as<cpx<T>>(), T(sc.f0), gpost);
    }
    JT_HIP(hipGetLastError());
}

template <typename T>
void r2r_axis(int family, const R2R1D<T>& r, T* data, LineMap lm, bool fwd, bool scale, cpx<T>* ws,
              hipStream_t st) {
    if (r.n == 1 || lm.nlines <= 0) return;
    if (family == 1) {
        if (fwd) r2r_axis_k<T, R2R_DCT, true>(r, data, lm, scale, ws, st);
        else r2r_axis_k<T, R2R_DCT, false>(r, data, lm, scale, ws, st);
    } else if (family == 2) {
        if (fwd) r2r_axis_k<T, R2R_DST, true>(r, data, lm, scale, ws, st);
        else r2r_axis_k<T, R2R_DST, false>(r, data, lm, scale, ws, st);
    } else {
        r2r_axis_k<T, R2R_DHT, true>(r, data, lm, fwd ? false : scale, ws, st);
    }
}

template <typename T>
void run_r2r_nd(int family, int nd, const int64_t* d, const std::shared_ptr<R2R1D<T>>* r, T* a, bool fwd,
                bool scale, cpx<T>* ws, hipStream_t st) {
    if (nd == 1) {
        r2r_axis<T>(family, *r[0], a, LineMap{1, 1, d[0], 0, 1}, fwd, scale, ws, st);
        return;
    }
    if (nd == 2) {
        const int64_t R = d[0], C = d[1];
        r2r_axis<T>(family, *r[1], a, LineMap{R, R, C, 0, 1}, fwd, scale, ws, st);
        r2r_axis<T>(family, *r[0], a, LineMap{C, C, 1, 0, C}, fwd, scale, ws, st);
        if (family == 3) {
            { JT_PROF(st, "ytransform2d");
            hipLaunchKernelGGL((ytransform2d<T>), ew_grid((R / 2 + 1) * (C / 2 + 1)), dim3(256), 0, st, a, R, C);
            }
            JT_HIP(hipGetLastError());
        }
        return;
    }
    const int64_t S = d[0], R = d[1], C = d[2];
    r2r_axis<T>(family, *r[2], a, LineMap{S * R, S * R, C, 0, 1}, fwd, scale, ws, st);
    r2r_axis<T>(family, *r[1], a, LineMap{S * C, C, 1, R * C, C}, fwd, scale, ws, st);
    r2r_axis<T>(family, *r[0], a, LineMap{R * C, R * C, 1, 0, R * C}, fwd, scale, ws, st);
    if (family == 3) {
        { JT_PROF(st, "ytransform3d");
        hipLaunchKernelGGL((ytransform3d<T>), ew_grid((S / 2 + 1) * (R / 2 + 1) * (C / 2 + 1)), dim3(256), 0, st,
                           a, S, R, C);
        }
        JT_HIP(hipGetLastError());
    }
}

}  // namespace jt
