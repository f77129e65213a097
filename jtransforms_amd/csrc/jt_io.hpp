// IO policies of the Stockham kernels: where a line's points come from, where
// the results go, and the O(N) pre/post passes fused around the FFT.
#pragma once
#include <string>

#include "jt_dft.hpp"

namespace jt {

__device__ __forceinline__ int lpos(int k) { return k + (k >> 5); }

// Lines l in [0, nlines) start at (l % n0)*s0 + (l / n0)*s1 (complex units);
// element k of a line sits es further per step.
struct LineMap {
    int64_t nlines, n0, s0, s1, es;
    __device__ __forceinline__ int64_t base(int64_t l) const { return (l % n0) * s0 + (l / n0) * s1; }
};

// Non-temporal (streaming) complex loads and stores: they skip the caches'
// allocation, which lets a pass that reads or writes each element once run
// closer to the HBM rate (a device copy: 6.4 TB/s with them, 5.7 without).
template <typename T>
__device__ __forceinline__ cpx<T> ld_nt(const cpx<T>* p) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return cpx<T>{v.x, v.y};
}
template <typename T>
__device__ __forceinline__ void st_nt(cpx<T>* p, cpx<T> c) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    v2 v = {c.x, c.y};
    __builtin_nontemporal_store(v, reinterpret_cast<v2*>(p));
}
enum { NT_LOAD = 1, NT_STORE = 2 };

// Plain complex lines, optionally scaled on store (complexInverse(scale)).
// NT: NT_LOAD / NT_STORE for a pass whose input is not read again / whose
// output is not read soon (contiguous lines of a large array).  A template
// argument, not a run-time flag: a branch per access costs the column
// kernels registers and load batching.
template <typename T, int NT = 0>
struct IOStrided {
    static constexpr bool pre_lds = false, post_lds = false;
    static constexpr int streaming = NT;
    static const char* tag() { return NT == 0 ? "c2c" : "c2c,nt"; }
    std::string desc() const { return ",es=" + std::to_string(m.es); }
    const cpx<T>* in;
    cpx<T>* out;
    LineMap m;
    int64_t nlines;
    T scale;
    bool xcd;
    __device__ __forceinline__ int64_t base(int64_t l) const { return m.base(l); }
    __device__ __forceinline__ cpx<T> load(int64_t b, int k) const {
        if constexpr ((NT & NT_LOAD) != 0) return ld_nt(in + b + k * m.es);
        else return in[b + k * m.es];
    }
    // one part (0 real, 1 imaginary) of element k
    __device__ __forceinline__ T loadf(int64_t b, int k, int part) const {
        return reinterpret_cast<const T*>(in + b + k * m.es)[part];
    }
    __device__ __forceinline__ void store(int64_t b, int k, cpx<T> v) const {
        const cpx<T> o{v.x * scale, v.y * scale};
        if constexpr ((NT & NT_STORE) != 0) st_nt(out + b + k * m.es, o);
        else out[b + k * m.es] = o;
    }
    template <int CN, int CTL>
    __device__ void pre(int64_t, int64_t, T*, T*, int, int) const {}
    template <int CN, int CTL>
    __device__ void post(int64_t, int64_t, const T*, const T*, int, int) const {}
};

// The column pass of a packed-real 2D transform with the rdft2d_sub fix of
// column 0 (reference: DoubleFFT_2D.java:2544-2574) done by the workgroup
// that transforms that column.  Column 0 holds the transform of
// Y[.][0] + i Y[.][cols/2]; for rows 0 < i < rows/2 and j = rows - i:
//   ISGN > 0 (realForward, after the FFT):  Z_j <- (Z_i - conj Z_j) / 2,
//                                           Z_i <- Z_i - Z_j(new);
//   ISGN < 0 (realInverse, before the FFT): Z_i <- Z_i + Z_j,
//                                           Z_j <- (Re(Z_i - Z_j), Im(Z_j - Z_i)).
template <typename T, int ISGN>
struct IOStridedFix2D : IOStrided<T> {
    static constexpr int line0_hook = ISGN;
    static const char* tag() { return ISGN > 0 ? "c2c,rdft2d" : "c2c,rdft2d_inv"; }
    int64_t rows;
    __device__ void line0(int tid, int nthreads) const {
        cpx<T>* col = this->out + this->m.base(0);
        const int64_t es = this->m.es;
        for (int64_t i = 1 + tid; i < rows / 2; i += nthreads) {
            const int64_t j = rows - i;
            const cpx<T> zi = col[i * es], zj = col[j * es];
            if constexpr (ISGN > 0) {
                const cpx<T> nj{T(0.5) * (zi.x - zj.x), T(0.5) * (zi.y + zj.y)};
                col[j * es] = nj;
                col[i * es] = cpx<T>{zi.x - nj.x, zi.y - nj.y};
            } else {
                col[i * es] = cpx<T>{zi.x + zj.x, zi.y + zj.y};
                col[j * es] = cpx<T>{zi.x - zj.x, zj.y - zi.y};
            }
        }
    }
};

// exp(DIR 2 pi i e / n) from two tables: hi[e >> 8] * lo[e & 255], both
// holding forward values exp(-2 pi i x/n).
template <int DIR, typename T>
__device__ __forceinline__ cpx<T> tw2(const cpx<T>* __restrict__ hi, const cpx<T>* __restrict__ lo,
                                     int64_t e) {
    cpx<T> w = cmul(hi[e >> 8], lo[e & 255]);
    return tw_dir<DIR>(w);
}

// Four-step pass A: transform length n = n1*n2 viewed as an n1 x n2 row-major
// matrix.  Line (b, j2) is column j2 of batch b (n1 points at stride n2); the
// result is multiplied by exp(DIR 2 pi i j2 k1 / n) and written to the same
// position of `out`.
template <typename T, int DIR>
struct IOFourA {
    static constexpr bool pre_lds = false, post_lds = false;
    static const char* tag() { return "fourA"; }
    std::string desc() const { return ",n1=" + std::to_string(n1) + ",n2=" + std::to_string(n2); }
    const cpx<T>* in;
    cpx<T>* out;
    const cpx<T>* hi;
    const cpx<T>* lo;
    int64_t nlines, n, n1, n2;
    bool xcd;
    __device__ __forceinline__ int64_t base(int64_t l) const { return (l / n2) * n + (l % n2); }
    __device__ __forceinline__ cpx<T> load(int64_t b, int k) const { return in[b + (int64_t)k * n2]; }
    __device__ __forceinline__ void store(int64_t b, int k, cpx<T> v) const {
        int64_t j2 = b % n;  // column index inside the batch entry
        int64_t e = (j2 * (int64_t)k) % n;
        out[b + (int64_t)k * n2] = cmul(v, tw2<DIR>(hi, lo, e));
    }
    template <int CN, int CTL>
    __device__ void pre(int64_t, int64_t, T*, T*, int, int) const {}
    template <int CN, int CTL>
    __device__ void post(int64_t, int64_t, const T*, const T*, int, int) const {}
};

// Four-step pass B: line (b, k1) is row k1 (n2 contiguous points); output
// point k2 is X[k1 + n1*k2], written with stride n1.
template <typename T>
struct IOFourB {
    static constexpr bool pre_lds = false, post_lds = false;
    static const char* tag() { return "fourB"; }
    std::string desc() const { return ",n1=" + std::to_string(n1) + ",n2=" + std::to_string(n2); }
    const cpx<T>* in;
    cpx<T>* out;
    int64_t nlines, n, n1, n2;
    T scale;
    bool xcd;
    // base encodes (b, k1) as b*n + k1*n2 (the input row start)
    __device__ __forceinline__ int64_t base(int64_t l) const { return (l / n1) * n + (l % n1) * n2; }
    __device__ __forceinline__ cpx<T> load(int64_t b, int k) const { return in[b + k]; }
    __device__ __forceinline__ void store(int64_t b, int k, cpx<T> v) const {
        int64_t bb = b / n, k1 = (b % n) / n2;
        out[bb * n + k1 + (int64_t)k * n1] = cpx<T>{v.x * scale, v.y * scale};
    }
    template <int CN, int CTL>
    __device__ void pre(int64_t, int64_t, T*, T*, int, int) const {}
    template <int CN, int CTL>
    __device__ void post(int64_t, int64_t, const T*, const T*, int, int) const {}
};

// realForward post pass (reference: CommonUtils.rftfsub :5750-5776 and the
// fold at DoubleFFT_1D.java:543-545).  The line holds the N-point complex FFT
// Z of the 2N reals taken pairwise; the packed result is
//   a[0] = X_0, a[1] = X_N, a[2k], a[2k+1] = Re, Im X_k  (0 < k < N),
// with X_k = (Z_k + conj Z_{N-k})/2 - i w^k (Z_k - conj Z_{N-k})/2,
// w = exp(-2 pi i/2N).  rtw[k] = w^k for k in [0, N/2].
template <typename T>
struct IOR2C {
    static constexpr bool pre_lds = false, post_lds = true;
    static const char* tag() { return "r2c"; }
    std::string desc() const { return ",es=" + std::to_string(m.es); }
    const cpx<T>* in;  // reals viewed as complex pairs
    T* out;            // packed reals
    const cpx<T>* rtw;
    LineMap m;  // in complex units; es must be 1
    int64_t nlines;
    int N;
    T scale;
    bool xcd;
    __device__ __forceinline__ int64_t base(int64_t l) const { return m.base(l); }
    __device__ __forceinline__ cpx<T> load(int64_t b, int k) const { return in[b + k]; }
    __device__ void store(int64_t, int, cpx<T>) const {}
    template <int CN, int CTL>
    __device__ void pre(int64_t, int64_t, T*, T*, int, int) const {}
    template <int CN, int CTL>
    __device__ void post(int64_t, int64_t b, const T* re, const T* im, int t, int nt) const {
        // a[2k], a[2k+1] = Re, Im X_k are one complex store at o[k]
        cpx<T>* o = reinterpret_cast<cpx<T>*>(out + 2 * b);
        const T h = T(0.5) * scale;
        if (t == 0) {
            T zr = re[lpos(0)], zi = im[lpos(0)];
            o[0] = cpx<T>{(zr + zi) * scale, (zr - zi) * scale};
        }
        const int n = CN > 0 ? CN : N;
        const int step = CTL > 0 ? CTL : nt;
        // k = 1 .. n/2, with compile-time trip counts when CN, CTL are known
        constexpr int ITER = (CN > 0 && CTL > 0) ? (CN / 2 + CTL - 1) / CTL : 1;
        for (int k0 = 1 + t; k0 <= n / 2; k0 += ITER * step) {
#pragma unroll
            for (int i = 0; i < ITER; i++) {
                const int k = k0 + i * step;
                if (k > n / 2) break;
                const int kk = n - k;
                cpx<T> A = {re[lpos(k)], im[lpos(k)]};
                cpx<T> B = {re[lpos(kk)], -im[lpos(kk)]};  // conj Z_{N-k}
                cpx<T> fe = cadd(A, B);
                cpx<T> d = csub(A, B);
                cpx<T> fo = {d.y, -d.x};  // -i * d
                cpx<T> wf = cmul(rtw[k], fo);
                // X_k = (fe + w^k fo)/2 ; X_{N-k} = conj(fe - w^k fo)/2
                o[k] = cpx<T>{(fe.x + wf.x) * h, (fe.y + wf.y) * h};
                if (kk != k) o[kk] = cpx<T>{(fe.x - wf.x) * h, -(fe.y - wf.y) * h};
            }
        }
    }
};

// realInverse pre pass (reference: CommonUtils.rftbsub :5806-5831 and
// DoubleFFT_1D.java:946-989).  From the packed half spectrum X of 2N reals it
// builds Z_k = E_k + i O_k with E_k = X_k + conj X_{N-k} and
// O_k = (X_k - conj X_{N-k}) exp(+2 pi i k/2N); the N-point backward FFT of Z
// is sum-normalised (x_{2m} + i x_{2m+1}) * 2N, stored times `scale`.
template <typename T>
struct IOC2R {
    static constexpr bool pre_lds = true, post_lds = false;
    static const char* tag() { return "c2r"; }
    std::string desc() const { return ",es=" + std::to_string(m.es); }
    const T* in;     // packed reals
    cpx<T>* out;     // reals viewed as complex pairs
    const cpx<T>* rtw;
    LineMap m;
    int64_t nlines;
    int N;
    T scale;
    bool xcd;
    __device__ __forceinline__ int64_t base(int64_t l) const { return m.base(l); }
    __device__ cpx<T> load(int64_t, int) const { return cpx<T>{T(0), T(0)}; }
    __device__ __forceinline__ void store(int64_t b, int k, cpx<T> v) const {
        out[b + k] = cpx<T>{v.x * scale, v.y * scale};
    }
    // Z_k and Z_{N-k} from X_k, X_{N-k} (k in (0, N/2]); Z_0 from X_0, X_N
    __device__ __forceinline__ void c2r_pair(int k, cpx<T> Xk, cpx<T> Xn, T* re, T* im) const {
        int kk = N - k;
        cpx<T> Xc = {Xn.x, -Xn.y};  // conj X_{N-k}
        cpx<T> E = cadd(Xk, Xc);
        cpx<T> D = csub(Xk, Xc);
        cpx<T> w = {rtw[k].x, -rtw[k].y};  // exp(+2 pi i k/2N)
        cpx<T> O = cmul(D, w);
        // Z_k = E + i O
        re[lpos(k)] = E.x - O.y;
        im[lpos(k)] = E.y + O.x;
        if (kk != k) {
            // E_{N-k} = conj E_k and O_{N-k} = conj O_k, so Z_{N-k} = conj E + i conj O
            re[lpos(kk)] = E.x + O.y;
            im[lpos(kk)] = O.x - E.y;
        }
    }
    template <int CN, int CTL>
    __device__ void pre(int64_t, int64_t b, T* re, T* im, int t, int nt) const {
        const T* a = in + 2 * b;
        if (t == 0) {
            T x0 = a[0], xn = a[1];
            // E_0 = X_0 + X_N, O_0 = X_0 - X_N
            re[lpos(0)] = x0 + xn;
            im[lpos(0)] = x0 - xn;
        }
        if constexpr (CN > 0) {
            // all loads first (compile-time trip count), then the LDS writes
            constexpr int IT = (CN / 2 + CTL - 1) / CTL;
            cpx<T> xk[IT], xn[IT];
#pragma unroll
            for (int q = 0; q < IT; q++) {
                int k = 1 + t + q * CTL;
                if (k <= CN / 2) {
                    xk[q] = cpx<T>{a[2 * k], a[2 * k + 1]};
                    xn[q] = cpx<T>{a[2 * (CN - k)], a[2 * (CN - k) + 1]};
                }
            }
#pragma unroll
            for (int q = 0; q < IT; q++) {
                int k = 1 + t + q * CTL;
                if (k <= CN / 2) c2r_pair(k, xk[q], xn[q], re, im);
            }
        } else {
            for (int k = 1 + t; k <= N / 2; k += nt)
                c2r_pair(k, cpx<T>{a[2 * k], a[2 * k + 1]}, cpx<T>{a[2 * (N - k)], a[2 * (N - k) + 1]}, re, im);
        }
    }
    template <int CN, int CTL>
    __device__ void post(int64_t, int64_t, const T*, const T*, int, int) const {}
};

}  // namespace jt
