// Host side of the 1D complex engine: plans (radix factorisation, FP64
// twiddle / chirp tables resident in HBM, four-step and Bluestein sub-plans)
// and the launchers that run a plan over a set of lines.
//
// Plan kinds, mirroring the reference's three plans (DoubleFFT_1D.java:56-60,
// 117-188) but chosen for the GPU:
//   TPL   n a power of two in [16, 4096]: compile-time stockham_kernel
//   GEN   n <= LDS limit with factors in {2,3,5,7,11,13}: generic_kernel
//   FOUR  smooth n above the LDS limit: n = n1*n2, two LDS passes
//   BLUE  any other n: Bluestein chirp-z over a smooth length m >= 2n-1
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "jt_generic.hpp"
#include "jt_prof.hpp"
#include "jt_io.hpp"
#include "jt_stockham.hpp"
#include "jt_mfstockham.hpp"

namespace jt {

// ----------------------------------------------------------------------------
// errors
// ----------------------------------------------------------------------------
struct JtError : std::runtime_error {
    int code;
    JtError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
enum { ST_OK = 0, ST_BAD_SIZE = 1, ST_NOT_POW2 = 2, ST_TOO_BIG = 3, ST_HIP = 4, ST_OOM = 5,
       ST_BAD_ARG = 6, ST_UNSUPPORTED = 7 };

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory) throw JtError(ST_OOM, std::string(what) + ": out of device memory");
        throw JtError(ST_HIP, std::string(what) + ": " + hipGetErrorString(e));
    }
}
#define JT_HIP(x) ::jt::hip_check((x), #x)

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t b) {
        if (p) {
            (void)hipFree(p);
            p = nullptr;
        }
        bytes = b;
        if (b) JT_HIP(hipMalloc(&p, b));
    }
    template <typename U>
    U* as() const { return static_cast<U*>(p); }
};

template <typename T>
inline void upload(DevBuf& d, const std::vector<cpx<T>>& h) {
    d.alloc(h.size() * sizeof(cpx<T>));
    if (!h.empty()) JT_HIP(hipMemcpy(d.p, h.data(), d.bytes, hipMemcpyHostToDevice));
}

// exp(-2 pi i num/den) in long double, num reduced exactly
template <typename T>
inline cpx<T> expm(int64_t num, int64_t den) {
    num %= den;
    if (num < 0) num += den;
    const long double PI2 = 6.283185307179586476925286766559005768L;
    long double a = PI2 * (long double)num / (long double)den;
    return cpx<T>{(T)cosl(a), (T)-sinl(a)};
}

inline bool is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

// ----------------------------------------------------------------------------
// factorisation
// ----------------------------------------------------------------------------
inline bool factor_radices(int64_t n, std::vector<int>& out) {
    std::vector<int> r;
    int64_t m = n;
    int e2 = 0;
    while (m % 2 == 0) {
        m /= 2;
        e2++;
    }
    while (e2 >= 4) {
        r.push_back(16);
        e2 -= 4;
    }
    if (e2 == 3) r.push_back(8);
    else if (e2 == 2) r.push_back(4);
    else if (e2 == 1) r.push_back(2);
    for (int p : {3, 5, 7, 11, 13})
        while (m % p == 0) {
            r.push_back(p);
            m /= p;
        }
    if (m != 1) return false;
    out = r;
    return true;
}

// radix lists of the compile-time kernels (must match dispatch_tpl)
inline bool tpl_radices(int64_t n, std::vector<int>& r) {
    switch (n) {
        case 4096: r = {16, 16, 16}; return true;
        case 2048: r = {16, 16, 8}; return true;
        case 1024: r = {16, 16, 4}; return true;
        case 512: r = {8, 8, 8}; return true;
        case 256: r = {16, 16}; return true;
        case 128: r = {16, 8}; return true;
        case 64: r = {8, 8}; return true;
        case 32: r = {8, 4}; return true;
        case 16: r = {16}; return true;
        default: return false;
    }
}

// radix lists of the split-FP16 MFMA kernels: radix-16 stages first, one
// closing radix 2/4/8 stage when N is not a power of 16
inline bool mf_radices(int64_t n, std::vector<int>& r) {
    switch (n) {
        case 4096: r = {16, 16, 16}; return true;
        case 2048: r = {16, 16, 8}; return true;
        case 1024: r = {16, 16, 4}; return true;
        case 512: r = {16, 16, 2}; return true;
        case 256: r = {16, 16}; return true;
        case 128: r = {16, 8}; return true;
        case 64: r = {16, 4}; return true;
        case 32: r = {16, 2}; return true;
        case 16: r = {16}; return true;
        default: return false;
    }
}

template <typename T>
inline int64_t gen_max() { return sizeof(T) == 8 ? 4096 : 8192; }

template <typename T>
std::vector<cpx<T>> stage_twiddles(const std::vector<int>& radices, std::vector<int>* offs) {
    std::vector<cpx<T>> tw;
    int64_t ns = 1;
    for (int R : radices) {
        if (offs) offs->push_back((int)tw.size());
        for (int64_t jm = 0; jm < ns; jm++)
            for (int r = 1; r < R; r++) tw.push_back(expm<T>(jm * r, ns * R));
        ns *= R;
    }
    if (tw.empty()) tw.push_back(cpx<T>{T(1), T(0)});
    return tw;
}

// Twiddles of the split-FP16 MFMA kernels (layout of mf_twoff): 16 factors
// per jm (r = 0..15) for every radix-16 stage, R-1 for a closing radix R.
// The radix-16 stages after the first carry the kernel's fixed 2^-4 range
// step (exact): their inputs stay below 2^14 (jt_mfma.hpp).
template <typename T>
std::vector<cpx<T>> stage_twiddles_mf(const std::vector<int>& radices) {
    std::vector<cpx<T>> tw;
    int64_t ns = 1;
    for (int R : radices) {
        const int r0 = R == 16 ? 0 : 1;
        const T f = (R == 16 && ns > 1) ? T(0.0625) : T(1);
        for (int64_t jm = 0; jm < ns; jm++)
            for (int r = r0; r < R; r++) {
                cpx<T> w = expm<T>(jm * r, ns * R);
                tw.push_back(cpx<T>{w.x * f, w.y * f});
            }
        ns *= R;
    }
    return tw;
}

// Host long-double DFT by recursive mixed radix (plan time only: Bluestein
// kernel spectrum).  sign -1 forward.
inline void host_fft_ld(std::vector<long double>& re, std::vector<long double>& im, int sign) {
    const size_t n = re.size();
    if (n <= 1) return;
    size_t p = 0;
    for (size_t q : {2, 3, 5, 7, 11, 13})
        if (n % q == 0) {
            p = q;
            break;
        }
    if (!p) p = n;  // direct DFT
    const long double PI2 = 6.283185307179586476925286766559005768L;
    if (p == n) {
        std::vector<long double> or_(n), oi(n);
        for (size_t k = 0; k < n; k++) {
            long double sr = 0, si = 0;
            for (size_t j = 0; j < n; j++) {
                long double a = sign * PI2 * (long double)((j * k) % n) / (long double)n;
                sr += re[j] * cosl(a) - im[j] * sinl(a);
                si += re[j] * sinl(a) + im[j] * cosl(a);
            }
            or_[k] = sr;
            oi[k] = si;
        }
        re = or_;
        im = oi;
        return;
    }
    const size_t m = n / p;
    std::vector<std::vector<long double>> sr(p, std::vector<long double>(m)), si(p, std::vector<long double>(m));
    for (size_t q = 0; q < p; q++)
        for (size_t j = 0; j < m; j++) {
            sr[q][j] = re[j * p + q];
            si[q][j] = im[j * p + q];
        }
    for (size_t q = 0; q < p; q++) host_fft_ld(sr[q], si[q], sign);
    for (size_t k = 0; k < n; k++) {
        long double accr = 0, acci = 0;
        for (size_t q = 0; q < p; q++) {
            long double a = sign * PI2 * (long double)((q * k) % n) / (long double)n;
            long double c = cosl(a), s = sinl(a);
            long double xr = sr[q][k % m], xi = si[q][k % m];
            accr += xr * c - xi * s;
            acci += xr * s + xi * c;
        }
        re[k] = accr;
        im[k] = acci;
    }
}

inline int64_t next_smooth(int64_t x) {
    // smallest 2^a 3^b 5^c >= x
    int64_t best = INT64_MAX;
    for (int64_t p2 = 1; p2 < 4 * x; p2 *= 2)
        for (int64_t p3 = p2; p3 < 4 * x; p3 *= 3)
            for (int64_t p5 = p3; p5 < 4 * x; p5 *= 5)
                if (p5 >= x && p5 < best) best = p5;
    return best;
}

// ----------------------------------------------------------------------------
// the plan
// ----------------------------------------------------------------------------
enum PlanKind { K_TPL = 0, K_GEN = 1, K_FOUR = 2, K_BLUE = 3 };

template <typename T>
struct Fft1D {
    int64_t n = 0;
    int kind = K_GEN;
    bool mf = false;  // split-FP16 MFMA radix-16 stages (FP32 compile-time plans)
    DevBuf tw_mf;     // stage twiddles of mf_radices(n), built when mf is first set
    std::vector<int> radices;
    GenPlan gp{};
    DevBuf tw;
    // four-step
    int64_t n1 = 0, n2 = 0;
    std::unique_ptr<Fft1D<T>> p1, p2;
    DevBuf hi, lo;
    // Bluestein
    int64_t m = 0;
    std::unique_ptr<Fft1D<T>> pm;
    DevBuf chirp, bspec;

    bool single() const { return kind == K_TPL || kind == K_GEN; }
    // workspace (complex elements) for `lines` contiguous lines
    int64_t workspace(int64_t lines) const {
        if (kind == K_FOUR) return lines * n + (p2->single() ? 0 : p2->workspace(lines * n1));
        if (kind == K_BLUE) return lines * m + pm->workspace(lines);
        return 0;
    }
};

template <typename T>
std::unique_ptr<Fft1D<T>> make_fft1d(int64_t n) {
    auto p = std::make_unique<Fft1D<T>>();
    p->n = n;
    std::vector<int> r;
    bool tpl = tpl_radices(n, r);
    if (tpl || (n <= gen_max<T>() && factor_radices(n, r))) {
        p->kind = tpl ? K_TPL : K_GEN;
        p->radices = r;
        std::vector<int> offs;
        std::vector<cpx<T>> tab = stage_twiddles<T>(r, &offs);
        if (n == 4096) {
            // the 4096-point complex row kernels run radices 16,8,8,4; their
            // table follows the 16,16,16 one (see dispatch_tpl_n)
            std::vector<cpx<T>> t2 = stage_twiddles<T>({16, 8, 8, 4}, nullptr);
            tab.insert(tab.end(), t2.begin(), t2.end());
        } else if (sizeof(T) == 4 && n == 512) {
            // FP32 512-point complex column kernels: radices 8,4,16
            std::vector<cpx<T>> t2 = stage_twiddles<T>({8, 4, 16}, nullptr);
            tab.insert(tab.end(), t2.begin(), t2.end());
        }
        upload(p->tw, tab);
        p->gp.n = (int)n;
        p->gp.nstages = (int)r.size();
        for (size_t i = 0; i < r.size(); i++) {
            p->gp.radix[i] = r[i];
            p->gp.twoff[i] = offs[i];
        }
        return p;
    }
    if (factor_radices(n, r)) {
        // four-step: n1 (column FFT) * n2 (row FFT), both single-kernel
        int64_t best1 = 0;
        for (int64_t a = 2; a <= gen_max<T>(); a++) {
            if (n % a) continue;
            int64_t b = n / a;
            if (b > gen_max<T>()) continue;
            std::vector<int> ra, rb;
            if (!factor_radices(a, ra) || !factor_radices(b, rb)) continue;
            // prefer balanced split, n1 a power of two when possible
            if (!best1 || std::llabs(a - b) < std::llabs(best1 - n / best1)) best1 = a;
        }
        if (!best1) {
            // n beyond gen_max^2: the largest single-kernel n1 whose cofactor
            // is smooth; the n2-point row FFTs are then a plan of their own
            // (four-step again), run in place on the workspace
            for (int64_t a = gen_max<T>(); a >= 2 && !best1; a--) {
                if (n % a) continue;
                std::vector<int> ra, rb;
                if (factor_radices(a, ra) && factor_radices(n / a, rb)) best1 = a;
            }
        }
        if (best1) {
            p->kind = K_FOUR;
            p->n1 = best1;
            p->n2 = n / best1;
            p->p1 = make_fft1d<T>(p->n1);
            p->p2 = make_fft1d<T>(p->n2);
            std::vector<cpx<T>> hi((n >> 8) + 1), lo(256);
            for (int64_t i = 0; i < (int64_t)hi.size(); i++) hi[i] = expm<T>(i << 8, n);
            for (int64_t i = 0; i < 256; i++) lo[i] = expm<T>(i, n);
            upload(p->hi, hi);
            upload(p->lo, lo);
            return p;
        }
    }
    // Bluestein (DoubleFFT_1D.java:1864-2107): chirp c_j = exp(i pi (j^2 mod 2n)/n)
    p->kind = K_BLUE;
    p->m = next_smooth(2 * n - 1);
    p->pm = make_fft1d<T>(p->m);
    std::vector<cpx<T>> ch(n);
    std::vector<long double> br(p->m, 0.0L), bi(p->m, 0.0L);
    const long double PI = 3.141592653589793238462643383279502884L;
    int64_t k = 0;
    for (int64_t j = 0; j < n; j++) {
        if (j) {
            k += 2 * j - 1;
            if (k >= 2 * n) k -= 2 * n;
        }
        long double a = PI * (long double)k / (long double)n;
        long double c = cosl(a), s = sinl(a);
        ch[j] = cpx<T>{(T)c, (T)s};
        br[j] = c;
        bi[j] = s;
        if (j) {
            br[p->m - j] = c;
            bi[p->m - j] = s;
        }
    }
    host_fft_ld(br, bi, -1);
    std::vector<cpx<T>> bs(p->m);
    for (int64_t i = 0; i < p->m; i++) bs[i] = cpx<T>{(T)(br[i] / p->m), (T)(bi[i] / p->m)};
    upload(p->chirp, ch);
    upload(p->bspec, bs);
    return p;
}

// ----------------------------------------------------------------------------
// launch: compile-time kernels
// ----------------------------------------------------------------------------
// Column mode reads W adjacent lines per workgroup, i.e. row segments of
// W * sizeof(cpx<T>) bytes: 128 B when the kernel has no LDS hooks (the
// exchange then runs in split re/im rounds), 64 B otherwise.
template <typename T, bool HOOKS>
constexpr int col_want() {
    return (HOOKS ? 64 : 128) / (int)sizeof(cpx<T>);
}

// Threads per line: TL0 for rows; for columns fewer, so that the wanted W
// lines fit 512 threads, down to 32 points per thread (FP32, and FP64 at
// N = 4096) or 16 (other FP64 lengths, whose 32-point form spills).
template <typename T, int N, int TL0, bool COL, bool HOOKS>
constexpr int line_threads() {
    if (!COL || HOOKS) return TL0;
    const int emax = (sizeof(T) == 4 || N == 4096) ? 32 : 16;
    const int lim = 512 / col_want<T, HOOKS>();
    const int tl = TL0 < lim ? TL0 : lim;
    const int tmin = N / emax > 0 ? N / emax : 1;
    return tl > tmin ? tl : (tmin < TL0 ? tmin : TL0);
}

// IO policies over lines of reals (IOR2R) declare real_lines = true
template <class IO, class = void>
struct io_real_lines : std::false_type {};
template <class IO>
struct io_real_lines<IO, std::void_t<decltype(IO::real_lines)>> : std::integral_constant<bool, IO::real_lines> {};

template <typename T, int N, int TL, bool COL, bool HOOKS = false, bool REAL = false>
struct TplShape {
    static constexpr int lines_row = (256 / TL) > 0 ? (256 / TL) : 1;
    // lines of reals: 256-B row segments while the two LDS arrays fit
    static constexpr int want_real = (256 / (int)sizeof(T)) * 2 * (N + N / 32 + 32) * (int)sizeof(T) <= 150 * 1024
                                         ? 256 / (int)sizeof(T)
                                         : 128 / (int)sizeof(T);
    static constexpr int want_col = REAL ? want_real : (col_want<T, HOOKS>() > 0 ? col_want<T, HOOKS>() : 1);
    static constexpr int col_threads_cap = 512 / TL > 0 ? 512 / TL : 1;
    static constexpr int w_col0 = want_col > col_threads_cap ? col_threads_cap : want_col;
    static constexpr int w_col = (w_col0 * TL < 256) ? (256 / TL) : w_col0;
    static constexpr int W = COL ? w_col : lines_row;
    static constexpr int bytes_full = 2 * LdsGeom<N, W>::size * (int)sizeof(T);
    static constexpr bool split = bytes_full > 65536;
};

template <typename T, int DIR, int N, int TL, int W, bool COL, bool SPLIT, class IO, int... Rs>
void launch_tpl(IO io, const cpx<T>* tw, hipStream_t st) {
    int64_t groups = (io.nlines + W - 1) / W;
    if (groups <= 0) return;
    JT_PROF(st, std::string("stockham<") + (sizeof(T) == 8 ? "f64" : "f32") + ",N=" + std::to_string(N) +
 (COL ? ",col" : ",row") + ",W=" + std::to_string(W) + ",lines=" + std::to_string(io.nlines) +
                    "," + IO::tag() + io.desc() + ">");
    hipLaunchKernelGGL((stockham_kernel<T, DIR, N, TL, W, COL, SPLIT, IO, Rs...>), dim3((unsigned)groups),
                       dim3(TL * W), 0, st, io, tw);
    JT_HIP(hipGetLastError());
}

template <typename T, int DIR, int N, int TL0, bool COL, class IO, int... Rs>
void go_tpl(IO io, const cpx<T>* tw, hipStream_t st) {
    // LDS hooks, or a store that computes twiddles (four-step pass A), keep
    // the narrower shape: their extra registers would spill
    constexpr bool HOOKS = IO::pre_lds || IO::post_lds || std::is_same<IO, IOFourA<T, DIR>>::value;
    constexpr int TL = line_threads<T, N, TL0, COL, HOOKS>();
    constexpr bool REAL = io_real_lines<IO>::value;
    using S = TplShape<T, N, TL, COL, HOOKS, false>;
    constexpr bool SPLIT = S::split && !IO::pre_lds && !IO::post_lds;
    if constexpr (COL && REAL) {
        // lines of reals: the wide group (256-B row segments) pays off only
        // for strides of 1 MiB and more (measured on the 512^3 DCT: 2 MiB
        // stride 647 -> 568 us, 4 KiB stride 478 -> 546 us)
        using SW = TplShape<T, N, TL, COL, HOOKS, true>;
        if (SW::W != S::W && io.m.es * (int64_t)sizeof(T) >= (int64_t(1) << 20)) {
            launch_tpl<T, DIR, N, TL, SW::W, COL, SPLIT, IO, Rs...>(io, tw, st);
            return;
        }
    }
    if constexpr (!COL && REAL && TL <= 16) {
        // short rows of reals (DCT/DST/DHT x axis): one wave per workgroup
        // (512^3 FP64 DCT row pass, 16 -> 8 -> 4 -> 2 lines per workgroup:
        // 469, 448, 439, 617 us)
        launch_tpl<T, DIR, N, TL, 64 / TL, COL, SPLIT, IO, Rs...>(io, tw, st);
        return;
    }
    launch_tpl<T, DIR, N, TL, S::W, COL, SPLIT, IO, Rs...>(io, tw, st);
}

template <class IO>
struct is_strided : std::false_type {};
template <typename T, int NT>
struct is_strided<IOStrided<T, NT>> : std::true_type {};

template <typename T, int DIR, bool COL, class IO>
bool dispatch_tpl_n(int64_t n, IO io, const cpx<T>* tw, hipStream_t st) {
    switch (n) {
        case 4096:
            if constexpr (!COL && is_strided<IO>::value) {
                // FP32 complex rows: 16,8,8,4 measured faster than 16,16,16
                // (16384 x 4096 batch, streaming IO: 181 vs 192-195 us)
                go_tpl<T, DIR, 4096, 256, COL, IO, 16, 8, 8, 4>(io, tw + RadixList<16, 16, 16>::twoff(3), st);
            } else {
                go_tpl<T, DIR, 4096, 256, COL, IO, 16, 16, 16>(io, tw, st);
            }
            return true;
        case 2048: go_tpl<T, DIR, 2048, 128, COL, IO, 16, 16, 8>(io, tw, st); return true;
        case 1024: go_tpl<T, DIR, 1024, 64, COL, IO, 16, 16, 4>(io, tw, st); return true;
        case 512:
            if constexpr (sizeof(T) == 4 && COL && is_strided<IO>::value) {
                // FP32 complex columns: 8,4,16 measured faster than 8,8,8
                // (FloatFFT_3D 512x512x256 row-axis pass: 202 vs 209 us)
                go_tpl<T, DIR, 512, 64, COL, IO, 8, 4, 16>(io, tw + RadixList<8, 8, 8>::twoff(3), st);
            } else {
                go_tpl<T, DIR, 512, 64, COL, IO, 8, 8, 8>(io, tw, st);
            }
            return true;
        case 256: go_tpl<T, DIR, 256, 16, COL, IO, 16, 16>(io, tw, st); return true;
        case 128: go_tpl<T, DIR, 128, 8, COL, IO, 16, 8>(io, tw, st); return true;
        case 64: go_tpl<T, DIR, 64, 8, COL, IO, 8, 8>(io, tw, st); return true;
        case 32: go_tpl<T, DIR, 32, 4, COL, IO, 8, 4>(io, tw, st); return true;
        case 16: go_tpl<T, DIR, 16, 1, COL, IO, 16>(io, tw, st); return true;
        default: return false;
    }
}

// split-FP16 MFMA kernels: W lines per workgroup, N/16 threads per line
template <int N, bool COL>
struct MfShape {
    static constexpr int TL = N / 16;
    // (columns: 64-B row segments; wider groups measured slower here)
    static constexpr int W = COL ? TplShape<float, N, TL, true, true>::W : (256 / TL > 0 ? 256 / TL : 1);
};

template <int DIR, int N, bool COL, class IO, int... Rs>
void go_mf(IO io, const cpx<float>* tw, hipStream_t st) {
    constexpr int W = MfShape<N, COL>::W;
    constexpr int TL = MfShape<N, COL>::TL;
    int64_t groups = (io.nlines + W - 1) / W;
    if (groups <= 0) return;
    JT_PROF(st, std::string("stockham<f32,N=") + std::to_string(N) + (COL ? ",col" : ",row") + ",mfma,W=" +
                    std::to_string(W) + ",lines=" + std::to_string(io.nlines) + "," + IO::tag() + io.desc() + ">");
    hipLaunchKernelGGL((mfma_stockham_kernel<DIR, N, W, COL, IO, Rs...>), dim3((unsigned)groups), dim3(TL * W), 0,
                       st, io, tw);
    JT_HIP(hipGetLastError());
}

template <int DIR, bool COL, class IO>
bool dispatch_mf(int64_t n, IO io, const cpx<float>* tw, hipStream_t st) {
    switch (n) {
        case 4096: go_mf<DIR, 4096, COL, IO, 16, 16, 16>(io, tw, st); return true;
        case 2048: go_mf<DIR, 2048, COL, IO, 16, 16, 8>(io, tw, st); return true;
        case 1024: go_mf<DIR, 1024, COL, IO, 16, 16, 4>(io, tw, st); return true;
        case 512: go_mf<DIR, 512, COL, IO, 16, 16, 2>(io, tw, st); return true;
        case 256: go_mf<DIR, 256, COL, IO, 16, 16>(io, tw, st); return true;
        case 128: go_mf<DIR, 128, COL, IO, 16, 8>(io, tw, st); return true;
        case 64: go_mf<DIR, 64, COL, IO, 16, 4>(io, tw, st); return true;
        case 32: go_mf<DIR, 32, COL, IO, 16, 2>(io, tw, st); return true;
        case 16: go_mf<DIR, 16, COL, IO, 16>(io, tw, st); return true;
        default: return false;
    }
}

template <typename T, int DIR, bool COL, class IO>
bool dispatch_tpl(int64_t n, IO io, const cpx<T>* tw, hipStream_t st, bool mf = false,
                  const cpx<T>* tw_mf = nullptr) {
    if constexpr (sizeof(T) == 4 && is_strided<IO>::value) {
        if (mf) return dispatch_mf<DIR, COL>(n, io, tw_mf, st);
    }
    return dispatch_tpl_n<T, DIR, COL>(n, io, tw, st);
}

template <typename T, int DIR, class IO>
void go_gen(const Fft1D<T>& p, IO io, hipStream_t st) {
    if (io.nlines <= 0) return;
    int n = (int)p.n;
    int np = n + (n >> 5) + 1;
    size_t shm = 4 * (size_t)np * sizeof(T);
    int nt = n / 4;
    nt = nt < 64 ? 64 : (nt > 256 ? 256 : ((nt + 63) / 64) * 64);
    // the LDS limit is raised once per instantiation and device
    static std::atomic<uint64_t> attr_devs{0};
    int dev = 0;
    JT_HIP(hipGetDevice(&dev));
    const uint64_t bit = dev < 64 ? (uint64_t(1) << dev) : 0;
    if (!bit || !(attr_devs.load(std::memory_order_relaxed) & bit)) {
        JT_HIP(hipFuncSetAttribute((const void*)generic_kernel<T, DIR, IO>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr_devs.fetch_or(bit);
    }
    int64_t lines = io.nlines;
    JT_PROF(st, std::string("generic<") + (sizeof(T) == 8 ? "f64" : "f32") + ",N=" + std::to_string(n) +
                    ",lines=" + std::to_string(lines) + "," + IO::tag() + io.desc() + ">");
    unsigned gx = (unsigned)(lines < 65535 ? lines : 65535);
    unsigned gy = (unsigned)((lines + gx - 1) / gx);
    hipLaunchKernelGGL((generic_kernel<T, DIR, IO>), dim3(gx, gy), dim3(nt), shm, st, io,
                       p.tw.template as<cpx<T>>(), p.gp);
    JT_HIP(hipGetLastError());
}

// Run a single-kernel plan over the lines described by io.
template <typename T, int DIR, bool COL, class IO>
void run_single(const Fft1D<T>& p, IO io, hipStream_t st) {
    if (p.kind == K_TPL) {
        if (dispatch_tpl<T, DIR, COL>(p.n, io, p.tw.template as<cpx<T>>(), st, p.mf,
                                      p.tw_mf.template as<cpx<T>>()))
            return;
    }
    if (p.kind == K_GEN || p.kind == K_TPL) {
        // generic kernel needs a radix plan
        go_gen<T, DIR>(p, io, st);
        return;
    }
    throw JtError(ST_UNSUPPORTED, "run_single on a multi-pass plan");
}

// ----------------------------------------------------------------------------
// elementwise kernels (Bluestein, real-transform helpers)
// ----------------------------------------------------------------------------
template <typename T, int DIR>
__global__ void blue_premul(const cpx<T>* __restrict__ x, cpx<T>* __restrict__ y,
                            const cpx<T>* __restrict__ ch, int64_t n, int64_t m, int64_t lines) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t tot = lines * m;
    for (; i < tot; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t b = i / m, j = i % m;
        cpx<T> v{T(0), T(0)};
        if (j < n) {
            cpx<T> a = x[b * n + j];
            v = DIR < 0 ? cmulc(a, ch[j]) : cmul(a, ch[j]);
        }
        y[i] = v;
    }
}

template <typename T, int DIR>
__global__ void blue_pointwise(cpx<T>* __restrict__ y, const cpx<T>* __restrict__ bs, int64_t m,
                               int64_t lines) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t tot = lines * m;
    for (; i < tot; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t j = i % m;
        y[i] = DIR < 0 ? cmul(y[i], bs[j]) : cmulc(y[i], bs[j]);
    }
}

template <typename T, int DIR>
__global__ void blue_postmul(const cpx<T>* __restrict__ y, cpx<T>* __restrict__ x,
                             const cpx<T>* __restrict__ ch, int64_t n, int64_t m, int64_t lines,
                             T scale) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t tot = lines * n;
    for (; i < tot; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t b = i / n, j = i % n;
        cpx<T> v = y[b * m + j];
        v = DIR < 0 ? cmulc(v, ch[j]) : cmul(v, ch[j]);
        x[i] = cpx<T>{v.x * scale, v.y * scale};
    }
}

// out[b*n + k1 + n1*k2] = ws[b*n + k1*n2 + k2] * s (the last step of a
// four-step transform whose n2-point rows ran as a plan of their own)
template <typename T>
__global__ void four_transpose(const cpx<T>* __restrict__ ws, cpx<T>* __restrict__ out, int64_t n1, int64_t n2,
                               int64_t lines, T s) {
    const int64_t n = n1 * n2;
    const int64_t tot = lines * n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = i / n, r = i - b * n;
        const int64_t k1 = r / n2, k2 = r - k1 * n2;
        const cpx<T> v = ws[i];
        out[b * n + k1 + n1 * k2] = cpx<T>{v.x * s, v.y * s};
    }
}

template <typename T>
__global__ void scale_kernel(cpx<T>* __restrict__ x, int64_t count, T s) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i < count; i += (int64_t)gridDim.x * blockDim.x) x[i] = cpx<T>{x[i].x * s, x[i].y * s};
}

inline dim3 ew_grid(int64_t count) {
    int64_t b = (count + 255) / 256;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

// ----------------------------------------------------------------------------
// run a plan over `lines` lines; complex data of line b at b*n ("contiguous")
// or through a LineMap for single-kernel plans.
// ----------------------------------------------------------------------------
template <typename T, int DIR>
void c2c_contig(const Fft1D<T>& p, const cpx<T>* in, cpx<T>* out, int64_t lines, T scale,
                cpx<T>* ws, hipStream_t st);

// arrays at least this large are streamed (non-temporal) by first/last passes
constexpr int64_t NT_MIN_BYTES = int64_t(64) << 20;

template <typename T, int DIR>
void c2c_mapped(const Fft1D<T>& p, const cpx<T>* in, cpx<T>* out, LineMap lm, T scale,
                hipStream_t st, int nt = 0) {
    // streaming (nt = NT_LOAD | NT_STORE) only for a single pass over
    // contiguous lines, i.e. a batch of 1D transforms whose result is not
    // read by a following pass: 16384 x 4096 FP32 218 -> 208 us.  In column
    // mode a workgroup stores partial row segments, and non-temporal stores
    // of those measured twice as slow.
    if (lm.es != 1) {
        IOStrided<T> io{in, out, lm, lm.nlines, scale, true};
        run_single<T, DIR, true>(p, io, st);
    } else if (nt == (NT_LOAD | NT_STORE)) {
        IOStrided<T, NT_LOAD | NT_STORE> io{in, out, lm, lm.nlines, scale, false};
        run_single<T, DIR, false>(p, io, st);
    } else {
        IOStrided<T> io{in, out, lm, lm.nlines, scale, false};
        run_single<T, DIR, false>(p, io, st);
    }
}

template <typename T, int DIR>
void c2c_contig(const Fft1D<T>& p, const cpx<T>* in, cpx<T>* out, int64_t lines, T scale,
                cpx<T>* ws, hipStream_t st) {
    if (lines <= 0) return;
    if (p.single()) {
        LineMap lm{lines, lines, p.n, 0, 1};
        // a single pass over a large batch reads and writes each element once
        const bool big = lines * p.n * (int64_t)sizeof(cpx<T>) >= NT_MIN_BYTES;
        c2c_mapped<T, DIR>(p, in, out, lm, scale, st, big ? (NT_LOAD | NT_STORE) : 0);
        return;
    }
    if (p.kind == K_FOUR) {
        if (!p.p1->single()) throw JtError(ST_UNSUPPORTED, "four-step column plan must be single-kernel");
        IOFourA<T, DIR> a{in, ws, p.hi.template as<cpx<T>>(), p.lo.template as<cpx<T>>(), lines * p.n2,
                          p.n, p.n1, p.n2, true};
        run_single<T, DIR, true>(*p.p1, a, st);
        if (p.p2->single()) {
            IOFourB<T> b{ws, out, lines * p.n1, p.n, p.n1, p.n2, scale, false};
            run_single<T, DIR, false>(*p.p2, b, st);
            return;
        }
        // long rows: FFT them in place in the workspace, then transpose
        c2c_contig<T, DIR>(*p.p2, ws, ws, lines * p.n1, T(1), ws + lines * p.n, st);
        { JT_PROF(st, "four_transpose");
        hipLaunchKernelGGL((four_transpose<T>), ew_grid(lines * p.n), dim3(256), 0, st, ws, out, p.n1, p.n2,
                           lines, scale);
        }
        JT_HIP(hipGetLastError());
        return;
    }
    // Bluestein
    cpx<T>* y = ws;
    cpx<T>* ws2 = ws + lines * p.m;
    const cpx<T>* ch = p.chirp.template as<cpx<T>>();
    { JT_PROF(st, "blue_premul");
    hipLaunchKernelGGL((blue_premul<T, DIR>), ew_grid(lines * p.m), dim3(256), 0, st, in, y, ch, p.n, p.m,
                       lines);
    }
    JT_HIP(hipGetLastError());
    c2c_contig<T, -1>(*p.pm, y, y, lines, T(1), ws2, st);
    { JT_PROF(st, "blue_pointwise");
    hipLaunchKernelGGL((blue_pointwise<T, DIR>), ew_grid(lines * p.m), dim3(256), 0, st, y,
                       p.bspec.template as<cpx<T>>(), p.m, lines);
    }
    JT_HIP(hipGetLastError());
    c2c_contig<T, 1>(*p.pm, y, y, lines, T(1), ws2, st);
    { JT_PROF(st, "blue_postmul");
    hipLaunchKernelGGL((blue_postmul<T, DIR>), ew_grid(lines * p.n), dim3(256), 0, st, y, out, ch, p.n,
                       p.m, lines, scale);
    }
    JT_HIP(hipGetLastError());
}

}  // namespace jt
