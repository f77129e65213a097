// Small-radix DFT codelets for the Stockham engine (gfx950).
//
// A codelet transforms R complex values held in registers in place:
//   X_k = sum_j x_j exp(DIR * 2*pi*i * j*k / R),  DIR = -1 forward, +1 backward.
// These are the butterflies that replace the reference's radix-4 split-radix
// kernels (CommonUtils.java:2844-5749) and FFTPACK's passf2/3/4/5
// (DoubleFFT_1D.java:6794-7613).
#pragma once
#include <hip/hip_runtime.h>

namespace jt {

template <typename T>
struct cpx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return {a.x * s, a.y * s}; }
// multiply by DIR*i  (DIR=-1: -i ; DIR=+1: +i)
template <int DIR, typename T>
__device__ __forceinline__ cpx<T> mul_i(cpx<T> a) {
    if constexpr (DIR < 0) return {a.y, -a.x};
    else return {-a.y, a.x};
}

// Twiddle w = exp(DIR*i*theta) given the forward table entry (cos, -sin).
template <int DIR, typename T>
__device__ __forceinline__ cpx<T> tw_dir(cpx<T> fwd) {
    if constexpr (DIR < 0) return fwd;
    else return {fwd.x, -fwd.y};
}

template <int DIR, typename T>
__device__ __forceinline__ void dft2(cpx<T>& a, cpx<T>& b) {
    cpx<T> t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <int DIR, typename T>
__device__ __forceinline__ void dft4(cpx<T>& v0, cpx<T>& v1, cpx<T>& v2, cpx<T>& v3) {
    cpx<T> a = cadd(v0, v2), b = csub(v0, v2);
    cpx<T> c = cadd(v1, v3), d = mul_i<DIR>(csub(v1, v3));
    v0 = cadd(a, c);
    v2 = csub(a, c);
    v1 = cadd(b, d);
    v3 = csub(b, d);
}

template <int DIR, typename T>
__device__ __forceinline__ void dft3(cpx<T>& v0, cpx<T>& v1, cpx<T>& v2) {
    const T c1 = T(-0.5);
    const T s1 = T(DIR) * T(0.86602540378443864676372317075294);  // DIR*sin(2pi/3)
    cpx<T> t1 = cadd(v1, v2);
    cpx<T> t2 = csub(v1, v2);
    cpx<T> m0 = cadd(v0, t1);
    cpx<T> m1 = {v0.x + c1 * t1.x, v0.y + c1 * t1.y};
    cpx<T> m2 = {-s1 * t2.y, s1 * t2.x};  // i*s1*t2
    v0 = m0;
    v1 = cadd(m1, m2);
    v2 = csub(m1, m2);
}

template <int DIR, typename T>
__device__ __forceinline__ void dft5(cpx<T>& v0, cpx<T>& v1, cpx<T>& v2, cpx<T>& v3, cpx<T>& v4) {
    const T c1 = T(0.30901699437494742410229341718282);   // cos(2pi/5)
    const T c2 = T(-0.80901699437494742410229341718282);  // cos(4pi/5)
    const T s1 = T(DIR) * T(0.95105651629515357211643933337938);  // sin(2pi/5)
    const T s2 = T(DIR) * T(0.58778525229247312916870595463907);  // sin(4pi/5)
    cpx<T> a1 = cadd(v1, v4), b1 = csub(v1, v4);
    cpx<T> a2 = cadd(v2, v3), b2 = csub(v2, v3);
    cpx<T> r1 = {v0.x + c1 * a1.x + c2 * a2.x, v0.y + c1 * a1.y + c2 * a2.y};
    cpx<T> r2 = {v0.x + c2 * a1.x + c1 * a2.x, v0.y + c2 * a1.y + c1 * a2.y};
    // i*(s1*b1 + s2*b2) and i*(s2*b1 - s1*b2)
    cpx<T> q1 = {s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y};
    cpx<T> q2 = {s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y};
    cpx<T> iq1 = {-q1.y, q1.x}, iq2 = {-q2.y, q2.x};
    v0 = {v0.x + a1.x + a2.x, v0.y + a1.y + a2.y};
    v1 = cadd(r1, iq1);
    v4 = csub(r1, iq1);
    v2 = cadd(r2, iq2);
    v3 = csub(r2, iq2);
}

template <int DIR, typename T>
__device__ __forceinline__ void dft7(cpx<T>* v) {
    const T c1 = T(0.62348980185873353052500488400424);
    const T c2 = T(-0.22252093395631440428890256449679);
    const T c3 = T(-0.90096886790241912623610231950745);
    const T s1 = T(DIR) * T(0.78183148246802980870844452667406);
    const T s2 = T(DIR) * T(0.97492791218182360701813168299393);
    const T s3 = T(DIR) * T(0.43388373911755812047576833284836);
    cpx<T> a1 = cadd(v[1], v[6]), b1 = csub(v[1], v[6]);
    cpx<T> a2 = cadd(v[2], v[5]), b2 = csub(v[2], v[5]);
    cpx<T> a3 = cadd(v[3], v[4]), b3 = csub(v[3], v[4]);
    cpx<T> x0 = v[0];
    cpx<T> r1 = {x0.x + c1 * a1.x + c2 * a2.x + c3 * a3.x, x0.y + c1 * a1.y + c2 * a2.y + c3 * a3.y};
    cpx<T> r2 = {x0.x + c2 * a1.x + c3 * a2.x + c1 * a3.x, x0.y + c2 * a1.y + c3 * a2.y + c1 * a3.y};
    cpx<T> r3 = {x0.x + c3 * a1.x + c1 * a2.x + c2 * a3.x, x0.y + c3 * a1.y + c1 * a2.y + c2 * a3.y};
    cpx<T> q1 = {s1 * b1.x + s2 * b2.x + s3 * b3.x, s1 * b1.y + s2 * b2.y + s3 * b3.y};
    cpx<T> q2 = {s2 * b1.x - s3 * b2.x - s1 * b3.x, s2 * b1.y - s3 * b2.y - s1 * b3.y};
    cpx<T> q3 = {s3 * b1.x - s1 * b2.x + s2 * b3.x, s3 * b1.y - s1 * b2.y + s2 * b3.y};
    v[0] = {x0.x + a1.x + a2.x + a3.x, x0.y + a1.y + a2.y + a3.y};
    cpx<T> iq1 = {-q1.y, q1.x}, iq2 = {-q2.y, q2.x}, iq3 = {-q3.y, q3.x};
    v[1] = cadd(r1, iq1);
    v[6] = csub(r1, iq1);
    v[2] = cadd(r2, iq2);
    v[5] = csub(r2, iq2);
    v[3] = cadd(r3, iq3);
    v[4] = csub(r3, iq3);
}

// radix 8 = 2 x 4 with internal twiddles exp(DIR*2pi*i*k/8)
template <int DIR, typename T>
__device__ __forceinline__ void dft8(cpx<T>* v) {
    const T h = T(0.70710678118654752440084436210485);
    // first: 4 radix-2 butterflies on (k, k+4)
    dft2<DIR>(v[0], v[4]);
    dft2<DIR>(v[1], v[5]);
    dft2<DIR>(v[2], v[6]);
    dft2<DIR>(v[3], v[7]);
    // twiddles on odd half: w^1, w^2, w^3 with w = exp(DIR*2pi i/8)
    {
        cpx<T> t = v[5];
        // (1 + DIR*i) * h
        if constexpr (DIR < 0) v[5] = {h * (t.x + t.y), h * (t.y - t.x)};
        else v[5] = {h * (t.x - t.y), h * (t.y + t.x)};
        v[6] = mul_i<DIR>(v[6]);
        t = v[7];
        // w^3 = (-1 + DIR*i) * h
        if constexpr (DIR < 0) v[7] = {h * (-t.x + t.y), h * (-t.y - t.x)};
        else v[7] = {h * (-t.x - t.y), h * (-t.y + t.x)};
    }
    dft4<DIR>(v[0], v[1], v[2], v[3]);
    dft4<DIR>(v[4], v[5], v[6], v[7]);
    // outputs: even k at v[0..3] -> X_{2m}, odd at v[4..7] -> X_{2m+1}
    cpx<T> e0 = v[0], e1 = v[1], e2 = v[2], e3 = v[3];
    cpx<T> o0 = v[4], o1 = v[5], o2 = v[6], o3 = v[7];
    v[0] = e0; v[2] = e1; v[4] = e2; v[6] = e3;
    v[1] = o0; v[3] = o1; v[5] = o2; v[7] = o3;
}

// radix 16 = 4 x 4 with internal twiddles exp(DIR*2pi*i*j*k/16)
template <int DIR, typename T>
__device__ __forceinline__ void dft16(cpx<T>* v) {
    const T c1 = T(0.92387953251128675612818318939679);  // cos(pi/8)
    const T s1 = T(0.38268343236508977172845998403040);  // sin(pi/8)
    const T h = T(0.70710678118654752440084436210485);
    // stage 1: for each q in 0..3, radix-4 over (q, q+4, q+8, q+12)
#pragma unroll
    for (int q = 0; q < 4; q++) dft4<DIR>(v[q], v[q + 4], v[q + 8], v[q + 12]);
    // twiddles w16^(q*m) on element v[q + 4m], m = 1..3
    auto tw = [&](cpx<T>& a, T c, T s) {  // multiply by c + DIR*i*s
        T sd = T(DIR) * s;
        a = {a.x * c - a.y * sd, a.x * sd + a.y * c};
    };
    // q=1: m=1 -> w^1, m=2 -> w^2, m=3 -> w^3
    tw(v[5], c1, s1);
    tw(v[9], h, h);
    tw(v[13], s1, c1);
    // q=2: w^2, w^4, w^6
    tw(v[6], h, h);
    v[10] = mul_i<DIR>(v[10]);
    tw(v[14], -h, h);
    // q=3: w^3, w^6, w^9
    tw(v[7], s1, c1);
    tw(v[11], -h, h);
    tw(v[15], -c1, -s1);
    // stage 2: for each m, radix-4 over q = 0..3 at (4m + q)
#pragma unroll
    for (int m = 0; m < 4; m++) dft4<DIR>(v[4 * m], v[4 * m + 1], v[4 * m + 2], v[4 * m + 3]);
    // output X_{m + 4p} sits at v[4m + p]; transpose 4x4
    cpx<T> t[16];
#pragma unroll
    for (int i = 0; i < 16; i++) t[i] = v[i];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int p = 0; p < 4; p++) v[m + 4 * p] = t[4 * m + p];
}

// Generic in-register DFT dispatch on a compile-time radix.
template <int R, int DIR, typename T>
__device__ __forceinline__ void dft(cpx<T>* v) {
    if constexpr (R == 2) dft2<DIR>(v[0], v[1]);
    else if constexpr (R == 3) dft3<DIR>(v[0], v[1], v[2]);
    else if constexpr (R == 4) dft4<DIR>(v[0], v[1], v[2], v[3]);
    else if constexpr (R == 5) dft5<DIR>(v[0], v[1], v[2], v[3], v[4]);
    else if constexpr (R == 7) dft7<DIR>(v);
    else if constexpr (R == 8) dft8<DIR>(v);
    else if constexpr (R == 16) dft16<DIR>(v);
    else static_assert(R == 2, "unsupported radix");
}

}  // namespace jt
