// Runtime-radix Stockham kernel: one workgroup per line, the line held in two
// LDS buffers (ping-pong), radices read from the kernel arguments.  Serves all
// lengths that have no compile-time kernel: the mixed-radix sizes of the
// reference's FFTPACK path (DoubleFFT_1D.java:6630-8008) and the sub-lengths
// of the four-step and Bluestein plans.
#pragma once
#include "jt_dft.hpp"

namespace jt {

constexpr int GEN_MAX_STAGES = 20;

struct GenPlan {
    int n;
    int nstages;
    int radix[GEN_MAX_STAGES];
    int twoff[GEN_MAX_STAGES];
};

__device__ __forceinline__ int gpos(int k) { return k + (k >> 5); }

// q = a / d for 0 <= a < 2^22, d >= 1, from a float reciprocal with one
// correction step (exact: the float quotient a*(1/d) is within 3*2^-24 a/d
// of the true one, i.e. within 1 for these ranges)
__device__ __forceinline__ int fdiv(int a, int d, float rcp) {
    int q = (int)((float)a * rcp);
    int r = a - q * d;
    if (r < 0) q--;
    else if (r >= d) q++;
    return q;
}

// cos / sin (2 pi m / P), m = 0..P-1, for the prime radices 11 and 13
template <int P>
struct PrimeTw;
template <>
struct PrimeTw<11> {
    static constexpr double C[11] = {1.0, 0.841253532831181168862, 0.415415013001886425529,
                                     -0.142314838273285140444, -0.654860733945285064057, -0.95949297361449738989,
                                     -0.95949297361449738989, -0.654860733945285064057, -0.142314838273285140444,
                                     0.415415013001886425529, 0.841253532831181168862};
    static constexpr double S[11] = {0.0, 0.540640817455597582108, 0.909631995354518371412,
                                     0.989821441880932732376, 0.755749574354258283774, 0.281732556841429697711,
                                     -0.281732556841429697711, -0.755749574354258283774, -0.989821441880932732376,
                                     -0.909631995354518371412, -0.540640817455597582108};
};
template <>
struct PrimeTw<13> {
    static constexpr double C[13] = {1.0, 0.8854560256532098959, 0.568064746731155802512,
                                     0.120536680255323053349, -0.35460488704253562597, -0.748510748171101098635,
                                     -0.970941817426052027157, -0.970941817426052027157, -0.748510748171101098635,
                                     -0.35460488704253562597, 0.120536680255323053349, 0.568064746731155802512,
                                     0.8854560256532098959};
    static constexpr double S[13] = {0.0, 0.464723172043768545656, 0.82298386589365639458,
                                     0.992708874098053992801, 0.93501624268541482344, 0.663122658240795202377,
                                     0.239315664287557767149, -0.239315664287557767149, -0.663122658240795202377,
                                     -0.93501624268541482344, -0.992708874098053992801, -0.82298386589365639458,
                                     -0.464723172043768545656};
};

// Prime radix P (11, 13), X_k = sum_j v_j exp(DIR 2 pi i jk/P), by pairing
// j with P-j: with a_j = v_j + v_{P-j}, b_j = v_j - v_{P-j} (j = 1..H,
// H = (P-1)/2) and m = jk mod P,
//   X_0     = v_0 + sum a_j
//   X_k     = v_0 + sum a_j C_m + i DIR sum b_j S_m
//   X_{P-k} = v_0 + sum a_j C_m - i DIR sum b_j S_m       (k = 1..H)
// All indices are compile-time constants after unrolling.
template <int P, int DIR, typename T>
__device__ __forceinline__ void dft_prime(cpx<T>* v) {
    constexpr int H = (P - 1) / 2;
    cpx<T> a[H + 1], b[H + 1];
    cpx<T> x0 = v[0];
#pragma unroll
    for (int j = 1; j <= H; j++) {
        a[j] = cadd(v[j], v[P - j]);
        b[j] = csub(v[j], v[P - j]);
        x0 = cadd(x0, a[j]);
    }
#pragma unroll
    for (int k = 1; k <= H; k++) {
        T re = v[0].x, im = v[0].y, sr = T(0), si = T(0);
#pragma unroll
        for (int j = 1; j <= H; j++) {
            const T c = (T)PrimeTw<P>::C[(j * k) % P];
            const T sn = (T)PrimeTw<P>::S[(j * k) % P];
            re += a[j].x * c;
            im += a[j].y * c;
            sr += b[j].y * sn;
            si += b[j].x * sn;
        }
        // i DIR (b S) = DIR (-(b.y S), b.x S)
        v[k] = cpx<T>{re - T(DIR) * sr, im + T(DIR) * si};
        v[P - k] = cpx<T>{re + T(DIR) * sr, im - T(DIR) * si};
    }
    v[0] = x0;
}

template <int R, int DIR, typename T>
__device__ __forceinline__ void dft_any(cpx<T>* v) {
    if constexpr (R == 11 || R == 13) dft_prime<R, DIR>(v);
    else dft<R, DIR>(v);
}

// One Stockham stage of compile-time radix R over LDS buffers a -> b.
template <int R, int DIR, typename T>
__device__ __forceinline__ void gen_stage(const T* are, const T* aim, T* bre, T* bim,
                                          const cpx<T>* __restrict__ tws, int n, int ns, int t,
                                          int nt) {
    const int m = n / R;
    const float rns = 1.0f / (float)ns;
    for (int j = t; j < m; j += nt) {
        const int jh = fdiv(j, ns, rns);
        const int jm = j - jh * ns;
        cpx<T> v[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            int k = gpos(j + r * m);
            v[r] = {are[k], aim[k]};
        }
        if (ns > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) v[r] = cmul(v[r], tw_dir<DIR>(tws[jm * (R - 1) + r - 1]));
        }
        dft_any<R, DIR>(v);
        const int d = jh * ns * R + jm;
#pragma unroll
        for (int r = 0; r < R; r++) {
            int k = gpos(d + r * ns);
            bre[k] = v[r].x;
            bim[k] = v[r].y;
        }
    }
}

template <int DIR, typename T>
__device__ __forceinline__ void gen_stage_rt(int R, const T* are, const T* aim, T* bre, T* bim,
                                             const cpx<T>* __restrict__ tws, int n, int ns, int t,
                                             int nt) {
    switch (R) {
        case 2: gen_stage<2, DIR>(are, aim, bre, bim, tws, n, ns, t, nt); break;
        case 3: gen_stage<3, DIR>(are, aim, bre, bim, tws, n, ns, t, nt); break;
        case 4: gen_stage<4, DIR>(are, aim, bre, bim, tws, n, ns, t, nt); break;
        case 5: gen_stage<5, DIR>(are, aim, bre, bim, tws, n, ns, t, nt); break;
        case 7: gen_stage<7, DIR>(are, aim, bre, bim, tws, n, ns, t, nt); break;
        case 8: gen_stage<8, DIR>(are, aim, bre, bim, tws, n, ns, t, nt); break;
        case 11: gen_stage<11, DIR>(are, aim, bre, bim, tws, n, ns, t, nt); break;
        case 13: gen_stage<13, DIR>(are, aim, bre, bim, tws, n, ns, t, nt); break;
        case 16: gen_stage<16, DIR>(are, aim, bre, bim, tws, n, ns, t, nt); break;
        default: break;
    }
}

// IO contract as for stockham_kernel, with W == 1 and pre/post hooks working
// on one line: pre(line, base, re, im, t, nthreads) / post(line, base, re, im,
// t, nthreads) address LDS through gpos(k).
template <typename T, int DIR, class IO>
__global__ void __launch_bounds__(256) generic_kernel(IO io, const cpx<T>* __restrict__ tw,
                                                     GenPlan plan) {
    extern __shared__ unsigned char smem_raw[];
    T* sm = reinterpret_cast<T*>(smem_raw);
    const int n = plan.n;
    const int np = n + (n >> 5) + 1;
    T* are = sm;
    T* aim = sm + np;
    T* bre = sm + 2 * np;
    T* bim = sm + 3 * np;
    const int t = threadIdx.x;
    const int nt = blockDim.x;
    const int64_t line = blockIdx.x + (int64_t)blockIdx.y * gridDim.x;
    if (line >= io.nlines) return;  // whole workgroup exits together
    const int64_t base = io.base(line);

    if constexpr (IO::pre_lds) {
        io.template pre<0, 0>(line, base, are, aim, t, nt);
    } else {
        for (int k = t; k < n; k += nt) {
            cpx<T> x = io.load(base, k);
            are[gpos(k)] = x.x;
            aim[gpos(k)] = x.y;
        }
    }
    int ns = 1;
    for (int s = 0; s < plan.nstages; s++) {
        __syncthreads();
        const int R = plan.radix[s];
        gen_stage_rt<DIR>(R, are, aim, bre, bim, tw + plan.twoff[s], n, ns, t, nt);
        ns *= R;
        T* x;
        x = are; are = bre; bre = x;
        x = aim; aim = bim; bim = x;
    }
    __syncthreads();
    if constexpr (IO::post_lds) {
        io.template post<0, 0>(line, base, are, aim, t, nt);
    } else {
        for (int k = t; k < n; k += nt) io.store(base, k, cpx<T>{are[gpos(k)], aim[gpos(k)]});
    }
}

}  // namespace jt
