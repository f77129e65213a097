// LDS-resident Stockham autosort FFT kernels for gfx950.
//
// One workgroup transforms W lines of N complex points.  Each line is served
// by TL threads; each thread keeps E = N/TL points in registers.  Stage s with
// radix R and span Ns (product of the earlier radices) runs, for every
// butterfly j in [0, N/R):
//     v_r   = in[j + r*N/R] * exp(DIR*2*pi*i*(j mod Ns)*r/(Ns*R))
//     v     = DFT_R(v)
//     out[(j/Ns)*Ns*R + (j mod Ns) + r*Ns] = v_r
// which leaves the result in natural order (no bit-reversal pass, unlike the
// reference's split-radix + bitrv2 in CommonUtils.java:708-822, 824-2843).
// Thread t owns butterflies j = t + TL*q, so the first stage reads and the last
// stage writes elements t + TL*i: consecutive lanes touch consecutive elements
// (row mode) or consecutive lines (column mode, lines adjacent in memory),
// i.e. every global access is coalesced.  Between stages the data goes through
// one padded LDS buffer; with SPLIT the real and imaginary halves are
// exchanged in two rounds so the buffer is half the size (more waves per CU).
#pragma once
#include <type_traits>

#include "jt_dft.hpp"

namespace jt {

// ----------------------------------------------------------------------------
// compile-time helpers
// ----------------------------------------------------------------------------
template <int... Rs>
struct RadixList {
    static constexpr int count = sizeof...(Rs);
    static constexpr int product() { return (1 * ... * Rs); }
    static constexpr int maxr() {
        int m = 1;
        ((m = Rs > m ? Rs : m), ...);
        return m;
    }
    // offset of stage S in the per-stage twiddle table
    static constexpr int twoff(int S) {
        int r[] = {Rs...};
        int off = 0, ns = 1;
        for (int i = 0; i < S; i++) {
            off += ns * (r[i] - 1);
            ns *= r[i];
        }
        return off;
    }
};

// LDS index with one pad element every 32 entries (bank spreading for the
// stride-R writes of the early stages).
template <int N>
__device__ __forceinline__ int lds_pos(int k) { return k + (k >> 5); }
template <int N, int W>
struct LdsGeom {
    static constexpr int base = N + (N >> 5);
    // line pitch chosen so that W lines fall into disjoint bank groups
    static constexpr int off = (16 / W) > 0 ? (16 / W) : 1;
    static constexpr int pitch = W == 1 ? base : base + (off - (base % 32) + 64) % 32;
    static constexpr int size = W * pitch;
};

// ----------------------------------------------------------------------------
// The per-line state handed to IO policies
// ----------------------------------------------------------------------------
template <typename T, int N, int W, bool SPLIT>
struct LdsBuf {
    using G = LdsGeom<N, W>;
    T* re;  // SPLIT: the single exchange array; else real parts
    T* im;  // !SPLIT: imaginary parts
};

template <typename T, int DIR, int N, int TL, int W, bool COLMODE, bool SPLIT>
struct Engine {
    static constexpr int E = N / TL;
    using G = LdsGeom<N, W>;

    // Write the register values of the butterflies of a stage (radix R, span
    // NS) to LDS at their output positions.
    template <int R, int NS, int PART>
    static __device__ __forceinline__ void put(T* buf, int w, int t, const cpx<T>* v) {
#pragma unroll
        for (int q = 0; q < E / R; q++) {
            int j = t + TL * q;
            int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
            for (int r = 0; r < R; r++) {
                int k = base + r * NS;
                T val = PART == 0 ? v[q * R + r].x : v[q * R + r].y;
                buf[w * G::pitch + lds_pos<N>(k)] = val;
            }
        }
    }
    // Read the inputs of the butterflies of a stage with radix R.
    template <int R, int PART>
    static __device__ __forceinline__ void get(const T* buf, int w, int t, cpx<T>* v) {
#pragma unroll
        for (int q = 0; q < E / R; q++) {
            int j = t + TL * q;
#pragma unroll
            for (int r = 0; r < R; r++) {
                int k = j + r * (N / R);
                T val = buf[w * G::pitch + lds_pos<N>(k)];
                if (PART == 0) v[q * R + r].x = val;
                else v[q * R + r].y = val;
            }
        }
    }

    // Exchange after a stage (RP, NSP) for the next stage with radix R.
    template <int RP, int NSP, int R>
    static __device__ __forceinline__ void exchange(LdsBuf<T, N, W, SPLIT> lds, int w, int t,
                                                    cpx<T>* v) {
        if constexpr (SPLIT) {
            __syncthreads();
            put<RP, NSP, 0>(lds.re, w, t, v);
            __syncthreads();
            get<R, 0>(lds.re, w, t, v);
            __syncthreads();
            put<RP, NSP, 1>(lds.re, w, t, v);
            __syncthreads();
            get<R, 1>(lds.re, w, t, v);
        } else {
            __syncthreads();
            put<RP, NSP, 0>(lds.re, w, t, v);
            put<RP, NSP, 1>(lds.im, w, t, v);
            __syncthreads();
            get<R, 0>(lds.re, w, t, v);
            get<R, 1>(lds.im, w, t, v);
        }
    }

    // Store the final stage's output (element order) into LDS for a post pass.
    template <int R, int NS>
    static __device__ __forceinline__ void put_final(LdsBuf<T, N, W, SPLIT> lds, int w, int t,
                                                     const cpx<T>* v) {
        __syncthreads();
        put<R, NS, 0>(lds.re, w, t, v);
        put<R, NS, 1>(lds.im, w, t, v);
        __syncthreads();
    }

    template <int R, int NS, int TWO>
    static __device__ __forceinline__ void twiddle(const cpx<T>* __restrict__ tw, int t, cpx<T>* v) {
        if constexpr (NS > 1) {
#pragma unroll
            for (int q = 0; q < E / R; q++) {
                int j = t + TL * q;
                int jm = j % NS;
                const cpx<T>* row = tw + TWO + jm * (R - 1);
#pragma unroll
                for (int r = 1; r < R; r++) v[q * R + r] = cmul(v[q * R + r], tw_dir<DIR>(row[r - 1]));
            }
        }
    }

    template <int R>
    static __device__ __forceinline__ void butterflies(cpx<T>* v) {
#pragma unroll
        for (int q = 0; q < E / R; q++) dft<R, DIR>(v + q * R);
    }

    // Run stages S.. (radix R, span NS) given that v holds the *inputs* of
    // stage S (already exchanged).  Returns with v holding the outputs of the
    // last stage, butterfly-major: v[q*RL + r] is element t + TL*(q + r*E/RL).
    template <class List, int S, int NS, int R, int... Rest>
    static __device__ __forceinline__ void stages(LdsBuf<T, N, W, SPLIT> lds, int w, int t,
                                                  const cpx<T>* __restrict__ tw, cpx<T>* v) {
        twiddle<R, NS, List::twoff(S)>(tw, t, v);
        butterflies<R>(v);
        if constexpr (sizeof...(Rest) > 0) {
            constexpr int RN = first<Rest...>();
            static_assert(E % RN == 0, "radix must divide the per-thread point count");
            exchange<R, NS, RN>(lds, w, t, v);
            stages<List, S + 1, NS * R, Rest...>(lds, w, t, tw, v);
        }
    }

    template <int A, int... B>
    static constexpr int first() { return A; }
};

template <int... Rs>
struct LastRadix;
template <int R>
struct LastRadix<R> {
    static constexpr int value = R;
};
template <int R, int... Rest>
struct LastRadix<R, Rest...> {
    static constexpr int value = LastRadix<Rest...>::value;
};
template <int R0, int... Rs>
struct FirstRadix {
    static constexpr int value = R0;
};

// Map workgroup id to line group so that consecutive groups (adjacent in
// memory) land on the same XCD (workgroups are dispatched round-robin over
// the 8 XCDs); keeps partial cache lines of neighbouring groups in one L2.
__device__ __forceinline__ int64_t xcd_group(int64_t bid, int64_t ngroups) {
    const int64_t nx = 8;
    int64_t per = ngroups / nx;
    int64_t full = per * nx;
    if (bid >= full) return bid;
    return (bid % nx) * per + bid / nx;
}

// IO policies that finish line 0 in place declare `static constexpr int
// line0_hook` (> 0: after the line's stores, < 0: before its loads) and
// `__device__ void line0(int tid, int nthreads)`.
template <class IO, class = void>
struct io_line0_hook : std::integral_constant<int, 0> {};
template <class IO>
struct io_line0_hook<IO, std::void_t<decltype(IO::line0_hook)>>
    : std::integral_constant<int, IO::line0_hook> {};

// ----------------------------------------------------------------------------
// The kernel.  IO supplies:
//   int64_t nlines;  bool xcd;
//   __device__ int64_t base(int64_t line)               element-0 offset
//   __device__ cpx<T> load(int64_t base, int k)          element k of the line
//   __device__ void store(int64_t base, int k, cpx<T>)   element k of the result
//   static constexpr bool pre_lds, post_lds;
//   pre_lds : __device__ void pre(int64_t line, int64_t base, T* re, T* im, int t)
//             fills the line's LDS slots (re/im arrays, index lds_pos(k)) with
//             the N transform inputs; threads t in [0,TL) share the work.
//   post_lds: __device__ void post(int64_t line, int64_t base, const T* re,
//             const T* im, int t) consumes the N outputs from LDS.
// ----------------------------------------------------------------------------
template <typename T, int DIR, int N, int TL, int W, bool COLMODE, bool SPLIT, class IO, int... Rs>
__global__ void __launch_bounds__(TL* W) stockham_kernel(IO io, const cpx<T>* __restrict__ tw) {
    using Eng = Engine<T, DIR, N, TL, W, COLMODE, SPLIT>;
    using G = LdsGeom<N, W>;
    constexpr int E = N / TL;
    constexpr int R0 = FirstRadix<Rs...>::value;
    constexpr int RL = LastRadix<Rs...>::value;
    constexpr int NL = N / RL;
    constexpr bool need_im = (!SPLIT) || IO::pre_lds || IO::post_lds;
    __shared__ T lds_re[G::size];
    __shared__ T lds_im[need_im ? G::size : 1];
    LdsBuf<T, N, W, SPLIT> lds{lds_re, lds_im};

    const int tid = threadIdx.x;
    const int w = COLMODE ? tid % W : tid / TL;
    const int t = COLMODE ? tid / W : tid % TL;
    const int64_t ngroups = (io.nlines + W - 1) / W;
    const int64_t g = io.xcd ? xcd_group(blockIdx.x, ngroups) : (int64_t)blockIdx.x;
    const int64_t line = g * W + w;
    const bool active = line < io.nlines;
    const int64_t base = active ? io.base(line) : 0;

    if constexpr (io_line0_hook<IO>::value < 0) {
        // the workgroup that holds line 0 rewrites it before anyone loads it
        if (g == 0) {
            io.line0(tid, TL * W);
            __syncthreads();
            __atomic_thread_fence(__ATOMIC_ACQ_REL);
        }
    }

    cpx<T> v[E];
    if constexpr (IO::pre_lds) {
        if (active) io.template pre<N, TL>(line, base, lds_re + w * G::pitch, lds_im + w * G::pitch, t, TL);
        __syncthreads();
        Eng::template get<R0, 0>(lds_re, w, t, v);
        Eng::template get<R0, 1>(lds_im, w, t, v);
    } else {
        // v[q*R0 + r] = element t + TL*(q + r*E/R0)
#pragma unroll
        for (int q = 0; q < E / R0; q++)
#pragma unroll
            for (int r = 0; r < R0; r++) {
                int k = t + TL * (q + r * (E / R0));
                v[q * R0 + r] = active ? io.load(base, k) : cpx<T>{T(0), T(0)};
            }
    }

    Eng::template stages<RadixList<Rs...>, 0, 1, Rs...>(lds, w, t, tw, v);

    if constexpr (IO::post_lds) {
        Eng::template put_final<RL, NL>(lds, w, t, v);
        if (active) io.template post<N, TL>(line, base, lds_re + w * G::pitch, lds_im + w * G::pitch, t, TL);
    } else {
        if (active) {
#pragma unroll
            for (int q = 0; q < E / RL; q++)
#pragma unroll
                for (int r = 0; r < RL; r++) {
                    int k = t + TL * (q + r * (E / RL));
                    io.store(base, k, v[q * RL + r]);
                }
        }
    }
    if constexpr (io_line0_hook<IO>::value > 0) {
        // the workgroup that holds line 0 finishes it once all its stores
        // are done; the fence orders them before the re-reads and drops L1
        // lines loaded at the start
        if (g == 0) {
            __syncthreads();
            __atomic_thread_fence(__ATOMIC_ACQ_REL);
            io.line0(tid, TL * W);
        }
    }
}

}  // namespace jt
