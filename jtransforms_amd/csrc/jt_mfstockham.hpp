// Stockham FFT with the radix-16 stages on MFMA in split FP16 (gfx950).
//
// Same algorithm and natural-order result as stockham_kernel (jt_stockham.hpp):
// stage (R, Ns) reads in[j + r*N/R], multiplies by exp(DIR 2 pi i (j mod Ns)
// r/(Ns R)), runs DFT_R and writes out[(j/Ns)*Ns*R + (j mod Ns) + r*Ns].
// Here every radix-16 stage is a 16x16 complex DFT-matrix product (MfmaDft16)
// and the data move in the matrix cores' own operand layouts:
//   * a wave holds 64 butterflies as 4 tiles of 16; for tile q, lane l reads
//     the B operand (8 reals: Re or Im of 8 of the 16 inputs) of butterfly
//     64*wave + 16q + l%16 straight from LDS, or from HBM for the first stage
//     of a contiguous line (each load instruction covers two full 128-B
//     segments: the Re and Im lanes of 16 adjacent points);
//   * the product leaves output i = 4*(l/16)+u of that butterfly, complex, in
//     lane l; the lane multiplies it by the NEXT radix-16 stage's twiddle for
//     its destination and writes it to LDS (or, after the last stage of a
//     contiguous line, straight to HBM);
//   * block floating point: the first stage's inputs of a line are scaled by
//     one power of two from the line's largest entry (an LDS max over the
//     workgroup's lines) into [2^13, 2^14); each later radix-16 stage's
//     twiddles carry 2^-4, so every B operand stays below 2^14; the store
//     undoes the product of these factors.  No per-tile range search.
// So the twiddle never has to be applied in the B layout (where Re and Im of
// a point sit in different lanes), and each stage costs one LDS write and one
// LDS read per point, as in the FP32 kernel.  A radix 2/4/8 stage may close
// the list; it runs on the vector ALUs with the FP32 engine's code.
#pragma once
#include "jt_mfma.hpp"
#include "jt_stockham.hpp"

namespace jt {

// Twiddle table of these kernels: radix-16 stage s (span 16^s) holds, for
// every jm < 16^s, the 16 factors exp(-2 pi i jm r/16^(s+1)), r = 0..15 (the
// r = 0 entry, 1, spares a branch); a closing radix-R stage follows in the
// FP32 engine's layout (R-1 factors per jm).
__host__ __device__ constexpr int mf_twoff(int s) {
    int off = 0, ns = 1;
    for (int i = 0; i < s; i++) {
        off += ns * 16;
        ns *= 16;
    }
    return off;
}

// Output v (0..15) of radix-16 butterfly j at span ns (a power of 16) is point
// p = a + v*ns with a = (j/ns)*16ns + j%ns, stored at lds_pos(p) = p + p/32.
// The padding term splits into a per-butterfly part and a part linear in v
// (so a lane computes one base per butterfly and adds constants):
//   ns = 1:    p lies in one 32-point block: lds_pos(a) + v;
//   ns = 16:   a%32 < 16, so p/32 = a/32 + v/2: lds_pos(a) + 16v + v/2;
//   ns >= 256: ns is a multiple of 32: lds_pos(a) + v*(ns + ns/32).
__device__ __forceinline__ int mf_slot_base(int j, int ns) {
    const int a = (j / ns) * ns * 16 + (j % ns);
    return a + (a >> 5);
}
__device__ __forceinline__ int mf_slot_step(int ns, int v) {
    return ns == 1 ? v : (ns == 16 ? 16 * v + (v >> 1) : v * (ns + (ns >> 5)));
}

template <int DIR, int N, int W, bool COLMODE, class IO, int... Rs>
__global__ void __launch_bounds__(N / 16 * W) mfma_stockham_kernel(IO io, const cpx<float>* __restrict__ tw) {
    using List = RadixList<Rs...>;
    constexpr int TL = N / 16;   // threads per line; one radix-16 butterfly each
    constexpr int NT = TL * W;   // threads (= radix-16 butterflies) per workgroup
    constexpr int NS_ALL = List::count;
    constexpr int RL = LastRadix<Rs...>::value;
    constexpr int NMF = RL == 16 ? NS_ALL : NS_ALL - 1;  // radix-16 stages (all first)
    static_assert(NT % 64 == 0, "radix-16 MFMA stages need whole waves");
    static_assert(NMF >= 1, "at least one radix-16 stage");
    static_assert(RL == 16 || RL == 8 || RL == 4 || RL == 2, "closing radix 2, 4, 8 or 16");
    using G = LdsGeom<N, W>;
    using Eng = Engine<float, DIR, N, TL, W, COLMODE, false>;
    __shared__ float lre[G::size];
    __shared__ float lim[G::size];
    __shared__ unsigned lmax[W];  // per line: bits of the largest |input|

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = tid >> 6;
    const int jcol = lane & 15;
    const int grp = lane >> 4;
    const int part = grp >> 1;       // B rows 0-15 Re, 16-31 Im
    const int rh = (grp & 1) * 8;    // first of this lane's 8 points
    MfmaDft16 mf;
    mf.template init<DIR>(lane);

    const int64_t ngroups = (io.nlines + W - 1) / W;
    const int64_t gid = io.xcd ? xcd_group(blockIdx.x, ngroups) : (int64_t)blockIdx.x;

    // the butterflies of this lane's 4 tiles: line in the group and index j
    int tl_line[4], tl_j[4];
    int64_t tl_base[4];
    bool tl_act[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int b = wv * 64 + q * 16 + jcol;
        tl_line[q] = b / TL;
        tl_j[q] = b % TL;
        const int64_t line = gid * W + tl_line[q];
        tl_act[q] = line < io.nlines;
        tl_base[q] = io.base(tl_act[q] ? line : gid * W);
    }
    // the FP32 engine's thread mapping (staging and the closing VALU stage)
    const int ew = COLMODE ? tid % W : tid / TL;
    const int et = COLMODE ? tid / W : tid % TL;
    const int64_t eline = gid * W + ew;
    const bool eact = eline < io.nlines;
    const int64_t ebase = io.base(eact ? eline : gid * W);

    if (tid < W) lmax[tid] = 0u;
    float b[4][8];
    // ---- inputs of the first stage in the B layout
    if constexpr (!COLMODE) {
        // contiguous lines (element stride 1); a line past the end reads
        // line 0 of the group instead and never stores
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float* src = reinterpret_cast<const float*>(io.in + tl_base[q]) + 2 * (tl_j[q] + rh * TL) + part;
#pragma unroll
            for (int u = 0; u < 8; u++) {
                if constexpr ((IO::streaming & NT_LOAD) != 0) b[q][u] = __builtin_nontemporal_load(src + 2 * TL * u);
                else b[q][u] = src[2 * TL * u];
            }
        }
    } else {
        // coalesced across adjacent lines, staged through LDS
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int k = et + TL * i;
            cpx<float> v = io.load(ebase, k);
            lre[ew * G::pitch + lds_pos<N>(k)] = v.x;
            lim[ew * G::pitch + lds_pos<N>(k)] = v.y;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float* src = (part ? lim : lre) + tl_line[q] * G::pitch;
#pragma unroll
            for (int u = 0; u < 8; u++) b[q][u] = src[lds_pos<N>(tl_j[q] + (rh + u) * TL)];
        }
    }

    // ---- one power-of-two scale per line (block floating point)
    if constexpr (!COLMODE) __syncthreads();  // lmax cleared
#pragma unroll
    for (int q = 0; q < 4; q++) {
        float m = 0.0f;
#pragma unroll
        for (int u = 0; u < 8; u++) m = fmaxf(m, fabsf(b[q][u]));
        atomicMax(&lmax[tl_line[q]], __float_as_uint(m));
    }
    __syncthreads();
    constexpr int STEP = 4 * (NMF - 1);  // the later stages' 2^-4 factors
    float undo[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int sh = mf_line_shift(lmax[tl_line[q]]);
        const float sc = mf_pow2(sh);
        undo[q] = io.scale * mf_pow2(STEP - sh);
#pragma unroll
        for (int u = 0; u < 8; u++) b[q][u] *= sc;
    }

    // ---- the radix-16 stages
    int ns = 1;
#pragma unroll
    for (int s = 0; s < NMF; s++) {
        float outr[4][4], outi[4][4];
#pragma unroll
        for (int q = 0; q < 4; q++) mf.tile(b[q], outr[q], outi[q]);
        const bool last_mf = s == NMF - 1;
        const bool direct_out = last_mf && NMF == NS_ALL && !COLMODE;
        if (direct_out) {
            // last stage: ns = N/16, output i of butterfly j is element j + i*TL
#pragma unroll
            for (int q = 0; q < 4; q++) {
                cpx<float>* dst = io.out + tl_base[q] + tl_j[q] + 4 * grp * TL;
                if (tl_act[q]) {
#pragma unroll
                    for (int u = 0; u < 4; u++)
                    {
                        const cpx<float> o{outr[q][u] * undo[q], outi[q][u] * undo[q]};
                        if constexpr ((IO::streaming & NT_STORE) != 0) st_nt(dst + TL * u, o);
                        else dst[TL * u] = o;
                    }
                }
            }
        } else {
            // the next radix-16 stage's twiddle is applied here, at the write
            const bool tw_next = !last_mf;
            const int two = mf_twoff(s + 1);
            const int ns2 = ns * 16;
            if (s > 0 || COLMODE) __syncthreads();  // everyone has read this stage's inputs
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int j = tl_j[q];
                const int slot = tl_line[q] * G::pitch + mf_slot_base(j, ns);
                // next stage's twiddle of point p: jm2 = p % ns2 = j%ns + v*ns and
                // r2 = p / TL = (j/ns) / (TL/ns2) (ns2 divides TL; j%ns + v*ns < ns2)
                const int twb = two + (j % ns) * 16 + (j / ns) / (TL / ns2 > 0 ? TL / ns2 : 1);
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int vv = 4 * grp + u;
                    cpx<float> v{outr[q][u], outi[q][u]};
                    if (tw_next) v = cmul(v, tw_dir<DIR>(tw[twb + vv * ns * 16]));
                    lre[slot + mf_slot_step(ns, vv)] = v.x;
                    lim[slot + mf_slot_step(ns, vv)] = v.y;
                }
            }
            __syncthreads();
            if (!last_mf) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const float* src = (part ? lim : lre) + tl_line[q] * G::pitch;
                    if constexpr (TL % 32 == 0) {
                        // j + (rh+u)*TL: the padding of the multiple of TL is exact
                        const float* s0 = src + lds_pos<N>(tl_j[q]) + rh * (TL + TL / 32);
#pragma unroll
                        for (int u = 0; u < 8; u++) b[q][u] = s0[u * (TL + TL / 32)];
                    } else {
#pragma unroll
                        for (int u = 0; u < 8; u++) b[q][u] = src[lds_pos<N>(tl_j[q] + (rh + u) * TL)];
                    }
                }
            }
        }
        ns *= 16;
    }

    // ---- closing stage / staged output (the engine's layout)
    if constexpr (NMF < NS_ALL || COLMODE) {
        constexpr int R = RL == 16 ? 1 : RL;  // 1: no closing stage, only the store
        IO eio = io;  // the store undoes the line's scaling
        eio.scale = io.scale * mf_pow2(STEP - mf_line_shift(lmax[ew]));
        cpx<float> v[16];
        if constexpr (R == 1) {
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int k = et + TL * i;
                v[i] = cpx<float>{lre[ew * G::pitch + lds_pos<N>(k)], lim[ew * G::pitch + lds_pos<N>(k)]};
            }
            if (eact) {
#pragma unroll
                for (int i = 0; i < 16; i++) eio.store(ebase, et + TL * i, v[i]);
            }
        } else {
            constexpr int NSL = N / R;
            Eng::template get<R, 0>(lre, ew, et, v);
            Eng::template get<R, 1>(lim, ew, et, v);
            Eng::template twiddle<R, NSL, mf_twoff(NMF)>(tw, et, v);
            Eng::template butterflies<R>(v);
            if (eact) {
#pragma unroll
                for (int q = 0; q < 16 / R; q++)
#pragma unroll
                    for (int r = 0; r < R; r++) eio.store(ebase, et + TL * (q + r * (16 / R)), v[q * R + r]);
            }
        }
    }
}

}  // namespace jt
