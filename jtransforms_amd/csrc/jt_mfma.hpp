// Radix-16 butterflies as a dense DFT-matrix product on MFMA (gfx950), in
// split FP16 with FP32 accumulation.
//
// A wave processes 64 butterflies (16 complex inputs each, twiddles already
// applied) as 4 tiles of 16 butterflies:
//     X = F16 * V,  F16[i][k] = exp(DIR 2 pi i ik/16)
// written over the reals as two 16x32 by 32x16 products:
//     Xr = [Fr | -Fi] [Vr; Vi],   Xi = [Fi | Fr] [Vr; Vi]
// with v_mfma_f32_16x16x32_f16.  Each operand is split as x = hi + lo with
// hi = x rounded toward zero to fp16's 11 significant bits (a bit mask on
// the float) and lo = fp16(x - hi); the product is the sum of the three
// terms hi*hi + hi*lo + lo*hi accumulated in FP32 (the lo*lo term, ~2^-22
// relative, is dropped).
//
// Range: the caller keeps every B entry below 2^14 in magnitude (block
// floating point: one power-of-two scale per line chosen from the line's
// input, then a fixed 2^-4 per radix-16 stage, folded into the twiddles, so
// a stage's outputs, at most 16 times its inputs, stay below 2^14 too).  All
// scalings are exact powers of two and are undone once at the store.
//
// Operand layouts of the 16x16x32 f16 MFMA (checked on the GPU): lane l
// holds A[i][k] with i = l%16, k = 8*(l/16)+v; B[k][j] with j = l%16,
// k = 8*(l/16)+v; D[i][j] with j = l%16, i = 4*(l/16)+v.
#pragma once
#include "jt_dft.hpp"

namespace jt {

// Power-of-two exponent sh that brings a line whose largest |entry| is m
// into [2^13, 2^14): m = f 2^e with f in [1, 2) gives sh = 13 - e, read from
// the float's exponent bits; clamped so that 2^sh and the store's undo factor
// 2^(4 (stages - 1) - sh) are normal floats.  m = 0 (or a subnormal m) gives
// the clamp.
__device__ __forceinline__ int mf_line_shift(unsigned mbits) {
    const int e = (int)((mbits >> 23) & 0xFFu) - 127;
    int sh = 13 - e;
    sh = sh < -100 ? -100 : sh;
    return sh > 100 ? 100 : sh;
}
__device__ __forceinline__ float mf_pow2(int k) { return __uint_as_float((unsigned)(k + 127) << 23); }

typedef _Float16 jt_half8 __attribute__((ext_vector_type(8)));
typedef float jt_float4 __attribute__((ext_vector_type(4)));

struct MfmaDft16 {
    jt_half8 arh, arl, aih, ail;  // A operands of Xr and Xi, hi/lo

    template <int DIR>
    __device__ __forceinline__ void init(int lane) {
        // exact cos/sin(2 pi m/16), m = 0..15
        const float c16[16] = {1.0f, 0.92387953251128675613f, 0.70710678118654752440f,
                               0.38268343236508977173f, 0.0f, -0.38268343236508977173f,
                               -0.70710678118654752440f, -0.92387953251128675613f, -1.0f,
                               -0.92387953251128675613f, -0.70710678118654752440f,
                               -0.38268343236508977173f, 0.0f, 0.38268343236508977173f,
                               0.70710678118654752440f, 0.92387953251128675613f};
        const int i = lane % 16;
#pragma unroll
        for (int v = 0; v < 8; v++) {
            int k = 8 * (lane / 16) + v;
            int kk = k & 15;
            int m = (i * kk) & 15;
            float fr = c16[m];
            float fi = float(DIR) * c16[(m + 12) & 15];  // sin(x) = cos(x - pi/2)
            float ar = k < 16 ? fr : -fi;
            float ai = k < 16 ? fi : fr;
            _Float16 h;
            h = (_Float16)ar;
            arh[v] = h;
            arl[v] = (_Float16)(ar - (float)h);
            h = (_Float16)ai;
            aih[v] = h;
            ail[v] = (_Float16)(ai - (float)h);
        }
    }

    // One 16-butterfly tile.  b: this lane's 8 entries of the B operand
    // (column lane%16, rows 8*(lane/16)+u: Re of points 8*((lane/16)&1)+u for
    // lane/16 < 2, Im of them otherwise), |b| < 2^14.  Returns this lane's 4
    // outputs i = 4*(lane/16)+u of butterfly lane%16, real parts in outr,
    // imaginary parts in outi, in FP32.
    __device__ __forceinline__ void tile(const float* b, float* outr, float* outi) const {
        jt_half8 bh, bl;
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const float h = __uint_as_float(__float_as_uint(b[u]) & 0xFFFFE000u);  // exact in fp16
            bh[u] = (_Float16)h;
            bl[u] = (_Float16)(b[u] - h);
        }
        jt_float4 dr = {0.f, 0.f, 0.f, 0.f}, di = {0.f, 0.f, 0.f, 0.f};
        dr = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, bh, dr, 0, 0, 0);
        di = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih, bh, di, 0, 0, 0);
        dr = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, bl, dr, 0, 0, 0);
        di = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih, bl, di, 0, 0, 0);
        dr = __builtin_amdgcn_mfma_f32_16x16x32_f16(arl, bh, dr, 0, 0, 0);
        di = __builtin_amdgcn_mfma_f32_16x16x32_f16(ail, bh, di, 0, 0, 0);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            outr[u] = dr[u];
            outi[u] = di[u];
        }
    }
};

}  // namespace jt
