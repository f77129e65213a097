"""Parity at the benchmark configurations' full sizes.

C1  DoubleFFT_1D(1024) on the reference test's own inputs
    (java.util.Random(20110602), 2*nextDouble()-1): complexForward,
    complexInverse(scale), realForward against the oracle.
C2  DoubleFFT_2D(4096, 4096) complexForward and realForward against the
    oracle at full size.
C5  DoubleDCT_3D / DoubleDST_3D 512^3 (scale=true) and FloatFFT_3D
    512x512x256 at full size through separable inputs: the transform of
    f(i) g(j) h(k) is the outer product of the 1D transforms, which the oracle
    computes exactly; plus the orthonormal transforms' norm preservation.
Tolerances: FP64 rel-l2 <= 1e-13*log2(N); FP32 2e-7*log2(N) + 1e-6.
"""
import math

import numpy as np
import pytest

import jtransforms_amd as jt

pytestmark = pytest.mark.gpu

SEED = 20110602


def tol(n, prec):
    lg = max(1.0, math.log2(max(n, 2)))
    return 1e-13 * lg if prec == "f64" else 2e-7 * lg + 1e-6


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_c1_fft1d_1024_reference_inputs(orc):
    n = 1024
    x = orc.jrandom_doubles(SEED, 2 * n)            # DoubleFFT_1DTest.java:248
    fft = jt.DoubleFFT_1D(n)
    a = x.copy()
    fft.complexForward(a)
    assert rel(a, orc.complex_forward(n, x)) <= tol(n, "f64")
    fft.complexInverse(a, True)
    assert np.sqrt(np.mean((a - x) ** 2)) <= 1e-12   # DoubleFFT_1DTest.java:242-255
    r = x[:n].copy()
    fft.realForward(r)
    assert rel(r, orc.real_forward(n, x[:n])) <= tol(n, "f64")


def test_c2_fft2d_4096_full(orc):
    import torch
    R = C = 4096
    x = orc.jrandom_doubles(SEED, 2 * R * C)
    t = torch.from_numpy(x.copy()).cuda()
    jt.DoubleFFT_2D(R, C).complexForward(t)
    torch.cuda.synchronize()
    e = rel(t.cpu().numpy(), orc.fftnd_complex([R, C], x, -1))
    print(f"4096x4096 complexForward rel={e:.3e}", flush=True)
    assert e <= tol(R * C, "f64")
    y = orc.jrandom_doubles(SEED + 1, R * C)
    u = torch.from_numpy(y.copy()).cuda()
    jt.DoubleFFT_2D(R, C).realForward(u)
    torch.cuda.synchronize()
    e = rel(u.cpu().numpy(), orc.fft2d_real_forward(R, C, y))
    print(f"4096x4096 realForward rel={e:.3e}", flush=True)
    assert e <= tol(R * C, "f64")


@pytest.mark.parametrize("kind", ["dct", "dst"])
def test_c5_r2r_3d_512_separable(orc, kind):
    import torch
    n = 512
    f, g, h = (orc.jrandom_doubles(SEED + k, n) for k in range(3))
    x = np.einsum("i,j,k->ijk", f, g, h)
    cls = jt.DoubleDCT_3D if kind == "dct" else jt.DoubleDST_3D
    t = torch.from_numpy(x.ravel().copy()).cuda()
    cls(n, n, n).forward(t, True)
    torch.cuda.synchronize()
    one = orc.dct_forward if kind == "dct" else orc.dst_forward
    F, G, H = one(n, f, True), one(n, g, True), one(n, h, True)
    ref = np.einsum("i,j,k->ijk", F, G, H)
    y = t.cpu().numpy()
    e = rel(y, ref)
    print(f"{kind} 512^3 separable rel={e:.3e}", flush=True)
    assert e <= tol(n ** 3, "f64")
    # orthonormal: the norm is preserved for any input
    z = torch.rand(n ** 3, dtype=torch.float64, device="cuda") * 2 - 1
    nz = float(torch.linalg.vector_norm(z))
    cls(n, n, n).forward(z, True)
    assert abs(float(torch.linalg.vector_norm(z)) / nz - 1) <= 1e-13


@pytest.mark.parametrize("split", [False, True])
def test_c5_floatfft3d_512x512x256_separable(orc, split):
    import torch
    S, R, C = 512, 512, 256
    f = orc.jrandom_doubles(SEED, 2 * S)
    g = orc.jrandom_doubles(SEED + 1, 2 * R)
    h = orc.jrandom_doubles(SEED + 2, 2 * C)
    fc, gc, hc = (v[0::2] + 1j * v[1::2] for v in (f, g, h))
    xc = np.einsum("i,j,k->ijk", fc, gc, hc).astype(np.complex64)
    x = np.empty(2 * xc.size, np.float32)
    x[0::2], x[1::2] = xc.real.ravel(), xc.imag.ravel()
    t = torch.from_numpy(x).cuda()
    p = jt.FloatFFT_3D(S, R, C)
    if split:
        p.set_split_fp16(True)
    p.complexForward(t)
    torch.cuda.synchronize()
    # the oracle on the float-cast factors (the separable input is their
    # product in float, so compare against the exact transform of that)
    def c2n(a):
        return a[0::2] + 1j * a[1::2]

    def n2c(z):
        out = np.empty(2 * z.size)
        out[0::2], out[1::2] = z.real, z.imag
        return out

    Fs = c2n(orc.complex_forward(S, n2c(fc)))
    Gs = c2n(orc.complex_forward(R, n2c(gc)))
    Hs = c2n(orc.complex_forward(C, n2c(hc)))
    ref = np.einsum("i,j,k->ijk", Fs, Gs, Hs)
    y = t.cpu().numpy().astype(np.float64)
    yc = y[0::2] + 1j * y[1::2]
    e = float(np.linalg.norm(yc - ref.ravel()) / np.linalg.norm(ref))
    print(f"FloatFFT_3D 512x512x256 split={split} separable rel={e:.3e}", flush=True)
    # the input itself carries float rounding of the product (<= 2^-24 rel)
    assert e <= (1e-6 if split else tol(S * R * C, "f32"))
