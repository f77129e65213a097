"""GPU parity of 2D/3D FFTs and of DCT/DST/DHT against the CPU oracle.
Tolerances as in test_gpu_fft: FP64 rel-l2 <= 1e-13*log2(N), FP32 <=
2e-7*log2(N) + 1e-6, N the total number of points."""
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
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d else 1.0))


DT = {"f64": np.float64, "f32": np.float32}
FFT2 = {"f64": jt.DoubleFFT_2D, "f32": jt.FloatFFT_2D}
FFT3 = {"f64": jt.DoubleFFT_3D, "f32": jt.FloatFFT_3D}


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("rows,cols", [(2, 2), (16, 16), (6, 10), (64, 128), (100, 36), (512, 4096),
                                       (8192, 16), (24, 10158)])
def test_fft2d_complex(orc, rows, cols, prec):
    x = orc.jrandom_doubles(SEED, 2 * rows * cols, signed=False).astype(DT[prec])
    ref = orc.fftnd_complex([rows, cols], x.astype(np.float64), -1)
    p = FFT2[prec](rows, cols)
    a = x.copy()
    p.complexForward(a)
    e = rel(a, ref)
    print(f"{prec} 2D {rows}x{cols} complexForward rel={e:.3e}", flush=True)
    assert e <= tol(rows * cols, prec)
    p.complexInverse(a, True)
    assert rel(a, x) <= 2 * tol(rows * cols, prec)


@pytest.mark.parametrize("prec", ["f64", "f32"])
# every compile-time column length (the column-0 split runs inside the column
# kernel) and lengths on the other plans (2, 4: generic; 8192: four-step)
@pytest.mark.parametrize("rows,cols", [(2, 2), (4, 8), (16, 4), (32, 8), (64, 64), (128, 16), (256, 1024),
                                       (512, 8), (1024, 4), (2048, 16), (4096, 64), (8192, 4)])
def test_fft2d_real(orc, rows, cols, prec):
    x = orc.jrandom_doubles(SEED, rows * cols, signed=False).astype(DT[prec])
    ref = orc.fft2d_real_forward(rows, cols, x.astype(np.float64))
    p = FFT2[prec](rows, cols)
    a = x.copy()
    p.realForward(a)
    e = rel(a, ref)
    print(f"{prec} 2D {rows}x{cols} realForward rel={e:.3e}", flush=True)
    assert e <= tol(rows * cols, prec)
    b = a.copy()
    p.realInverse(b, True)
    assert rel(b, x) <= 2 * tol(rows * cols, prec)
    c = a.copy()
    p.realInverse(c, False)
    assert rel(c, orc.fft2d_real_inverse(rows, cols, a.astype(np.float64), False)) <= 2 * tol(rows * cols, prec)
    full = np.zeros(2 * rows * cols, DT[prec])
    full[:rows * cols] = x
    p.realForwardFull(full)
    assert rel(full, orc.fftnd_real_full([rows, cols], np.concatenate([x, np.zeros_like(x)]).astype(np.float64),
                                         -1)) <= tol(rows * cols, prec)


def test_fft2d_real_not_pow2():
    with pytest.raises(jt.IllegalArgumentException, match="rows and columns must be power of two numbers"):
        jt.DoubleFFT_2D(6, 8).realForward(np.zeros(48))


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("dims", [(2, 2, 2), (4, 8, 16), (5, 6, 7), (16, 32, 64), (64, 64, 64)])
def test_fft3d_complex(orc, dims, prec):
    tot = int(np.prod(dims))
    x = orc.jrandom_doubles(20110625, 2 * tot, signed=False).astype(DT[prec])
    ref = orc.fftnd_complex(list(dims), x.astype(np.float64), -1)
    p = FFT3[prec](*dims)
    a = x.copy()
    p.complexForward(a)
    e = rel(a, ref)
    print(f"{prec} 3D {dims} complexForward rel={e:.3e}", flush=True)
    assert e <= tol(tot, prec)
    p.complexInverse(a, True)
    assert rel(a, x) <= 2 * tol(tot, prec)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("dims", [(2, 2, 2), (4, 8, 16), (16, 4, 8), (32, 32, 32)])
def test_fft3d_real(orc, dims, prec):
    s, r, c = dims
    tot = s * r * c
    x = orc.jrandom_doubles(20110625, tot, signed=False).astype(DT[prec])
    ref = orc.fft3d_real_forward(s, r, c, x.astype(np.float64))
    p = FFT3[prec](*dims)
    a = x.copy()
    p.realForward(a)
    e = rel(a, ref)
    print(f"{prec} 3D {dims} realForward rel={e:.3e}", flush=True)
    assert e <= tol(tot, prec)
    b = a.copy()
    p.realInverse(b, True)
    assert rel(b, x) <= 2 * tol(tot, prec)
    # unscaled realInverse (DoubleFFT_3D.java:1629-1645) against the oracle,
    # whose unscaled factor is pinned by the reference's call sequence
    u = a.copy()
    p.realInverse(u, False)
    want = orc.fft3d_real_inverse(s, r, c, a.astype(np.float64), False)
    e = rel(u, want)
    print(f"{prec} 3D {dims} realInverse(scale=false) rel={e:.3e}", flush=True)
    assert e <= 2 * tol(tot, prec)


R2R = {
    ("dct", "f64"): [jt.DoubleDCT_1D, jt.DoubleDCT_2D, jt.DoubleDCT_3D],
    ("dct", "f32"): [jt.FloatDCT_1D, jt.FloatDCT_2D, jt.FloatDCT_3D],
    ("dst", "f64"): [jt.DoubleDST_1D, jt.DoubleDST_2D, jt.DoubleDST_3D],
    ("dst", "f32"): [jt.FloatDST_1D, jt.FloatDST_2D, jt.FloatDST_3D],
    ("dht", "f64"): [jt.DoubleDHT_1D, jt.DoubleDHT_2D, jt.DoubleDHT_3D],
    ("dht", "f32"): [jt.FloatDHT_1D, jt.FloatDHT_2D, jt.FloatDHT_3D],
}
R2R_DIMS = [(2,), (3,), (4,), (7,), (16,), (100,), (512,), (1024,), (4096,), (8192,), (10158,),
            (4, 8), (6, 10), (128, 64), (4, 8, 16), (5, 6, 7), (32, 32, 32)]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("kind", ["dct", "dst", "dht"])
@pytest.mark.parametrize("dims", R2R_DIMS)
def test_r2r(orc, kind, dims, prec):
    tot = int(np.prod(dims))
    x = orc.jrandom_doubles(SEED, tot).astype(DT[prec])
    cls = R2R[(kind, prec)][len(dims) - 1]
    p = cls(*dims)
    scales = [False] if kind == "dht" else [True, False]
    for scale in scales:
        a = x.copy()
        if kind == "dht":
            p.forward(a)
        else:
            p.forward(a, scale)
        ref = orc.realnd(kind, list(dims), x.astype(np.float64), True, scale)
        e = rel(a, ref)
        print(f"{prec} {kind} {dims} forward scale={scale} rel={e:.3e}", flush=True)
        assert e <= tol(tot, prec)
        b = x.copy()
        p.inverse(b, scale)
        refi = orc.realnd(kind, list(dims), x.astype(np.float64), False, scale)
        e = rel(b, refi)
        print(f"{prec} {kind} {dims} inverse scale={scale} rel={e:.3e}", flush=True)
        assert e <= tol(tot, prec)
    # round trip with scale=true (the reference's own test form)
    a = x.copy()
    if kind == "dht":
        p.forward(a)
    else:
        p.forward(a, True)
    p.inverse(a, True)
    assert rel(a, x) <= 2 * tol(tot, prec)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("dims", [(6, 10), (100, 36), (5, 6, 7), (3, 12, 9), (16, 20)])
def test_real_full_any_size(orc, dims, prec):
    """realForwardFull / realInverseFull on 2D/3D sizes that are not powers of
    two (the reference's mixed-radix full branch, DoubleFFT_2D.java:956-975):
    the full complex spectrum of the real input, and back."""
    tot = int(np.prod(dims))
    x = orc.jrandom_doubles(SEED + 7, tot, signed=False).astype(DT[prec])
    p = (FFT2 if len(dims) == 2 else FFT3)[prec](*dims)
    full = np.zeros(2 * tot, DT[prec])
    full[:tot] = x
    p.realForwardFull(full)
    src = np.concatenate([x, np.zeros_like(x)]).astype(np.float64)
    ref = orc.fftnd_real_full(list(dims), src, -1)
    e = rel(full, ref)
    print(f"{prec} {dims} realForwardFull rel={e:.3e}", flush=True)
    assert e <= tol(tot, prec)
    back = np.zeros(2 * tot, DT[prec])
    back[:tot] = x
    p.realInverseFull(back, True)
    e2 = rel(back, orc.fftnd_real_full(list(dims), src, 1, True))
    print(f"{prec} {dims} realInverseFull(scale) rel={e2:.3e}", flush=True)
    assert e2 <= tol(tot, prec)


# RealFFTUtils_2DTest / RealFFTUtils_3DTest (fft/RealFFTUtils_2DTest.java:142-166): every mode unpacked from
# the packed realForward output equals the realForwardFull output
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("dims", [(16, 32), (64, 8), (8, 16, 32), (32, 4, 16)])
def test_realfftutils_unpack_matches_full(orc, dims, prec):
    tot = int(np.prod(dims))
    x = orc.jrandom_doubles(SEED + tot, tot).astype(DT[prec])
    p = (FFT2 if len(dims) == 2 else FFT3)[prec](*dims)
    u = (jt.RealFFTUtils_2D if len(dims) == 2 else jt.RealFFTUtils_3D)(*dims)
    packed = x.copy()
    p.realForward(packed)
    full = np.zeros(2 * tot, DT[prec])
    full[:tot] = x
    p.realForwardFull(full)
    full = full.reshape(dims[:-1] + (2 * dims[-1],))
    got = np.zeros_like(full)
    for idx in np.ndindex(*full.shape):
        got[idx] = u.unpack(*idx, packed)
    e = rel(got, full)
    print(f"{prec} RealFFTUtils {dims} rel={e:.3e}", flush=True)
    assert e <= tol(tot, prec)
