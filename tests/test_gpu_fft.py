"""GPU parity of the HIP FFT path against the CPU oracle and the reference's
FFTW golden vectors.

Tolerances: FP64 rel-l2 <= 1e-13 * log2(n) against the oracle (and the
reference's own RMSE bar 1e-12 against the FFTW fixtures); FP32 rel-l2 <=
2e-7 * log2(n) + 1e-6 against the FP64 oracle on the same (float-cast) input.
"""
import math
import os

import numpy as np
import pytest

import jtransforms_amd as jt
from tests import jt_layout as L

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


FFT1 = {"f64": jt.DoubleFFT_1D, "f32": jt.FloatFFT_1D}
DT = {"f64": np.float64, "f32": np.float32}

FIXTURE_SIZES = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 16, 32, 64, 100, 120, 128, 256, 310,
                 512, 1024, 1056, 2048, 8192, 10158]


@pytest.mark.parametrize("n", FIXTURE_SIZES)
def test_complex_forward_fftw_fixture(n):
    x = L.read_fftw(n, "in")
    y = L.read_fftw(n, "out")
    a = x.copy()
    jt.DoubleFFT_1D(n).complexForward(a)
    print(f"n={n} rmse={np.sqrt(np.mean((a - y) ** 2)):.3e}", flush=True)
    assert np.sqrt(np.mean((a - y) ** 2)) <= 1e-12      # DoubleFFT_1DTest.java:85
    assert rel(a, y) <= tol(n, "f64")


# sizes covering every plan kind: compile-time (pow2 <= 4096), generic
# (smooth), four-step (smooth > LDS limit), Bluestein (large prime factor)
C2C_SIZES = [2, 3, 16, 17, 64, 100, 240, 512, 1000, 1024, 2048, 4096, 4725, 8192, 10158, 12800,
             16384, 65536, 65530]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("n", C2C_SIZES)
def test_complex_roundtrip_vs_oracle(orc, n, prec):
    x = orc.jrandom_doubles(SEED + n, 2 * n).astype(DT[prec])
    ref = orc.complex_forward(n, x.astype(np.float64))
    fft = FFT1[prec](n)
    a = x.copy()
    fft.complexForward(a)
    e = rel(a, ref)
    print(f"{prec} n={n} fwd rel={e:.3e}", flush=True)
    assert e <= tol(n, prec)
    b = a.copy()
    fft.complexInverse(b, True)
    assert rel(b, x) <= 2 * tol(n, prec)
    c = x.copy()
    fft.complexInverse(c, False)
    assert rel(c, orc.complex_inverse(n, x.astype(np.float64), False)) <= tol(n, prec)


REAL_SIZES = [2, 3, 4, 5, 8, 9, 16, 30, 64, 100, 128, 512, 1024, 2048, 4096, 6000, 8192, 10158,
              16384, 131072]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("n", REAL_SIZES)
def test_real_forward_inverse_vs_oracle(orc, n, prec):
    x = orc.jrandom_doubles(SEED, n).astype(DT[prec])
    fft = FFT1[prec](n)
    a = x.copy()
    fft.realForward(a)
    ref = orc.real_forward(n, x.astype(np.float64))
    e = rel(a, ref)
    print(f"{prec} n={n} realForward rel={e:.3e}", flush=True)
    assert e <= tol(n, prec)
    b = a.copy()
    fft.realInverse(b, True)
    assert rel(b, x) <= 2 * tol(n, prec)
    c = a.copy()
    fft.realInverse(c, False)
    # unscaled factor n/2 (power of two) or n (DoubleFFT_1DTest.java:583-588)
    assert rel(c, orc.real_inverse(n, a.astype(np.float64), False)) <= 2 * tol(n, prec)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("n", [1, 2, 5, 16, 100, 1024, 4097])
def test_real_full_vs_oracle(orc, n, prec):
    x = np.zeros(2 * n, DT[prec])
    x[:n] = orc.jrandom_doubles(SEED, n)
    fft = FFT1[prec](n)
    a = x.copy()
    fft.realForwardFull(a)
    assert rel(a, orc.real_forward_full(n, x.astype(np.float64))) <= tol(n, prec)
    b = x.copy()
    fft.realInverseFull(b, True)
    assert rel(b, orc.real_inverse_full(n, x.astype(np.float64), True)) <= tol(n, prec)


def test_offset_argument(orc):
    n = 1024
    x = orc.jrandom_doubles(SEED, 2 * n + 10)
    a = x.copy()
    jt.DoubleFFT_1D(n).complexForward(a, 10)
    assert np.array_equal(a[:10], x[:10])
    assert rel(a[10:], orc.complex_forward(n, x[10:])) <= tol(n, "f64")


def test_device_tensor_path(orc):
    import torch
    n = 4096
    x = orc.jrandom_doubles(SEED, 2 * n)
    t = torch.from_numpy(x.copy()).cuda()
    jt.DoubleFFT_1D(n).complexForward(t)
    torch.cuda.synchronize()
    assert rel(t.cpu().numpy(), orc.complex_forward(n, x)) <= tol(n, "f64")


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("n,batch", [(4096, 64), (1024, 33), (240, 7), (65536, 3), (10158, 2)])
def test_batch_plan(orc, n, batch, prec):
    x = orc.jrandom_doubles(SEED, 2 * n * batch).astype(DT[prec])
    p = jt.FFTBatch1D(n, batch, precision=prec)
    a = x.copy()
    p.complexForward(a)
    for b in (0, batch - 1):
        seg = slice(2 * n * b, 2 * n * (b + 1))
        assert rel(a[seg], orc.complex_forward(n, x[seg].astype(np.float64))) <= tol(n, prec)


@pytest.mark.parametrize("n", [1 << 25, 3 << 22, 1000003])
def test_large_1d_f64_vs_oracle(orc, n):
    """Lengths beyond one four-step level (2^25 = 4096 x 8192, the 8192-point
    rows a four-step plan of their own), a large balanced four-step, and a
    large Bluestein length, against the oracle on the reference test's inputs."""
    import torch
    x = orc.jrandom_doubles(SEED, 2 * n)
    t = torch.from_numpy(x).cuda()
    jt.DoubleFFT_1D(n).complexForward(t)
    torch.cuda.synchronize()
    e = rel(t.cpu().numpy(), orc.complex_forward(n, x))
    print(f"f64 n={n} complexForward rel={e:.3e}", flush=True)
    assert e <= tol(n, "f64")


def test_large_1d_f32_properties():
    """FloatFFT_1D(2^27) (two four-step levels at FP32's 8192-point limit):
    Parseval and the scaled round trip, properties that hold at any size."""
    import torch
    n = 1 << 27
    g = torch.Generator(device="cuda")
    g.manual_seed(SEED)
    x = torch.randn(2 * n, dtype=torch.float32, device="cuda", generator=g)
    a = x.clone()
    p = jt.FloatFFT_1D(n)
    p.complexForward(a)
    ex = float(torch.sum(x.double() ** 2))
    ea = float(torch.sum(a.double() ** 2))
    par = abs(ea / (n * ex) - 1.0)
    print(f"f32 n=2^27 Parseval deviation {par:.3e}", flush=True)
    assert par <= 1e-5
    p.complexInverse(a, True)
    e = float(torch.linalg.vector_norm((a - x).double()) / torch.linalg.vector_norm(x.double()))
    print(f"f32 n=2^27 round trip rel={e:.3e}", flush=True)
    assert e <= 2 * tol(n, "f32")
