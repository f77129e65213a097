"""Split-FP16 MFMA form of the radix-16 stages against the FP64 oracle.

Bar (from the precision form's definition): rel-l2 <= 1e-6 against the FP64
oracle on the same (float-cast) input, forward and inverse; the plain FP32
path is checked on the same data for comparison.
"""
import numpy as np
import pytest

import jtransforms_amd as jt

pytestmark = pytest.mark.gpu

SPLIT_TOL = 1e-6


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def std_normal(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.standard_normal(n).astype(np.float32)


@pytest.mark.parametrize("n,batch", [(16, 64), (256, 32), (1024, 8), (2048, 8), (4096, 8)])
def test_split_fp16_batch_vs_oracle(orc, n, batch):
    x = std_normal(2 * n * batch, 20110602 + n)
    split = jt.FFTBatch1D(n, batch, precision="f32", split_fp16=True)
    plain = jt.FFTBatch1D(n, batch, precision="f32")
    a = x.copy()
    split.complexForward(a)
    p = x.copy()
    plain.complexForward(p)
    errs, perrs = [], []
    for b in range(batch):
        seg = slice(2 * n * b, 2 * n * (b + 1))
        ref = orc.complex_forward(n, x[seg].astype(np.float64))
        errs.append(rel(a[seg], ref))
        perrs.append(rel(p[seg], ref))
    e, pe = max(errs), max(perrs)
    print(f"n={n} split-FP16 fwd rel={e:.3e} (plain FP32 {pe:.3e})", flush=True)
    assert e <= SPLIT_TOL
    # inverse (scaled) back to the input
    split.complexInverse(a, True)
    e2 = rel(a, x)
    print(f"n={n} split-FP16 round trip rel={e2:.3e}", flush=True)
    assert e2 <= 2 * SPLIT_TOL


def test_split_fp16_full_batch_sampled(orc):
    """The 16384 x 4096 FP32 batch: every line transformed on the GPU, 64
    sampled lines checked against the FP64 oracle."""
    import torch
    n, batch = 4096, 16384
    gen = torch.Generator(device="cuda")
    gen.manual_seed(20110602)
    x = torch.randn(2 * n * batch, dtype=torch.float32, device="cuda", generator=gen)
    a = x.clone()
    p = jt.FFTBatch1D(n, batch, precision="f32", split_fp16=True)
    p.complexForward(a)
    torch.cuda.synchronize()
    rng = np.random.default_rng(7)
    lines = sorted(rng.choice(batch, 64, replace=False).tolist())
    errs = []
    for b in lines:
        seg = slice(2 * n * b, 2 * n * (b + 1))
        xi = x[seg].cpu().numpy().astype(np.float64)
        errs.append(rel(a[seg].cpu().numpy(), orc.complex_forward(n, xi)))
    print(f"16384x4096 split-FP16 sampled max rel={max(errs):.3e} mean={np.mean(errs):.3e}", flush=True)
    assert max(errs) <= SPLIT_TOL


@pytest.mark.parametrize("dims", [(16, 32, 64), (8, 256, 16)])
def test_split_fp16_fft3d(orc, dims):
    tot = int(np.prod(dims))
    x = std_normal(2 * tot, 20110625)
    p = jt.FloatFFT_3D(*dims)
    p.set_split_fp16(True)
    a = x.copy()
    p.complexForward(a)
    ref = orc.fftnd_complex(list(dims), x.astype(np.float64), -1)
    e = rel(a, ref)
    print(f"FloatFFT_3D{dims} split-FP16 rel={e:.3e}", flush=True)
    assert e <= SPLIT_TOL


def test_split_fp16_fft2d_real(orc):
    # the column pass runs on the matrix-core kernel; the column-0 split
    # then runs as its own launch
    rows, cols = 1024, 64
    x = std_normal(rows * cols, 20110626)
    p = jt.FloatFFT_2D(rows, cols)
    p.set_split_fp16(True)
    a = x.copy()
    p.realForward(a)
    e = rel(a, orc.fft2d_real_forward(rows, cols, x.astype(np.float64)))
    print(f"FloatFFT_2D({rows},{cols}) realForward split-FP16 rel={e:.3e}", flush=True)
    assert e <= SPLIT_TOL
