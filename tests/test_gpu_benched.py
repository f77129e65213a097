"""Parity of the exact launches the benchmark times, at their full sizes.

- The plain FP32 batch of 16384 x 4096-point transforms (512 MiB, so it runs
  the streaming-IO instantiation of the row kernel): Parseval on every line,
  64 sampled lines against the FP64 oracle, including the first and the last,
  and a scaled complexInverse round trip.  Reference bar for FloatFFT_1D:
  FloatFFT_1DTest.java:84 (RMSE 1e-4); here rel-l2 <= 2e-7*log2(n) + 1e-6.
- The 20000-transform ragged batch under the schedule the autotuner picks,
  captured in a hipGraph on a non-default stream and replayed: Parseval on
  every transform, one transform of every distinct length against the oracle
  (every four-step and Bluestein length among them) plus 200 sampled others.
  Reference: DoubleFFT_1D.java:117-188 (one plan per length).
"""
import math

import numpy as np
import pytest

import jtransforms_amd as jt
from jtransforms_amd.workloads import ragged_lengths

pytestmark = pytest.mark.gpu


def tol(n, prec):
    lg = max(1.0, math.log2(max(n, 2)))
    return 1e-13 * lg if prec == "f64" else 2e-7 * lg + 1e-6


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d else 1.0))


def test_plain_f32_batch_16384x4096(orc):
    import torch
    n, batch = 4096, 16384
    gen = torch.Generator(device="cuda")
    gen.manual_seed(20110602)
    x = torch.randn(2 * n * batch, dtype=torch.float32, device="cuda", generator=gen)
    p = jt.FFTBatch1D(n, batch, precision="f32")
    assert "split" not in p.describe()
    y = x.clone()
    p.complexForward(y)
    torch.cuda.synchronize()
    # Parseval on every line: sum |X_k|^2 = n sum |x_j|^2
    ex = (x.double().view(batch, 2 * n) ** 2).sum(dim=1)
    ey = (y.double().view(batch, 2 * n) ** 2).sum(dim=1)
    pars = float(((ey / (n * ex)) - 1.0).abs().max())
    print(f"16384x4096 FP32 plain: Parseval max dev {pars:.3e}", flush=True)
    assert pars <= 1e-5
    rng = np.random.default_rng(11)
    lines = sorted(set([0, batch - 1] + rng.choice(batch, 62, replace=False).tolist()))
    worst = 0.0
    for b in lines:
        seg = slice(2 * n * b, 2 * n * (b + 1))
        e = rel(y[seg].cpu().numpy(), orc.complex_forward(n, x[seg].cpu().numpy().astype(np.float64)))
        worst = max(worst, e)
    print(f"16384x4096 FP32 plain: {len(lines)} sampled lines, max rel-l2 {worst:.3e} "
          f"(tolerance {tol(n, 'f32'):.3e})", flush=True)
    assert worst <= tol(n, "f32")
    p.complexInverse(y, True)
    torch.cuda.synchronize()
    e2 = float(torch.linalg.vector_norm((y - x).double()) / torch.linalg.vector_norm(x.double()))
    print(f"16384x4096 FP32 plain: scaled round trip rel-l2 {e2:.3e}", flush=True)
    assert e2 <= 2 * tol(n, "f32")


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_ragged_full_batch_autotuned_graph(orc, prec):
    import torch
    lens = ragged_lengths(20000)
    p = jt.RaggedBatch1D(lens, precision=prec, graph=True, autotune=True)
    print(p.describe()[:300], flush=True)
    tot = int(2 * lens.sum())
    dt = torch.float64 if prec == "f64" else torch.float32
    g = torch.Generator(device="cuda")
    g.manual_seed(20110607)
    x = torch.rand(tot, dtype=dt, device="cuda", generator=g) * 2 - 1
    y = torch.empty_like(x)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for rep in range(3):  # capture, then two replays of the captured graph
            y.copy_(x)
            p.complexForward(y)
    s.synchronize()
    p.sync_status()
    xs, ys = x.cpu().numpy().astype(np.float64), y.cpu().numpy().astype(np.float64)
    starts = p.offsets[:-1]
    ex = np.add.reduceat(xs * xs, starts)
    ey = np.add.reduceat(ys * ys, starts)
    pars = np.abs(ey / (lens * ex) - 1.0)
    print(f"{prec}: Parseval max dev {pars.max():.3e} over {len(lens)} transforms", flush=True)
    assert pars.max() <= (1e-12 if prec == "f64" else 1e-5)
    _, first = np.unique(lens, return_index=True)
    rng = np.random.default_rng(5)
    pick = sorted(set(first.tolist()) | set(rng.choice(len(lens), 200, replace=False).tolist()))
    worst = 0.0
    for i in pick:
        n = int(lens[i])
        seg = slice(int(p.offsets[i]), int(p.offsets[i + 1]))
        e = rel(ys[seg], orc.complex_forward(n, xs[seg]))
        worst = max(worst, e / tol(n, prec))
        assert e <= tol(n, prec), (i, n, e)
    print(f"{prec}: {len(pick)} transforms ({len(first)} distinct lengths) against the oracle, "
          f"worst error / tolerance = {worst:.3f}", flush=True)
