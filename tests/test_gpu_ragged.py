"""Ragged batches (one persistent launch) against the oracle.
Tolerances as for single transforms: FP64 rel-l2 <= 1e-13*log2(n), FP32
<= 2e-7*log2(n) + 1e-6, per transform."""
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


DT = {"f64": np.float64, "f32": np.float32}
# every item kind: LDS singles, four-step (n > 4096 FP64 / 8192 FP32),
# Bluestein in LDS and Bluestein over a four-step convolution length
MIXED = [1, 2, 3, 7, 16, 17, 100, 211, 1000, 1024, 2048, 2049, 3000, 4096, 5000, 8192, 10007, 12800,
         65536, 65521, 4093, 240, 6561]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("graph", [False, True])
def test_ragged_mixed(orc, prec, graph):
    import torch
    lens = np.array(MIXED, dtype=np.int64)
    p = jt.RaggedBatch1D(lens, precision=prec, graph=graph)
    tot = int(2 * lens.sum())
    x = orc.jrandom_doubles(20110602, tot).astype(DT[prec])
    t = torch.from_numpy(x.copy()).cuda()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        for rep in range(2 if graph else 1):  # graph: capture, then replay
            t.copy_(torch.from_numpy(x))
            p.complexForward(t)
    s.synchronize()
    p.sync_status()
    y = t.cpu().numpy()
    for i, n in enumerate(lens):
        seg = slice(int(p.offsets[i]), int(p.offsets[i + 1]))
        e = rel(y[seg], orc.complex_forward(int(n), x[seg].astype(np.float64)))
        assert e <= tol(n, prec), (n, e)
    # scaled inverse back to the input (host path)
    a = y.copy()
    p.complexInverse(a, True)
    for i, n in enumerate(lens):
        seg = slice(int(p.offsets[i]), int(p.offsets[i + 1]))
        assert rel(a[seg], x[seg]) <= 2 * tol(n, prec), n


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_ragged_repeated_lengths(orc, prec):
    """Transforms of one length share work items (several lines per item):
    repeated in-LDS lengths, in-LDS Bluestein lengths (17, 211, 2049) and
    four-step lengths, interleaved, forward and scaled inverse."""
    import torch
    lens = np.array([17, 100, 211, 17, 12800, 2049, 100, 17, 211, 12800, 2049, 17, 100, 211, 17, 4096, 4096],
                    dtype=np.int64)
    p = jt.RaggedBatch1D(lens, precision=prec)
    x = orc.jrandom_doubles(20110604, int(2 * lens.sum())).astype(DT[prec])
    t = torch.from_numpy(x.copy()).cuda()
    p.complexForward(t)
    torch.cuda.synchronize()
    p.sync_status()
    y = t.cpu().numpy()
    for i, n in enumerate(lens):
        seg = slice(int(p.offsets[i]), int(p.offsets[i + 1]))
        e = rel(y[seg], orc.complex_forward(int(n), x[seg].astype(np.float64)))
        assert e <= tol(n, prec), (i, n, e)
    p.complexInverse(t, True)
    torch.cuda.synchronize()
    p.sync_status()
    z = t.cpu().numpy()
    for i, n in enumerate(lens):
        seg = slice(int(p.offsets[i]), int(p.offsets[i + 1]))
        assert rel(z[seg], x[seg]) <= 2 * tol(n, prec), (i, n)
    # unscaled inverse of the forward result: n times the input
    t.copy_(torch.from_numpy(y))
    p.complexInverse(t, False)
    torch.cuda.synchronize()
    p.sync_status()
    w = t.cpu().numpy()
    for i, n in enumerate(lens):
        seg = slice(int(p.offsets[i]), int(p.offsets[i + 1]))
        assert rel(w[seg], float(n) * x[seg].astype(np.float64)) <= 2 * tol(n, prec), (i, n)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("sched", ["lag=0", "lag=1,chunk=2", "lag=16,chunk=8", "wgs=1,nt=1024,lag=2",
                                   "order=1,tplx=1,lag=4"])
def test_ragged_schedules(orc, prec, sched, monkeypatch):
    """Explicit schedules (queue order by phase or transform by transform,
    item size, workgroup shape, radix order) give the same transforms."""
    import torch
    monkeypatch.setenv("JT_RAGGED_SCHEDULE", sched)
    lens = np.array(MIXED + [12800, 65536, 10007, 6561, 100, 17, 2049], dtype=np.int64)
    p = jt.RaggedBatch1D(lens, precision=prec)
    assert all(kv in p.describe() for kv in sched.split(",")), p.describe()
    x = orc.jrandom_doubles(20110605, int(2 * lens.sum())).astype(DT[prec])
    t = torch.from_numpy(x.copy()).cuda()
    p.complexForward(t)
    torch.cuda.synchronize()
    p.sync_status()
    y = t.cpu().numpy()
    for i, n in enumerate(lens):
        seg = slice(int(p.offsets[i]), int(p.offsets[i + 1]))
        e = rel(y[seg], orc.complex_forward(int(n), x[seg].astype(np.float64)))
        assert e <= tol(n, prec), (i, n, e)


def test_ragged_lengths_spec():
    L = ragged_lengths(20000)
    assert L.min() >= 16 and L.max() <= 65536
    assert len(L) == 20000


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_ragged_full_batch(orc, prec):
    """The 20000-transform batch: Parseval on every transform, 24 sampled
    transforms against the oracle."""
    import torch
    lens = ragged_lengths(20000)
    p = jt.RaggedBatch1D(lens, precision=prec)
    tot = int(2 * lens.sum())
    g = torch.Generator(device="cuda")
    g.manual_seed(20110602)
    dt = torch.float64 if prec == "f64" else torch.float32
    x = torch.rand(tot, dtype=dt, device="cuda", generator=g) * 2 - 1
    y = x.clone()
    p.complexForward(y)
    torch.cuda.synchronize()
    p.sync_status()
    xs, ys = x.cpu().numpy().astype(np.float64), y.cpu().numpy().astype(np.float64)
    starts = p.offsets[:-1]
    ex = np.add.reduceat(xs * xs, starts)
    ey = np.add.reduceat(ys * ys, starts)
    pars = np.abs(ey / (lens * ex) - 1.0)
    print(f"{prec}: Parseval max dev {pars.max():.3e} over {len(lens)} transforms", flush=True)
    assert pars.max() <= (1e-12 if prec == "f64" else 1e-5)
    rng = np.random.default_rng(3)
    pick = rng.choice(len(lens), 24, replace=False)
    worst = 0.0
    for i in pick:
        n = int(lens[i])
        seg = slice(int(p.offsets[i]), int(p.offsets[i + 1]))
        e = rel(ys[seg], orc.complex_forward(n, xs[seg]))
        worst = max(worst, e / tol(n, prec))
        assert e <= tol(n, prec), (n, e)
    print(f"{prec}: sampled worst error / tolerance = {worst:.3f}", flush=True)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_ragged_autotune(orc, prec):
    """The autotuned schedule gives the same transforms (to tolerance) and is
    remembered for the same length list."""
    import torch
    lens = np.array(MIXED, dtype=np.int64)
    p = jt.RaggedBatch1D(lens, precision=prec, autotune=True)
    d = p.describe()
    print(d, flush=True)
    assert "autotuned" in d and "schedule wgs=" in d
    tot = int(2 * lens.sum())
    x = orc.jrandom_doubles(20110625, tot).astype(DT[prec])
    t = torch.from_numpy(x.copy()).cuda()
    p.complexForward(t)
    torch.cuda.synchronize()
    p.sync_status()
    y = t.cpu().numpy()
    for i, n in enumerate(lens):
        seg = slice(int(p.offsets[i]), int(p.offsets[i + 1]))
        e = rel(y[seg], orc.complex_forward(int(n), x[seg].astype(np.float64)))
        assert e <= tol(n, prec), (n, e)
    q = jt.RaggedBatch1D(lens, precision=prec, autotune=True)
    assert "remembered" in q.describe()
    assert q.describe().split(" autotuned")[0] == d.split(" autotuned")[0]


def test_ragged_wait_timeout_is_reported(orc, monkeypatch):
    """Failure path of the bounded dependency waits, forced from the host:
    every consumer item waits for a count no producer reaches, gives up after
    1 ms, and the plan reports the failure, also when a clean launch ran
    after the failed one; once reported, the next launch is clean."""
    import torch
    lens = np.array([12800, 65536, 12800], dtype=np.int64)   # four-step items
    p = jt.RaggedBatch1D(lens, precision="f64")
    x = orc.jrandom_doubles(20110603, int(2 * lens.sum()))
    t = torch.from_numpy(x.copy()).cuda()
    monkeypatch.setenv("JT_RAGGED_FORCE_TIMEOUT", "1")
    p.complexForward(t)
    with pytest.raises(jt.JTransformsError, match="timed out"):
        p.sync_status()
    # the flag is sticky: a clean launch queued before the check does not
    # hide the earlier failure
    p.complexForward(t)
    monkeypatch.delenv("JT_RAGGED_FORCE_TIMEOUT")
    p.complexForward(t)
    with pytest.raises(jt.JTransformsError, match="timed out"):
        p.sync_status()
    t.copy_(torch.from_numpy(x))
    p.complexForward(t)
    p.sync_status()
    y = t.cpu().numpy()
    for i, n in enumerate(lens):
        seg = slice(int(p.offsets[i]), int(p.offsets[i + 1]))
        assert rel(y[seg], orc.complex_forward(int(n), x[seg])) <= tol(n, "f64")
