"""Arrays beyond the 2^31-element limit of a Java array: the reference serves
them through its DoubleLargeArray overloads (e.g. fft/DoubleFFT_1D.java:
280-304, long offa).  Here every index is 64-bit and a device buffer of any
size is one pointer, so the same entry points cover them.

The checks are size-independent properties that pin the indexing:
* a unit impulse at position p transforms to exp(-2 pi i p k / n); sampled
  bins are compared with that closed form (exact integer phase p*k mod n);
* the scaled inverse returns the impulse, checked at every one of the n
  points on the device;
* a batch of 4096-point transforms longer than 2^31 complex points: lines
  on both sides of the 2^31 boundary against the oracle.
"""
import math

import numpy as np
import pytest

import jtransforms_amd as jt

pytestmark = pytest.mark.gpu


def _need_gib(gib):
    import torch
    free, _ = torch.cuda.mem_get_info()
    if free < gib * (1 << 30):
        pytest.skip(f"needs {gib} GiB of free device memory")


def test_1d_impulse_beyond_2_31_points():
    import torch
    n = 1 << 31  # 2^31 complex points: 2^32 doubles, 32 GiB
    # data 32 GiB; the plan's buffer holds the largest op's workspace plus its
    # staging area: realForwardFull, 96 + 32 GiB
    _need_gib(168)
    p = n - 123457
    plan = jt.DoubleFFT_1D(n)
    print(f"DoubleFFT_1D({n}) plan: {plan.describe()}", flush=True)
    a = torch.zeros(2 * n, dtype=torch.float64, device="cuda")
    a[2 * p] = 1.0
    plan.complexForward(a)
    torch.cuda.synchronize()
    rng = np.random.Generator(np.random.PCG64(5))
    ks = np.unique(np.concatenate([[0, 1, 2, n // 2, n - 1, (1 << 30) + 7],
                                   rng.integers(0, n, 2000)])).astype(np.int64)
    got = a.view(-1, 2)[torch.from_numpy(ks).cuda()].cpu().numpy()
    ph = np.array([(p * int(k)) % n for k in ks], dtype=np.float64) * (2 * math.pi / n)
    want = np.stack([np.cos(ph), -np.sin(ph)], axis=1)
    err = float(np.abs(got - want).max())
    print(f"impulse at {p}: max |X_k - exp(-2 pi i p k/n)| over {len(ks)} bins = {err:.3e}", flush=True)
    assert err <= 1e-13 * 31
    plan.complexInverse(a, True)
    a[2 * p] -= 1.0
    chunk = 1 << 28
    resid = max(float(a[i:i + chunk].abs().max()) for i in range(0, 2 * n, chunk))
    print(f"scaled inverse: max |x - impulse| over all {n} points = {resid:.3e}", flush=True)
    assert resid <= 1e-13 * 31
    del a
    torch.cuda.empty_cache()


def test_batch_beyond_2_31_points(orc):
    import torch
    n, lines = 4096, (1 << 19) + 8  # 2^31 + 32768 complex points, 32 GiB
    _need_gib(72)  # data 32 GiB + the plan's staging area 32 GiB
    plan = jt.FFTBatch1D(n, lines, precision="f64")
    x = torch.empty(2 * n * lines, dtype=torch.float64, device="cuda")
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    x.uniform_(-1.0, 1.0, generator=g)
    check = [0, 1, (1 << 19) - 1, 1 << 19, lines - 1]  # lines 2^19.. start beyond 2^31 points
    before = {l: x[2 * n * l:2 * n * (l + 1)].cpu().numpy() for l in check}
    plan.complexForward(x)
    torch.cuda.synchronize()
    worst = 0.0
    for l in check:
        got = x[2 * n * l:2 * n * (l + 1)].cpu().numpy()
        want = orc.complex_forward(n, before[l])
        e = float(np.linalg.norm(got - want) / np.linalg.norm(want))
        worst = max(worst, e)
    print(f"FFTBatch1D({n}, {lines}): worst line rel error {worst:.3e}", flush=True)
    assert worst <= 1e-13 * 12
    del x
    torch.cuda.empty_cache()
