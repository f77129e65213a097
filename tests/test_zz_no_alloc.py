"""jt_execute_noalloc against jt_execute, through the C ABI.

jt_execute_noalloc runs the same kernels with the op's workspace (and the
staging of a misaligned offset) in the library's static 64 MiB device array
instead of the plan's own buffer, so its results must equal jt_execute's
bit for bit; both are also checked against the oracle.  Two calls in a row,
the second on a smaller input, exercise the array's reuse (it keeps the last
call's contents).
"""
import ctypes
import math

import numpy as np
import pytest

import jtransforms_amd as jt

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def tol(n):
    return 1e-13 * max(1.0, math.log2(n))


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def run_both(plan, op, x, offa=0, scale=0):
    """(jt_execute result, jt_execute_noalloc result) on copies of x, as
    numpy arrays; both on the current stream."""
    import torch
    L = jt.lib()
    out = []
    for fn in (L.jt_execute, L.jt_execute_noalloc):
        t = torch.from_numpy(x.copy()).cuda()
        torch.cuda.synchronize()
        st = torch.cuda.current_stream().cuda_stream
        rc = fn(plan._h, op, ctypes.c_void_p(t.data_ptr()), offa, scale, ctypes.c_void_p(st))
        assert rc == jt.JT_OK, L.jt_last_error().decode()
        torch.cuda.synchronize()
        out.append(t.cpu().numpy())
    return out


def test_noalloc_two_calls_second_smaller(orc):
    # four-step, 65536 points: 1 MiB of workspace
    n1 = 65536
    p1 = jt.DoubleFFT_1D(n1)
    x1 = orc.jrandom_doubles(101, 2 * n1)
    a, b = run_both(p1, jt.COMPLEX_FORWARD, x1)
    print(f"n={n1}: noalloc == execute: {np.array_equal(a, b)}, "
          f"rel vs oracle {rel(b, orc.complex_forward(n1, x1)):.3e}", flush=True)
    assert np.array_equal(a, b)
    assert rel(b, orc.complex_forward(n1, x1)) <= tol(n1)
    # second call, smaller input: Bluestein, 10007 points
    n2 = 10007
    p2 = jt.DoubleFFT_1D(n2)
    x2 = orc.jrandom_doubles(102, 2 * n2)
    a, b = run_both(p2, jt.COMPLEX_INVERSE, x2, scale=1)
    print(f"n={n2}: noalloc == execute: {np.array_equal(a, b)}", flush=True)
    assert np.array_equal(a, b)
    assert rel(b, orc.complex_inverse(n2, x2, True)) <= tol(n2)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_misaligned_offset_and_full(orc, prec):
    """An odd offset stages the range in the scratch array; realForwardFull
    of a 2D plan expands the reals through it."""
    dt = np.float64 if prec == "f64" else np.float32
    n = 12800  # four-step
    p = (jt.DoubleFFT_1D if prec == "f64" else jt.FloatFFT_1D)(n)
    x = orc.jrandom_doubles(103, 2 * n + 3).astype(dt)
    a, b = run_both(p, jt.COMPLEX_FORWARD, x, offa=3)
    assert np.array_equal(a, b)
    assert np.array_equal(b[:3], x[:3])
    t = 1e-13 * 14 if prec == "f64" else 2e-7 * 14 + 1e-6
    assert rel(b[3:], orc.complex_forward(n, x[3:].astype(np.float64))) <= t
    rows, cols = 48, 40
    q = (jt.DoubleFFT_2D if prec == "f64" else jt.FloatFFT_2D)(rows, cols)
    y = np.zeros(2 * rows * cols, dt)
    y[:rows * cols] = orc.jrandom_doubles(104, rows * cols).astype(dt)
    a, b = run_both(q, jt.REAL_FORWARD_FULL, y)
    assert np.array_equal(a, b)
    want = orc.fftnd_real_full([rows, cols], y.astype(np.float64), -1)
    assert rel(b, want) <= (1e-13 * 11 if prec == "f64" else 2e-7 * 11 + 1e-6)


def test_noalloc_rejects_oversized_and_ragged():
    import torch
    L = jt.lib()
    n = 1 << 23  # four-step workspace of n complex doubles: 128 MiB > 64 MiB
    p = jt.DoubleFFT_1D(n)
    t = torch.zeros(2 * n, dtype=torch.float64, device="cuda")
    rc = L.jt_execute_noalloc(p._h, jt.COMPLEX_FORWARD, ctypes.c_void_p(t.data_ptr()), 0, 0, None)
    assert rc == jt.JT_TOO_BIG
    assert b"64 MiB" in L.jt_last_error() or str(SCRATCH_BYTES).encode() in L.jt_last_error()
    r = jt.RaggedBatch1D(np.array([16, 32], dtype=np.int64))
    u = torch.zeros(96, dtype=torch.float64, device="cuda")
    rc = L.jt_execute_noalloc(r._h, jt.COMPLEX_FORWARD, ctypes.c_void_p(u.data_ptr()), 0, 0, None)
    assert rc == jt.JT_UNSUPPORTED
    torch.cuda.synchronize()
