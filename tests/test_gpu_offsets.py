"""The reference's offa overloads (e.g. DoubleFFT_1D.complexForward(double[],
int), fft/DoubleFFT_1D.java:243-263): every method family on a sub-range of
a larger array, from host arrays (jt_execute_host) and device tensors
(jt_execute), at even and odd offsets.  An odd offset leaves the device range
misaligned for complex pairs; the library then runs on an aligned copy.  The
elements outside the range must stay untouched."""
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


# (name, plan factory, call(plan, array, offa), elements touched, oracle(x) on those elements)
CASES = [
    ("complexForward 1D", lambda P: P["fft1"](1000), lambda p, a, o: p.complexForward(a, o), 2000,
     lambda orc, x: orc.complex_forward(1000, x)),
    ("complexInverse 1D scaled", lambda P: P["fft1"](1000), lambda p, a, o: p.complexInverse(a, o, True), 2000,
     lambda orc, x: orc.complex_inverse(1000, x, True)),
    ("realForward 1D", lambda P: P["fft1"](1024), lambda p, a, o: p.realForward(a, o), 1024,
     lambda orc, x: orc.real_forward(1024, x)),
    ("realInverse 1D scaled", lambda P: P["fft1"](1024), lambda p, a, o: p.realInverse(a, o, True), 1024,
     lambda orc, x: orc.real_inverse(1024, x, True)),
    ("realForwardFull 1D", lambda P: P["fft1"](600), lambda p, a, o: p.realForwardFull(a, o), 1200,
     lambda orc, x: orc.real_forward_full(600, x)),
    ("complexForward 2D", lambda P: P["fft2"](16, 24), lambda p, a, o: p.complexForward(a, o), 768,
     lambda orc, x: orc.fftnd_complex([16, 24], x, -1)),
    ("realForward 2D", lambda P: P["fft2"](16, 32), lambda p, a, o: p.realForward(a, o), 512,
     lambda orc, x: orc.fft2d_real_forward(16, 32, x)),
    ("DCT forward 1D scaled", lambda P: P["dct1"](100), lambda p, a, o: p.forward(a, o, True), 100,
     lambda orc, x: orc.dct_forward(100, x, True)),
    ("DST inverse 1D scaled", lambda P: P["dst1"](64), lambda p, a, o: p.inverse(a, o, True), 64,
     lambda orc, x: orc.dst_inverse(64, x, True)),
    ("DHT forward 2D", lambda P: P["dht2"](8, 12), lambda p, a, o: p.forward(a, o), 96,
     lambda orc, x: orc.realnd("dht", [8, 12], x, True, False)),
]

PLANS = {
    "f64": {"fft1": jt.DoubleFFT_1D, "fft2": jt.DoubleFFT_2D, "dct1": jt.DoubleDCT_1D, "dst1": jt.DoubleDST_1D,
            "dht2": jt.DoubleDHT_2D},
    "f32": {"fft1": jt.FloatFFT_1D, "fft2": jt.FloatFFT_2D, "dct1": jt.FloatDCT_1D, "dst1": jt.FloatDST_1D,
            "dht2": jt.FloatDHT_2D},
}
DT = {"f64": np.float64, "f32": np.float32}


@pytest.mark.parametrize("device", [False, True])
@pytest.mark.parametrize("offa", [0, 3, 4])
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_offset(orc, case, prec, offa, device):
    name, make, call, ne, ref = case
    tail = 5
    x = orc.jrandom_doubles(SEED + offa, offa + ne + tail).astype(DT[prec])
    p = make(PLANS[prec])
    if device:
        import torch
        t = torch.from_numpy(x.copy()).cuda()
        call(p, t, offa)
        torch.cuda.synchronize()
        y = t.cpu().numpy()
    else:
        y = x.copy()
        call(p, y, offa)
    assert np.array_equal(y[:offa], x[:offa]), "prefix changed"
    assert np.array_equal(y[offa + ne:], x[offa + ne:]), "suffix changed"
    e = rel(y[offa:offa + ne], ref(orc, x[offa:offa + ne].astype(np.float64)))
    print(f"{name} {prec} offa={offa} device={device} rel={e:.3e}", flush=True)
    assert e <= 2 * tol(ne, prec)
