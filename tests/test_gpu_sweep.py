"""Seeded sweep of 1D lengths against the oracle: every n <= 64 and 120 lengths
log-uniform in [65, 100000], FP64 and FP32, complexForward, the scaled
complexInverse round trip and realForward.  Exercises every plan kind
(compile-time and runtime-radix single kernels, four-step splits with odd
factors, Bluestein) and the reference's plan boundaries (radices 2, 3, 4, 5,
7, 11, 13; Bluestein for larger prime factors)."""
import math

import numpy as np
import pytest

import jtransforms_amd as jt

pytestmark = pytest.mark.gpu


def _lengths():
    rng = np.random.Generator(np.random.PCG64(20110602))
    big = np.unique(np.exp(rng.uniform(math.log(65), math.log(100000), 120)).astype(np.int64))
    return list(range(1, 65)) + [int(v) for v in big]


LENGTHS = _lengths()


def tol(n, prec):
    lg = max(1.0, math.log2(max(n, 2)))
    return 1e-13 * lg if prec == "f64" else 2e-7 * lg + 1e-6


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d else 1.0))


CHUNKS = 8


@pytest.mark.parametrize("chunk", range(CHUNKS))
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_length_sweep(orc, prec, chunk):
    dt = np.float64 if prec == "f64" else np.float32
    cls = jt.DoubleFFT_1D if prec == "f64" else jt.FloatFFT_1D
    worst = (0.0, 0)
    lengths = LENGTHS[chunk::CHUNKS]
    for n in lengths:
        x = orc.jrandom_doubles(1000 + n, 2 * n).astype(dt)
        p = cls(n)
        a = x.copy()
        p.complexForward(a)
        e = rel(a, orc.complex_forward(n, x.astype(np.float64)))
        assert e <= tol(n, prec), (n, "complexForward", e, orc.plan_kind(n))
        p.complexInverse(a, True)
        e2 = rel(a, x)
        assert e2 <= 2 * tol(n, prec), (n, "round trip", e2)
        r = x[:n].copy()
        p.realForward(r)
        e3 = rel(r, orc.real_forward(n, x[:n].astype(np.float64)))
        assert e3 <= tol(n, prec), (n, "realForward", e3, orc.plan_kind(n))
        worst = max(worst, (e / tol(n, prec), n))
    print(f"{prec} chunk {chunk}: {len(lengths)} lengths, worst complexForward error / tolerance = {worst[0]:.3f} at n={worst[1]}",
          flush=True)
