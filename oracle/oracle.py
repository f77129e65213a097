"""ctypes front end of the CPU oracle (oracle/jt_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``.  The product library never uses
it.  Every function works in float64 on numpy arrays and follows the
JTransforms method it is named after (file:line citations live in
jt_oracle.c).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libjt_oracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_I64 = ctypes.c_int64


def build() -> str:
    """Compile the oracle with gcc (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_jrandom_doubles.argtypes = [_I64, ctypes.c_int, _D, _I64]
        L.orc_jrandom_floats.argtypes = [_I64, ctypes.c_int, _F, _I64]
        L.orc_get_reminder.argtypes = [_I64]
        L.orc_get_reminder.restype = _I64
        L.orc_next_pow2.argtypes = [_I64]
        L.orc_next_pow2.restype = _I64
        L.orc_plan_kind.argtypes = [_I64]
        L.orc_bluestein_len.argtypes = [_I64]
        L.orc_bluestein_len.restype = _I64
        for name in ("orc_complex_forward", "orc_real_forward", "orc_real_forward_full",
                     "orc_dht_forward"):
            getattr(L, name).argtypes = [_I64, _D]
        for name in ("orc_complex_inverse", "orc_real_inverse", "orc_real_inverse_full",
                     "orc_dct_forward", "orc_dct_inverse", "orc_dst_forward",
                     "orc_dst_inverse", "orc_dht_inverse"):
            getattr(L, name).argtypes = [_I64, _D, ctypes.c_int]
        L.orc_cfft.argtypes = [_I64, _D, ctypes.c_int]
        L.orc_dft_brute_ld.argtypes = [_I64, _D, _D, ctypes.c_int]
        L.orc_fftnd_complex.argtypes = [ctypes.c_int, ctypes.POINTER(_I64), _D, ctypes.c_int,
                                        ctypes.c_int]
        L.orc_fftnd_real_full.argtypes = L.orc_fftnd_complex.argtypes
        L.orc_fft2d_real_forward.argtypes = [_I64, _I64, _D]
        L.orc_fft2d_real_inverse.argtypes = [_I64, _I64, _D, ctypes.c_int]
        L.orc_fft3d_real_forward.argtypes = [_I64, _I64, _I64, _D]
        L.orc_fft3d_real_inverse.argtypes = [_I64, _I64, _I64, _D, ctypes.c_int]
        L.orc_realnd.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_I64),
                                 _D, ctypes.c_int]
        L.orc_num_threads.argtypes = []
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


_STATUS = {1: "bad size", 2: "sizes must be powers of two", 3: "out of memory"}


def _check(st: int):
    if st:
        raise OracleError(_STATUS.get(st, f"status {st}"))


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def _dims(d):
    arr = (_I64 * len(d))(*d)
    return arr


# --- inputs -----------------------------------------------------------------
def jrandom_doubles(seed: int, count: int, signed: bool = True) -> np.ndarray:
    """java.util.Random(seed): 2*nextDouble()-1 (signed) or nextDouble()."""
    out = np.empty(count, np.float64)
    lib().orc_jrandom_doubles(seed, 1 if signed else 0, _dp(out), count)
    return out


def jrandom_floats(seed: int, count: int, signed: bool = True) -> np.ndarray:
    out = np.empty(count, np.float32)
    lib().orc_jrandom_floats(seed, 1 if signed else 0, out.ctypes.data_as(_F), count)
    return out


# --- plan choice --------------------------------------------------------------
def plan_kind(n: int) -> str:
    return ("SPLIT_RADIX", "MIXED_RADIX", "BLUESTEIN")[lib().orc_plan_kind(n)]


def get_reminder(n: int) -> int:
    return lib().orc_get_reminder(n)


def bluestein_len(n: int) -> int:
    return lib().orc_bluestein_len(n)


# --- 1D ----------------------------------------------------------------------
def _inplace(fn, n, a, *extra):
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    _check(fn(n, _dp(a), *extra))
    return a


def complex_forward(n, a):
    return _inplace(lib().orc_complex_forward, n, a)


def complex_inverse(n, a, scale):
    return _inplace(lib().orc_complex_inverse, n, a, int(bool(scale)))


def real_forward(n, a):
    return _inplace(lib().orc_real_forward, n, a)


def real_inverse(n, a, scale):
    return _inplace(lib().orc_real_inverse, n, a, int(bool(scale)))


def real_forward_full(n, a):
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    assert a.size >= 2 * n
    _check(lib().orc_real_forward_full(n, _dp(a)))
    return a


def real_inverse_full(n, a, scale):
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    assert a.size >= 2 * n
    _check(lib().orc_real_inverse_full(n, _dp(a), int(bool(scale))))
    return a


def dft_brute_ld(n, a, sign=-1):
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.empty(2 * n, np.float64)
    lib().orc_dft_brute_ld(n, _dp(a), _dp(out), sign)
    return out


def dct_forward(n, a, scale):
    return _inplace(lib().orc_dct_forward, n, a, int(bool(scale)))


def dct_inverse(n, a, scale):
    return _inplace(lib().orc_dct_inverse, n, a, int(bool(scale)))


def dst_forward(n, a, scale):
    return _inplace(lib().orc_dst_forward, n, a, int(bool(scale)))


def dst_inverse(n, a, scale):
    return _inplace(lib().orc_dst_inverse, n, a, int(bool(scale)))


def dht_forward(n, a):
    return _inplace(lib().orc_dht_forward, n, a)


def dht_inverse(n, a, scale):
    return _inplace(lib().orc_dht_inverse, n, a, int(bool(scale)))


# --- multi-D -------------------------------------------------------------------
def fftnd_complex(dims, a, sign, scale=False):
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    _check(lib().orc_fftnd_complex(len(dims), _dims(dims), _dp(a), sign, int(bool(scale))))
    return a


def fftnd_real_full(dims, a, sign, scale=False):
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    _check(lib().orc_fftnd_real_full(len(dims), _dims(dims), _dp(a), sign, int(bool(scale))))
    return a


def fft2d_real_forward(rows, cols, a):
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    _check(lib().orc_fft2d_real_forward(rows, cols, _dp(a)))
    return a


def fft2d_real_inverse(rows, cols, a, scale):
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    _check(lib().orc_fft2d_real_inverse(rows, cols, _dp(a), int(bool(scale))))
    return a


def fft3d_real_forward(s, r, c, a):
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    _check(lib().orc_fft3d_real_forward(s, r, c, _dp(a)))
    return a


def fft3d_real_inverse(s, r, c, a, scale):
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    _check(lib().orc_fft3d_real_inverse(s, r, c, _dp(a), int(bool(scale))))
    return a


_KIND = {"dct": 0, "dst": 1, "dht": 2}


def realnd(kind, dims, a, forward, scale):
    """Separable DCT/DST/DHT over every axis (DHT adds the yTransform)."""
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    _check(lib().orc_realnd(_KIND[kind], -1 if forward else 1, len(dims), _dims(dims), _dp(a),
                            int(bool(scale))))
    return a


def num_threads() -> int:
    return lib().orc_num_threads()


def rel_l2(y, ref) -> float:
    y = np.asarray(y, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(y - ref) / (den if den > 0 else 1.0))


def rmse(y, ref) -> float:
    """IOUtils.computeRMSE (IOUtils.java:185-197)."""
    y = np.asarray(y, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    return float(np.sqrt(np.mean((y - ref) ** 2)))
