"""Independent statements of JTransforms' packed layouts, used by the tests.

These follow reference files other than the ones the oracle's packing code
restates, so that a mistake in either side shows up as a mismatch:

* ``unpack_real_1d_like_test`` follows the unpacking procedure of the
  reference's own test (DoubleFFT_1DTest.java:325-365).
* ``getindex_2d`` / ``getindex_3d`` follow RealFFTUtils_2D.getIndex
  (RealFFTUtils_2D.java:191-242) and RealFFTUtils_3D.getIndex
  (RealFFTUtils_3D.java:232-330).
"""
from __future__ import annotations

import os

import numpy as np

MIN = None  # marks a Fourier mode that the packed layout stores as zero

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_RESOURCES = "/root/reference/src/test/resources"


def is_pow2(n: int) -> bool:
    return n > 0 and (n & (n - 1)) == 0


def read_fftw(n: int, which: str, allow_reference: bool = False) -> np.ndarray:
    """FFTW fixture (little-endian float64, 2n values), DoubleFFT_1DTest.java:156-172."""
    p = os.path.join(GOLDEN, "fftw", f"fftw{n}.{which}")
    if not os.path.exists(p) and allow_reference:
        p = os.path.join(REF_RESOURCES, f"fftw{n}.{which}")
    return np.fromfile(p, dtype="<f8")


def unpack_real_1d_like_test(n: int, packed: np.ndarray) -> np.ndarray:
    """Full 2n-double spectrum from realForward's packed output, as
    DoubleFFT_1DTest.testRealForward rebuilds it."""
    a = np.zeros(2 * n)
    a[:n] = packed[:n]
    if not is_pow2(n):
        m = n // 2 if n % 2 == 0 else (n + 1) // 2
        a[n] = a[1]
        a[1] = 0
        for k in range(1, m):
            a[2 * n - 2 * k] = a[2 * k]
            a[2 * n - 2 * k + 1] = -a[2 * k + 1]
    else:
        for k in range(1, n // 2):
            a[2 * n - 2 * k] = a[2 * k]
            a[2 * n - 2 * k + 1] = -a[2 * k + 1]
        a[n] = a[1]
        a[1] = 0
    return a


def getindex_2d(rows: int, columns: int, r: int, c: int):
    """RealFFTUtils_2D.getIndex: index into the packed array of the real
    (c even) or imaginary (c odd) part of mode [r][c//2]; a negative result
    means the value is -packed[-index]; None means the value is 0."""
    cmod2 = c & 1
    rmul2 = r << 1
    if r != 0:
        if c <= 1:
            if rmul2 == rows:
                if cmod2 == 1:
                    return MIN
                return (rows * columns) >> 1
            elif rmul2 < rows:
                return columns * r + cmod2
            elif cmod2 == 0:
                return columns * (rows - r)
            else:
                return -(columns * (rows - r) + 1)
        elif c == columns or c == columns + 1:
            if rmul2 == rows:
                if cmod2 == 1:
                    return MIN
                return ((rows * columns) >> 1) + 1
            elif rmul2 < rows:
                if cmod2 == 0:
                    return columns * (rows - r) + 1
                else:
                    return -(columns * (rows - r))
            else:
                return columns * r + 1 - cmod2
        elif c < columns:
            return columns * r + c
        elif cmod2 == 0:
            return columns * (rows + 2 - r) - c
        else:
            return -(columns * (rows + 2 - r) - c + 2)
    elif c == 1 or c == columns + 1:
        return MIN
    elif c == columns:
        return 1
    elif c < columns:
        return c
    elif cmod2 == 0:
        return (columns << 1) - c
    else:
        return -((columns << 1) - c + 2)


def getindex_3d(slices: int, rows: int, columns: int, s: int, r: int, c: int):
    """RealFFTUtils_3D.getIndex (same conventions as getindex_2d)."""
    slice_stride = rows * columns
    row_stride = columns
    cmod2 = c & 1
    rmul2 = r << 1
    smul2 = s << 1
    ss = 0 if s == 0 else slices - s
    rr = 0 if r == 0 else rows - r
    if c <= 1:
        if r == 0:
            if s == 0:
                return 0 if c == 0 else MIN
            elif smul2 < slices:
                return s * slice_stride + c
            elif smul2 > slices:
                index = ss * slice_stride
                return index if cmod2 == 0 else -(index + 1)
            else:
                return s * slice_stride if cmod2 == 0 else MIN
        elif rmul2 < rows:
            return s * slice_stride + r * row_stride + c
        elif rmul2 > rows:
            index = ss * slice_stride + rr * row_stride
            return index if cmod2 == 0 else -(index + 1)
        elif s == 0:
            return r * row_stride if cmod2 == 0 else MIN
        elif smul2 < slices:
            return s * slice_stride + r * row_stride + c
        elif smul2 > slices:
            index = ss * slice_stride + r * row_stride
            return index if cmod2 == 0 else -(index + 1)
        else:
            index = s * slice_stride + r * row_stride
            return index if cmod2 == 0 else MIN
    elif c < columns:
        return s * slice_stride + r * row_stride + c
    elif c > columns + 1:
        cc = (columns << 1) - c
        index = ss * slice_stride + rr * row_stride + cc
        return index if cmod2 == 0 else -(index + 2)
    elif r == 0:
        if s == 0:
            return 1 if cmod2 == 0 else MIN
        elif smul2 < slices:
            index = ss * slice_stride
            return index + 1 if cmod2 == 0 else -index
        elif smul2 > slices:
            index = s * slice_stride
            return index + 1 if cmod2 == 0 else index
        else:
            index = s * slice_stride
            return index + 1 if cmod2 == 0 else MIN
    elif rmul2 < rows:
        index = ss * slice_stride + rr * row_stride
        return index + 1 if cmod2 == 0 else -index
    elif rmul2 > rows:
        index = s * slice_stride + r * row_stride
        return index + 1 if cmod2 == 0 else index
    elif s == 0:
        index = r * row_stride + 1
        return index if cmod2 == 0 else MIN
    elif smul2 < slices:
        index = ss * slice_stride + r * row_stride
        return index + 1 if cmod2 == 0 else -index
    elif smul2 > slices:
        index = s * slice_stride + r * row_stride
        return index + 1 if cmod2 == 0 else index
    else:
        index = s * slice_stride + r * row_stride
        return index + 1 if cmod2 == 0 else MIN


def _fetch(packed, idx):
    if idx is MIN:
        return 0.0
    if idx >= 0:
        return packed[idx]
    return -packed[-idx]


def unpack_2d_getindex(rows: int, columns: int, packed: np.ndarray) -> np.ndarray:
    """Full complex spectrum (rows x columns) from the packed 2D layout."""
    out = np.zeros((rows, columns), np.complex128)
    for r in range(rows):
        for k in range(columns):
            re = _fetch(packed, getindex_2d(rows, columns, r, 2 * k))
            im = _fetch(packed, getindex_2d(rows, columns, r, 2 * k + 1))
            out[r, k] = re + 1j * im
    return out


def unpack_3d_getindex(slices: int, rows: int, columns: int, packed: np.ndarray) -> np.ndarray:
    out = np.zeros((slices, rows, columns), np.complex128)
    for s in range(slices):
        for r in range(rows):
            for k in range(columns):
                re = _fetch(packed, getindex_3d(slices, rows, columns, s, r, 2 * k))
                im = _fetch(packed, getindex_3d(slices, rows, columns, s, r, 2 * k + 1))
                out[s, r, k] = re + 1j * im
    return out


def pack_1d_from_spectrum(n: int, X: np.ndarray) -> np.ndarray:
    """Packed realForward layout from a complex spectrum, following the
    javadoc of DoubleFFT_1D.realForward (DoubleFFT_1D.java:436-451)."""
    a = np.zeros(n)
    if n % 2 == 0:
        for k in range(n // 2):
            a[2 * k] = X[k].real
            if k > 0:
                a[2 * k + 1] = X[k].imag
        a[1] = X[n // 2].real
    else:
        for k in range((n + 1) // 2):
            a[2 * k] = X[k].real
        for k in range(1, (n - 1) // 2):
            a[2 * k + 1] = X[k].imag
        if n > 1:
            a[1] = X[(n - 1) // 2].imag
    return a
