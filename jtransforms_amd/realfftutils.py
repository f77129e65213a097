"""Read/write access to single Fourier modes of the packed output of
`realForward` on 2D and 3D data: the Python side of the reference's
`RealFFTUtils_2D` (fft/RealFFTUtils_2D.java:97-665) and `RealFFTUtils_3D`
(fft/RealFFTUtils_3D.java:98-794).

`getIndex(..., c)` addresses the full spectrum the way `realForwardFull`
lays it out: column index c in [0, 2*columns), even c the real part and odd c
the imaginary part of mode c // 2.  The result is an index into the packed
array: `index >= 0` means the mode is `packed[index]`, a negative index means
it is `-packed[-index]`, and `ZERO_MODE` means the packed layout does not
store it because it is always zero (the imaginary part of a self-conjugate
mode).  The reference returns Integer.MIN_VALUE / Long.MIN_VALUE for that
case; here it is Long.MIN_VALUE for every size.

The index is derived from the packed layout documented for realForward
(fft/DoubleFFT_2D.java:794-811, fft/DoubleFFT_3D.java:1298-1330):

* modes 0 < k < columns/2 of the last axis sit in place, re/im interleaved;
* modes k > columns/2 follow from Hermitian symmetry,
  X[p][k] = conj X[-p][columns - k];
* the two real-axis planes k = 0 and k = columns/2, X0 and Xh, share the
  first two reals of every line.  Over the leading indices p, each pair of
  conjugate positions {p, -p} (indices modulo their lengths) holds two
  complex numbers: the slot of the "upper" position p holds X0[p], the slot
  of -p holds i*Xh[p].  A self-conjugate p (every leading index 0 or half
  its length) holds the two real numbers (X0[p], Xh[p]).  Upper means, for
  2D rows r: 0 < r < rows/2; for 3D (s, r): 0 < r < rows/2, or r in
  {0, rows/2} with 0 < s < slices/2.

The Java overloads taking `double[]`/`float[]`/`*LargeArray` with `pos` and
`double[][]`/`[][][]` map to one method each here: `packed` is any
C-contiguous numpy array or torch tensor (of any shape, read as its flat
element sequence).
"""
from __future__ import annotations

ZERO_MODE = -(1 << 63)

__all__ = ["RealFFTUtils_2D", "RealFFTUtils_3D", "ZERO_MODE"]


class _PackedModes:
    """The leading shape `lead` (rows, or (slices, rows)) and `columns`."""

    def __init__(self, lead, columns):
        self._lead = tuple(int(d) for d in lead)
        self.columns = int(columns)
        self._line_strides = []
        st = self.columns
        for d in reversed(self._lead):
            self._line_strides.insert(0, st)
            st *= d

    # -- layout of the packed array -----------------------------------------
    def _base(self, p):
        return sum(i * s for i, s in zip(p, self._line_strides))

    def _neg(self, p):
        return tuple((-i) % d for i, d in zip(p, self._lead))

    def _self_conjugate(self, p):
        return all(2 * i % d == 0 for i, d in zip(p, self._lead))

    def _upper(self, p):
        """Is p the upper member of its (non-self-conjugate) pair {p, -p}."""
        *outer, r = p
        rows = self._lead[-1]
        if 2 * r != 0 and 2 * r != rows:
            return 2 * r < rows
        # r in {0, rows/2}: the next leading index decides
        return bool(outer) and 2 * outer[0] < self._lead[0]

    def _index(self, p, c):
        C = self.columns
        if not 0 <= c < 2 * C or any(not 0 <= i < d for i, d in zip(p, self._lead)):
            raise IndexError(f"mode {tuple(p) + (c,)} outside {self._lead + (2 * C,)}")
        k, part = c >> 1, c & 1
        sign = 1
        if 2 * k > C:  # Hermitian mirror into 0 <= k <= columns/2
            p, k = self._neg(p), C - k
            sign = -1 if part else 1
        if 0 < 2 * k < C:
            return sign * (self._base(p) + 2 * k + part)
        h = 1 if 2 * k == C else 0  # plane X0 (h = 0) or Xh (h = 1)
        if self._self_conjugate(p):
            return ZERO_MODE if part else self._base(p) + h
        if not self._upper(p):  # X_e[p] = conj X_e[-p]
            p = self._neg(p)
            if part:
                sign = -sign
        if not h:
            return sign * (self._base(p) + part)
        # the partner slot -p holds i*Xh[p] = (-Im Xh[p], Re Xh[p])
        q = self._base(self._neg(p))
        return sign * (q + 1) if part == 0 else -sign * q

    # -- access -------------------------------------------------------------
    @staticmethod
    def _flat(packed):
        return packed.reshape(-1)

    def _pack(self, val, idx, mode, packed, pos):
        flat = self._flat(packed)
        if idx == ZERO_MODE:
            from . import IllegalArgumentException
            raise IllegalArgumentException(
                "".join(f"[{i}]" for i in mode) + " component cannot be modified (always zero)")
        if idx >= 0:
            flat[pos + idx] = val
        else:
            flat[pos - idx] = -val

    def _unpack(self, idx, packed, pos):
        flat = self._flat(packed)
        if idx == ZERO_MODE:
            return 0.0 * flat[pos]
        return flat[pos + idx] if idx >= 0 else -flat[pos - idx]


class RealFFTUtils_2D(_PackedModes):
    """Mode access for `DoubleFFT_2D/FloatFFT_2D.realForward` output
    (fft/RealFFTUtils_2D.java:159-665)."""

    def __init__(self, rows: int, columns: int):
        super().__init__((rows,), columns)
        self.rows = int(rows)

    def getIndex(self, r: int, c: int) -> int:
        """Index of mode [r][c] (fft/RealFFTUtils_2D.java:191-242)."""
        return self._index((int(r),), int(c))

    def pack(self, val, r: int, c: int, packed, pos: int = 0) -> None:
        """Set mode [r][c] (fft/RealFFTUtils_2D.java:335-350); raises
        IllegalArgumentException for a mode that is always zero."""
        self._pack(val, self.getIndex(r, c), (r, c), packed, pos)

    def unpack(self, r: int, c: int, packed, pos: int = 0):
        """Mode [r][c] (fft/RealFFTUtils_2D.java:517-531)."""
        return self._unpack(self.getIndex(r, c), packed, pos)


class RealFFTUtils_3D(_PackedModes):
    """Mode access for `DoubleFFT_3D/FloatFFT_3D.realForward` output
    (fft/RealFFTUtils_3D.java:192-794)."""

    def __init__(self, slices: int, rows: int, columns: int):
        super().__init__((slices, rows), columns)
        self.slices, self.rows = int(slices), int(rows)

    def getIndex(self, s: int, r: int, c: int) -> int:
        """Index of mode [s][r][c] (fft/RealFFTUtils_3D.java:232-330)."""
        return self._index((int(s), int(r)), int(c))

    def pack(self, val, s: int, r: int, c: int, packed, pos: int = 0) -> None:
        """Set mode [s][r][c] (fft/RealFFTUtils_3D.java:425-443)."""
        self._pack(val, self.getIndex(s, r, c), (s, r, c), packed, pos)

    def unpack(self, s: int, r: int, c: int, packed, pos: int = 0):
        """Mode [s][r][c] (fft/RealFFTUtils_3D.java:624-640)."""
        return self._unpack(self.getIndex(s, r, c), packed, pos)
