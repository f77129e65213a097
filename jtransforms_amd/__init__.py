"""MI355X (gfx950) implementation of JTransforms' transform API.

The classes mirror ``org.jtransforms.{fft,dct,dst,dht}`` (same class names,
method names, argument meaning and in-place layouts) over the C ABI of
``libjtamd.so`` (include/jtransforms_amd.h).  Arrays may be numpy arrays
(host memory, like the reference's ``double[]``; the call copies to HBM,
transforms there and copies back) or torch tensors on a ROCm device (the
transform runs in HBM on the current stream, asynchronously).

There is no CPU fallback: if the HIP library cannot be loaded every call
raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

__all__ = [
    "DoubleFFT_1D", "FloatFFT_1D", "DoubleFFT_2D", "FloatFFT_2D", "DoubleFFT_3D", "FloatFFT_3D",
    "DoubleDCT_1D", "FloatDCT_1D", "DoubleDCT_2D", "FloatDCT_2D", "DoubleDCT_3D", "FloatDCT_3D",
    "DoubleDST_1D", "FloatDST_1D", "DoubleDST_2D", "FloatDST_2D", "DoubleDST_3D", "FloatDST_3D",
    "DoubleDHT_1D", "FloatDHT_1D", "DoubleDHT_2D", "FloatDHT_2D", "DoubleDHT_3D", "FloatDHT_3D",
    "FFTBatch1D", "RaggedBatch1D", "RealFFTUtils_2D", "RealFFTUtils_3D", "IllegalArgumentException",
    "JTransformsError", "lib", "LIB_PATH",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libjtamd.so")

# status codes (include/jtransforms_amd.h)
JT_OK, JT_BAD_SIZE, JT_NOT_POW2, JT_TOO_BIG, JT_HIP_ERROR, JT_OOM, JT_BAD_ARG, JT_UNSUPPORTED = range(8)
JT_F64, JT_F32 = 0, 1
JT_FLAG_SPLIT_FP16 = 1
JT_FLAG_GRAPH = 2
JT_FLAG_AUTOTUNE = 4
(COMPLEX_FORWARD, COMPLEX_INVERSE, REAL_FORWARD, REAL_INVERSE, REAL_FORWARD_FULL,
 REAL_INVERSE_FULL, FORWARD, INVERSE) = range(8)

EXPORTED_SYMBOLS = [
    "jt_plan_create_fft1d", "jt_plan_create_fft2d", "jt_plan_create_fft3d", "jt_plan_create_dct",
    "jt_plan_create_dst", "jt_plan_create_dht", "jt_plan_create_fft1d_batch", "jt_plan_destroy",
    "jt_execute", "jt_execute_host", "jt_plan_elements", "jt_status_string", "jt_last_error",
    "jt_version", "jt_profile_begin", "jt_profile_end", "jt_plan_set_flags",
    "jt_plan_create_ragged", "jt_plan_sync_status", "jt_plan_describe", "jt_execute_noalloc",
]


class JTransformsError(RuntimeError):
    """A HIP/runtime failure reported by the library."""


class IllegalArgumentException(ValueError):
    """Raised where the reference throws IllegalArgumentException."""


_lib = None


def lib():
    """Load libjtamd.so (built by ``__graft_entry__.build()``); raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise JTransformsError(f"HIP library not built: {LIB_PATH} is missing")
        # PyTorch-ROCm ships its own libamdhip64.so.7.  Loading torch first
        # makes the dynamic linker bind this library's libamdhip64.so.7
        # dependency to that same runtime, so device pointers and streams
        # from torch tensors are valid here (two HIP runtimes in one process
        # cannot share a device).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(ctypes.c_void_p)
        i64 = ctypes.c_int64
        L.jt_plan_create_fft1d.argtypes = [P, ctypes.c_int, i64]
        L.jt_plan_create_fft2d.argtypes = [P, ctypes.c_int, i64, i64]
        L.jt_plan_create_fft3d.argtypes = [P, ctypes.c_int, i64, i64, i64]
        for f in ("jt_plan_create_dct", "jt_plan_create_dst", "jt_plan_create_dht"):
            getattr(L, f).argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(i64)]
        L.jt_plan_create_fft1d_batch.argtypes = [P, ctypes.c_int, i64, i64, ctypes.c_uint32]
        L.jt_plan_destroy.argtypes = [ctypes.c_void_p]
        L.jt_plan_set_flags.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        L.jt_plan_create_ragged.argtypes = [P, ctypes.c_int, i64, ctypes.POINTER(i64), ctypes.c_uint32]
        L.jt_plan_sync_status.argtypes = [ctypes.c_void_p]
        L.jt_plan_describe.argtypes = [ctypes.c_void_p, ctypes.c_char_p, i64]
        L.jt_plan_describe.restype = i64
        L.jt_execute.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, i64, ctypes.c_int,
                                 ctypes.c_void_p]
        L.jt_execute_host.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, i64,
                                      ctypes.c_int]
        L.jt_execute_noalloc.argtypes = L.jt_execute.argtypes
        L.jt_plan_elements.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.jt_plan_elements.restype = i64
        L.jt_status_string.argtypes = [ctypes.c_int]
        L.jt_status_string.restype = ctypes.c_char_p
        L.jt_last_error.argtypes = []
        L.jt_last_error.restype = ctypes.c_char_p
        L.jt_version.argtypes = []
        L.jt_version.restype = ctypes.c_char_p
        L.jt_profile_begin.argtypes = []
        L.jt_profile_end.argtypes = [ctypes.c_char_p, i64]
        L.jt_profile_end.restype = i64
        _lib = L
    return _lib


def _raise(rc: int):
    L = lib()
    msg = L.jt_last_error().decode() or L.jt_status_string(rc).decode()
    if rc in (JT_BAD_SIZE, JT_NOT_POW2, JT_TOO_BIG):
        raise IllegalArgumentException(msg)
    if rc == JT_BAD_ARG:
        raise ValueError(msg)
    if rc == JT_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise JTransformsError(f"{msg} (status {rc})")


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


class _Plan:
    """One C-ABI plan; subclasses name the reference class they replace."""

    _precision = JT_F64

    def __init__(self, creator, *args):
        L = lib()
        h = ctypes.c_void_p()
        rc = creator(ctypes.byref(h), self._precision, *args)
        if rc:
            _raise(rc)
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.jt_plan_destroy(h)
            self._h = None

    @property
    def dtype(self):
        return np.float64 if self._precision == JT_F64 else np.float32

    def describe(self) -> str:
        """The library's one-line description of this plan (schedule, tuning)."""
        n = lib().jt_plan_describe(self._h, None, 0)
        buf = ctypes.create_string_buffer(int(n) + 1)
        lib().jt_plan_describe(self._h, buf, n + 1)
        return buf.value.decode()

    def elements(self, op: int) -> int:
        cache = self.__dict__.setdefault("_elements", {})
        n = cache.get(op)
        if n is None:
            n = cache[op] = int(lib().jt_plan_elements(self._h, op))
        return n

    # -- the array-of-arrays overloads (double[][] / double[][][]) ----------
    def _nested_shape(self, op: int):
        """Shape of the reference's double[][] / double[][][] argument for op,
        or None where the class has no such overload."""
        return None

    @staticmethod
    def _is_nested(a) -> bool:
        if isinstance(a, (list, tuple)):
            return True
        if isinstance(a, np.ndarray):
            return a.ndim > 1
        return _is_torch(a) and a.dim() > 1

    def _run_nested(self, op: int, a, scale: bool):
        """Runs op on a 2D/3D array: a C-contiguous numpy array or device tensor
        of the overload's shape (transformed in place through a flat view), or
        nested lists of 1D rows, like Java's double[][] (gathered, transformed,
        and written back into every row in place).  For realForwardFull /
        realInverseFull each row holds 2*columns values of which the first
        columns are the real input (DoubleFFT_2D.java:1010-1041,
        DoubleFFT_3D.java realForwardFull(double[][][])); the flat overloads
        take the rows*columns reals compactly instead."""
        shape = self._nested_shape(op)
        if shape is None:
            raise TypeError(f"{type(self).__name__} has no array-of-arrays overload")
        shape = tuple(int(d) for d in shape)
        rows_of = None
        if isinstance(a, (list, tuple)):
            rows_of = a
            work = np.array(a, dtype=self.dtype)
        else:
            work = a
        if tuple(work.shape) != shape:
            raise IndexError(f"expected an array of shape {shape}, got {tuple(work.shape)}")
        if _is_torch(work):
            if not work.is_contiguous():
                raise TypeError("tensor must be contiguous")
            flat = work.view(-1)
        else:
            if not work.flags.c_contiguous or work.dtype != self.dtype:
                raise TypeError(f"expected a C-contiguous {np.dtype(self.dtype).name} array")
            flat = work.reshape(-1)
        if op in (REAL_FORWARD_FULL, REAL_INVERSE_FULL):
            cols = shape[-1] // 2
            buf = flat.clone() if _is_torch(flat) else flat.copy()
            reals = work[..., :cols].reshape(-1)
            buf[: reals.shape[0]] = reals
            self._run(op, buf, 0, scale)
            flat[:] = buf
        else:
            self._run(op, flat, 0, scale)
        if rows_of is not None:
            _write_back(rows_of, work)

    def _run(self, op: int, a, offa: int = 0, scale: bool = False):
        if self._is_nested(a):
            if offa:
                raise TypeError("the array-of-arrays overloads take no offset")
            self._run_nested(op, a, scale)
            return
        L = lib()
        need = self.elements(op)
        if _is_torch(a):
            import torch
            want = torch.float64 if self._precision == JT_F64 else torch.float32
            if a.dtype != want:
                raise TypeError(f"expected a {want} tensor, got {a.dtype}")
            if not a.is_cuda:
                raise TypeError("torch tensors must live on the GPU (use numpy for host arrays)")
            if not a.is_contiguous():
                raise TypeError("tensor must be contiguous")
            if a.numel() - offa < need:
                raise IndexError(f"array too short: need {need} elements from offset {offa}")
            if a.device.index == torch.cuda.current_device():
                # the common case: no device switch around the call
                stream = torch.cuda.current_stream().cuda_stream
                rc = L.jt_execute(self._h, op, ctypes.c_void_p(a.data_ptr()), offa, int(bool(scale)),
                                  ctypes.c_void_p(stream))
            else:
                with torch.cuda.device(a.device):
                    stream = torch.cuda.current_stream().cuda_stream
                    rc = L.jt_execute(self._h, op, ctypes.c_void_p(a.data_ptr()), offa, int(bool(scale)),
                                      ctypes.c_void_p(stream))
        else:
            if not isinstance(a, np.ndarray) or a.dtype != self.dtype or not a.flags.c_contiguous:
                raise TypeError(f"expected a contiguous numpy {np.dtype(self.dtype).name} array")
            if a.size - offa < need:
                raise IndexError(f"array too short: need {need} elements from offset {offa}")
            rc = L.jt_execute_host(self._h, op, a.ctypes.data_as(ctypes.c_void_p), offa,
                                   int(bool(scale)))
        if rc:
            _raise(rc)


def _write_back(rows_of, values):
    """Copies values (an array of the nested shape) into the caller's nested
    lists of rows, row by row and in place."""
    for i, row in enumerate(rows_of):
        if isinstance(row, (list, tuple)):
            _write_back(row, values[i])
        elif isinstance(row, np.ndarray):
            row[...] = values[i]
        else:
            raise TypeError("rows of an array-of-arrays argument must be numpy arrays")


def _offa_scale(args, name):
    # Java overloads: m(a, scale) and m(a, offa, scale)
    if len(args) == 1:
        return 0, args[0]
    if len(args) == 2:
        return args[0], args[1]
    raise TypeError(f"{name}(a, [offa,] scale)")


class _FFTMethods(_Plan):
    def set_split_fp16(self, on: bool = True):
        """Select the split-FP16 MFMA form of the radix-16 stages (FP32 plans)."""
        rc = lib().jt_plan_set_flags(self._h, JT_FLAG_SPLIT_FP16 if on else 0)
        if rc:
            _raise(rc)

    def complexForward(self, a, offa: int = 0):
        self._run(COMPLEX_FORWARD, a, offa)

    def complexInverse(self, a, *args):
        offa, scale = _offa_scale(args, "complexInverse")
        self._run(COMPLEX_INVERSE, a, offa, scale)

    def realForward(self, a, offa: int = 0):
        self._run(REAL_FORWARD, a, offa)

    def realInverse(self, a, *args):
        offa, scale = _offa_scale(args, "realInverse")
        self._run(REAL_INVERSE, a, offa, scale)

    def realForwardFull(self, a, offa: int = 0):
        self._run(REAL_FORWARD_FULL, a, offa)

    def realInverseFull(self, a, *args):
        offa, scale = _offa_scale(args, "realInverseFull")
        self._run(REAL_INVERSE_FULL, a, offa, scale)


class DoubleFFT_1D(_FFTMethods):
    """org.jtransforms.fft.DoubleFFT_1D (fft/DoubleFFT_1D.java:117)."""
    _precision = JT_F64

    def __init__(self, n: int):
        self.n = int(n)
        super().__init__(lib().jt_plan_create_fft1d, self.n)


class FloatFFT_1D(DoubleFFT_1D):
    """org.jtransforms.fft.FloatFFT_1D."""
    _precision = JT_F32


class DoubleFFT_2D(_FFTMethods):
    """org.jtransforms.fft.DoubleFFT_2D (fft/DoubleFFT_2D.java:75)."""
    _precision = JT_F64

    def __init__(self, rows: int, columns: int):
        self.rows, self.columns = int(rows), int(columns)
        super().__init__(lib().jt_plan_create_fft2d, self.rows, self.columns)

    def _nested_shape(self, op):
        # double[][] overloads (DoubleFFT_2D.java:345 complexForward, :924
        # realForward, ...): rows x 2*columns complex, rows x columns packed real
        c = self.columns if op in (REAL_FORWARD, REAL_INVERSE) else 2 * self.columns
        return (self.rows, c)


class FloatFFT_2D(DoubleFFT_2D):
    _precision = JT_F32


class DoubleFFT_3D(_FFTMethods):
    """org.jtransforms.fft.DoubleFFT_3D (fft/DoubleFFT_3D.java:88)."""
    _precision = JT_F64

    def __init__(self, slices: int, rows: int, columns: int):
        self.slices, self.rows, self.columns = int(slices), int(rows), int(columns)
        super().__init__(lib().jt_plan_create_fft3d, self.slices, self.rows, self.columns)

    def _nested_shape(self, op):
        # double[][][] overloads (DoubleFFT_3D.java:542 complexForward, :1467
        # realForward, ...)
        c = self.columns if op in (REAL_FORWARD, REAL_INVERSE) else 2 * self.columns
        return (self.slices, self.rows, c)


class FloatFFT_3D(DoubleFFT_3D):
    _precision = JT_F32


class FFTBatch1D(_FFTMethods):
    """`batch` independent 1D transforms of length n stored back to back
    (no reference counterpart: it replaces a loop of DoubleFFT_1D calls with
    offa = b*2n).  split_fp16 selects the split-FP16 MFMA form (FP32 only)."""

    def __init__(self, n: int, batch: int, precision: str = "f32", split_fp16: bool = False):
        self._precision = JT_F64 if precision in ("f64", "double") else JT_F32
        self.n, self.batch = int(n), int(batch)
        flags = JT_FLAG_SPLIT_FP16 if split_fp16 else 0
        super().__init__(lib().jt_plan_create_fft1d_batch, self.n, self.batch, flags)


class RaggedBatch1D(_Plan):
    """Many 1D complex transforms of unequal lengths stored back to back, run
    in one persistent launch (no reference counterpart: it replaces one
    DoubleFFT_1D per length).  graph=True replays a captured hipGraph."""

    def __init__(self, lengths, precision: str = "f64", graph: bool = False, autotune: bool = False):
        self._precision = JT_F64 if precision in ("f64", "double") else JT_F32
        self.lengths = np.ascontiguousarray(lengths, dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(2 * self.lengths)])
        arr = self.lengths.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        flags = (JT_FLAG_GRAPH if graph else 0) | (JT_FLAG_AUTOTUNE if autotune else 0)
        super().__init__(lib().jt_plan_create_ragged, len(self.lengths), arr, flags)

    def complexForward(self, a):
        self._run(COMPLEX_FORWARD, a, 0)

    def complexInverse(self, a, scale: bool):
        self._run(COMPLEX_INVERSE, a, 0, scale)

    def sync_status(self):
        rc = lib().jt_plan_sync_status(self._h)
        if rc:
            _raise(rc)


class _R2R(_Plan):
    _creator = "jt_plan_create_dct"
    _has_scale_fwd = True

    def __init__(self, *dims):
        self.dims = tuple(int(d) for d in dims)
        arr = (ctypes.c_int64 * len(dims))(*self.dims)
        super().__init__(getattr(lib(), self._creator), len(dims), arr)

    def _nested_shape(self, op):
        # double[][] / double[][][] overloads of the 2D/3D classes
        # (DoubleDCT_2D.java:280 forward, :544 inverse; DoubleDHT_2D.java:286)
        return self.dims if len(self.dims) > 1 else None

    def forward(self, a, *args):
        if self._has_scale_fwd:
            offa, scale = _offa_scale(args, "forward")
        else:
            offa, scale = (args[0] if args else 0), False
        self._run(FORWARD, a, offa, scale)

    def inverse(self, a, *args):
        offa, scale = _offa_scale(args, "inverse")
        self._run(INVERSE, a, offa, scale)


def _r2r_class(name, creator, precision, has_scale_fwd=True, doc=""):
    return type(name, (_R2R,), {"_creator": creator, "_precision": precision,
                                "_has_scale_fwd": has_scale_fwd, "__doc__": doc})


DoubleDCT_1D = _r2r_class("DoubleDCT_1D", "jt_plan_create_dct", JT_F64,
                          doc="org.jtransforms.dct.DoubleDCT_1D (dct/DoubleDCT_1D.java:88)")
FloatDCT_1D = _r2r_class("FloatDCT_1D", "jt_plan_create_dct", JT_F32)
DoubleDCT_2D = _r2r_class("DoubleDCT_2D", "jt_plan_create_dct", JT_F64,
                          doc="org.jtransforms.dct.DoubleDCT_2D (dct/DoubleDCT_2D.java:73)")
FloatDCT_2D = _r2r_class("FloatDCT_2D", "jt_plan_create_dct", JT_F32)
DoubleDCT_3D = _r2r_class("DoubleDCT_3D", "jt_plan_create_dct", JT_F64,
                          doc="org.jtransforms.dct.DoubleDCT_3D (dct/DoubleDCT_3D.java:85)")
FloatDCT_3D = _r2r_class("FloatDCT_3D", "jt_plan_create_dct", JT_F32)
DoubleDST_1D = _r2r_class("DoubleDST_1D", "jt_plan_create_dst", JT_F64,
                          doc="org.jtransforms.dst.DoubleDST_1D (dst/DoubleDST_1D.java:59)")
FloatDST_1D = _r2r_class("FloatDST_1D", "jt_plan_create_dst", JT_F32)
DoubleDST_2D = _r2r_class("DoubleDST_2D", "jt_plan_create_dst", JT_F64)
FloatDST_2D = _r2r_class("FloatDST_2D", "jt_plan_create_dst", JT_F32)
DoubleDST_3D = _r2r_class("DoubleDST_3D", "jt_plan_create_dst", JT_F64,
                          doc="org.jtransforms.dst.DoubleDST_3D (dst/DoubleDST_3D.java:84)")
FloatDST_3D = _r2r_class("FloatDST_3D", "jt_plan_create_dst", JT_F32)
DoubleDHT_1D = _r2r_class("DoubleDHT_1D", "jt_plan_create_dht", JT_F64, has_scale_fwd=False,
                          doc="org.jtransforms.dht.DoubleDHT_1D (dht/DoubleDHT_1D.java:60)")
FloatDHT_1D = _r2r_class("FloatDHT_1D", "jt_plan_create_dht", JT_F32, has_scale_fwd=False)
DoubleDHT_2D = _r2r_class("DoubleDHT_2D", "jt_plan_create_dht", JT_F64, has_scale_fwd=False)
FloatDHT_2D = _r2r_class("FloatDHT_2D", "jt_plan_create_dht", JT_F32, has_scale_fwd=False)
DoubleDHT_3D = _r2r_class("DoubleDHT_3D", "jt_plan_create_dht", JT_F64, has_scale_fwd=False)
FloatDHT_3D = _r2r_class("FloatDHT_3D", "jt_plan_create_dht", JT_F32, has_scale_fwd=False)


def profile_begin():
    """Start recording per-kernel HIP-event timings (jt_profile_begin)."""
    rc = lib().jt_profile_begin()
    if rc:
        _raise(rc)


def profile_end():
    """Stop recording; returns {kernel: (launches, total_ms)}."""
    L = lib()
    n = L.jt_profile_end(None, 0)
    buf = ctypes.create_string_buffer(int(n) + 1)
    L.jt_profile_end(buf, n + 1)
    out = {}
    for line in buf.value.decode().splitlines():
        name, cnt, ms = line.split("\t")
        out[name] = (int(cnt), float(ms))
    return out


def build():
    """Compile libjtamd.so for gfx950 (make -C jtransforms_amd/csrc)."""
    import subprocess
    jobs = os.environ.get("MAX_JOBS", "8")
    subprocess.run(["make", "-s", "-j", jobs, "-C", os.path.join(_HERE, "csrc")], check=True)
    return LIB_PATH


from .realfftutils import RealFFTUtils_2D, RealFFTUtils_3D  # noqa: E402  (host-side index helpers)
