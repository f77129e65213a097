"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and maps the reference's IllegalArgumentException cases
before touching the device."""
import ctypes
import os
import re

import pytest

import jtransforms_amd as jt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "jtransforms_amd.h")


def header_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(jt_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_python_list():
    assert header_symbols() == sorted(jt.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(jt.LIB_PATH)
    for name in header_symbols():
        assert hasattr(L, name), name


def test_version_and_status_strings():
    L = jt.lib()
    assert b"gfx950" in L.jt_version()
    assert L.jt_status_string(jt.JT_NOT_POW2) == b"sizes must be power of two numbers"


@pytest.mark.parametrize("ctor,args,msg", [
    (jt.DoubleFFT_1D, (0,), "n must be greater than 0"),          # DoubleFFT_1D.java:119-121
    (jt.FloatFFT_1D, (-3,), "n must be greater than 0"),
    (jt.DoubleFFT_2D, (1, 8), "rows and columns must be greater than 1"),   # DoubleFFT_2D.java:77-79
    (jt.DoubleFFT_3D, (4, 1, 4), "slices, rows and columns must be greater than 1"),
    (jt.DoubleDCT_1D, (0,), "n must be greater than 0"),
    (jt.DoubleDCT_1D, ((1 << 28) + 1,), "n must be smaller or equal to 268435456"),  # DoubleDCT_1D.java:98-100
    (jt.FloatDCT_1D, ((1 << 31) - 1,), "when useLargeArrays argument is set to false"),
    # DoubleDST_1D builds a DoubleDCT_1D(n) (DoubleDST_1D.java:59-65)
    (jt.DoubleDST_1D, ((1 << 28) + 1,), "n must be smaller or equal to 268435456"),
    (jt.FloatDST_1D, ((1 << 30),), "n must be smaller or equal to 268435456"),
])
def test_illegal_argument(ctor, args, msg):
    with pytest.raises(jt.IllegalArgumentException, match=re.escape(msg)):
        ctor(*args)


def test_split_fp16_needs_f32():
    with pytest.raises(ValueError):
        jt.FFTBatch1D(4096, 4, precision="f64", split_fp16=True)


def test_null_plan_rejected():
    L = jt.lib()
    assert L.jt_execute(None, 0, None, 0, 0, None) == jt.JT_BAD_ARG
    assert L.jt_execute_host(None, 0, None, 0, 0) == jt.JT_BAD_ARG
    assert L.jt_plan_elements(None, 0) == -1


@pytest.mark.parametrize("lens,msg", [([], "count must be greater than 0"),
                                      ([0], "n must be greater than 0"),
                                      ([5, -1], "n must be greater than 0")])
def test_ragged_illegal_lengths(lens, msg):
    """Ragged batches take the reference's 1D size rule for every length
    (DoubleFFT_1D.java:119-121) and need at least one transform."""
    import numpy as np
    with pytest.raises(jt.IllegalArgumentException, match=re.escape(msg)):
        jt.RaggedBatch1D(np.array(lens, dtype=np.int64))


def test_ragged_schedule_override_rejects_unknown_key(monkeypatch):
    import numpy as np
    monkeypatch.setenv("JT_RAGGED_SCHEDULE", "wgs=2,bogus=1")
    with pytest.raises(ValueError, match="unknown key bogus"):
        jt.RaggedBatch1D(np.array([16, 32], dtype=np.int64))


@pytest.mark.parametrize("sched", ["tplx=40", "wgs=0", "nt=300", "chunk=0", "order=2", "lag=-1"])
def test_ragged_schedule_override_range_checked(monkeypatch, sched):
    """Out-of-range schedule values are refused before any launch."""
    import numpy as np
    monkeypatch.setenv("JT_RAGGED_SCHEDULE", sched)
    with pytest.raises(ValueError, match="ragged schedule out of range"):
        jt.RaggedBatch1D(np.array([16, 32], dtype=np.int64))
