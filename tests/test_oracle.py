"""Pins the CPU oracle (oracle/jt_oracle.c) against the reference's golden
vectors and against closed forms / independent libraries (no GPU)."""
import os

import numpy as np
import pytest
import scipy.fft as sfft

from tests import jt_layout as L

FIXTURE_SIZES = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 16, 32, 64, 100, 120, 128, 256, 310,
                 512, 1024, 1056, 2048, 8192, 10158]
ALL_REF_SIZES = [int(x) for x in open(os.path.join(L.GOLDEN, "fftw", "sizes.txt")).read().split()]
SEED = 20110602  # DoubleFFT_1DTest.java:83
EPS = 1e-12      # DoubleFFT_1DTest.java:85


def c2n(a):
    return a[0::2] + 1j * a[1::2]


def n2c(z):
    out = np.empty(2 * z.size)
    out[0::2] = z.real
    out[1::2] = z.imag
    return out


# ---------------------------------------------------------------------------
# java.util.Random
# ---------------------------------------------------------------------------
def test_jrandom_known_values(orc):
    # Published values of java.util.Random: new Random(0).nextDouble() and
    # new Random(42).nextDouble().
    assert orc.jrandom_doubles(0, 1, signed=False)[0] == pytest.approx(0.730967787376657, abs=1e-15)
    assert orc.jrandom_doubles(42, 1, signed=False)[0] == pytest.approx(0.7275636800328681,
                                                                        abs=1e-16)
    x = orc.jrandom_doubles(SEED, 100000)
    assert x.min() >= -1 and x.max() < 1
    assert abs(x.mean()) < 0.01
    # 53-bit grid: 2*d-1 with d = m * 2^-53
    assert np.all(np.floor((x + 1) * 2.0 ** 52) == (x + 1) * 2.0 ** 52)


# ---------------------------------------------------------------------------
# complexForward against the reference's FFTW fixtures
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", FIXTURE_SIZES)
def test_complex_forward_fftw_fixture(orc, n):
    x = L.read_fftw(n, "in")
    y = L.read_fftw(n, "out")
    assert x.size == 2 * n
    got = orc.complex_forward(n, x)
    assert orc.rmse(got, y) <= EPS
    assert orc.rel_l2(got, y) <= 1e-13 * max(1.0, np.log2(n))


@pytest.mark.skipif(not os.path.isdir(L.REF_RESOURCES), reason="reference tree not mounted")
@pytest.mark.parametrize("n", [n for n in ALL_REF_SIZES if n not in FIXTURE_SIZES])
def test_complex_forward_fftw_fixture_large(orc, n):
    x = L.read_fftw(n, "in", allow_reference=True)
    y = L.read_fftw(n, "out", allow_reference=True)
    got = orc.complex_forward(n, x)
    assert orc.rmse(got, y) <= EPS


def test_plan_choice(orc):
    # DoubleFFT_1D.java:117-188 with factors {4,2,3,5}, Bluestein at cofactor >= 211
    assert orc.plan_kind(1024) == "SPLIT_RADIX"
    assert orc.plan_kind(1056) == "MIXED_RADIX"
    assert orc.plan_kind(310) == "MIXED_RADIX"
    assert orc.plan_kind(199) == "MIXED_RADIX"
    assert orc.plan_kind(211) == "BLUESTEIN"
    assert orc.plan_kind(10158) == "BLUESTEIN"   # 2*3*1693
    assert orc.plan_kind(65530) == "BLUESTEIN"   # 2*5*6553
    assert orc.bluestein_len(10158) == 32768
    assert orc.get_reminder(1056) == 11


@pytest.mark.parametrize("n", [7, 17, 100, 211, 422, 1024, 1125])
def test_complex_forward_vs_long_double(orc, n):
    x = orc.jrandom_doubles(SEED + n, 2 * n)
    ref = orc.dft_brute_ld(n, x, -1)
    assert orc.rel_l2(orc.complex_forward(n, x), ref) <= 2e-15 * np.log2(n) + 1e-15


@pytest.mark.parametrize("n", [1, 2, 5, 12, 64, 211, 1000, 4096])
def test_complex_inverse(orc, n):
    x = orc.jrandom_doubles(SEED, 2 * n)
    ref = n2c(np.fft.ifft(c2n(x)) * n)
    assert orc.rel_l2(orc.complex_inverse(n, x, False), ref) <= 1e-14
    rt = orc.complex_inverse(n, orc.complex_forward(n, x), True)
    assert orc.rmse(rt, x) <= EPS


# ---------------------------------------------------------------------------
# real transforms: packed layout through the test's own unpacking
# ---------------------------------------------------------------------------
REAL_SIZES = [2, 3, 4, 5, 6, 7, 8, 9, 16, 100, 120, 211, 310, 1024, 1056]


@pytest.mark.parametrize("n", REAL_SIZES)
def test_real_forward_packed(orc, n):
    x = orc.jrandom_doubles(SEED, n)
    packed = orc.real_forward(n, x)
    full = L.unpack_real_1d_like_test(n, packed)
    expected = n2c(np.fft.fft(x))
    assert orc.rmse(full, expected) <= EPS
    # javadoc map (DoubleFFT_1D.java:436-451)
    assert np.allclose(packed, L.pack_1d_from_spectrum(n, np.fft.fft(x)), atol=1e-12)


@pytest.mark.parametrize("n", REAL_SIZES)
def test_real_inverse(orc, n):
    x = orc.jrandom_doubles(SEED, n)
    packed = orc.real_forward(n, x)
    assert orc.rmse(orc.real_inverse(n, packed, True), x) <= EPS
    # DoubleFFT_1DTest.java:583-588: unscaled factor n/2 for powers of two, n otherwise
    s = 2.0 / n if (L.is_pow2(n) and n > 1) else 1.0 / n
    assert orc.rmse(orc.real_inverse(n, packed, False) * s, x) <= EPS


@pytest.mark.parametrize("n", REAL_SIZES)
def test_real_full(orc, n):
    x = orc.jrandom_doubles(SEED, n)
    a = np.zeros(2 * n)
    a[:n] = x
    assert orc.rmse(orc.real_forward_full(n, a), n2c(np.fft.fft(x))) <= EPS
    assert orc.rmse(orc.real_inverse_full(n, a, True), n2c(np.fft.ifft(x))) <= EPS
    assert orc.rmse(orc.real_inverse_full(n, a, False), n2c(np.fft.ifft(x) * n)) <= EPS * n


# ---------------------------------------------------------------------------
# 2D / 3D
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("rows,cols", [(2, 2), (4, 8), (16, 16), (6, 10), (32, 64), (12, 211)])
def test_fft2d_complex(orc, rows, cols):
    x = orc.jrandom_doubles(SEED, 2 * rows * cols, signed=False)
    z = c2n(x).reshape(rows, cols)
    got = orc.fftnd_complex([rows, cols], x, -1)
    assert orc.rel_l2(got, n2c(np.fft.fft2(z).ravel())) <= 1e-14
    inv = orc.fftnd_complex([rows, cols], got, +1, scale=True)
    assert orc.rmse(inv, x) <= EPS


@pytest.mark.parametrize("rows,cols", [(2, 2), (2, 8), (4, 4), (8, 16), (16, 4), (32, 32)])
def test_fft2d_real_forward_layout(orc, rows, cols):
    x = orc.jrandom_doubles(SEED, rows * cols, signed=False)
    packed = orc.fft2d_real_forward(rows, cols, x)
    full = L.unpack_2d_getindex(rows, cols, packed)
    assert np.abs(full - np.fft.fft2(x.reshape(rows, cols))).max() <= 1e-12
    assert orc.rmse(orc.fft2d_real_inverse(rows, cols, packed, True), x) <= EPS
    un = orc.fft2d_real_inverse(rows, cols, packed, False)
    assert orc.rmse(un / (rows * cols / 2), x) <= EPS


def test_fft2d_real_requires_pow2(orc):
    with pytest.raises(orc.OracleError):
        orc.fft2d_real_forward(6, 8, np.zeros(48))


@pytest.mark.parametrize("dims", [(2, 2, 2), (4, 8, 2), (8, 4, 16), (16, 16, 16)])
def test_fft3d_real_forward_layout(orc, dims):
    s, r, c = dims
    x = orc.jrandom_doubles(20110625, s * r * c, signed=False)
    packed = orc.fft3d_real_forward(s, r, c, x)
    full = L.unpack_3d_getindex(s, r, c, packed)
    assert np.abs(full - np.fft.fftn(x.reshape(s, r, c))).max() <= 1e-12
    assert orc.rmse(orc.fft3d_real_inverse(s, r, c, packed, True), x) <= EPS
    un = orc.fft3d_real_inverse(s, r, c, packed, False)
    assert orc.rmse(un / (s * r * c / 2), x) <= EPS


# Unscaled 2D/3D realInverse, pinned by following the reference's call
# sequence with independently pinned 1D pieces (no constant restated here):
#   DoubleFFT_2D.realInverse (fft/DoubleFFT_2D.java:1077-1095):
#       rdft2d_sub(-1); cdft2d_sub(1, scale) over the complex columns of the
#       packed rows; fftColumns.realInverse(row, scale) for every row
#   DoubleFFT_3D.realInverse (fft/DoubleFFT_3D.java:1629-1645):
#       rdft3d_sub(-1); cdft3db_sub(1, scale) along slices; xdft3da_sub1(1, 1):
#       complexInverse along rows, then realInverse of every row
# The forward path ends with rdft2d_sub(+1) / rdft3d_sub(+1), pairwise maps
# (u, v) -> ((u+v)/2, (u-v)/2) whose (-1) forms (u, v) -> (u+v, u-v) are their
# exact inverses (DoubleFFT_2D.java:2544-2574, DoubleFFT_3D.java:6909-7021),
# so the state the inverse's column passes see is the forward state before
# that fix: the row realForward followed by the column complexForward passes.
# The 1D pieces are pinned above: real_forward / real_inverse(scale=false) by
# numpy and DoubleFFT_1DTest.java:583-588, complex_forward / complex_inverse
# (scale=false) by the FFTW fixtures and numpy.
def _rows_real(orc, fn, x, cols, *extra):
    m = x.reshape(-1, cols)
    return np.stack([fn(cols, m[i].copy(), *extra) for i in range(m.shape[0])])


def _complex_axis(orc, fn, z, axis, *extra):
    z = np.moveaxis(z, axis, -1)
    n = z.shape[-1]
    flat = z.reshape(-1, n)
    out = np.empty_like(flat)
    for i in range(flat.shape[0]):
        out[i] = c2n(fn(n, n2c(flat[i]), *extra))
    return np.moveaxis(out.reshape(z.shape), -1, axis)


@pytest.mark.parametrize("rows,cols", [(4, 8), (8, 16), (16, 4), (32, 32)])
def test_fft2d_real_inverse_unscaled_by_call_sequence(orc, rows, cols):
    x = orc.jrandom_doubles(SEED, rows * cols)
    packed = orc.fft2d_real_forward(rows, cols, x)
    # forward state before rdft2d_sub(+1)
    z = c2n(_rows_real(orc, orc.real_forward, x, cols).ravel()).reshape(rows, cols // 2)
    y = _complex_axis(orc, orc.complex_forward, z, 0)
    # the reference's inverse sequence after rdft2d_sub(-1), unscaled
    u = _complex_axis(orc, orc.complex_inverse, y, 0, False)
    want = _rows_real(orc, orc.real_inverse, n2c(u.ravel()), cols, False).ravel()
    got = orc.fft2d_real_inverse(rows, cols, packed, False)
    assert orc.rel_l2(got, want) <= 1e-14
    assert orc.rel_l2(orc.fft2d_real_inverse(rows, cols, packed, True), x) <= 1e-14


@pytest.mark.parametrize("dims", [(2, 4, 8), (4, 8, 4), (8, 4, 16), (16, 16, 16)])
def test_fft3d_real_inverse_unscaled_by_call_sequence(orc, dims):
    s, r, c = dims
    x = orc.jrandom_doubles(20110625, s * r * c)
    packed = orc.fft3d_real_forward(s, r, c, x)
    z = c2n(_rows_real(orc, orc.real_forward, x, c).ravel()).reshape(s, r, c // 2)
    y = _complex_axis(orc, orc.complex_forward, _complex_axis(orc, orc.complex_forward, z, 1), 0)
    u = _complex_axis(orc, orc.complex_inverse, y, 0, False)
    u = _complex_axis(orc, orc.complex_inverse, u, 1, False)
    want = _rows_real(orc, orc.real_inverse, n2c(u.ravel()), c, False).ravel()
    got = orc.fft3d_real_inverse(s, r, c, packed, False)
    assert orc.rel_l2(got, want) <= 1e-14
    assert orc.rel_l2(orc.fft3d_real_inverse(s, r, c, packed, True), x) <= 1e-14


@pytest.mark.parametrize("dims", [(2, 3, 4), (8, 8, 8), (5, 6, 7)])
def test_fft3d_complex(orc, dims):
    x = orc.jrandom_doubles(20110625, 2 * int(np.prod(dims)), signed=False)
    z = c2n(x).reshape(dims)
    got = orc.fftnd_complex(list(dims), x, -1)
    assert orc.rel_l2(got, n2c(np.fft.fftn(z).ravel())) <= 1e-14
    full = orc.fftnd_real_full(list(dims), np.concatenate([x[:x.size // 2], np.zeros(x.size // 2)]),
                               -1)
    assert orc.rel_l2(full, n2c(np.fft.fftn(x[:x.size // 2].reshape(dims)).ravel())) <= 1e-14


# ---------------------------------------------------------------------------
# DCT / DST / DHT
# ---------------------------------------------------------------------------
DCT_SIZES = [2, 3, 4, 5, 8, 12, 16, 100, 128, 310, 512]


@pytest.mark.parametrize("n", DCT_SIZES)
def test_dct(orc, n):
    x = orc.jrandom_doubles(SEED, n)
    # scale=true: orthonormal DCT-II / DCT-III
    assert orc.rmse(orc.dct_forward(n, x, True), sfft.dct(x, 2, norm="ortho")) <= EPS
    assert orc.rmse(orc.dct_inverse(n, x, True), sfft.dct(x, 3, norm="ortho")) <= EPS
    # scale=false: power of two -> C_k (Ooura ddct), otherwise 2*C_k
    un = sfft.dct(x, 2)  # = 2*C_k
    f = 0.5 if L.is_pow2(n) else 1.0
    assert orc.rmse(orc.dct_forward(n, x, False), f * un) <= EPS * n
    # inverse unscaled: power of two -> a_0 + sum a_k cos(.), otherwise
    # (1/n)(a_0/2 + sum a_k cos(.)); scipy dct type 3 = a_0 + 2 sum a_k cos(.)
    d3 = sfft.dct(x, 3)
    exp = 0.5 * (d3 + x[0]) if L.is_pow2(n) else d3 / (2 * n)
    assert orc.rmse(orc.dct_inverse(n, x, False), exp) <= EPS * n
    rt = orc.dct_inverse(n, orc.dct_forward(n, x, True), True)
    assert orc.rmse(rt, x) <= EPS


@pytest.mark.parametrize("n", DCT_SIZES)
def test_dst(orc, n):
    x = orc.jrandom_doubles(SEED, n)
    assert orc.rmse(orc.dst_forward(n, x, True), sfft.dst(x, 2, norm="ortho")) <= EPS
    assert orc.rmse(orc.dst_inverse(n, x, True), sfft.dst(x, 3, norm="ortho")) <= EPS
    rt = orc.dst_inverse(n, orc.dst_forward(n, x, True), True)
    assert orc.rmse(rt, x) <= EPS


@pytest.mark.parametrize("n", [2, 3, 4, 7, 16, 100, 128])
def test_dht(orc, n):
    x = orc.jrandom_doubles(SEED, n)
    j = np.arange(n)
    ang = 2 * np.pi * np.outer(j, j) / n
    cas = np.cos(ang) + np.sin(ang)
    assert orc.rmse(orc.dht_forward(n, x), cas @ x) <= EPS
    assert orc.rmse(orc.dht_inverse(n, orc.dht_forward(n, x), True), x) <= EPS


@pytest.mark.parametrize("dims", [(4, 8), (6, 10), (4, 4, 8), (3, 5, 6)])
def test_multid_real(orc, dims):
    x = orc.jrandom_doubles(SEED, int(np.prod(dims)))
    X = x.reshape(dims)
    got = orc.realnd("dct", list(dims), x, True, True)
    assert orc.rmse(got, sfft.dctn(X, 2, norm="ortho").ravel()) <= EPS
    got = orc.realnd("dst", list(dims), x, True, True)
    assert orc.rmse(got, sfft.dstn(X, 2, norm="ortho").ravel()) <= EPS
    F = np.fft.fftn(X)
    got = orc.realnd("dht", list(dims), x, True, False)
    assert orc.rmse(got, (F.real - F.imag).ravel()) <= EPS
    inv = orc.realnd("dht", list(dims), got, False, True)
    assert orc.rmse(inv, x) <= EPS
