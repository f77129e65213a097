// Multi-dimensional transforms as sequences of batched axis passes
// (reference: DoubleFFT_2D.java:115-213 row pass + cdft2d_sub column pass;
// DoubleFFT_3D.java:145-325).  Axis passes run directly on the strided data
// when the axis length fits one LDS kernel (coalesced column mode for
// non-contiguous axes); longer axes are gathered into the workspace.
#pragma once
#include "jt_real1d.hpp"

namespace jt {

template <typename T>
__global__ void gather_lines(const cpx<T>* __restrict__ x, cpx<T>* __restrict__ y, LineMap lm, int64_t n) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i < lm.nlines * n; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t l = i / n, k = i % n;
        y[i] = x[lm.base(l) + k * lm.es];
    }
}
template <typename T>
__global__ void scatter_lines(const cpx<T>* __restrict__ y, cpx<T>* __restrict__ x, LineMap lm, int64_t n) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i < lm.nlines * n; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t l = i / n, k = i % n;
        x[lm.base(l) + k * lm.es] = y[i];
    }
}

// One axis pass of a complex transform over lines lm (in place).
template <typename T, int DIR>
void axis_c2c(const Fft1D<T>& p, cpx<T>* data, LineMap lm, T scale, cpx<T>* ws, hipStream_t st) {
    if (p.n == 1 || lm.nlines <= 0) {
        if (scale != T(1) && lm.nlines > 0) {
            // scaling only (n == 1 axis never occurs for 2D/3D plans)
        }
        return;
    }
    if (p.single()) {
        c2c_mapped<T, DIR>(p, data, data, lm, scale, st);
        return;
    }
    bool contig = lm.es == 1 && lm.s0 == p.n && lm.n0 >= lm.nlines;
    if (contig) {
        c2c_contig<T, DIR>(p, data, data, lm.nlines, scale, ws, st);
        return;
    }
    cpx<T>* g = ws;
    cpx<T>* ws2 = ws + lm.nlines * p.n;
    { JT_PROF(st, "gather_lines");
    hipLaunchKernelGGL((gather_lines<T>), ew_grid(lm.nlines * p.n), dim3(256), 0, st, data, g, lm, p.n);
    }
    JT_HIP(hipGetLastError());
    c2c_contig<T, DIR>(p, g, g, lm.nlines, scale, ws2, st);
    { JT_PROF(st, "scatter_lines");
    hipLaunchKernelGGL((scatter_lines<T>), ew_grid(lm.nlines * p.n), dim3(256), 0, st, g, data, lm, p.n);
    }
    JT_HIP(hipGetLastError());
}

template <typename T>
int64_t axis_ws(const Fft1D<T>& p, int64_t lines) {
    if (p.single()) return 0;
    return lines * p.n + p.workspace(lines);
}

// Column pass of a 2D realForward (ISGN > 0, DIR -1) or realInverse (ISGN < 0,
// DIR +1) with the rdft2d_sub fix of column 0 done by the workgroup that
// transforms it (one launch instead of two).  Only for a column plan that is
// one compile-time Stockham kernel; returns false, having launched nothing,
// for any other plan.
template <typename T, int DIR, int ISGN>
bool col_fft_fix2d(const Fft1D<T>& p, cpx<T>* z, int64_t rows, int64_t ch, T scale, hipStream_t st) {
    if (p.kind != K_TPL || p.mf || p.n != rows || ch <= 0) return false;
    IOStridedFix2D<T, ISGN> io{{z, z, LineMap{ch, ch, 1, 0, ch}, ch, scale, true}, rows};
    return dispatch_tpl_n<T, DIR, true>(rows, io, p.tw.template as<cpx<T>>(), st);
}

// rdft2d_sub (DoubleFFT_2D.java:2544-2574) on the packed column pair 0/1:
// isgn > 0 splits the column-0 transform of (Y[.][0] + i Y[.][c/2]) into the
// documented layout; isgn < 0 recombines it.
template <typename T>
__global__ void rdft2d_fix(T* __restrict__ a, int64_t rows, int64_t cols, int isgn) {
    int64_t i = 1 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i < rows / 2; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t j = rows - i;
        T* ai = a + i * cols;
        T* aj = a + j * cols;
        if (isgn > 0) {
            T ir = ai[0], ii = ai[1], jr = aj[0], ji = aj[1];
            aj[0] = T(0.5) * (ir - jr);
            ai[0] = ir - aj[0];
            aj[1] = T(0.5) * (ii + ji);
            ai[1] = ii - aj[1];
        } else {
            T ir = ai[0], ii = ai[1], jr = aj[0], ji = aj[1];
            ai[0] = ir + jr;
            aj[0] = ir - jr;
            ai[1] = ii + ji;
            aj[1] = ji - ii;
        }
    }
}

// rdft3d_sub (DoubleFFT_3D.java:6909-7021): the same split/recombination for
// the packed column pair of a (slices, rows) plane pair.
template <typename T>
__global__ void rdft3d_fix(T* __restrict__ a, int64_t slices, int64_t rows, int64_t cols, int isgn) {
    // work items: (i, k) with i in [0, slices), k in [0, rows/2]; each pairs
    // (i, k) with ((slices-i)%slices, (rows-k)%rows) once.
    const int64_t ss = rows * cols, rs = cols;
    const int64_t n1h = slices / 2, n2h = rows / 2;
    int64_t tot = slices * (n2h + 1);
    int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = idx / (n2h + 1), k = idx % (n2h + 1);
        int64_t j = (slices - i) % slices, l = (rows - k) % rows;
        // each unordered pair visited once: skip the mirror half
        bool self = (i == j) && (k == l);
        if (self) continue;
        if (k == 0 || k == n2h) {
            if (i > n1h) continue;  // (i,k) with (j,k): handle i < j only
            if (i == 0 || i == n1h) continue;
        }
        T* p = a + i * ss + k * rs;
        T* q = a + j * ss + l * rs;
        T pr = p[0], pi = p[1], qr = q[0], qi = q[1];
        if (isgn > 0) {
            q[0] = T(0.5) * (pr - qr);
            p[0] = pr - q[0];
            q[1] = T(0.5) * (pi + qi);
            p[1] = pi - q[1];
        } else {
            p[0] = pr + qr;
            q[0] = pr - qr;
            p[1] = pi + qi;
            q[1] = qi - pi;
        }
    }
}

}  // namespace jt
