// Ragged batches: many 1D complex transforms of unequal lengths, stored back
// to back, in ONE persistent launch.
//
// The reference transforms such a set with one DoubleFFT_1D per length, each
// picking SPLIT_RADIX / MIXED_RADIX / BLUESTEIN (DoubleFFT_1D.java:117-188).
// Here every length gets a per-length descriptor (radix plan, four-step split
// or Bluestein convolution length) and the whole set becomes a queue of work
// items executed by a persistent kernel:
//   SINGLE     transforms of one length that fits one workgroup's LDS
//   FOUR_A/B   the column / row halves of a four-step length (n > LDS limit)
//   BLUE_S     Bluestein transforms of one length whose convolution length m
//              fits LDS (premul, FFT_m, spectrum product, inverse FFT_m,
//              postmul in one residency)
//   BLUE_1..3  Bluestein with a four-step m: premul+columns; rows (forward
//              FFT, spectrum product, inverse FFT, twiddle in one residency);
//              inverse columns+postmul
// Every item runs the same round loop (lines into LDS, the stages, lines
// out), so the kernel holds one copy of the stage code.
// Workgroups take items in queue order with an atomic counter.  An item that
// consumes another phase's output waits on a per-transform counter; every
// producer of that phase sits earlier in the queue and was therefore taken by
// a running workgroup before the consumer, so the wait always ends.  Each wait
// is also bounded in time (an error flag records a timeout and the item is
// abandoned) so the grid always drains.  Counters are cleared by a kernel, not
// a memset, so the pair of launches can be captured in a hipGraph.
#pragma once
#include <algorithm>
#include <cstdlib>
#include <deque>
#include <functional>
#include <map>

#include "jt_generic.hpp"
#include "jt_fft1d.hpp"

namespace jt {

enum RKind { RK_SINGLE = 0, RK_FOUR = 1, RK_BLUE_S = 2, RK_BLUE_L = 3 };
enum RType { RT_SINGLE = 0, RT_FOUR_A = 1, RT_FOUR_B = 2, RT_BLUE_S = 3, RT_BLUE_1 = 4, RT_BLUE_2 = 5,
             RT_BLUE_3 = 6 };

struct RLen {
    int n;
    int kind;
    int nst;
    int radix[GEN_MAX_STAGES];
    int twoff[GEN_MAX_STAGES];  // offsets into the shared table (complex units)
    int n1, n2;                 // four-step split (of n, or of m for Bluestein)
    int l1, l2;                 // descriptor indices of the n1 / n2 lengths
    int64_t hioff, looff;       // four-step twiddles exp(-2 pi i e / N): hi[e>>8] lo[e&255]
    int m, lm;                  // Bluestein convolution length and its descriptor
    int64_t choff, bsoff;       // chirp (n entries), kernel spectrum (m entries)
    int tpl;                    // fewest threads per line for the in-place stages
};

// Loads in flight per thread when a round's lines are read into LDS, and
// points per thread per step of the stores.
constexpr int RLB = 8;
constexpr int RSB = 4;

// In-place stages hold at most 16 points per thread.
template <typename T>
constexpr int ragged_pts() { return 16; }

struct RItem {
    int type;
    int len;     // descriptor index
    int c0, cnt; // first column/row and count (four-step items)
    int64_t tr;  // transform index
};

template <typename T>
struct RArgs {
    cpx<T>* data;
    const int64_t* off;   // complex offset of each transform
    const int64_t* wsoff; // workspace offset of each transform (multi-phase ones)
    cpx<T>* ws;
    const RLen* lens;
    const cpx<T>* tw;
    const RItem* items;
    const int64_t* tlist;  // transform indices of the grouped whole-transform items
    int nitems;
    int* next;     // queue head
    int* cnt;      // 3 phase counters per transform
    int* err;      // set to 1 if a wait timed out
    int dir;       // -1 forward, +1 backward
    int scaled;    // multiply each transform by 1/n (scaled inverse)
    int tplx;      // threads per line = (the length's minimum, RLen::tpl) << tplx
    int wait_extra;          // 0; 1 makes every dependency target unreachable (failure-path test)
    uint64_t wait_ticks;     // bound of one wait, in 100 MHz wall-clock ticks
    int64_t ntr;
};

// The tunable schedule of a ragged plan (chosen by ragged_autotune):
//   wgs    persistent workgroups per CU, i.e. how the CU's LDS is split
//          (each workgroup holds lines in its share)
//   nt     tile shape: threads per workgroup (256, 512 or 1024)
//   tplx   tile shape: threads per line, (the length's minimum) << tplx,
//          so lines per round = nt / threads per line (within the LDS share)
//   chunk  lines per four-step work item, chunk * (single-kernel limit) points
//   order  radix order of every runtime plan: 0 = 16, 8/4/2, then odd primes
//          ascending; 1 = the reverse
//   lag    queue order of the multi-phase transforms (four-step, Bluestein
//          over a four-step m).  0: phase by phase (every first phase, then
//          every second, ...).  > 0: transform by transform, largest first,
//          each later phase queued once lag * 2^20 points of other work
//          follow the phase before it, so that it reads the workspace
//          written shortly before, from the last-level cache, rather than
//          from HBM, and its wait for that phase has normally ended
struct RTune {
    int wgs = 2, nt = 0, tplx = 0, chunk = 4, order = 0, lag = 0;  // nt 0: smallest that fits
    std::string str() const {
        return "wgs=" + std::to_string(wgs) + ",nt=" + std::to_string(nt) + ",tplx=" + std::to_string(tplx) +
               ",chunk=" + std::to_string(chunk) + ",order=" + std::to_string(order) + ",lag=" + std::to_string(lag);
    }
    bool operator==(const RTune& o) const {
        return wgs == o.wgs && nt == o.nt && tplx == o.tplx && chunk == o.chunk && order == o.order && lag == o.lag;
    }
};

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

// The index of the queue item, read from LDS, is the same in every lane; the
// compiler cannot tell, so it is moved to a scalar register explicitly.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Read-only tables of the plan (descriptors, items, offsets, twiddles) are
// read through the constant address space: the compiler then knows that the
// kernel's stores never change them, so their loads are hoisted and batched
// across stores, and loads at a wave-uniform address become scalar loads.
template <class X>
using ro_ptr = const __attribute__((address_space(4))) X*;
template <class X>
__device__ __forceinline__ ro_ptr<X> ro(const X* p) { return (ro_ptr<X>)p; }
template <typename T>
__device__ __forceinline__ cpx<T> rd(ro_ptr<cpx<T>> p, int64_t i) { return cpx<T>{p[i].x, p[i].y}; }

// One Stockham stage of radix R (span ns) in place on a line's LDS arrays:
// thread tl of the line's tpl threads gathers the inputs of its butterflies
// j = tl + b*tpl (b < PM/R; the host sizes tpl so that this covers all n/R)
// into registers, the workgroup synchronises, and the outputs overwrite the
// same arrays.  Every thread of the workgroup calls this (the barrier).
template <int R, int DIR, int PM, typename T>
__device__ __forceinline__ void stage_ip(T* re, T* im, ro_ptr<cpx<T>> tws, int n, int ns, int tl,
                                         int tpl, bool act) {
    constexpr int NB = PM / R;
    const int m = n / R;
    cpx<T> v[NB][R];
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int j = tl + b * tpl;
        if (act && j < m) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int k = gpos(j + r * m);
                v[b][r] = cpx<T>{re[k], im[k]};
            }
        }
    }
    __syncthreads();
    const float rns = 1.0f / (float)ns;
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int j = tl + b * tpl;
        if (act && j < m) {
            const int jh = fdiv(j, ns, rns);
            const int jm = j - jh * ns;
            if (ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) v[b][r] = cmul(v[b][r], tw_dir<DIR>(rd(tws, jm * (R - 1) + r - 1)));
            }
            dft_any<R, DIR>(v[b]);
            const int d = jh * ns * R + jm;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int k = gpos(d + r * ns);
                re[k] = v[b][r].x;
                im[k] = v[b][r].y;
            }
        }
    }
}

template <int DIR, int PM, typename T>
__device__ __forceinline__ void stage_ip_rt(int R, T* re, T* im, ro_ptr<cpx<T>> tws, int n, int ns,
                                            int tl, int tpl, bool act) {
    switch (R) {
        case 2: stage_ip<2, DIR, PM>(re, im, tws, n, ns, tl, tpl, act); break;
        case 3: stage_ip<3, DIR, PM>(re, im, tws, n, ns, tl, tpl, act); break;
        case 4: stage_ip<4, DIR, PM>(re, im, tws, n, ns, tl, tpl, act); break;
        case 5: stage_ip<5, DIR, PM>(re, im, tws, n, ns, tl, tpl, act); break;
        case 7: stage_ip<7, DIR, PM>(re, im, tws, n, ns, tl, tpl, act); break;
        case 8: stage_ip<8, DIR, PM>(re, im, tws, n, ns, tl, tpl, act); break;
        case 11: stage_ip<11, DIR, PM>(re, im, tws, n, ns, tl, tpl, act); break;
        case 13: stage_ip<13, DIR, PM>(re, im, tws, n, ns, tl, tpl, act); break;
        default: stage_ip<16, DIR, PM>(re, im, tws, n, ns, tl, tpl, act); break;
    }
}

// All stages of L in place, forward direction; the workgroup's threads all
// call this.  An inverse transform runs the same stages on conjugated data
// (the inverse DFT of x is conj(DFT(conj x))): the callers conjugate on the
// way into LDS and on the way out, so the kernel holds one copy of the stage
// code for both directions.
template <typename T>
__device__ __forceinline__ void stages_fwd(ro_ptr<RLen> L, ro_ptr<cpx<T>> tw, T* re, T* im, int tl,
                                           int tpl, bool act) {
    constexpr int PM = ragged_pts<T>();
    const int nst = L->nst, n = L->n;
    int ns = 1;
    for (int s = 0; s < nst; s++) {
        __syncthreads();
        const int R = L->radix[s];
        // an empty asm that "changes" the lane index: the LDS addresses of a
        // stage are then computed inside the stage, not hoisted out of the
        // loops and held in registers for every stage and round (which took
        // the kernel to 256 VGPRs plus scratch; with it, 127 and none)
        int tls = tl;
        asm volatile("" : "+v"(tls));
        stage_ip_rt<-1, PM>(R, re, im, tw + L->twoff[s], n, ns, tls, tpl, act);
        ns *= R;
    }
    __syncthreads();
}

// Loading lines into LDS.  Line `lg` of the round (lg < nl) occupies
// sm[lg*npl + gpos(k)] (real parts) and sm[np + lg*npl + gpos(k)] (imaginary
// parts).  cj conjugates on the way in.
// Lanes across lines: consecutive threads take consecutive lines of one
// point (adjacent columns in memory); RLB loads in flight per thread.
template <typename T, class F>
__device__ __forceinline__ void rload_across(T* sm, int np, int npl, int n, int nl, int base, bool cj, F load) {
    const int t = threadIdx.x, NT = blockDim.x;
    const int tot = nl * n;
    const float rnl = 1.0f / (float)nl;
    for (int i0 = t; i0 < tot; i0 += RLB * NT) {
        cpx<T> v[RLB];
#pragma unroll
        for (int u = 0; u < RLB; u++) {
            const int idx = i0 + u * NT;
            if (idx < tot) {
                const int k = fdiv(idx, nl, rnl);
                v[u] = load(base + idx - k * nl, k);
            }
        }
#pragma unroll
        for (int u = 0; u < RLB; u++) {
            const int idx = i0 + u * NT;
            if (idx < tot) {
                const int k = fdiv(idx, nl, rnl), lg = idx - k * nl;
                sm[lg * npl + gpos(k)] = v[u].x;
                sm[np + lg * npl + gpos(k)] = cj ? -v[u].y : v[u].y;
            }
        }
    }
}

// Lanes along a line: the line's tpl threads take consecutive points.
template <typename T, class F>
__device__ __forceinline__ void rload_along(T* are, T* aim, int n, int line, int tl, int tpl, bool act, bool cj,
                                            F load) {
    if (!act) return;
    for (int k0 = tl; k0 < n; k0 += RLB * tpl) {
        cpx<T> v[RLB];
#pragma unroll
        for (int u = 0; u < RLB; u++) {
            const int k = k0 + u * tpl;
            if (k < n) v[u] = load(line, k);
        }
#pragma unroll
        for (int u = 0; u < RLB; u++) {
            const int k = k0 + u * tpl;
            if (k < n) {
                are[gpos(k)] = v[u].x;
                aim[gpos(k)] = cj ? -v[u].y : v[u].y;
            }
        }
    }
}

// Stores, RSB per thread at a time: the table loads of a store epilogue
// (twiddles, chirp) of the RSB points are then issued together.
template <typename T, class F>
__device__ __forceinline__ void rstore_across(const T* sm, int np, int npl, int n, int nl, int base, bool cj,
                                              F store) {
    const int t = threadIdx.x, NT = blockDim.x;
    const int tot = nl * n;
    const float rnl = 1.0f / (float)nl;
    for (int i0 = t; i0 < tot; i0 += RSB * NT) {
#pragma unroll
        for (int u = 0; u < RSB; u++) {
            const int idx = i0 + u * NT;
            if (idx < tot) {
                const int k = fdiv(idx, nl, rnl), lg = idx - k * nl;
                const T y = sm[np + lg * npl + gpos(k)];
                store(base + lg, k, cpx<T>{sm[lg * npl + gpos(k)], cj ? -y : y});
            }
        }
    }
}

template <typename T, class F>
__device__ __forceinline__ void rstore_along(const T* are, const T* aim, int n, int line, int tl, int tpl, bool act,
                                             bool cj, F store) {
    if (!act) return;
    for (int k0 = tl; k0 < n; k0 += RSB * tpl) {
#pragma unroll
        for (int u = 0; u < RSB; u++) {
            const int k = k0 + u * tpl;
            if (k < n) {
                const T y = aim[gpos(k)];
                store(line, k, cpx<T>{are[gpos(k)], cj ? -y : y});
            }
        }
    }
}

template <typename T>
__device__ __forceinline__ cpx<T> rtw2(int dir, ro_ptr<cpx<T>> tw, ro_ptr<RLen> L, int64_t e) {
    // e = j2*k1 < n2*n1: no reduction needed
    cpx<T> w = cmul(rd(tw, L->hioff + (e >> 8)), rd(tw, L->looff + (e & 255)));
    return dir < 0 ? w : cpx<T>{w.x, -w.y};
}

// Wait until *c >= target (thread 0 polls, the workgroup then proceeds).
// Polls are relaxed; one agent-scope acquire fence by thread 0 after the
// condition holds makes the producers' data visible to the whole workgroup
// (the caches it invalidates are shared by the workgroup's waves).
// Bounded: after `ticks` of wall clock (2 s normally) the wait gives up and
// flags an error.
__device__ __forceinline__ bool rwait(int* c, int target, int* err, uint64_t ticks) {
    __shared__ int ok;
    if (threadIdx.x == 0) {
        int good = 1;
        uint64_t t0 = wall_clock64();
        while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(8);
            if (wall_clock64() - t0 > ticks) {  // 100 MHz counter
                good = 0;
                atomicExch(err, 1);
                break;
            }
        }
        __atomic_thread_fence(__ATOMIC_ACQUIRE);  // agent scope by default in HIP
        ok = good;
    }
    __syncthreads();
    return ok != 0;
}

// Every thread waits for its own stores, then after the barrier one
// agent-scope release fence and the counter increment by thread 0.
__device__ __forceinline__ void rsignal(int* c, int add) {
    __builtin_amdgcn_s_waitcnt(0);  // this thread's stores have reached L2
    __syncthreads();
    if (threadIdx.x == 0) {
        __atomic_thread_fence(__ATOMIC_RELEASE);
        __hip_atomic_fetch_add(c, add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <typename T>
__global__ void ragged_reset(RArgs<T> a) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i < 3 * a.ntr; i += (int64_t)gridDim.x * blockDim.x) a.cnt[i] = 0;
    // (the failure flag is not cleared here: it accumulates over launches
    // until the host reads and clears it, so a timeout is never lost to a
    // later clean launch)
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.next = 0;
}

template <typename T, int NTMAX>
__global__ void __launch_bounds__(NTMAX) ragged_kernel(RArgs<T> a, int np) {
    extern __shared__ unsigned char smem_raw[];
    T* sm = reinterpret_cast<T*>(smem_raw);
    __shared__ int cur;
    const int t = threadIdx.x, NT = blockDim.x;
    const int dir = a.dir;
    for (;;) {
        __syncthreads();
        if (t == 0) cur = atomicAdd(a.next, 1);
        __syncthreads();
        const int it = uni(cur);
        if (it >= a.nitems) break;  // uniform across the workgroup
        ro_ptr<RItem> item = ro(a.items) + it;
        const int type = item->type;
        const int64_t tr = item->tr;
        const int c0 = item->c0, cnt = item->cnt;
        ro_ptr<RLen> lens = ro(a.lens);
        ro_ptr<RLen> L = lens + item->len;
        // B: the four-step length (of the transform, or of its Bluestein
        // convolution); for RT_BLUE_S the in-LDS convolution length
        ro_ptr<RLen> B = type >= RT_BLUE_S ? lens + L->lm : L;
        const int n1 = B->n1, n2 = B->n2;
        int* c3 = a.cnt + 3 * tr;
        // an item that consumes another phase's output waits for all of it
        if (type == RT_FOUR_B || type == RT_BLUE_2) {
            if (!rwait(c3 + 0, n2 + a.wait_extra, a.err, a.wait_ticks)) continue;
        } else if (type == RT_BLUE_3) {
            if (!rwait(c3 + 1, n1 + a.wait_extra, a.err, a.wait_ticks)) continue;
        }
        // the lines this item transforms: SINGLE / BLUE_S whole transforms
        // (listed from tlist[c0]); four-step items columns (FFT_n1) or rows
        // (FFT_n2) c0 .. c0+cnt-1 of the n1 x n2 view
        ro_ptr<RLen> LL = type == RT_SINGLE                         ? L
                          : type == RT_BLUE_S                       ? B
                          : (type == RT_FOUR_B || type == RT_BLUE_2) ? lens + B->l2
                                                                    : lens + B->l1;
        // Bluestein's convolution: forward pass, spectrum product, inverse pass
        const int npass = (type == RT_BLUE_2 || type == RT_BLUE_S) ? 2 : 1;
        const int d1 = (type == RT_BLUE_1 || npass == 2) ? -1 : type == RT_BLUE_3 ? 1 : dir;
        const bool cj_in = d1 > 0, cj_out = (npass == 2) || d1 > 0;
        const int nreal = L->n;
        const T scale = a.scaled ? T(1) / T(nreal) : T(1);
        ro_ptr<cpx<T>> tw = ro(a.tw);
        ro_ptr<int64_t> off = ro(a.off);
        ro_ptr<int64_t> tl0 = ro(a.tlist) + c0;
        cpx<T>* x = a.data + off[tr];
        cpx<T>* wsA = a.ws + ro(a.wsoff)[tr];
        ro_ptr<cpx<T>> chirp = tw + L->choff;
        ro_ptr<cpx<T>> bspec = tw + L->bsoff;
        cpx<T>* data = a.data;

        // G lines at a time, each owned by tpl threads and a slice of the two
        // LDS arrays (re, im; np reals each); when LDS or the item caps the
        // line count, the spare threads widen each line
        const int n = LL->n;
        const int npl = n + (n >> 5) + 1;
        int tpl = min(NT, LL->tpl << a.tplx);
        int G = min(min(NT / tpl, np / npl), cnt);
        if (G < 1) G = 1;
        while (2 * tpl * G <= NT && tpl < n) tpl <<= 1;
        const int g = t / tpl, tl_ = t % tpl;
        T* are = sm + g * npl;
        T* aim = sm + np + g * npl;
        for (int base = 0; base < cnt; base += G) {
            int tl = tl_;
            asm volatile("" : "+v"(tl));  // as in stages_fwd: no addresses held across rounds
            const int line = base + g;
            const bool act = g < G && line < cnt;
            const int nl = min(G, cnt - base);
            __syncthreads();
            switch (type) {
                case RT_SINGLE:
                    rload_along<T>(are, aim, n, line, tl, tpl, act, cj_in,
                                   [=](int ln, int k) { return data[off[tl0[ln]] + k]; });
                    break;
                case RT_FOUR_A:
                    rload_across<T>(sm, np, npl, n, nl, base, cj_in,
                                    [=](int ln, int j1) { return x[j1 * n2 + (c0 + ln)]; });
                    break;
                case RT_BLUE_1:
                    // chirp premultiplied, zero-padded input
                    rload_across<T>(sm, np, npl, n, nl, base, false, [=](int ln, int j1) {
                        const int e = j1 * n2 + (c0 + ln);
                        if (e >= nreal) return cpx<T>{T(0), T(0)};
                        return dir < 0 ? cmulc(x[e], rd(chirp, e)) : cmul(x[e], rd(chirp, e));
                    });
                    break;
                case RT_FOUR_B:
                case RT_BLUE_2:
                    rload_along<T>(are, aim, n, line, tl, tpl, act, cj_in,
                                   [=](int ln, int j2) { return wsA[(c0 + ln) * n2 + j2]; });
                    break;
                case RT_BLUE_3:
                    rload_across<T>(sm, np, npl, n, nl, base, cj_in,
                                    [=](int ln, int k1) { return wsA[k1 * n2 + (c0 + ln)]; });
                    break;
                default:  // RT_BLUE_S: m points, the first n chirp-premultiplied
                    rload_along<T>(are, aim, n, line, tl, tpl, act, false, [=](int ln, int k) {
                        if (k >= nreal) return cpx<T>{T(0), T(0)};
                        const cpx<T> v = data[off[tl0[ln]] + k];
                        return dir < 0 ? cmulc(v, rd(chirp, k)) : cmul(v, rd(chirp, k));
                    });
                    break;
            }
#pragma clang loop unroll(disable)
            for (int p = 0; p < npass; p++) {
                if (p == 1 && act) {
                    // kernel spectrum product (RT_BLUE_2: point k2 of row k1
                    // is element k1 + n1*k2 of the spectrum), conjugated for
                    // the inverse pass
                    for (int k = tl; k < n; k += tpl) {
                        cpx<T> v = {are[gpos(k)], aim[gpos(k)]};
                        const cpx<T> bb = rd(bspec, type == RT_BLUE_2 ? (c0 + line) + n1 * k : k);
                        v = dir < 0 ? cmul(v, bb) : cmulc(v, bb);
                        are[gpos(k)] = v.x;
                        aim[gpos(k)] = -v.y;
                    }
                }
                stages_fwd<T>(LL, tw, are, aim, tl, tpl, act);
            }
            switch (type) {
                case RT_SINGLE:
                    rstore_along<T>(are, aim, n, line, tl, tpl, act, cj_out, [=](int ln, int k, cpx<T> v) {
                        data[off[tl0[ln]] + k] = cpx<T>{v.x * scale, v.y * scale};
                    });
                    break;
                case RT_FOUR_A:
                case RT_BLUE_1:
                    // times exp(d1 2 pi i j2 k1 / N), into the workspace
                    rstore_across<T>(sm, np, npl, n, nl, base, cj_out, [=](int ln, int k1, cpx<T> v) {
                        const int j2 = c0 + ln;
                        wsA[k1 * n2 + j2] = cmul(v, rtw2<T>(d1, tw, B, (int64_t)j2 * k1));
                    });
                    break;
                case RT_FOUR_B:
                    // output point k2 of row k1 is element k1 + n1*k2
                    rstore_across<T>(sm, np, npl, n, nl, base, cj_out, [=](int ln, int k2, cpx<T> v) {
                        x[(c0 + ln) + n1 * k2] = cpx<T>{v.x * scale, v.y * scale};
                    });
                    break;
                case RT_BLUE_2:
                    // back to j2, times exp(+2 pi i j2 k1 / N), in place
                    rstore_along<T>(are, aim, n, line, tl, tpl, act, cj_out, [=](int ln, int j2, cpx<T> v) {
                        const int k1 = c0 + ln;
                        wsA[k1 * n2 + j2] = cmul(v, rtw2<T>(1, tw, B, (int64_t)j2 * k1));
                    });
                    break;
                case RT_BLUE_3:
                    // point j1*n2 + j2 of the convolution; chirp postmul of
                    // the first n points into the transform
                    rstore_across<T>(sm, np, npl, n, nl, base, cj_out, [=](int ln, int j1, cpx<T> v) {
                        const int K = j1 * n2 + (c0 + ln);
                        if (K < nreal) {
                            const cpx<T> c = rd(chirp, K);
                            v = dir < 0 ? cmulc(v, c) : cmul(v, c);
                            x[K] = cpx<T>{v.x * scale, v.y * scale};
                        }
                    });
                    break;
                default:  // RT_BLUE_S: chirp postmul of the first n points
                    rstore_along<T>(are, aim, nreal, line, tl, tpl, act, cj_out, [=](int ln, int k, cpx<T> v) {
                        const cpx<T> c = rd(chirp, k);
                        v = dir < 0 ? cmulc(v, c) : cmul(v, c);
                        data[off[tl0[ln]] + k] = cpx<T>{v.x * scale, v.y * scale};
                    });
                    break;
            }
        }
        if (type == RT_FOUR_A || type == RT_BLUE_1) rsignal(c3 + 0, cnt);
        else if (type == RT_BLUE_2) rsignal(c3 + 1, cnt);
    }
}

// ---------------------------------------------------------------------------
// host plan
// ---------------------------------------------------------------------------
template <typename T>
struct RaggedPlan {
    int64_t ntr = 0, total = 0;  // transforms, total complex points
    int maxn = 0;                // largest LDS line length
    int nitems = 0;
    std::vector<int64_t> lens_of;  // per transform
    DevBuf d_lens, d_tw, d_items, d_tlist, d_off, d_wsoff, d_cnt, d_next, d_err, ws;
    int grid = 0;
    size_t shm = 0;
    int np = 0;
    RTune tune;
    // hipGraph cache: one executable graph per (data, stream, dir, scale)
    struct GKey {
        void* data;
        hipStream_t st;
        int dir;
        bool scale;
    };
    std::vector<std::pair<GKey, hipGraphExec_t>> graphs;
    ~RaggedPlan() {
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    }
};

// four-step split n = n1*n2 with both factors <= lim and smooth, balanced
inline bool ragged_split(int64_t n, int64_t lim, int64_t& n1, int64_t& n2) {
    int64_t best = 0;
    for (int64_t a = 2; a <= lim; a++) {
        if (n % a) continue;
        int64_t b = n / a;
        if (b > lim) continue;
        std::vector<int> ra, rb;
        if (!factor_radices(a, ra) || !factor_radices(b, rb)) continue;
        if (!best || std::llabs(a - b) < std::llabs(best - n / best)) best = a;
    }
    if (!best) return false;
    n1 = best;
    n2 = n / best;
    return true;
}

// Kernel spectrum of Bluestein: FFT_m of the symmetric chirp, / m, on the
// GPU in FP64 (the same engine), returned to the host.
inline std::vector<cpx<double>> bluestein_spectrum(int64_t n, int64_t m, const std::vector<cpx<double>>& chirp) {
    std::vector<cpx<double>> b(m, cpx<double>{0.0, 0.0});
    for (int64_t j = 0; j < n; j++) {
        b[j] = chirp[j];
        if (j) b[m - j] = chirp[j];
    }
    auto p = make_fft1d<double>(m);
    DevBuf d, w;
    upload(d, b);
    w.alloc((size_t)std::max<int64_t>(p->workspace(1), 1) * sizeof(cpx<double>));
    c2c_contig<double, -1>(*p, d.as<cpx<double>>(), d.as<cpx<double>>(), 1, 1.0 / (double)m,
                           w.as<cpx<double>>(), 0);
    JT_HIP(hipDeviceSynchronize());
    JT_HIP(hipMemcpy(b.data(), d.p, m * sizeof(cpx<double>), hipMemcpyDeviceToHost));
    return b;
}

// the kernel instantiation for a workgroup size (each has its own register budget)
template <typename T>
const void* ragged_kernel_ptr(int nt) {
    if (nt == 1024) return (const void*)ragged_kernel<T, 1024>;
    if (nt == 512) return (const void*)ragged_kernel<T, 512>;
    if (nt == 256) return (const void*)ragged_kernel<T, 256>;
    throw JtError(ST_BAD_ARG, "ragged workgroup size must be 256, 512 or 1024");
}

// Bluestein kernel spectra by length, shared by the candidate plans of a tuning run
// (chirp, spectrum) per Bluestein length
using SpecCache = std::map<int64_t, std::pair<std::vector<cpx<double>>, std::vector<cpx<double>>>>;

inline void radix_order(std::vector<int>& r, int order) {
    if (order == 1) std::reverse(r.begin(), r.end());
}

template <typename T>
std::unique_ptr<RaggedPlan<T>> make_ragged(const int64_t* lengths, int64_t count, const RTune& tune = RTune(),
                                           SpecCache* spec_cache = nullptr) {
    auto P = std::make_unique<RaggedPlan<T>>();
    P->ntr = count;
    P->tune = tune;
    // longest line kept in one workgroup's LDS (re and im arrays)
    const int64_t lim = sizeof(T) == 8 ? 4096 : 8192;
    if (tune.wgs < 1 || tune.wgs > 8 || !(tune.nt == 0 || tune.nt == 256 || tune.nt == 512 || tune.nt == 1024) ||
        tune.tplx < 0 || tune.tplx > 4 || tune.chunk < 1 || tune.chunk > 1024 || tune.order < 0 || tune.order > 1 ||
        tune.lag < 0 || tune.lag > (1 << 20))
        throw JtError(ST_BAD_ARG, "ragged schedule out of range (wgs 1..8, nt 0/256/512/1024, tplx 0..4, "
                                  "chunk 1..1024, order 0/1, lag 0..2^20): " + tune.str());
    std::vector<RLen> lens;
    std::vector<cpx<T>> tab;
    int max_tpl = 1;                     // largest threads-per-line requirement
    std::map<int64_t, int> idx_single;  // single-kernel descriptor by n
    std::map<int64_t, int> idx_len;     // top-level descriptor by n
    auto add_twiddles = [&](const std::vector<int>& r, RLen& L) {
        std::vector<int> offs;
        auto tw = stage_twiddles<T>(r, &offs);
        int64_t base = (int64_t)tab.size();
        tab.insert(tab.end(), tw.begin(), tw.end());
        L.nst = (int)r.size();
        for (size_t i = 0; i < r.size(); i++) {
            L.radix[i] = r[i];
            L.twoff[i] = (int)(base + offs[i]);
        }
    };
    std::function<int(int64_t)> single_of = [&](int64_t n) -> int {
        auto it = idx_single.find(n);
        if (it != idx_single.end()) return it->second;
        RLen L{};
        L.n = (int)n;
        L.kind = RK_SINGLE;
        std::vector<int> r;
        if (!factor_radices(n, r)) throw JtError(ST_UNSUPPORTED, "internal: non-smooth single length");
        radix_order(r, tune.order);
        add_twiddles(r, L);
        // threads per line so that every stage's butterflies fit the
        // in-place stage's register budget (ragged_pts points per thread)
        L.tpl = 1;
        for (int R : r) {
            const int per = std::max(1, ragged_pts<T>() / R);  // butterflies per thread
            const int m = (int)(n / R);
            L.tpl = std::max(L.tpl, (m + per - 1) / per);
        }
        max_tpl = std::max(max_tpl, L.tpl);
        lens.push_back(L);
        P->maxn = std::max<int>(P->maxn, (int)n);
        return idx_single[n] = (int)lens.size() - 1;
    };
    auto four_of = [&](int64_t n, RLen& L) {
        int64_t n1, n2;
        if (!ragged_split(n, lim, n1, n2)) throw JtError(ST_UNSUPPORTED, "length too large for the ragged plan");
        L.n1 = (int)n1;
        L.n2 = (int)n2;
        int a1 = single_of(n1), a2 = single_of(n2);
        L.l1 = a1;
        L.l2 = a2;
        L.hioff = (int64_t)tab.size();
        for (int64_t i = 0; i <= (n >> 8); i++) tab.push_back(expm<T>(i << 8, n));
        L.looff = (int64_t)tab.size();
        for (int64_t i = 0; i < 256; i++) tab.push_back(expm<T>(i, n));
    };
    std::function<int(int64_t)> len_of = [&](int64_t n) -> int {
        auto it = idx_len.find(n);
        if (it != idx_len.end()) return it->second;
        std::vector<int> r;
        int id;
        if (factor_radices(n, r) && n <= lim) {
            id = single_of(n);
        } else if (factor_radices(n, r)) {
            RLen L{};
            L.n = (int)n;
            L.kind = RK_FOUR;
            four_of(n, L);
            lens.push_back(L);
            id = (int)lens.size() - 1;
        } else {
            RLen L{};
            L.n = (int)n;
            int64_t m = next_smooth(2 * n - 1);
            L.m = (int)m;
            L.kind = m <= lim ? RK_BLUE_S : RK_BLUE_L;
            std::vector<cpx<double>> ch, bs;
            if (spec_cache && spec_cache->count(n)) {
                ch = (*spec_cache)[n].first;
                bs = (*spec_cache)[n].second;
            } else {
                // chirp exp(i pi (j^2 mod 2n)/n) by the integer recurrence
                ch.resize(n);
                const long double PI = 3.141592653589793238462643383279502884L;
                int64_t k = 0;
                for (int64_t j = 0; j < n; j++) {
                    if (j) {
                        k += 2 * j - 1;
                        if (k >= 2 * n) k -= 2 * n;
                    }
                    long double ang = PI * (long double)k / (long double)n;
                    ch[j] = cpx<double>{(double)cosl(ang), (double)sinl(ang)};
                }
                bs = bluestein_spectrum(n, m, ch);
                if (spec_cache) (*spec_cache)[n] = std::make_pair(ch, bs);
            }
            L.choff = (int64_t)tab.size();
            for (auto& c : ch) tab.push_back(cpx<T>{(T)c.x, (T)c.y});
            L.bsoff = (int64_t)tab.size();
            for (auto& c : bs) tab.push_back(cpx<T>{(T)c.x, (T)c.y});
            if (L.kind == RK_BLUE_S) {
                L.lm = single_of(m);
            } else {
                RLen M{};
                M.n = (int)m;
                M.kind = RK_FOUR;
                four_of(m, M);
                lens.push_back(M);
                L.lm = (int)lens.size() - 1;
            }
            lens.push_back(L);
            id = (int)lens.size() - 1;
        }
        return idx_len[n] = id;
    };

    // per-transform descriptors, offsets, workspace
    std::vector<int> tlen(count);
    std::vector<int64_t> off(count), wsoff(count, 0);
    int64_t tot = 0, wtot = 0;
    for (int64_t i = 0; i < count; i++) {
        if (lengths[i] < 1) throw JtError(ST_BAD_SIZE, "n must be greater than 0");
        tlen[i] = len_of(lengths[i]);
        off[i] = tot;
        tot += lengths[i];
        const RLen& L = lens[tlen[i]];
        if (L.kind == RK_FOUR) {
            wsoff[i] = wtot;
            wtot += L.n;
        } else if (L.kind == RK_BLUE_L) {
            wsoff[i] = wtot;
            wtot += (int64_t)L.m;
        }
    }
    P->total = tot;
    P->lens_of.assign(lengths, lengths + count);

    // work items: the phases of the multi-phase transforms, and whole-transform
    // items ordered by cost
    std::vector<RItem> ph[3];
    struct Multi {
        std::vector<RItem> ph[3];
        int nph;
        int64_t pts;  // points one phase reads or writes (n, or m for Bluestein)
    };
    std::vector<Multi> multi;
    // four-step items carry ~4 LDS loads of lines each
    const int CHUNK = (int)(tune.chunk * lim);
    auto split_items = [&](std::vector<RItem>& v, int type, int len, int64_t tr, int total_lines, int line_len) {
        int per = std::max(1, CHUNK / std::max(1, line_len));
        for (int c = 0; c < total_lines; c += per) v.push_back(RItem{type, len, c, std::min(per, total_lines - c), tr});
    };
    std::vector<std::pair<double, RItem>> sgl;
    std::map<int, std::vector<int64_t>> by_len;       // single-kernel transforms grouped by length
    std::map<int, std::vector<int64_t>> by_len_blue;  // in-LDS Bluestein transforms grouped by length
    std::vector<int64_t> tlist;
    for (int64_t i = 0; i < count; i++) {
        const RLen& L = lens[tlen[i]];
        if (L.kind == RK_SINGLE) {
            by_len[tlen[i]].push_back(i);
        } else if (L.kind == RK_BLUE_S) {
            by_len_blue[tlen[i]].push_back(i);
        } else if (L.kind == RK_FOUR) {
            Multi mt;
            mt.nph = 2;
            mt.pts = L.n;
            split_items(mt.ph[0], RT_FOUR_A, tlen[i], i, L.n2, L.n1);
            split_items(mt.ph[1], RT_FOUR_B, tlen[i], i, L.n1, L.n2);
            multi.push_back(std::move(mt));
        } else {
            const RLen& M = lens[L.lm];
            Multi mt;
            mt.nph = 3;
            mt.pts = M.n;
            split_items(mt.ph[0], RT_BLUE_1, tlen[i], i, M.n2, M.n1);
            split_items(mt.ph[1], RT_BLUE_2, tlen[i], i, M.n1, M.n2);
            split_items(mt.ph[2], RT_BLUE_3, tlen[i], i, M.n2, M.n1);
            multi.push_back(std::move(mt));
        }
    }
    // whole-transform items: up to CHUNK points of one length (the line
    // length m for Bluestein), listed from tlist[c0]
    auto group_items = [&](const std::map<int, std::vector<int64_t>>& groups, int type) {
        for (auto& kv : groups) {
            const RLen& L = lens[kv.first];
            const int ln = type == RT_BLUE_S ? L.m : L.n;
            const double w = type == RT_BLUE_S ? 3.0 : 1.0;  // two FFT_m passes plus the chirp products
            const int per = std::max(1, (int)(CHUNK / std::max(1, ln)));
            const auto& v = kv.second;
            for (size_t c = 0; c < v.size(); c += per) {
                int cnt = (int)std::min<size_t>(per, v.size() - c);
                RItem it{type, kv.first, (int)tlist.size(), cnt, v[c]};
                for (int q = 0; q < cnt; q++) tlist.push_back(v[c + q]);
                sgl.push_back({w * cnt * ln * std::log2((double)ln + 1), it});
            }
        }
    };
    group_items(by_len, RT_SINGLE);
    group_items(by_len_blue, RT_BLUE_S);
    std::stable_sort(sgl.begin(), sgl.end(), [](auto& a, auto& b) { return a.first > b.first; });
    // Every later phase of a transform sits after its earlier phases in the
    // queue (the waits of the kernel rely on it).
    std::vector<RItem> items;
    if (tune.lag <= 0) {
        // phase by phase, a third of the whole-transform items after each
        for (auto& mt : multi)
            for (int p = 0; p < mt.nph; p++) ph[p].insert(ph[p].end(), mt.ph[p].begin(), mt.ph[p].end());
        const size_t q = (sgl.size() + 2) / 3;
        for (int p = 0; p < 3; p++) {
            items.insert(items.end(), ph[p].begin(), ph[p].end());
            for (size_t s = p * q; s < std::min(sgl.size(), (p + 1) * q); s++) items.push_back(sgl[s].second);
        }
    } else {
        const int64_t LAG = (int64_t)tune.lag << 20;
        std::vector<size_t> ord(multi.size());
        for (size_t k = 0; k < ord.size(); k++) ord[k] = k;
        std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return multi[a].pts > multi[b].pts; });
        struct Pend {
            int64_t mark;  // points queued when the phase before it was queued
            size_t idx;
            int phase;
        };
        std::deque<Pend> pend;
        int64_t queued = 0;
        size_t next_sgl = 0;
        auto emit = [&](size_t idx, int p) {
            const Multi& mt = multi[idx];
            items.insert(items.end(), mt.ph[p].begin(), mt.ph[p].end());
            queued += mt.pts;
            if (p + 1 < mt.nph) pend.push_back(Pend{queued, idx, p + 1});
        };
        auto ready = [&]() {
            while (!pend.empty() && queued - pend.front().mark >= LAG) {
                const Pend q = pend.front();
                pend.pop_front();
                emit(q.idx, q.phase);
            }
        };
        for (size_t k = 0; k < ord.size(); k++) {
            emit(ord[k], 0);
            ready();
        }
        // the remaining later phases, with whole-transform items as filler
        while (!pend.empty()) {
            while (next_sgl < sgl.size() && queued - pend.front().mark < LAG) {
                const RItem& it = sgl[next_sgl++].second;
                items.push_back(it);
                queued += (int64_t)it.cnt * (it.type == RT_BLUE_S ? lens[it.len].m : lens[it.len].n);
            }
            const Pend q = pend.front();
            pend.pop_front();
            emit(q.idx, q.phase);
        }
        for (; next_sgl < sgl.size(); next_sgl++) items.push_back(sgl[next_sgl].second);
    }
    P->nitems = (int)items.size();

    // upload
    {
        DevBuf& d = P->d_lens;
        d.alloc(lens.size() * sizeof(RLen));
        JT_HIP(hipMemcpy(d.p, lens.data(), d.bytes, hipMemcpyHostToDevice));
    }
    upload(P->d_tw, tab);
    P->d_items.alloc(std::max<size_t>(1, items.size()) * sizeof(RItem));
    if (!items.empty()) JT_HIP(hipMemcpy(P->d_items.p, items.data(), items.size() * sizeof(RItem), hipMemcpyHostToDevice));
    P->d_tlist.alloc(std::max<size_t>(1, tlist.size()) * sizeof(int64_t));
    if (!tlist.empty())
        JT_HIP(hipMemcpy(P->d_tlist.p, tlist.data(), tlist.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    P->d_off.alloc(count * sizeof(int64_t));
    JT_HIP(hipMemcpy(P->d_off.p, off.data(), count * sizeof(int64_t), hipMemcpyHostToDevice));
    P->d_wsoff.alloc(count * sizeof(int64_t));
    JT_HIP(hipMemcpy(P->d_wsoff.p, wsoff.data(), count * sizeof(int64_t), hipMemcpyHostToDevice));
    P->d_cnt.alloc(3 * count * sizeof(int));
    P->d_next.alloc(sizeof(int));
    P->d_err.alloc(sizeof(int));
    JT_HIP(hipMemset(P->d_err.p, 0, sizeof(int)));
    P->ws.alloc(std::max<int64_t>(wtot, 1) * sizeof(cpx<T>));

    // launch shape: LDS for two buffers of the longest line; one persistent
    // workgroup per resident slot
    // (the LDS is sized to the budget of two workgroups per CU, not to the
    // longest line: short lines then run many to a workgroup)
    int mx = std::max(P->maxn, 16);
    const int need = mx + (mx >> 5) + 1;
    if (P->tune.nt == 0) {
        P->tune.nt = 256;
        while (P->tune.nt < max_tpl && P->tune.nt < 1024) P->tune.nt *= 2;
    }
    if (max_tpl > P->tune.nt) throw JtError(ST_UNSUPPORTED, "workgroup too small for the longest line");
    const int budget = (160 * 1024 - 2048) / std::max(1, tune.wgs) / (2 * (int)sizeof(T));
    if (budget < need) throw JtError(ST_UNSUPPORTED, "LDS share too small for the longest line");
    P->np = budget;
    P->shm = 2 * (size_t)P->np * sizeof(T);
    const void* kfn = ragged_kernel_ptr<T>(P->tune.nt);
    JT_HIP(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)P->shm));
    int dev = 0, cus = 0, per = 0;
    JT_HIP(hipGetDevice(&dev));
    JT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    JT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kfn, P->tune.nt, P->shm));
    if (per < 1) throw JtError(ST_UNSUPPORTED, "ragged schedule does not fit a CU");
    P->grid = std::max(1, cus * std::max(1, std::min(per, tune.wgs)));
    return P;
}

template <typename T>
RArgs<T> ragged_args(RaggedPlan<T>& P, cpx<T>* data, int dir, bool scaled) {
    RArgs<T> a;
    a.data = data;
    a.off = P.d_off.template as<int64_t>();
    a.wsoff = P.d_wsoff.template as<int64_t>();
    a.ws = P.ws.template as<cpx<T>>();
    a.lens = P.d_lens.template as<RLen>();
    a.tw = P.d_tw.template as<cpx<T>>();
    a.items = P.d_items.template as<RItem>();
    a.tlist = P.d_tlist.template as<int64_t>();
    a.nitems = P.nitems;
    a.next = P.d_next.template as<int>();
    a.cnt = P.d_cnt.template as<int>();
    a.err = P.d_err.template as<int>();
    a.dir = dir;
    a.scaled = scaled ? 1 : 0;
    a.tplx = P.tune.tplx;
    // JT_RAGGED_FORCE_TIMEOUT=1 (tests of the failure path only): every
    // dependency wait targets a count no producer reaches and gives up after
    // 1 ms, so each consumer item is abandoned, the error flag is set and the
    // kernel still exits normally
    const char* ft = std::getenv("JT_RAGGED_FORCE_TIMEOUT");
    const bool force = ft && ft[0] == '1';
    a.wait_extra = force ? 1 : 0;
    a.wait_ticks = force ? 100000ull : 200000000ull;
    a.ntr = P.ntr;
    return a;
}

template <typename T>
void ragged_launch(RaggedPlan<T>& P, cpx<T>* data, int dir, bool scaled, hipStream_t st) {
    RArgs<T> a = ragged_args(P, data, dir, scaled);
    {
        JT_PROF(st, "ragged_reset");
        hipLaunchKernelGGL((ragged_reset<T>), ew_grid(3 * P.ntr + 1), dim3(256), 0, st, a);
    }
    JT_HIP(hipGetLastError());
    {
        JT_PROF(st, std::string("ragged<") + (sizeof(T) == 8 ? "f64" : "f32") + ",transforms=" +
                        std::to_string(P.ntr) + ",points=" + std::to_string(P.total) + "," + P.tune.str() + ">");
        if (P.tune.nt == 1024)
            hipLaunchKernelGGL((ragged_kernel<T, 1024>), dim3(P.grid), dim3(1024), P.shm, st, a, P.np);
        else if (P.tune.nt == 512)
            hipLaunchKernelGGL((ragged_kernel<T, 512>), dim3(P.grid), dim3(512), P.shm, st, a, P.np);
        else
            hipLaunchKernelGGL((ragged_kernel<T, 256>), dim3(P.grid), dim3(256), P.shm, st, a, P.np);
    }
    JT_HIP(hipGetLastError());
}

// Runs through a cached hipGraph when the stream is not the null stream.
template <typename T>
void ragged_execute(RaggedPlan<T>& P, cpx<T>* data, int dir, bool scale, hipStream_t st, bool use_graph) {
    const bool scaled = dir > 0 && scale;
    if (!use_graph || st == nullptr) {
        ragged_launch(P, data, dir, scaled, st);
        return;
    }
    for (auto& g : P.graphs)
        if (g.first.data == data && g.first.st == st && g.first.dir == dir && g.first.scale == scale) {
            JT_HIP(hipGraphLaunch(g.second, st));
            return;
        }
    hipGraph_t graph;
    JT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        ragged_launch(P, data, dir, scaled, st);
    } catch (...) {
        (void)hipStreamEndCapture(st, &graph);
        throw;
    }
    JT_HIP(hipStreamEndCapture(st, &graph));
    hipGraphExec_t exec;
    JT_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    (void)hipGraphDestroy(graph);
    P.graphs.push_back({typename RaggedPlan<T>::GKey{data, st, dir, scale}, exec});
    JT_HIP(hipGraphLaunch(exec, st));
}

// ---------------------------------------------------------------------------
// autotuner
// ---------------------------------------------------------------------------
// Coordinate descent over the RTune knobs, timing whole forward launches of
// candidate plans on a scratch buffer of the plan's size (the running time
// does not depend on the values).  Each knob is swept with the others held at
// their best so far; candidates whose LDS share cannot hold the longest line
// are skipped.  Returns the fastest plan; `log` receives one line per
// candidate (settings and milliseconds).
template <typename T>
std::unique_ptr<RaggedPlan<T>> ragged_autotune(const int64_t* lengths, int64_t count, std::string* log = nullptr) {
    SpecCache cache;
    hipStream_t st;
    JT_HIP(hipStreamCreate(&st));
    hipEvent_t e0, e1;
    JT_HIP(hipEventCreate(&e0));
    JT_HIP(hipEventCreate(&e1));
    DevBuf scratch;
    std::unique_ptr<RaggedPlan<T>> best;
    float best_ms = 0.0f;
    std::vector<RTune> tried;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(st);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(st);
    };
    auto time_plan = [&](RaggedPlan<T>& P) -> float {
        if (!scratch.p) {
            scratch.alloc((size_t)std::max<int64_t>(P.total, 1) * sizeof(cpx<T>));
            JT_HIP(hipMemsetAsync(scratch.p, 0, scratch.bytes, st));
        }
        cpx<T>* d = scratch.as<cpx<T>>();
        ragged_launch(P, d, -1, false, st);  // warm-up
        JT_HIP(hipEventRecord(e0, st));
        const int reps = 3;
        for (int i = 0; i < reps; i++) ragged_launch(P, d, -1, false, st);
        JT_HIP(hipEventRecord(e1, st));
        JT_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        JT_HIP(hipEventElapsedTime(&ms, e0, e1));
        int err = 0;
        JT_HIP(hipMemcpy(&err, P.d_err.p, sizeof(int), hipMemcpyDeviceToHost));
        if (err) throw JtError(ST_HIP, "ragged autotune: a dependency wait timed out");
        return ms / reps;
    };
    auto consider = [&](const RTune& t) {
        for (auto& u : tried)
            if (u == t) return;
        tried.push_back(t);
        std::unique_ptr<RaggedPlan<T>> P;
        try {
            P = make_ragged<T>(lengths, count, t, &cache);
        } catch (const JtError& e) {
            if (e.code != ST_UNSUPPORTED) throw;
            if (log) *log += t.str() + " skipped\n";
            return;
        }
        if (!(P->tune == t)) {  // an automatic setting resolved to a concrete one
            for (auto& u : tried)
                if (u == P->tune) return;
            tried.push_back(P->tune);
        }
        float ms = time_plan(*P);
        if (log) *log += P->tune.str() + " " + std::to_string(ms) + " ms\n";
        if (!best || ms < best_ms) {
            best = std::move(P);
            best_ms = ms;
        }
    };
    try {
        consider(RTune());
        // LDS split and workgroup size together, then the rest one by one
        const RTune base0 = best->tune;
        for (int w : {1, 2, 3, 4})
            for (int nt : {256, 512, 1024}) {
                RTune t = base0;
                t.wgs = w;
                t.nt = nt;
                if (w * nt <= 2048) consider(t);
            }
        for (int knob = 0; knob < 4; knob++) {
            const std::vector<int> vals = knob == 0 ? std::vector<int>{0, 1, 2}
                                        : knob == 1 ? std::vector<int>{2, 4, 8, 16, 32}
                                        : knob == 2 ? std::vector<int>{0, 1}
                                                    : std::vector<int>{0, 2, 8, 32, 128};
            const RTune base = best->tune;
            for (int v : vals) {
                RTune t = base;
                (knob == 0 ? t.tplx : knob == 1 ? t.chunk : knob == 2 ? t.order : t.lag) = v;
                consider(t);
            }
        }
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
    return best;
}

}  // namespace jt
