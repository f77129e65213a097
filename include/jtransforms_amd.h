/*
 * jtransforms_amd.h -- C ABI of the MI355X (gfx950) transform library.
 *
 * The reference (JTransforms) exposes plan objects with in-place methods on
 * Java arrays; it has no native interface.  Each entry point below is what a
 * JNI (or Panama FFM) binding of the corresponding Java class binds: the
 * constructor becomes jt_plan_create_*, every method becomes one jt_execute /
 * jt_execute_host call with an op code.  Paths are relative to the reference
 * source tree src/main/java/org/jtransforms/.
 *
 * Conventions kept from the reference:
 *   - transforms are in place on the caller's array, starting at `offset`
 *     (the Java `offa` argument, counted in elements of the array type);
 *   - forward sign is exp(-2 pi i jk/n); inverse is unscaled unless `scale`;
 *   - packed real layouts of realForward (1D/2D/3D) and the unscaled
 *     realInverse factors are the reference's;
 *   - size errors return JT_BAD_SIZE / JT_NOT_POW2 where the reference throws
 *     IllegalArgumentException (the binding maps them to that exception with
 *     the text of jt_last_error()).
 * Re-entrancy: one plan may be used from several host threads and streams at
 * once, as the reference's 1D plans may (they allocate their scratch per
 * call, fft/DoubleFFT_1D.java:1922,6639).  The plan owns one device buffer,
 * sized at creation, holding each op's workspace and staging area; calls
 * that use it are serialised (the host side under a per-plan lock, the
 * device side by an event chain across streams), calls that do not (most
 * single-pass transforms on device memory) run concurrently.  Under the
 * caller's stream capture the captured graph orders the work instead.
 * Execution never allocates device memory.
 */
#ifndef JTRANSFORMS_AMD_H
#define JTRANSFORMS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct jt_plan jt_plan;

/* status codes */
enum {
    JT_OK = 0,
    JT_BAD_SIZE = 1,    /* IllegalArgumentException: size < 1 (1D) or <= 1 (2D/3D) */
    JT_NOT_POW2 = 2,    /* IllegalArgumentException: packed real 2D/3D needs powers of two */
    JT_TOO_BIG = 3,     /* IllegalArgumentException: "data array is too big" */
    JT_HIP_ERROR = 4,   /* a HIP runtime call failed; see jt_last_error() */
    JT_OOM = 5,         /* device allocation failed */
    JT_BAD_ARG = 6,     /* null pointer, unknown op, op not defined for this plan */
    JT_UNSUPPORTED = 7  /* valid in the reference but not provided by this build */
};

/* precision */
enum { JT_F64 = 0, JT_F32 = 1 };

/* plan flags */
enum {
    JT_FLAG_NONE = 0,
    /* radix-16 stages as split-FP16 MFMA products with FP32 accumulation
     * (FP32 plans only; no counterpart in the reference) */
    JT_FLAG_SPLIT_FP16 = 1,
    /* ragged batches: capture the launch sequence in a hipGraph on the first
     * call for a given (data, stream, op) and replay it afterwards */
    JT_FLAG_GRAPH = 2,
    /* ragged batches: choose the schedule (LDS split, tile shape, work-item
     * size, radix order) by timing candidate plans at creation; the choice is
     * remembered for the same length list within the process */
    JT_FLAG_AUTOTUNE = 4
};

/* operations (one per reference method family) */
enum {
    JT_COMPLEX_FORWARD = 0,   /* complexForward(a, offa)                          */
    JT_COMPLEX_INVERSE = 1,   /* complexInverse(a, offa, scale)                   */
    JT_REAL_FORWARD = 2,      /* realForward(a, offa)                             */
    JT_REAL_INVERSE = 3,      /* realInverse(a, offa, scale)                      */
    JT_REAL_FORWARD_FULL = 4, /* realForwardFull(a, offa)                         */
    JT_REAL_INVERSE_FULL = 5, /* realInverseFull(a, offa, scale)                  */
    JT_FORWARD = 6,           /* DCT/DST forward(a, offa, scale); DHT forward(a, offa) */
    JT_INVERSE = 7            /* DCT/DST/DHT inverse(a, offa, scale)              */
};

/* Replaces DoubleFFT_1D(long n) (fft/DoubleFFT_1D.java:117-188) and
 * FloatFFT_1D(long n) (fft/FloatFFT_1D.java:117-...). */
int jt_plan_create_fft1d(jt_plan** out, int precision, int64_t n);

/* Replaces DoubleFFT_2D(long rows, long columns) (fft/DoubleFFT_2D.java:75-98)
 * and FloatFFT_2D. */
int jt_plan_create_fft2d(jt_plan** out, int precision, int64_t rows, int64_t columns);

/* Replaces DoubleFFT_3D(long slices, long rows, long columns)
 * (fft/DoubleFFT_3D.java:88-124) and FloatFFT_3D. */
int jt_plan_create_fft3d(jt_plan** out, int precision, int64_t slices, int64_t rows,
                         int64_t columns);

/* Replaces DoubleDCT_1D/2D/3D(...) (dct/DoubleDCT_1D.java:88-138,
 * dct/DoubleDCT_2D.java:73-95, dct/DoubleDCT_3D.java:85-120) and the Float
 * twins; ndim in 1..3, dims in the constructor's argument order. */
int jt_plan_create_dct(jt_plan** out, int precision, int ndim, const int64_t* dims);

/* Replaces DoubleDST_1D/2D/3D(...) (dst/DoubleDST_1D.java:59-65,
 * dst/DoubleDST_3D.java:84) and the Float twins. */
int jt_plan_create_dst(jt_plan** out, int precision, int ndim, const int64_t* dims);

/* Replaces DoubleDHT_1D/2D/3D(...) (dht/DoubleDHT_1D.java:60-66,
 * dht/DoubleDHT_2D.java:72, dht/DoubleDHT_3D.java:84) and the Float twins. */
int jt_plan_create_dht(jt_plan** out, int precision, int ndim, const int64_t* dims);

/* Many independent 1D transforms of one length n laid out back to back
 * (transform b at element offset b*2n for complex data).  No counterpart in
 * the reference, whose callers loop over DoubleFFT_1D.complexForward(a, offa)
 * with offa = b*2n (the form DoubleFFT_2D's row pass uses,
 * fft/DoubleFFT_2D.java:3100-3140). */
int jt_plan_create_fft1d_batch(jt_plan** out, int precision, int64_t n, int64_t batch,
                               uint32_t flags);

/* `count` independent 1D complex transforms of the given lengths, stored
 * back to back (transform i starts at element 2*sum_{j<i} lengths[j]), run in
 * one persistent launch (JT_COMPLEX_FORWARD / JT_COMPLEX_INVERSE, scale = 1/n
 * per transform).  No counterpart in the reference, whose callers build one
 * DoubleFFT_1D per length (fft/DoubleFFT_1D.java:117) and call it per
 * transform. */
int jt_plan_create_ragged(jt_plan** out, int precision, int64_t count, const int64_t* lengths,
                          uint32_t flags);

/* Writes a one-line description of the plan (its kernels' plan kinds; for a
 * ragged batch its schedule and, when autotuned, the timed candidates) into
 * buf (NUL-terminated, truncated to len bytes) and returns the full length.
 * No counterpart in the reference. */
int64_t jt_plan_describe(const jt_plan* plan, char* buf, int64_t len);

/* Waits for the device and reports a failure recorded on it by the plan's
 * last launches (a ragged batch whose dependency wait timed out). */
int jt_plan_sync_status(jt_plan* plan);

/* Sets plan flags after creation (JT_FLAG_SPLIT_FP16 on FP32 FFT plans of
 * any rank, e.g. a FloatFFT_3D).  No counterpart in the reference: selects
 * the split-FP16 MFMA form of the radix-16 stages. */
int jt_plan_set_flags(jt_plan* plan, uint32_t flags);

/* Frees the plan's tables and workspace (the reference's plans are garbage
 * collected). */
int jt_plan_destroy(jt_plan* plan);

/* Runs `op` in place on device memory `data` (+ `offset` elements of the
 * precision's type) on HIP stream `stream` (a hipStream_t; NULL = default
 * stream).  Asynchronous with respect to the host.  Replaces the reference
 * method named by `op` (see the enum above) on the plan's class; `scale` is
 * the method's boolean argument where it has one and is ignored otherwise. */
int jt_execute(jt_plan* plan, int op, void* data, int64_t offset, int scale, void* stream);

/* Same as jt_execute, allocating nothing: the op's workspace (four-step,
 * Bluestein and *Full intermediates) and the staging of a misaligned offset
 * live in a static device array of 64 MiB per GPU (held until the process
 * ends) instead of the plan's own buffer.  No counterpart in the reference;
 * for callers that cannot allocate device memory, e.g. code under HIP graph
 * capture or a serving loop.  An op that needs more than 64 MiB of scratch
 * returns JT_TOO_BIG; ragged batches return JT_UNSUPPORTED.  All calls on one
 * GPU share its array: calls that overlap in time on one GPU (other streams,
 * other threads) race, so the caller orders them. */
int jt_execute_noalloc(jt_plan* plan, int op, void* data, int64_t offset, int scale, void* stream);

/* Same on a host array (the Java double[]/float[] a JNI binding pins with
 * GetPrimitiveArrayCritical): copies the touched range to the device, runs,
 * copies back, and returns when the result is in `data`. */
int jt_execute_host(jt_plan* plan, int op, void* data, int64_t offset, int scale);

/* Number of array elements (of the plan's precision) an op reads and writes
 * starting at `offset`, i.e. the minimal array length minus offset. */
int64_t jt_plan_elements(const jt_plan* plan, int op);

/* Human-readable text of a status code. */
const char* jt_status_string(int status);

/* Text of the last error raised on the calling host thread ("" if none). */
const char* jt_last_error(void);

/* Per-kernel timing with HIP events recorded on each launch's stream (no
 * counterpart in the reference; used by the benchmark).  jt_profile_begin
 * clears and enables the recorder; jt_profile_end disables it, waits for the
 * recorded events and writes "kernel<TAB>launches<TAB>total_ms" lines into
 * buf (NUL-terminated, truncated to len); returns the full text length.
 * Later calls return the same text until the next jt_profile_begin. */
int jt_profile_begin(void);
int64_t jt_profile_end(char* buf, int64_t len);

/* Library version string. */
const char* jt_version(void);

#ifdef __cplusplus
}
#endif

#endif /* JTRANSFORMS_AMD_H */
