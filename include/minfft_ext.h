/*
 * minfft_ext.h -- extensions of the MI355X minfft drop-in.  None of these
 * has a counterpart in the reference (aimukhin/minfft); they never change the
 * 27 prototypes of minfft.h.
 *
 *  - error channel: the reference reports errors only as a NULL aux
 *    (minfft.c:1062-1064); minfft_last_error() says why.
 *  - device-pointer asynchronous execution on a caller's HIP stream.
 *  - minfft_*_batch: many transforms of mixed kinds and power-of-two sizes in
 *    one graph launch (semantics: the single-transform routine applied to
 *    each entry).
 */
#ifndef MINFFT_EXT_H
#define MINFFT_EXT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reason for the last failure on this thread ("" if none).  Exported by both
   libminfft.so and libminfft_single.so (via libminfft_engine.so). */
const char *minfft_last_error(void);

/* transform kinds, in the order of minfft.h */
enum minfft_kind {
    MINFFT_DFT = 0, MINFFT_INVDFT = 1, MINFFT_REALDFT = 2, MINFFT_INVREALDFT = 3,
    MINFFT_DCT2 = 4, MINFFT_DST2 = 5, MINFFT_DCT3 = 6, MINFFT_DST3 = 7,
    MINFFT_DCT4 = 8, MINFFT_DST4 = 9
};

/* arithmetic of a batch: FP64 throughout; FP32 butterflies with twiddles
   generated in FP64 and rounded; FP32 data with radix-16 stages computed as
   split-FP16 (hi+lo) MFMA products accumulated in FP32. */
enum minfft_prec { MINFFT_FP64 = 0, MINFFT_FP32 = 1, MINFFT_SPLIT_FP16 = 2 };

/* One entry of a batch: `kind` on a d-dimensional (1..3) array of dims n[0..d-1]
   (C order).  Input and output live at the given byte offsets of the buffers
   passed to minfft_batch_run(); they must be aligned to the item size
   (complex or real of the batch precision). */
typedef struct minfft_batch_entry {
    int kind;
    int d;
    int n[3];
    long long in_offset;
    long long out_offset;
} minfft_batch_entry;

typedef struct minfft_batch_plan minfft_batch_plan;

/* Build a batch plan: validates entries (NULL on error), groups identical
   shapes, builds the twiddle tables on the device, gathers the short 1D
   complex groups (2^4..2^12 points) of each direction into one launch
   (off with MINFFT_MULTI=0) and picks each other group's launch
   schedule by timing candidates on the device (autotuner; off with
   MINFFT_AUTOTUNE=0; results cached per shape, and in the file named by
   MINFFT_TUNE_CACHE if set). */
minfft_batch_plan *minfft_mkaux_batch(int count, const minfft_batch_entry *entries, int prec);

/* Run the batch on device buffers `in` / `out` (device pointers) on the HIP
   stream `stream` (hipStream_t; NULL = the plan's own stream, which is
   ordered after work queued earlier on the legacy default stream).  A plan with
   several units of work (groups launched on their own, persistent launches)
   captures a HIP graph on the first run with a given (in, out, stream) and
   replays it later, its units on concurrent branches; a plan with one unit
   launches its kernels straight into the stream (MINFFT_GRAPH=1 / 0 forces
   the graph / direct launches).  On a stream that the caller is capturing
   into a graph of its own, the launches go into that capture (no graph of
   the plan's own; run the plan once on the same buffers beforehand, so that
   the capture makes no pointer queries or stream creations).  Asynchronous:
   returns 0 once enqueued, -1 on error. */
int minfft_batch_run(minfft_batch_plan *p, const void *in, void *out, void *stream);

/* Same, but synchronous and accepting host or device pointers; in_bytes and
   out_bytes give the extents of the two buffers. */
int minfft_batch(minfft_batch_plan *p, const void *in, size_t in_bytes, void *out, size_t out_bytes);

void minfft_free_batch(minfft_batch_plan *p);

/* Replace group `group`'s schedule (groups are numbered as listed by
   minfft_batch_describe) by the process-wide one with the "key=value,..."
   changes in `spec` (keys tpw, cfb, cfcore, wave, xcd, e32, e4, split,
   split2, fused, wpc, falt, rsplit, coop, coopla, mall, rlong, mstages,
   csplit;
   NULL or "" = the process-wide schedule).  The next run re-captures the
   graph.  A group that runs in the batch's persistent launch of short 1D
   complex groups (described as "persistent") keeps the schedule unused.
   Returns 0, or -1 on an unknown key or group. */
int minfft_batch_schedule(minfft_batch_plan *p, int group, const char *spec);

/* Text description of a plan's groups and their chosen schedules (one line
   per group; "persistent" for the short 1D complex groups that share one
   persistent launch per direction) into buf (at most len bytes,
   NUL-terminated).  Returns the length needed including the NUL, or -1. */
int minfft_batch_describe(const minfft_batch_plan *p, char *buf, int len);

/* Enqueue one transform of an aux on device pointers and a caller's stream
   (hipStream_t; NULL = the aux's own stream, which is ordered after work
   queued earlier on the legacy default stream) without waiting.  Returns 0 or -1
   (also for host pointers, which the kernels cannot read). */
int minfft_enqueue(int kind, const void *in, void *out, const void *aux, void *stream);

/* Make the 27-symbol calls on `aux` (minfft_dft ... minfft_dst4) and its
   NULL-stream enqueues run on the caller's HIP stream `stream` (hipStream_t
   of the aux's device; NULL = back to the aux's own stream).  The calls stay
   synchronous: they return once the result is written, and their kernels are
   ordered after the work queued earlier on that stream.  The stream must
   outlive its use by the aux.  Returns 0, or -1 for a bad aux or a stream of
   another device. */
int minfft_set_stream(const void *aux, void *stream);

/* Largest scratch minfft_enqueue_noalloc may use: 64 MiB, held by each GPU
   the library runs on until the process ends. */
#define MINFFT_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* minfft_enqueue that allocates nothing.  No counterpart in the reference:
   it is for callers that cannot allocate device memory during the call
   (code under HIP graph capture, tight serving loops).  The kernels' scratch
   (the two-step intermediate of long axes, the complex staging of long
   real/DCT/DST axes, the multi-dimensional inverse real DFT's temp) comes
   from a fixed device array of MINFFT_NOALLOC_SCRATCH_BYTES instead of the
   aux's own buffers; results equal minfft_enqueue's.  The aux's device must
   be current.  Returns -1 with a reason (minfft_last_error) when the
   transform needs more scratch than that (e.g. FP32 complex 1D beyond 2^22
   points, FP64 beyond 2^21), for host pointers and for a bad aux or kind.
   Every call on one GPU shares that GPU's array: calls that overlap in time
   on one GPU (other threads, or streams running at once) race on it. */
int minfft_enqueue_noalloc(int kind, const void *in, void *out, const void *aux, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MINFFT_EXT_H */
