/*
 * s512.h -- C ABI of the MI355X sort library (libs512.so).
 *
 * Drop-in boundary for the reference's public sort API (berenger-eu/avx-512-sort,
 * header-only C++).  Every reference entry point on the sort path has one C
 * function here; the C++ headers sort512.hpp / sort512kv.hpp in this directory
 * keep the reference's namespaces and template signatures and forward to them.
 *
 * Conventions
 *   - Every function returns s512_status (0 = success).  On failure
 *     s512_last_error() describes the cause (thread-local string).
 *   - Plain entry points accept host OR device pointers.  Device (or managed)
 *     memory is sorted in place on the GPU; host memory is copied to the GPU,
 *     sorted and copied back.  They return when the result is in place.
 *   - `_dev` entry points require device pointers and run on `stream` (a
 *     hipStream_t, NULL = the default stream).  They may synchronise that
 *     stream internally (the multiway partition reads bucket counts between
 *     levels) and return with the result complete on the stream.
 *   - Keys are compared exactly.  int32: signed order.  double: IEEE order for
 *     non-NaN keys; -0.0 is placed before +0.0 (they compare equal).  NaN keys
 *     are accepted and placed by their bit pattern (negative NaNs first,
 *     positive NaNs last); the reference leaves NaN undefined.
 *   - n <= 1 is a no-op (the reference underflows size-1 at n == 0).
 *   - Sizes are size_t, as the reference's IndexType (sort512.hpp:5960).
 *     Plain sorts and the merge sort take any n that fits in device memory
 *     (2^32 keys or more: pieces of 2^31 keys, each sorted by the partition
 *     levels, then merged by merge-path rounds; about 2n keys of scratch).
 *     Key/value sorts and segmented sorts take n < 2^32; argsort and the
 *     8-byte-value sorts n < 2^31 (their permutation is int32).
 *   - Calls on one device are serialised by the library (it keeps one
 *     grow-only workspace per device).
 */
#ifndef S512_H
#define S512_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int s512_status;
enum {
    S512_OK = 0,
    S512_ERR_INVALID = 1,   /* bad argument (null pointer, range, length)   */
    S512_ERR_NOMEM = 2,     /* device or pinned allocation failed           */
    S512_ERR_HIP = 3,       /* HIP runtime error                            */
    S512_ERR_NODEVICE = 4,  /* no usable GPU                                */
    S512_ERR_TOO_LARGE = 5, /* n beyond the supported maximum               */
    S512_ERR_INTERNAL = 6   /* invariant violated (reported, never silent)  */
};

/* Human-readable description of the last failure on this thread. */
const char* s512_last_error(void);

/* Library version string. */
const char* s512_version(void);

/* Frees the calling device's cached workspaces (scratch key arrays, bucket
 * ids, staging buffers, selection scratch); the next call allocates again.
 * No reference counterpart (the reference allocates nothing between calls):
 * for long-running callers that need the memory back.  Other calls on the
 * device wait for it.                                                      */
s512_status s512_release(void);

/* ------------------------------------------------------------------ sort */
/* Sort512::Sort<int,size_t>      sort512.hpp:5960-5963 (CoreSort 5947-5958) */
s512_status s512_sort_i32(int32_t* a, size_t n);
/* Sort512::Sort<double,size_t>   sort512.hpp:5960-5963                      */
s512_status s512_sort_f64(double* a, size_t n);
s512_status s512_sort_i32_dev(int32_t* d_a, size_t n, void* stream);
s512_status s512_sort_f64_dev(double* d_a, size_t n, void* stream);

/* SortOmp / Sort512::SortOmpPartition<T,size_t>  sort512.hpp:5992-6004 (the
 * README's "SortOmp", sort512.hpp:11, which the header does not define) map to
 * the same GPU sort as Sort512::Sort: s512_sort_i32 / s512_sort_f64.        */

/* Sort512::SortOmpMerge / SortOmpMergeDeps / SortOmpParMerge
 * sort512.hpp:6006-6058, 6060-6102, 6104-6174: sort chunks, then merge runs
 * pairwise.  GPU merge sort: bitonic tiles + merge-path rounds.             */
s512_status s512_sort_merge_i32(int32_t* a, size_t n);
s512_status s512_sort_merge_f64(double* a, size_t n);
s512_status s512_sort_merge_i32_dev(int32_t* d_a, size_t n, void* stream);
s512_status s512_sort_merge_f64_dev(double* d_a, size_t n, void* stream);

/* ------------------------------------------------ 64-bit and unsigned keys */
/* No reference counterpart (the reference sorts int and double keys): keys-only
 * ascending sorts of int64 (signed order), uint64 and uint32 by the same
 * engine, for callers whose keys are wider or unsigned.  The _ex variants take
 * the options and statistics of s512_sort_*_dev_ex below (the learned
 * classifier is not offered for these types).                              */
s512_status s512_sort_i64(int64_t* a, size_t n);
s512_status s512_sort_u64(uint64_t* a, size_t n);
s512_status s512_sort_u32(uint32_t* a, size_t n);
s512_status s512_sort_i64_dev(int64_t* d_a, size_t n, void* stream);
s512_status s512_sort_u64_dev(uint64_t* d_a, size_t n, void* stream);
s512_status s512_sort_u32_dev(uint32_t* d_a, size_t n, void* stream);

/* ---------------------------------- key/value with 64-bit keys; argsort */
/* No reference counterpart (Sort512kv covers int keys, sort512kv.hpp:6132):
 * keys ascending, each int32 value moves with its key; equal keys keep their
 * values in no particular order, as in Sort512kv.  Host or device pointers. */
s512_status s512kv_sort_f64(double* k, int32_t* v, size_t n);
s512_status s512kv_sort_i64(int64_t* k, int32_t* v, size_t n);
s512_status s512kv_sort_u64(uint64_t* k, int32_t* v, size_t n);
s512_status s512kv_sort_f64_dev(double* d_k, int32_t* d_v, size_t n, void* stream);
s512_status s512kv_sort_i64_dev(int64_t* d_k, int32_t* d_v, size_t n, void* stream);
s512_status s512kv_sort_u64_dev(uint64_t* d_k, int32_t* d_v, size_t n, void* stream);

/* Key/value sorts with an 8-byte value payload (int64, or any 8-byte value
 * passed as int64): keys ascending in place, each value follows its key; n <
 * 2^31.  Host or device pointers; device work runs on `stream`.          */
s512_status s512kv_sort_i32_v64(int32_t* k, int64_t* v, size_t n, void* stream);
s512_status s512kv_sort_f64_v64(double* k, int64_t* v, size_t n, void* stream);
s512_status s512kv_sort_i64_v64(int64_t* k, int64_t* v, size_t n, void* stream);
s512_status s512kv_sort_u64_v64(uint64_t* k, int64_t* v, size_t n, void* stream);

/* argsort: idx[i] = position in keys of the i-th smallest key (equal keys in
 * no particular order).  keys are not modified; n < 2^31.  keys and idx may
 * each be host or device memory; device work runs on `stream`.             */
s512_status s512_argsort_i32(const int32_t* keys, int32_t* idx, size_t n, void* stream);
s512_status s512_argsort_f64(const double* keys, int32_t* idx, size_t n, void* stream);
s512_status s512_argsort_i64(const int64_t* keys, int32_t* idx, size_t n, void* stream);
s512_status s512_argsort_u64(const uint64_t* keys, int32_t* idx, size_t n, void* stream);

/* ------------------------------------------------------ top-k, selection */
/* No reference counterpart (built from the multiway sort for callers that need
 * a prefix of the order).  topk: out[0..k) = the k smallest keys, ascending
 * (0 <= k <= n).  select: *kth = the k-th smallest key, 0-based (k < n).
 * Keys are not modified; keys and out may each be host or device memory.  */
s512_status s512_topk_i32(const int32_t* keys, size_t n, size_t k, int32_t* out, void* stream);
s512_status s512_topk_f64(const double* keys, size_t n, size_t k, double* out, void* stream);
s512_status s512_topk_i64(const int64_t* keys, size_t n, size_t k, int64_t* out, void* stream);
s512_status s512_topk_u64(const uint64_t* keys, size_t n, size_t k, uint64_t* out, void* stream);
s512_status s512_select_i32(const int32_t* keys, size_t n, size_t k, int32_t* kth, void* stream);
s512_status s512_select_f64(const double* keys, size_t n, size_t k, double* kth, void* stream);
s512_status s512_select_i64(const int64_t* keys, size_t n, size_t k, int64_t* kth, void* stream);
s512_status s512_select_u64(const uint64_t* keys, size_t n, size_t k, uint64_t* kth, void* stream);

/* ------------------------------------------------------------- partition */
/* Sort512::Partition512<size_t>(int[], left, right, pivot)    sort512.hpp:5729
 * Sort512::Partition512<size_t>(double[], left, right, pivot) sort512.hpp:5821
 * Rearranges the inclusive range [left, right] of a (indices from a) so that
 * keys <= pivot come first; *out_bound = left + #{keys <= pivot}.  Both sides
 * keep their input order.  right = left - 1 is an empty range.             */
s512_status s512_partition_i32(int32_t* a, size_t left, size_t right, int32_t pivot, size_t* out_bound);
s512_status s512_partition_f64(double* a, size_t left, size_t right, double pivot, size_t* out_bound);
s512_status s512_partition_i32_dev(int32_t* d_a, size_t left, size_t right, int32_t pivot, size_t* out_bound,
                                   void* stream);
s512_status s512_partition_f64_dev(double* d_a, size_t left, size_t right, double pivot, size_t* out_bound,
                                   void* stream);
/* No reference counterpart: s512_partition_f64_dev for callers that cannot
 * allocate (code under hipGraph capture, serving loops, harnesses that forbid
 * hipMalloc).  Same result; its workspace is a fixed 64 MiB array that every
 * GPU the library runs on holds until the process ends, so it partitions at
 * most 8,384,480 keys per call (larger ranges return S512_ERR_TOO_LARGE with
 * nothing changed) and calls no allocator.  Calls
 * on one GPU share that array: calls that overlap on one GPU race.  Device
 * pointers; runs on `stream` and reads the bound back.                     */
s512_status s512_partition_f64_dev_noalloc(double* d_a, size_t left, size_t right, double pivot,
                                           size_t* out_bound, void* stream);

/* ------------------------------------------------------------ small sort */
/* Sort512::SmallSort16V(double*, len <= 128)  sort512.hpp:2255
 * Sort512::SmallSort16V(int*, len <= 256)     sort512.hpp:5312
 * One wave64 bitonic network.  Longer lengths return S512_ERR_INVALID.     */
s512_status s512_smallsort16v_i32(int32_t* a, size_t len);
s512_status s512_smallsort16v_f64(double* a, size_t len);

/* ----------------------------------------------------------------- merge */
/* ParallelInplace::parallelMergeInPlace(array, sizeArray, centerPosition, ...)
 * parallelInplace.hpp:209-306: merge the sorted runs [0, center) and
 * [center, size) of a in place.  Stable (ties keep the left run first).    */
s512_status s512_merge_inplace_i32(int32_t* a, size_t size, size_t center);
s512_status s512_merge_inplace_f64(double* a, size_t size, size_t center);
s512_status s512_merge_inplace_i32_dev(int32_t* d_a, size_t size, size_t center, void* stream);
s512_status s512_merge_inplace_f64_dev(double* d_a, size_t size, size_t center, void* stream);

/* ------------------------------------------------------ options and stats */
/* Tuning knobs and per-call observability.  Zero fields select the defaults:
 * the autotuned table compiled into the library (avx-512-sort_amd/s512_tune.py
 * writes it per dtype and size class), which an explicit field overrides.  */
enum {
    S512_PHASE_SAMPLE = 0,      /* draw splitter samples                    */
    S512_PHASE_SAMPLE_SORT = 1, /* bitonic sort of the samples              */
    S512_PHASE_TREE = 2,        /* splitter tree (and model splitters)     */
    S512_PHASE_COUNT = 3,       /* classify + per-chunk bucket histogram    */
    S512_PHASE_SCAN = 4,        /* bucket offsets                           */
    S512_PHASE_SCATTER = 5,     /* classify + LDS-staged scatter            */
    S512_PHASE_PLAN = 6,        /* route buckets                            */
    S512_PHASE_COPY = 7,        /* finished runs back to the caller array   */
    S512_PHASE_TILE = 8,        /* bitonic tile sorts of the leaves         */
    S512_PHASE_TRAIN = 9,       /* learned model training                   */
    S512_NPHASE = 10
};

enum {
    S512_CLASSIFY_SPLITTERS = 0, /* sampled splitters at every level (default) */
    S512_CLASSIFY_LEARNED = 1    /* learned CDF model for the first two levels,
                                    sampled splitters as the fallback when the
                                    model's training loss is too high or a
                                    first-level bucket is too large          */
};

typedef struct s512_options {
    uint32_t target_leaf;   /* mean leaf size the fan-out aims for (0: 1024)     */
    uint32_t oversample;    /* samples per splitter (0: 16)                  */
    uint32_t max_fanout_log2; /* cap on log2 of the per-level fan-out, 1..9 (0: 9) */
    uint32_t profile;       /* nonzero: time every kernel with HIP events    */
    uint64_t seed;          /* sample positions (0: fixed default)           */
    uint32_t classifier;    /* S512_CLASSIFY_*; learned applies to keys-only int32/double sorts */
    uint32_t sample_permille; /* learned: model sample, per mille of n (0: 10 = 1%) */
    uint32_t train_steps;   /* learned: Adam steps (0: 10)                  */
    uint32_t imbalance_pct; /* learned: fall back to splitters when a first-level bucket
                               exceeds this percentage of the mean (0: 400) */
    uint32_t scatter_variant; /* 0: default (1); else 1 + shape + 4 * reclassify, where shape
                                 0..3 = scatter workgroup 512x16, 256x16, 1024x8 (threads x
                                 keys per thread), 512x16 with write combining (only with
                                 stored ids; reclassifying runs it as shape 0), and
                                 reclassify = 1 classifies every key in the scatter again
                                 instead of reading the bucket ids that the count pass
                                 stored (2 bytes per key); 9 and 10: the line-buffered
                                 scatter (every 128-byte line of a bucket's run stored
                                 whole; keys-only sorts, key/value sorts run shape 0),
                                 9 classifying again, 10 reading the stored ids; 11: the
                                 full-line scatter (every store covers whole 128-byte
                                 lines, a bucket's unfinished line waits in LDS for the
                                 next round; persistent workgroups over contiguous chunk
                                 ranges; stored ids; keys-only sorts, key/value sorts run
                                 shape 0), launched as a pair with shape 0 reading the
                                 stored ids: per level, a device-side gate runs the
                                 plain one when under a quarter of the (chunk, bucket)
                                 pairs hold keys (sorted or reversed input), the
                                 full-line one otherwise.  The tuned default for
                                 double keys from 2^27 keys on                          */
    uint32_t count_variant;   /* 0: default (shape 2); else 1 + count shape 0..4 = (threads,
                                 keys per thread per batch, keys classified together):
                                 256x16x8, 256x16x4, 256x8x8, 512x8x8, 256x8x4          */
    uint32_t tile_variant;    /* 0: default (plan 2); else 1 + tile plan 0..2 = keys per
                                 thread of the bitonic tile sorts: 8, 16 or 32 (up to 16
                                 on 64-bit tiles of plan 0)                             */
    uint32_t engine;          /* S512_ENGINE_*: how partition levels read the keys      */
    uint32_t region_slack;    /* one-pass root level: margin of each capacity region over
                                 its sampled estimate, in tenths of the estimate's standard
                                 deviation (0: 50).  Smaller saves scratch memory and
                                 overflows more often; an overflow redoes the level with
                                 two passes                                            */
} s512_options;

enum {
    S512_ENGINE_AUTO = 0,     /* the two-pass levels (measured faster on MI355X)          */
    S512_ENGINE_TWO_PASS = 1, /* every level counts (one read), then scatters (read+write) */
    S512_ENGINE_ONE_PASS = 2  /* the first two levels read and write every key once: the
                                 root level into sample-sized regions of the scratch
                                 array, the second level straight to final positions; the
                                 root level falls back to two passes when a region
                                 overflows (the input is unchanged by then)            */
};

typedef struct s512_stats {
    int32_t levels;                     /* partition levels run              */
    uint64_t tiles;                     /* leaf tiles sorted                 */
    double kernel_ms[S512_NPHASE];      /* device time per phase (profile)   */
    uint32_t launches[S512_NPHASE];     /* kernel launches per phase         */
    uint64_t keys[S512_NPHASE];         /* keys processed per phase          */
    int32_t classifier;                 /* classifier that produced the result */
    int32_t learned_fallback;           /* 1: the learned model was rejected  */
    int32_t learned_levels;             /* levels partitioned by the model    */
    float learned_max_bucket_ratio;     /* largest first-level bucket / mean  */
    float model_loss;                   /* final training loss (MSE on CDF)   */
    int32_t engine;                     /* S512_ENGINE_ONE_PASS or _TWO_PASS: how the
                                           partition levels ran (0: no level) */
    int32_t region_overflow;            /* 1: a one-pass root region overflowed and
                                           the sort ran with two passes        */
    float learned_leaf_max_ratio;       /* largest second-level bucket placed by the
                                           learned model / their mean          */
} s512_stats;

/* Sort512::Sort with options and statistics (device pointers). */
s512_status s512_sort_i32_dev_ex(int32_t* d_a, size_t n, const s512_options* opt, s512_stats* stats, void* stream);
s512_status s512_sort_f64_dev_ex(double* d_a, size_t n, const s512_options* opt, s512_stats* stats, void* stream);
/* The same for the 64-bit and unsigned keys above (no reference counterpart). */
s512_status s512_sort_i64_dev_ex(int64_t* d_a, size_t n, const s512_options* opt, s512_stats* stats, void* stream);
s512_status s512_sort_u64_dev_ex(uint64_t* d_a, size_t n, const s512_options* opt, s512_stats* stats, void* stream);
s512_status s512_sort_u32_dev_ex(uint32_t* d_a, size_t n, const s512_options* opt, s512_stats* stats, void* stream);

/* Learned partitioner model check: full-batch loss and gradient (fp32) of the
 * monotone 1-32-32-1 CDF model at parameters d_params (s512_mlp_nparams()
 * floats) over m points (d_u features in [0,1], d_t targets); device pointers.
 * No reference counterpart (the reference has no learned path).            */
uint32_t s512_mlp_nparams(void);
s512_status s512_mlp_loss_grad_f32(const float* d_params, const float* d_u, const float* d_t, uint32_t m,
                                   float* d_grad, float* d_loss, void* stream);

/* -------------------------------------------------------------- key/value */
/* Sort512kv::Sort<int,size_t>    sort512kv.hpp:6132-6135.  Keys ascending;
 * each value stays with its key; the order of values among equal keys is
 * unspecified (the reference is unstable too).                            */
s512_status s512kv_sort_i32(int32_t* keys, int32_t* values, size_t n);
s512_status s512kv_sort_i32_dev(int32_t* d_keys, int32_t* d_values, size_t n, void* stream);
s512_status s512kv_sort_i32_dev_ex(int32_t* d_keys, int32_t* d_values, size_t n, const s512_options* opt,
                                   s512_stats* stats, void* stream);
/* Sort512kv::Partition512<size_t>   sort512kv.hpp:5958-6078                 */
s512_status s512kv_partition_i32(int32_t* keys, int32_t* values, size_t left, size_t right, int32_t pivot,
                                 size_t* out_bound);
s512_status s512kv_partition_i32_dev(int32_t* d_keys, int32_t* d_values, size_t left, size_t right, int32_t pivot,
                                     size_t* out_bound, void* stream);
/* Sort512kv::SmallSort16V(int*, int*, len <= 256)   sort512kv.hpp:5235     */
s512_status s512kv_smallsort16v_i32(int32_t* keys, int32_t* values, size_t len);

/* --------------------------------------------------------- segmented sort */
/* Many independent arrays in one call.  No single reference entry point: it
 * replaces a caller's loop of SmallSort16V (sort512perf.cpp:656-661 sorts many
 * small arrays that way) or of Sort512::Sort (sort512.hpp:5960) per array.
 * Segment i is keys[offsets[i] .. offsets[i+1]) (offsets has nseg + 1
 * entries, non-decreasing); each segment is sorted in place, independently,
 * with the same key order as s512_sort_*.  Keys outside every segment are
 * left alone.  Offsets that decrease return S512_ERR_INVALID with the keys
 * unchanged.  At most 2^31-1 segments and 2^32-1 keys per call.
 * One fixed kernel sequence sorts every segment up to the largest tile (16384
 * int32 / 8192 double keys) and partitions larger ones once; only buckets
 * still above a tile then need the host-driven engine.                     */
enum { S512_DTYPE_I32 = 0, S512_DTYPE_F64 = 1 };
s512_status s512_segsort_i32(int32_t* keys, const uint64_t* offsets, size_t nseg);
s512_status s512_segsort_f64(double* keys, const uint64_t* offsets, size_t nseg);
/* Device pointers on `stream`; stats (may be NULL): levels = 1 when a segment
 * was partitioned, tiles / keys[S512_PHASE_TILE] = tiles sorted and their keys,
 * learned_levels = buckets finished by the multi-level engine.             */
s512_status s512_segsort_i32_dev(int32_t* d_keys, const uint64_t* d_offsets, size_t nseg, s512_stats* stats,
                                 void* stream);
s512_status s512_segsort_f64_dev(double* d_keys, const uint64_t* d_offsets, size_t nseg, s512_stats* stats,
                                 void* stream);
/* Same with tuning options (NULL or zero fields: the autotuned defaults). */
s512_status s512_segsort_i32_dev_ex(int32_t* d_keys, const uint64_t* d_offsets, size_t nseg, const s512_options* opt,
                                    s512_stats* stats, void* stream);
s512_status s512_segsort_f64_dev_ex(double* d_keys, const uint64_t* d_offsets, size_t nseg, const s512_options* opt,
                                    s512_stats* stats, void* stream);
/* Reusable plan: the kernel sequence for fixed (d_keys, d_offsets, nseg,
 * nkeys = length of d_keys) captured once in a hipGraph, with its own
 * workspace.  Launch replays the graph on `stream` (the offsets and keys may
 * change between launches, the pointers and sizes may not), then reads one
 * control block back and finishes oversized buckets, if any.  A plan is used
 * by one thread at a time.                                                 */
typedef struct s512_segsort_plan_t* s512_segsort_plan;
s512_status s512_segsort_plan_create(int dtype, void* d_keys, const uint64_t* d_offsets, size_t nseg, size_t nkeys,
                                     s512_segsort_plan* out);
s512_status s512_segsort_plan_launch(s512_segsort_plan plan, s512_stats* stats, void* stream);
s512_status s512_segsort_plan_destroy(s512_segsort_plan plan);

#ifdef __cplusplus
}
#endif
#endif /* S512_H */
