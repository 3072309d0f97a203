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
 *   - Sizes are limited to n < 2^32 keys per call.
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

/* ------------------------------------------------------------------ sort */
/* Sort512::Sort<int,size_t>      sort512.hpp:5960-5963 (CoreSort 5947-5958) */
s512_status s512_sort_i32(int32_t* a, size_t n);
/* Sort512::Sort<double,size_t>   sort512.hpp:5960-5963                      */
s512_status s512_sort_f64(double* a, size_t n);
s512_status s512_sort_i32_dev(int32_t* d_a, size_t n, void* stream);
s512_status s512_sort_f64_dev(double* d_a, size_t n, void* stream);

/* -------------------------------------------------------------- key/value */
/* Sort512kv::Sort<int,size_t>    sort512kv.hpp:6132-6135.  Keys ascending;
 * each value stays with its key; the order of values among equal keys is
 * unspecified (the reference is unstable too).                            */
s512_status s512kv_sort_i32(int32_t* keys, int32_t* values, size_t n);
s512_status s512kv_sort_i32_dev(int32_t* d_keys, int32_t* d_values, size_t n, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* S512_H */
