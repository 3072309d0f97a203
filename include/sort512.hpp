// sort512.hpp -- source-compatible drop-in for the reference header of the same
// name (berenger-eu/avx-512-sort, namespace Sort512), backed by the MI355X
// library through the C ABI in s512.h.  Link with -ls512.
//
// Same namespaces, names and template signatures as the reference:
//   Sort512::Sort<SortType,IndexType>              sort512.hpp:5960
//   Sort512::SortOmp / SortOmpPartition            sort512.hpp:11, 5992
//   Sort512::SortOmpMerge / SortOmpMergeDeps / SortOmpParMerge   6006/6060/6104
//   Sort512::Partition512<IndexType>(int|double[], left, right, pivot)  5729/5821
//   Sort512::SmallSort16V(double*|int*, length)    2255/5312
//   ParallelInplace::parallelMergeInPlace          parallelInplace.hpp:209
// Arrays may live in host or device memory.  Where the reference has
// undefined behaviour this header defines it: size 0 and 1 are no-ops,
// SmallSort16V rejects lengths above 16 vectors, an empty partition range
// (right == left - 1) returns left.  A failure prints the library's message
// and aborts (the reference functions return void, so there is no status to
// hand back); Sort512::lastError() gives the message of the last failure.
// The register-level kernels (CoreSmallSort*, taking __m512d/__m512i) have no
// counterpart: an x86 register type cannot cross into GPU code.
#ifndef SORT512_HPP
#define SORT512_HPP

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <type_traits>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "s512.h"

namespace Sort512 {

namespace detail {
inline void check(s512_status st, const char* what) {
    if (st != S512_OK) {
        std::fprintf(stderr, "Sort512::%s failed (status %d): %s\n", what, st, s512_last_error());
        std::abort();
    }
}
inline s512_status sort(int* a, size_t n) { return s512_sort_i32(a, n); }
inline s512_status sort(double* a, size_t n) { return s512_sort_f64(a, n); }
inline s512_status merge_sort(int* a, size_t n) { return s512_sort_merge_i32(a, n); }
inline s512_status merge_sort(double* a, size_t n) { return s512_sort_merge_f64(a, n); }
template <class SortType>
inline void require_type() {
    static_assert(std::is_same<SortType, int>::value || std::is_same<SortType, double>::value,
                  "Sort512 sorts int and double arrays");
}
template <class IndexType>
inline size_t as_size(IndexType v) {
    return v < IndexType(0) ? 0 : static_cast<size_t>(v);
}
}  // namespace detail

inline const char* lastError() { return s512_last_error(); }

template <class SortType, class IndexType = size_t>
static inline void Sort(SortType array[], const IndexType size) {
    detail::require_type<SortType>();
    detail::check(detail::sort(array, detail::as_size(size)), "Sort");
}

template <class SortType, class IndexType = size_t>
static inline void SortOmp(SortType array[], const IndexType size) {
    detail::require_type<SortType>();
    detail::check(detail::sort(array, detail::as_size(size)), "SortOmp");
}

template <class SortType, class IndexType = size_t>
static inline void SortOmpPartition(SortType array[], const IndexType size) {
    detail::require_type<SortType>();
    detail::check(detail::sort(array, detail::as_size(size)), "SortOmpPartition");
}

template <class SortType, class IndexType = size_t>
static inline void SortOmpMerge(SortType array[], const IndexType size) {
    detail::require_type<SortType>();
    detail::check(detail::merge_sort(array, detail::as_size(size)), "SortOmpMerge");
}

template <class SortType, class IndexType = size_t>
static inline void SortOmpMergeDeps(SortType array[], const IndexType size) {
    detail::require_type<SortType>();
    detail::check(detail::merge_sort(array, detail::as_size(size)), "SortOmpMergeDeps");
}

template <class SortType, class IndexType = size_t>
static inline void SortOmpParMerge(SortType array[], const IndexType size) {
    detail::require_type<SortType>();
    detail::check(detail::merge_sort(array, detail::as_size(size)), "SortOmpParMerge");
}

template <class IndexType>
static inline IndexType Partition512(int array[], IndexType left, IndexType right, const int pivot) {
    if (right < left) return left;
    size_t bound = 0;
    detail::check(s512_partition_i32(array, static_cast<size_t>(left), static_cast<size_t>(right), pivot, &bound),
                  "Partition512");
    return static_cast<IndexType>(bound);
}

template <class IndexType>
static inline IndexType Partition512(double array[], IndexType left, IndexType right, const double pivot) {
    if (right < left) return left;
    size_t bound = 0;
    detail::check(s512_partition_f64(array, static_cast<size_t>(left), static_cast<size_t>(right), pivot, &bound),
                  "Partition512");
    return static_cast<IndexType>(bound);
}

inline void SmallSort16V(double* __restrict__ ptr, const size_t length) {
    detail::check(s512_smallsort16v_f64(ptr, length), "SmallSort16V");
}

inline void SmallSort16V(int* __restrict__ ptr, const size_t length) {
    detail::check(s512_smallsort16v_i32(ptr, length), "SmallSort16V");
}

}  // namespace Sort512

namespace ParallelInplace {
// Merges the sorted runs [0, centerPosition) and [centerPosition, sizeArray)
// of array in place: one GPU merge per call (no reference counterpart of this
// three-argument form; the reference's form follows).
template <class NumType>
inline void parallelMergeInPlace(NumType array[], const int sizeArray, int centerPosition) {
    static_assert(std::is_same<NumType, int>::value || std::is_same<NumType, double>::value,
                  "parallelMergeInPlace merges int and double arrays");
    s512_status st;
    if (std::is_same<NumType, int>::value)
        st = s512_merge_inplace_i32(reinterpret_cast<int32_t*>(array), (size_t)sizeArray, (size_t)centerPosition);
    else
        st = s512_merge_inplace_f64(reinterpret_cast<double*>(array), (size_t)sizeArray, (size_t)centerPosition);
    Sort512::detail::check(st, "parallelMergeInPlace");
}

// The reference's bookkeeping record (parallelInplace.hpp:138-145), kept so
// that callers declaring arrays of it compile unchanged.
template <class NumType>
struct WorkingInterval {
    NumType* array;
    int currentStart;
    int currentMiddle;
    int currentEnd;
    int level;
    int depthLimite;
};

// parallelInplace.hpp:209-211: called by each of the numThreadsInvolved
// threads from firstThread of an OpenMP team, with the team's shared
// intervals and barrier arrays.  A GPU merge is one call: the first involved
// thread runs it and returns once the merge is complete; the other involved
// threads have no share of the work and return at once (in the reference
// every involved thread returns after the merge).  Callers that publish a
// merge's completion from its first thread, as the reference's
// SortOmpParMerge does (sort512.hpp:6160-6166), work unchanged.  intervals
// and barrier are not touched.
template <class NumType>
inline void parallelMergeInPlace(NumType array[], const int sizeArray, int centerPosition,
                                 const long int numThreadsInvolved, const long int firstThread,
                                 volatile WorkingInterval<NumType> intervals[], volatile int barrier[]) {
    (void)numThreadsInvolved;
    (void)intervals;
    (void)barrier;
    long int me = 0;
#ifdef _OPENMP
    me = omp_get_thread_num();
#endif
    if (me == firstThread) parallelMergeInPlace(array, sizeArray, centerPosition);
}
}  // namespace ParallelInplace

#endif  // SORT512_HPP
