// sort512kv.hpp -- source-compatible drop-in for the reference header of the
// same name (namespace Sort512kv: int keys with an int value array), backed by
// the MI355X library through s512.h.  Link with -ls512.
//   Sort512kv::Sort<SortType,IndexType>(keys[], values[], size)   sort512kv.hpp:6132
//   Sort512kv::SortOmpPartition                                   sort512kv.hpp:6164
//   Sort512kv::Partition512<IndexType>(keys[], values[], l, r, p) sort512kv.hpp:5958
//   Sort512kv::SmallSort16V(keys*, values*, length <= 256)        sort512kv.hpp:5235
// Keys come out ascending and every value stays with its key; the order of
// values among equal keys is unspecified (the reference is unstable too).
// Unlike the reference, a real key equal to INT_MAX in a short leaf keeps its
// own value (the reference pads leaves with (INT_MAX, INT_MAX) pairs).
// The include guard is the one the reference meant (SORT512KV_HPP).
#ifndef SORT512KV_HPP
#define SORT512KV_HPP

#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <type_traits>

#include "s512.h"

namespace Sort512kv {

namespace detail {
inline void check(s512_status st, const char* what) {
    if (st != S512_OK) {
        std::fprintf(stderr, "Sort512kv::%s failed (status %d): %s\n", what, st, s512_last_error());
        std::abort();
    }
}
}  // namespace detail

template <class SortType, class IndexType = size_t>
static inline void Sort(SortType array[], SortType values[], const IndexType size) {
    static_assert(std::is_same<SortType, int>::value, "Sort512kv sorts int keys with int values");
    detail::check(s512kv_sort_i32(array, values, size < IndexType(0) ? 0 : (size_t)size), "Sort");
}

template <class SortType, class IndexType = size_t>
static inline void SortOmpPartition(SortType array[], SortType values[], const IndexType size) {
    static_assert(std::is_same<SortType, int>::value, "Sort512kv sorts int keys with int values");
    detail::check(s512kv_sort_i32(array, values, size < IndexType(0) ? 0 : (size_t)size), "SortOmpPartition");
}

template <class IndexType>
static inline IndexType Partition512(int array[], int values[], IndexType left, IndexType right, const int pivot) {
    if (right < left) return left;
    size_t bound = 0;
    detail::check(s512kv_partition_i32(array, values, (size_t)left, (size_t)right, pivot, &bound), "Partition512");
    return static_cast<IndexType>(bound);
}

inline void SmallSort16V(int* __restrict__ ptr, int* __restrict__ values, const size_t length) {
    detail::check(s512kv_smallsort16v_i32(ptr, values, length), "SmallSort16V");
}

}  // namespace Sort512kv

#endif  // SORT512KV_HPP
