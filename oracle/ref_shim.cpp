// ref_shim.cpp -- C entry points over the reference headers, for the checker.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
// reference's own headers where they lie (REF_DIR, read-only), into
// oracle/_ref/libs512ref.so.  No reference source is copied into this
// repository: this file only instantiates the reference templates.
//
// Entry points wrap, one to one:
//   sort512.hpp:5960  Sort<T,size_t>              ref_sort_{i32,f64}
//   sort512.hpp:5993  SortOmpPartition<T,size_t>  ref_sort_omp_partition_{i32,f64}
//   sort512.hpp:6007  SortOmpMerge<T,size_t>      ref_sort_omp_merge_{i32,f64}
//   sort512.hpp:6061  SortOmpMergeDeps<T,size_t>  ref_sort_omp_mergedeps_{i32,f64}
//   sort512.hpp:6105  SortOmpParMerge<T,size_t>   ref_sort_omp_parmerge_{i32,f64}
//   sort512.hpp:5729/5821 Partition512<size_t>    ref_partition_{i32,f64}
//   sort512.hpp:5312/2255 SmallSort16V            ref_smallsort_{i32,f64}
//   sort512kv.hpp:6132 Sort                       ref_kv_sort
//   sort512kv.hpp:6164 SortOmpPartition           ref_kv_sort_omp_partition
//   sort512kv.hpp:5958 Partition512               ref_kv_partition
//   sort512kv.hpp:5235 SmallSort16V               ref_kv_smallsort
//   parallelInplace.hpp:78 FindMedian             ref_findmedian_{i32,f64}
//   sort512.hpp:5960  Sort<T,size_t> per segment  ref_segsort_loop_{i32,f64}
#include <cstddef>
#include <cstdint>
#include <omp.h>
#include "sort512.hpp"
#include "sort512kv.hpp"

// A batch of independent arrays on the CPU: Sort512::Sort over every segment
// [off[j], off[j+1]) of a, in order on one thread (parallel == 0) or spread
// over the OpenMP threads with schedule(dynamic).  Segments of fewer than two
// keys are skipped (the reference's Sort computes size-1).
template <class T>
static void segsort_loop(T* a, const uint64_t* off, size_t nseg, int parallel) {
    if (parallel) {
#pragma omp parallel for schedule(dynamic)
        for (long j = 0; j < (long)nseg; ++j)
            if (off[j + 1] - off[j] > 1) Sort512::Sort<T, size_t>(a + off[j], off[j + 1] - off[j]);
    } else {
        for (size_t j = 0; j < nseg; ++j)
            if (off[j + 1] - off[j] > 1) Sort512::Sort<T, size_t>(a + off[j], off[j + 1] - off[j]);
    }
}

extern "C" {
void ref_sort_i32(int* a, size_t n) { Sort512::Sort<int, size_t>(a, n); }
void ref_sort_f64(double* a, size_t n) { Sort512::Sort<double, size_t>(a, n); }
void ref_sort_omp_partition_i32(int* a, size_t n) { Sort512::SortOmpPartition<int, size_t>(a, n); }
void ref_sort_omp_partition_f64(double* a, size_t n) { Sort512::SortOmpPartition<double, size_t>(a, n); }
void ref_sort_omp_merge_i32(int* a, size_t n) { Sort512::SortOmpMerge<int, size_t>(a, n); }
void ref_sort_omp_merge_f64(double* a, size_t n) { Sort512::SortOmpMerge<double, size_t>(a, n); }
void ref_sort_omp_mergedeps_i32(int* a, size_t n) { Sort512::SortOmpMergeDeps<int, size_t>(a, n); }
void ref_sort_omp_mergedeps_f64(double* a, size_t n) { Sort512::SortOmpMergeDeps<double, size_t>(a, n); }
void ref_sort_omp_parmerge_i32(int* a, size_t n) { Sort512::SortOmpParMerge<int, size_t>(a, n); }
void ref_sort_omp_parmerge_f64(double* a, size_t n) { Sort512::SortOmpParMerge<double, size_t>(a, n); }
size_t ref_partition_i32(int* a, size_t l, size_t r, int p) { return Sort512::Partition512<size_t>(a, l, r, p); }
size_t ref_partition_f64(double* a, size_t l, size_t r, double p) { return Sort512::Partition512<size_t>(a, l, r, p); }
void ref_smallsort_i32(int* a, size_t n) { Sort512::SmallSort16V(a, n); }
void ref_smallsort_f64(double* a, size_t n) { Sort512::SmallSort16V(a, n); }
void ref_kv_sort(int* k, int* v, size_t n) { Sort512kv::Sort<int, size_t>(k, v, n); }
void ref_kv_sort_omp_partition(int* k, int* v, size_t n) { Sort512kv::SortOmpPartition<int, size_t>(k, v, n); }
size_t ref_kv_partition(int* k, int* v, size_t l, size_t r, int p) { return Sort512kv::Partition512<size_t>(k, v, l, r, p); }
void ref_kv_smallsort(int* k, int* v, size_t n) { Sort512kv::SmallSort16V(k, v, n); }
void ref_findmedian_i32(int* a, int center, int size, int* mA, int* mB) { ParallelInplace::FindMedian(a, center, size, mA, mB); }
void ref_findmedian_f64(double* a, int center, int size, int* mA, int* mB) { ParallelInplace::FindMedian(a, center, size, mA, mB); }
int ref_max_threads(void) { return omp_get_max_threads(); }
void ref_segsort_loop_i32(int* a, const uint64_t* off, size_t nseg, int parallel) { segsort_loop(a, off, nseg, parallel); }
void ref_segsort_loop_f64(double* a, const uint64_t* off, size_t nseg, int parallel) { segsort_loop(a, off, nseg, parallel); }
}
