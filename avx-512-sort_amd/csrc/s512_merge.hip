// s512_merge.hip -- merge-path merging: the counterpart of the reference's
// OpenMP in-place merge (ParallelInplace::parallelMergeInPlace,
// parallelInplace.hpp:209-306, whose FindMedian at 77-134 splits a merge into
// independent pieces) and of the merge-based sorts SortOmpMerge /
// SortOmpMergeDeps / SortOmpParMerge (sort512.hpp:6006-6174: sort chunks, then
// merge pairs of runs level by level).
//
// Merge path (Green, McColl, Bader 2012): output position p of the stable
// merge of A and B takes co-rank(p) keys from A, found by binary search along
// the cross diagonal.  One workgroup produces D = P*E consecutive outputs:
//   k_merge_split  co-rank at both ends of every output tile
//   k_merge_tile   tile's inputs into LDS; each thread finds its own co-rank
//                  inside the tile and merges E keys; staged, coalesced store
// Ties take the key from A first (stable, as std::inplace_merge).  Keys are
// compared in the encoded order (see s512_common.hpp).
#include "s512_kernels.hpp"
#include "s512_runtime.hpp"
#include "s512_hostio.hpp"
#include "../../include/s512.h"
#include <algorithm>
#include <vector>

namespace s512 {

constexpr int MRG_P = 256;
constexpr int MRG_E = 8;
constexpr int MRG_D = MRG_P * MRG_E;

// Which merge does output position p belong to?  Either one merge of
// [0,center) with [center,n), or (center == 0) pairs of runs of width w.
struct MergeShape {
    uint64_t n, w, center;
    __host__ __device__ void pair_of(uint64_t p, uint64_t& ps, uint64_t& pa, uint64_t& pe) const {
        if (center) {
            ps = 0;
            pa = center;
            pe = n;
        } else {
            ps = (p / (2 * w)) * (2 * w);
            pa = ps + w < n ? ps + w : n;
            pe = ps + 2 * w < n ? ps + 2 * w : n;
        }
    }
};

template <class T>
__device__ __forceinline__ uint64_t corank_dev(const T* A, uint64_t na, const T* B, uint64_t nb, uint64_t diag) {
    using Tr = KeyTraits<T>;
    uint64_t lo = diag > nb ? diag - nb : 0, hi = diag < na ? diag : na;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (Tr::enc(B[diag - mid - 1]) < Tr::enc(A[mid])) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

template <class T>
__global__ __launch_bounds__(256) void k_merge_split(const T* src, MergeShape sh, uint64_t ntiles, uint64_t* split) {
    const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (k >= ntiles) return;
    const uint64_t p0 = k * MRG_D, p1 = p0 + MRG_D < sh.n ? p0 + MRG_D : sh.n;
    uint64_t ps, pa, pe;
    sh.pair_of(p0, ps, pa, pe);
    const T* A = src + ps;
    const T* B = src + pa;
    split[2 * k] = corank_dev(A, pa - ps, B, pe - pa, p0 - ps);
    split[2 * k + 1] = corank_dev(A, pa - ps, B, pe - pa, p1 - ps);
}

template <class T>
__global__ __launch_bounds__(MRG_P) void k_merge_tile(const T* src, T* dst, MergeShape sh, const uint64_t* split) {
    using Tr = KeyTraits<T>;
    using K = typename Tr::K;
    __shared__ K sin[MRG_D];
    __shared__ K sout[MRG_D];
    const uint64_t k = blockIdx.x;
    const uint64_t p0 = k * MRG_D, p1 = p0 + MRG_D < sh.n ? p0 + MRG_D : sh.n;
    uint64_t ps, pa, pe;
    sh.pair_of(p0, ps, pa, pe);
    const uint64_t a0 = split[2 * k], a1 = split[2 * k + 1];
    const uint64_t b0 = (p0 - ps) - a0, b1 = (p1 - ps) - a1;
    const uint32_t la = (uint32_t)(a1 - a0), lb = (uint32_t)(b1 - b0), tot = la + lb;
    const T* A = src + ps + a0;
    const T* B = src + pa + b0;
    for (uint32_t i = threadIdx.x; i < tot; i += MRG_P) sin[i] = i < la ? Tr::enc(A[i]) : Tr::enc(B[i - la]);
    __syncthreads();
    const K* sA = sin;
    const K* sB = sin + la;
    const uint32_t ld = threadIdx.x * MRG_E < tot ? threadIdx.x * MRG_E : tot;
    // co-rank of the thread's first output inside the tile
    uint32_t lo = ld > lb ? ld - lb : 0, hi = ld < la ? ld : la;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sB[ld - mid - 1] < sA[mid]) hi = mid;
        else lo = mid + 1;
    }
    uint32_t i = lo, j = ld - lo;
#pragma unroll
    for (int e = 0; e < MRG_E; ++e) {
        if (ld + e >= tot) break;
        const bool takeA = i < la && (j >= lb || !(sB[j] < sA[i]));
        sout[ld + e] = takeA ? sA[i] : sB[j];
        i += takeA ? 1u : 0u;
        j += takeA ? 0u : 1u;
    }
    __syncthreads();
    for (uint32_t x = threadIdx.x; x < tot; x += MRG_P) dst[p0 + x] = Tr::dec(sout[x]);
}

template <class T>
static s512_status merge_round(const T* src, T* dst, MergeShape sh, uint64_t* split, hipStream_t st) {
    const uint64_t ntiles = (sh.n + MRG_D - 1) / MRG_D;
    k_merge_split<T><<<(unsigned)((ntiles + 255) / 256), 256, 0, st>>>(src, sh, ntiles, split);
    S512_CHECK_LAUNCH();
    k_merge_tile<T><<<(unsigned)ntiles, MRG_P, 0, st>>>(src, dst, sh, split);
    S512_CHECK_LAUNCH();
    return S512_OK;
}

template <class T>
static s512_status merge_inplace_dev(T* a, size_t size, size_t center, hipStream_t st) {
    if (size <= 1 || center == 0 || center >= size) return S512_OK;
    if (!a) { set_error("null array"); return S512_ERR_INVALID; }
    int dev;
    s512_status rc = current_device(&dev);
    if (rc) return rc;
    DeviceState& ds = device_state(dev);
    std::lock_guard<std::mutex> lk(ds.mu);
    const uint64_t ntiles = (size + MRG_D - 1) / MRG_D;
    if ((rc = ds.data.reserve(size * sizeof(T), st))) return rc;
    if ((rc = ds.arena.reserve(2 * ntiles * sizeof(uint64_t), st))) return rc;
    T* tmp = (T*)ds.data.ptr;
    MergeShape sh{size, 0, center};
    if ((rc = merge_round<T>(a, tmp, sh, (uint64_t*)ds.arena.ptr, st))) return rc;
    S512_HIP(hipMemcpyAsync(a, tmp, size * sizeof(T), hipMemcpyDeviceToDevice, st));
    S512_HIP(hipStreamSynchronize(st));
    return S512_OK;
}

// Sorted runs of width w (the last one may be shorter) of a[0, n) merged,
// round after round, into one; ds.data holds the other buffer of each round.
// Caller holds ds.mu.
template <class T>
s512_status merge_runs_locked(T* a, size_t n, uint64_t w, hipStream_t st, DeviceState& ds) {
    if (n <= 1 || w >= n) return S512_OK;
    const uint64_t ntiles = (n + MRG_D - 1) / MRG_D;
    s512_status rc;
    if ((rc = ds.data.reserve(n * sizeof(T), st))) return rc;
    if ((rc = ds.arena.reserve(2 * ntiles * sizeof(uint64_t), st))) return rc;
    uint64_t* split = (uint64_t*)ds.arena.ptr;
    T* src = a;
    T* dst = (T*)ds.data.ptr;
    for (; w < n; w *= 2) {
        if ((rc = merge_round<T>(src, dst, MergeShape{n, w, 0}, split, st))) return rc;
        std::swap(src, dst);
    }
    if (src != a) S512_HIP(hipMemcpyAsync(a, src, n * sizeof(T), hipMemcpyDeviceToDevice, st));
    S512_HIP(hipStreamSynchronize(st));
    return S512_OK;
}
template s512_status merge_runs_locked<int32_t>(int32_t*, size_t, uint64_t, hipStream_t, DeviceState&);
template s512_status merge_runs_locked<double>(double*, size_t, uint64_t, hipStream_t, DeviceState&);
template s512_status merge_runs_locked<int64_t>(int64_t*, size_t, uint64_t, hipStream_t, DeviceState&);
template s512_status merge_runs_locked<uint64_t>(uint64_t*, size_t, uint64_t, hipStream_t, DeviceState&);
template s512_status merge_runs_locked<uint32_t>(uint32_t*, size_t, uint64_t, hipStream_t, DeviceState&);

// Merge sort: bitonic tiles of the largest tile class, then merge-path rounds.
template <class T>
static s512_status merge_sort_dev(T* a, size_t n, hipStream_t st) {
    using K = typename KeyTraits<T>::K;
    if (n <= 1) return S512_OK;
    if (!a) { set_error("null array"); return S512_ERR_INVALID; }
    int dev;
    s512_status rc = current_device(&dev);
    if (rc) return rc;
    DeviceState& ds = device_state(dev);
    std::lock_guard<std::mutex> lk(ds.mu);
    constexpr uint32_t T_TILE = sizeof(K) == 8 ? 8192 : 16384;
    const uint64_t nt = (n + T_TILE - 1) / T_TILE;
    const uint64_t ntiles = (n + MRG_D - 1) / MRG_D;
    SizeCounter sc;
    sc.take<TileSeg>(nt);
    sc.take<uint32_t>(1);
    const size_t up = sc.off;
    sc.take<uint64_t>(2 * ntiles);
    if ((rc = ds.arena.reserve(sc.off, st))) return rc;
    if ((rc = ds.pinned.reserve(up, st))) return rc;
    if ((rc = ds.data.reserve(n * sizeof(T), st))) return rc;
    Arena ar{(char*)ds.arena.ptr, 0}, hp{(char*)ds.pinned.ptr, 0};
    TileSeg* dseg = ar.take<TileSeg>(nt);
    uint32_t* dcnt = ar.take<uint32_t>(1);
    uint64_t* split = ar.take<uint64_t>(2 * ntiles);
    TileSeg* hseg = hp.take<TileSeg>(nt);
    uint32_t* hcnt = hp.take<uint32_t>(1);
    for (uint64_t i = 0; i < nt; ++i) {
        const uint64_t b = i * T_TILE;
        hseg[i] = TileSeg{b, (uint32_t)std::min<uint64_t>(T_TILE, n - b), 0};
    }
    *hcnt = (uint32_t)nt;
    S512_HIP(hipMemcpyAsync(ds.arena.ptr, ds.pinned.ptr, up, hipMemcpyHostToDevice, st));
    if constexpr (sizeof(K) == 8) k_tile_sort<T, 512, 16><<<(unsigned)nt, 512, 0, st>>>(dseg, dcnt, a, a);
    else k_tile_sort<T, 1024, 16><<<(unsigned)nt, 1024, 0, st>>>(dseg, dcnt, a, a);
    S512_CHECK_LAUNCH();
    T* src = a;
    T* dst = (T*)ds.data.ptr;
    for (uint64_t w = T_TILE; w < n; w *= 2) {
        if ((rc = merge_round<T>(src, dst, MergeShape{n, w, 0}, split, st))) return rc;
        std::swap(src, dst);
    }
    if (src != a) S512_HIP(hipMemcpyAsync(a, src, n * sizeof(T), hipMemcpyDeviceToDevice, st));
    S512_HIP(hipStreamSynchronize(st));  // pinned staging is reused by the next call
    return S512_OK;
}

template <class T, class Fn>
static s512_status on_device(T* a, size_t n, Fn fn) {
    return with_staged_array(a, n, fn);
}

}  // namespace s512

using namespace s512;

extern "C" s512_status s512_merge_inplace_i32_dev(int32_t* d_a, size_t size, size_t center, void* stream) {
    return merge_inplace_dev<int32_t>(d_a, size, center, (hipStream_t)stream);
}
extern "C" s512_status s512_merge_inplace_f64_dev(double* d_a, size_t size, size_t center, void* stream) {
    return merge_inplace_dev<double>(d_a, size, center, (hipStream_t)stream);
}
extern "C" s512_status s512_merge_inplace_i32(int32_t* a, size_t size, size_t center) {
    if (size <= 1 || center == 0 || center >= size) return S512_OK;
    if (!a) { set_error("null array"); return S512_ERR_INVALID; }
    return on_device(a, size, [&](int32_t* d, hipStream_t st) { return merge_inplace_dev<int32_t>(d, size, center, st); });
}
extern "C" s512_status s512_merge_inplace_f64(double* a, size_t size, size_t center) {
    if (size <= 1 || center == 0 || center >= size) return S512_OK;
    if (!a) { set_error("null array"); return S512_ERR_INVALID; }
    return on_device(a, size, [&](double* d, hipStream_t st) { return merge_inplace_dev<double>(d, size, center, st); });
}
extern "C" s512_status s512_sort_merge_i32_dev(int32_t* d_a, size_t n, void* stream) {
    return merge_sort_dev<int32_t>(d_a, n, (hipStream_t)stream);
}
extern "C" s512_status s512_sort_merge_f64_dev(double* d_a, size_t n, void* stream) {
    return merge_sort_dev<double>(d_a, n, (hipStream_t)stream);
}
extern "C" s512_status s512_sort_merge_i32(int32_t* a, size_t n) {
    if (n <= 1) return S512_OK;
    if (!a) { set_error("null array"); return S512_ERR_INVALID; }
    return on_device(a, n, [&](int32_t* d, hipStream_t st) { return merge_sort_dev<int32_t>(d, n, st); });
}
extern "C" s512_status s512_sort_merge_f64(double* a, size_t n) {
    if (n <= 1) return S512_OK;
    if (!a) { set_error("null array"); return S512_ERR_INVALID; }
    return on_device(a, n, [&](double* d, hipStream_t st) { return merge_sort_dev<double>(d, n, st); });
}
