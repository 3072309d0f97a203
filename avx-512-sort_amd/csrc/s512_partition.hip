// s512_partition.hip -- Partition512 (sort512.hpp:5729-5910, sort512kv.hpp:
// 5958-6078) as a GPU kernel pair.
//
// Contract kept from the reference: the inclusive range [left, right] is
// rearranged so that every key `<= pivot` (int32: signed compare; double: the
// IEEE `<=`, so NaN goes right, as _CMP_LE_OQ does) precedes every other key,
// and the function returns left + #{keys <= pivot}.  The arrangement inside
// each side is unspecified in the reference; here it is stable (both sides keep
// their input order), which is one of the arrangements the reference allows.
//
// Kernels:
//   k_part_count    per chunk: number of keys <= pivot
//   k_part_offsets  one block: exclusive scan of the chunk counts
//   k_part_scatter  per chunk, in rounds of P*E keys: wave ballot -> lane rank,
//                   per-(slot, wave) counts scanned in LDS, keys written to the
//                   scratch array (low side from 0, high side from #low)
// followed by one device-to-device copy of the range back into the caller's
// array.  Keys are read twice and written twice.
#include "s512_common.hpp"
#include "s512_runtime.hpp"
#include "s512_hostio.hpp"
#include "../../include/s512.h"

namespace s512 {

constexpr int PART_P = 256;
constexpr int PART_E = 16;
constexpr int PART_ROUND = PART_P * PART_E;
constexpr int PART_NW = PART_P / 64;

template <class T>
__device__ __forceinline__ bool le_pivot(T x, T p) { return x <= p; }

template <class T>
__global__ __launch_bounds__(PART_P) void k_part_count(const T* a, uint64_t n, uint64_t chunk, T pivot, uint64_t* cnt) {
    __shared__ uint32_t wsum[PART_NW];
    const uint64_t lo = blockIdx.x * chunk;
    const uint64_t hi = lo + chunk < n ? lo + chunk : n;
    uint32_t c = 0;
    uint64_t i = lo + threadIdx.x;
    for (; i + 3 * PART_P < hi; i += 4 * PART_P) {
        const T x0 = a[i], x1 = a[i + PART_P], x2 = a[i + 2 * PART_P], x3 = a[i + 3 * PART_P];
        c += (le_pivot(x0, pivot) ? 1u : 0u) + (le_pivot(x1, pivot) ? 1u : 0u) + (le_pivot(x2, pivot) ? 1u : 0u) +
             (le_pivot(x3, pivot) ? 1u : 0u);
    }
    for (; i < hi; i += PART_P) c += le_pivot(a[i], pivot) ? 1u : 0u;
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t s = 0;
        for (int w = 0; w < PART_NW; ++w) s += wsum[w];
        cnt[blockIdx.x] = s;
    }
}

// cnt[c] -> exclusive prefix of low counts (in place); total low in cnt[nc].
__global__ __launch_bounds__(1024) void k_part_offsets(uint64_t* cnt, uint32_t nc) {
    __shared__ uint64_t part[1024];
    // sequential per thread over a strip, then a scan of the strip sums
    const uint32_t per = (nc + 1023) / 1024;
    const uint32_t b = threadIdx.x * per, e = b + per < nc ? b + per : nc;
    uint64_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += cnt[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t run = 0;
        for (int i = 0; i < 1024; ++i) {
            const uint64_t v = part[i];
            part[i] = run;
            run += v;
        }
        cnt[nc] = run;
    }
    __syncthreads();
    uint64_t run = part[threadIdx.x];
    for (uint32_t i = b; i < e; ++i) {
        const uint64_t v = cnt[i];
        cnt[i] = run;
        run += v;
    }
}

template <class T, bool KV>
__global__ __launch_bounds__(PART_P) void k_part_scatter(const T* a, const int32_t* va, uint64_t n, uint64_t chunk,
                                                         T pivot, const uint64_t* off, T* out, int32_t* vout) {
    __shared__ uint32_t tab[PART_E][PART_NW];  // per (slot, wave) low counts, then their prefix
    __shared__ uint32_t round_low;
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    const uint64_t lo = blockIdx.x * chunk;
    const uint64_t hi = lo + chunk < n ? lo + chunk : n;
    const uint64_t total_low = off[gridDim.x];
    uint64_t low_run = off[blockIdx.x];   // low keys before this point
    uint64_t high_run = lo - low_run;     // high keys before this point
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (uint64_t base = lo; base < hi; base += PART_ROUND) {
        T x[PART_E];
        int32_t xv[PART_E];
        uint64_t bal[PART_E];
        bool f[PART_E];
#pragma unroll
        for (int e = 0; e < PART_E; ++e) {
            const uint64_t i = base + (uint64_t)e * PART_P + t;
            const bool in = i < hi;
            const uint64_t ic = in ? i : hi - 1;  // in bounds, unconditional load
            x[e] = a[ic];
            if constexpr (KV) xv[e] = va[ic];
            f[e] = in && le_pivot(x[e], pivot);
        }
#pragma unroll
        for (int e = 0; e < PART_E; ++e) {
            bal[e] = __ballot(f[e]);
            if (lane == 0) tab[e][w] = (uint32_t)__popcll(bal[e]);
        }
        __syncthreads();
        if (t == 0) {
            uint32_t run = 0;
            for (int e = 0; e < PART_E; ++e)
                for (int ww = 0; ww < PART_NW; ++ww) {
                    const uint32_t v = tab[e][ww];
                    tab[e][ww] = run;
                    run += v;
                }
            round_low = run;
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < PART_E; ++e) {
            const uint64_t i = base + (uint64_t)e * PART_P + t;
            if (i >= hi) continue;
            const uint32_t lr = tab[e][w] + (uint32_t)__popcll(bal[e] & lt_mask);  // low keys before i in round
            const uint64_t in_round = (uint64_t)e * PART_P + t;                     // keys before i in round
            uint64_t d;
            if (f[e]) d = low_run + lr;
            else d = total_low + high_run + (in_round - lr);
            out[d] = x[e];
            if constexpr (KV) vout[d] = xv[e];
        }
        const uint32_t rl = round_low;
        const uint64_t cnt = (hi - base) < (uint64_t)PART_ROUND ? (hi - base) : (uint64_t)PART_ROUND;
        low_run += rl;
        high_run += cnt - rl;
        __syncthreads();
    }
}

template <class T, bool KV>
static s512_status partition_dev(T* a, int32_t* va, size_t left, size_t right, T pivot, size_t* out_bound,
                                 hipStream_t st) {
    if (!out_bound) { set_error("null out_bound"); return S512_ERR_INVALID; }
    if (right + 1 <= left) { *out_bound = left; return S512_OK; }  // empty range (right = left-1)
    if (!a || (KV && !va)) { set_error("null array"); return S512_ERR_INVALID; }
    const uint64_t n = (uint64_t)right - left + 1;
    int dev;
    s512_status rc = current_device(&dev);
    if (rc) return rc;
    DeviceState& ds = device_state(dev);
    std::lock_guard<std::mutex> lk(ds.mu);
    uint64_t chunk = (n + 2047) / 2048;
    chunk = (chunk + PART_ROUND - 1) / PART_ROUND * PART_ROUND;
    const uint32_t nc = (uint32_t)((n + chunk - 1) / chunk);
    const size_t dbytes = n * sizeof(T) + (KV ? n * sizeof(int32_t) : 0);
    if ((rc = ds.data.reserve(dbytes, st))) return rc;
    if ((rc = ds.arena.reserve((nc + 1) * sizeof(uint64_t), st))) return rc;
    if ((rc = ds.pinned2.reserve(4096, st))) return rc;
    T* out = (T*)ds.data.ptr;
    int32_t* vout = KV ? (int32_t*)((char*)ds.data.ptr + n * sizeof(T)) : nullptr;
    uint64_t* cnt = (uint64_t*)ds.arena.ptr;
    T* src = a + left;
    int32_t* vsrc = KV ? va + left : nullptr;
    k_part_count<T><<<nc, PART_P, 0, st>>>(src, n, chunk, pivot, cnt);
    S512_CHECK_LAUNCH();
    k_part_offsets<<<1, 1024, 0, st>>>(cnt, nc);
    S512_CHECK_LAUNCH();
    k_part_scatter<T, KV><<<nc, PART_P, 0, st>>>(src, vsrc, n, chunk, pivot, cnt, out, vout);
    S512_CHECK_LAUNCH();
    S512_HIP(hipMemcpyAsync(src, out, n * sizeof(T), hipMemcpyDeviceToDevice, st));
    if (KV) S512_HIP(hipMemcpyAsync(vsrc, vout, n * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
    uint64_t* h = (uint64_t*)ds.pinned2.ptr;
    S512_HIP(hipMemcpyAsync(h, cnt + nc, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    S512_HIP(hipStreamSynchronize(st));
    *out_bound = left + (size_t)h[0];
    return S512_OK;
}

// host arrays: stage [0, right] (the reference indexes from the base pointer)
template <class T, class Fn>
static s512_status on_device_range(T* a, size_t right, Fn fn) {
    return with_staged_array(a, right + 1, fn);
}

}  // namespace s512

using namespace s512;

extern "C" s512_status s512_partition_i32_dev(int32_t* d_a, size_t left, size_t right, int32_t pivot,
                                              size_t* out_bound, void* stream) {
    return partition_dev<int32_t, false>(d_a, nullptr, left, right, pivot, out_bound, (hipStream_t)stream);
}
extern "C" s512_status s512_partition_f64_dev(double* d_a, size_t left, size_t right, double pivot,
                                              size_t* out_bound, void* stream) {
    return partition_dev<double, false>(d_a, nullptr, left, right, pivot, out_bound, (hipStream_t)stream);
}
extern "C" s512_status s512kv_partition_i32_dev(int32_t* d_k, int32_t* d_v, size_t left, size_t right, int32_t pivot,
                                                size_t* out_bound, void* stream) {
    return partition_dev<int32_t, true>(d_k, d_v, left, right, pivot, out_bound, (hipStream_t)stream);
}
extern "C" s512_status s512_partition_i32(int32_t* a, size_t left, size_t right, int32_t pivot, size_t* out_bound) {
    if (!out_bound) { set_error("null out_bound"); return S512_ERR_INVALID; }
    if (right + 1 <= left) { *out_bound = left; return S512_OK; }
    if (!a) { set_error("null array"); return S512_ERR_INVALID; }
    return on_device_range(a, right, [&](int32_t* d, hipStream_t st) {
        return s512_partition_i32_dev(d, left, right, pivot, out_bound, st);
    });
}
extern "C" s512_status s512_partition_f64(double* a, size_t left, size_t right, double pivot, size_t* out_bound) {
    if (!out_bound) { set_error("null out_bound"); return S512_ERR_INVALID; }
    if (right + 1 <= left) { *out_bound = left; return S512_OK; }
    if (!a) { set_error("null array"); return S512_ERR_INVALID; }
    return on_device_range(a, right, [&](double* d, hipStream_t st) {
        return s512_partition_f64_dev(d, left, right, pivot, out_bound, st);
    });
}
extern "C" s512_status s512kv_partition_i32(int32_t* k, int32_t* v, size_t left, size_t right, int32_t pivot,
                                            size_t* out_bound) {
    if (!out_bound) { set_error("null out_bound"); return S512_ERR_INVALID; }
    if (right + 1 <= left) { *out_bound = left; return S512_OK; }
    if (!k || !v) { set_error("null array"); return S512_ERR_INVALID; }
    return with_staged_pair(k, v, right + 1, [&](int32_t* dk, int32_t* dv, hipStream_t st) {
        return s512kv_partition_i32_dev(dk, dv, left, right, pivot, out_bound, st);
    });
}
