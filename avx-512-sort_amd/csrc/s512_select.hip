// s512_select.hip -- top-k (the k smallest keys, ascending) and selection
// (the k-th smallest key).  No reference counterpart: built from this
// library's sort for callers that need only a prefix of the order.
//
// 1. A band [lo, hi] of keys from a sorted sample: sample quantiles a little
//    outside the wanted ranks (top-k: lo = the smallest key, hi just above
//    rank k; selection: around rank k), so the band likely holds them.
// 2. k_count_band counts, per block of 4096, the keys in the band and below
//    it (encoded order, one pass).  A band that misses a wanted rank is
//    widened (finally to the whole key range).
// 3. k_filter_band copies every key of the band into scratch at the block's offset
//    (exclusive scan of the per-block counts, k_sel_scan) plus its rank in the
//    block (wave ballots): input order, no atomics.  The library's sort orders
//    those c keys; the first k are the answer.  Equal keys make c larger than
//    k, never wrong.
#include "s512_common.hpp"
#include "s512_runtime.hpp"
#include "s512_engine.hpp"
#include "s512_hostio.hpp"
#include "../../include/s512.h"
#include <algorithm>
#include <cmath>

namespace s512 {
namespace {

constexpr uint32_t SEL_SAMPLE = 1u << 16;

// Sample i: key at position floor((2i+1) n / 2S) (evenly spaced, deterministic).
template <class T>
__global__ void k_sel_sample(const T* a, uint64_t n, uint32_t s, typename KeyTraits<T>::K* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s) return;
    const uint64_t pos = (((uint64_t)2 * i + 1) * n) / (2ull * s);
    out[i] = KeyTraits<T>::enc(a[pos]);
}

// Block b owns keys [b * SEL_TILE, (b + 1) * SEL_TILE): 16 rounds of 256.
constexpr uint32_t SEL_TILE = 4096;

// Per block: the number of its keys in the band [t[0], t[1]] (encoded order)
// and, when below != nullptr, the number below t[0].
template <class T>
__global__ __launch_bounds__(256) void k_count_band(const T* a, uint64_t n, const typename KeyTraits<T>::K* t,
                                                    uint32_t* bcnt, uint32_t* below) {
    using K = typename KeyTraits<T>::K;
    __shared__ uint32_t wsum[2][4];
    const K lo_t = t[0], hi_t = t[1];
    const uint64_t lo = (uint64_t)blockIdx.x * SEL_TILE;
    uint32_t c = 0, b = 0;
#pragma unroll 4
    for (uint32_t r = 0; r < SEL_TILE / 256; ++r) {
        const uint64_t i = lo + r * 256 + threadIdx.x;
        if (i < n) {
            const K e = KeyTraits<T>::enc(a[i]);
            c += (e >= lo_t && e <= hi_t) ? 1u : 0u;
            b += e < lo_t ? 1u : 0u;
        }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        c += __shfl_xor(c, d);
        b += __shfl_xor(b, d);
    }
    if ((threadIdx.x & 63u) == 0) {
        wsum[0][threadIdx.x >> 6] = c;
        wsum[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        bcnt[blockIdx.x] = wsum[0][0] + wsum[0][1] + wsum[0][2] + wsum[0][3];
        if (below) below[blockIdx.x] = wsum[1][0] + wsum[1][1] + wsum[1][2] + wsum[1][3];
    }
}

// One workgroup: bstart = exclusive prefix of bcnt[0..m), bstart[m] = total.
__global__ __launch_bounds__(1024) void k_sel_scan(const uint32_t* bcnt, uint32_t m, unsigned long long* bstart) {
    __shared__ unsigned long long part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (m + 1023) / 1024;
    const uint32_t b0 = t * per, b1 = min(m, b0 + per);
    unsigned long long s = 0;
    for (uint32_t b = b0; b < b1; ++b) s += bcnt[b];
    part[t] = s;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        const unsigned long long y = t >= d ? part[t - d] : 0ull;
        __syncthreads();
        part[t] += y;
        __syncthreads();
    }
    unsigned long long run = part[t] - s;
    for (uint32_t b = b0; b < b1; ++b) {
        bstart[b] = run;
        run += bcnt[b];
    }
    if (t == 1023) bstart[m] = part[1023];
}

// Every key in the band [t[0], t[1]] to out[bstart[block] + its rank among
// the block's kept keys] (input order; no atomics).
template <class T>
__global__ __launch_bounds__(256) void k_filter_band(const T* a, uint64_t n, const typename KeyTraits<T>::K* t,
                                                     const unsigned long long* bstart, T* out, uint64_t cap) {
    using K = typename KeyTraits<T>::K;
    __shared__ uint32_t wcnt[4];
    const K lo_t = t[0], hi_t = t[1];
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    const uint64_t lo = (uint64_t)blockIdx.x * SEL_TILE;
    uint64_t cur = bstart[blockIdx.x];
    for (uint32_t r = 0; r < SEL_TILE / 256; ++r) {
        const uint64_t i = lo + r * 256 + threadIdx.x;
        const bool in = i < n;
        const T x = in ? a[i] : T(0);
        const K e = KeyTraits<T>::enc(x);
        const bool keep = in && e >= lo_t && e <= hi_t;
        const uint64_t m = __ballot(keep);
        if (lane == 0) wcnt[wid] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0, total = 0;
#pragma unroll
        for (uint32_t w = 0; w < 4; ++w) {
            before += w < wid ? wcnt[w] : 0u;
            total += wcnt[w];
        }
        const uint64_t slot = cur + before + __popcll(m & ((1ull << lane) - 1ull));
        if (keep && slot < cap) out[slot] = x;
        cur += total;
        __syncthreads();  // wcnt is rewritten by the next round
    }
}

template <class T>
struct SelScratch {
    std::mutex mu;
    DevBuf smp, bcnt, bstart, keys;
};

template <class T>
SelScratch<T>& sel_scratch(int dev) {
    static SelScratch<T> s[64];
    return s[dev];
}

template <class T>
void release_sel(int dev) {
    SelScratch<T>& sc = sel_scratch<T>(dev);
    std::lock_guard<std::mutex> lk(sc.mu);
    (void)hipDeviceSynchronize();
    sc.smp.release();
    sc.bcnt.release();
    sc.bstart.release();
    sc.keys.release();
}

// Keys of ranks [r0, r1) of a (0 <= r0 < r1 <= n) sorted into scratch: the
// keys of a band [lo, hi] holding those ranks, sorted, starting at rank
// *below.  The band comes from sample quantiles a little outside r0/n and
// r1/n, widened and retried when the counts show it misses a rank; the last
// try is the whole key range (below = 0, c = n).  Top-k is r0 = 0 (lo = the
// smallest encoded value), selection r1 = r0 + 1.
template <class T>
s512_status select_band(const T* a, size_t n, size_t r0, size_t r1, hipStream_t st, T** outp, size_t* cout,
                        size_t* below) {
    using K = typename KeyTraits<T>::K;
    int dev;
    if (s512_status rc = current_device(&dev)) return rc;
    if (dev < 0 || dev >= 64) { set_error("device %d out of range", dev); return S512_ERR_INVALID; }
    SelScratch<T>& sc = sel_scratch<T>(dev);
    std::lock_guard<std::mutex> lk(sc.mu);
    const uint32_t S = (uint32_t)std::min<uint64_t>(n, SEL_SAMPLE);
    const uint64_t nblk = (n + SEL_TILE - 1) / SEL_TILE;
    if (nblk >= (1ull << 31)) { set_error("n = %zu too large", n); return S512_ERR_TOO_LARGE; }
    if (s512_status rc = sc.smp.reserve((size_t)S * sizeof(K) + 256, st)) return rc;
    if (s512_status rc = sc.bcnt.reserve(2 * nblk * sizeof(uint32_t), st)) return rc;
    if (s512_status rc = sc.bstart.reserve(2 * (nblk + 1) * sizeof(unsigned long long), st)) return rc;
    K* smp = (K*)sc.smp.ptr;
    K* band = smp + S;  // two keys after the sample: the band [lo, hi]
    uint32_t* bcnt = (uint32_t*)sc.bcnt.ptr;
    uint32_t* bbelow = bcnt + nblk;
    unsigned long long* bstart = (unsigned long long*)sc.bstart.ptr;
    unsigned long long* sbelow = bstart + (nblk + 1);
    k_sel_sample<T><<<(S + 255) / 256, 256, 0, st>>>(a, n, S, smp);
    S512_CHECK_LAUNCH();
    EngineParams sp = params_for(nullptr, sizeof(K) == 8 ? TUNE_F64 : TUNE_I32, S);
    if (s512_status rc = sort_engine<K, false>(smp, nullptr, S, st, sp, nullptr, false)) return rc;
    const unsigned grid = (unsigned)nblk;
    const double p0 = (double)r0 / (double)n, p1 = (double)r1 / (double)n;
    size_t c = 0, nb = 0;
    for (int attempt = 0;; ++attempt) {
        // sample ranks: 4 sigma (+5% of the rank) outside, 8x wider per retry;
        // the fourth try is the whole range
        const bool whole = attempt >= 3;
        const double w = (double)(1 << (3 * attempt));
        const double s0 = std::sqrt(std::max(p0 * (1.0 - p0) * S, 1.0));
        const double s1 = std::sqrt(std::max(p1 * (1.0 - p1) * S, 1.0));
        const double q0 = p0 * S * 0.95 - 4.0 * w * s0 - 2.0;
        const double q1 = p1 * S * 1.05 + 4.0 * w * s1 + 2.0;
        const bool lo_min = whole || r0 == 0 || q0 <= 0.0;
        const bool hi_max = whole || r1 == n || q1 >= (double)(S - 1);
        // band ends: sample keys, or the extreme encoded values
        if (lo_min) S512_HIP(hipMemsetAsync(band, 0x00, sizeof(K), st));
        else S512_HIP(hipMemcpyAsync(band, smp + (uint32_t)q0, sizeof(K), hipMemcpyDeviceToDevice, st));
        if (hi_max) S512_HIP(hipMemsetAsync(band + 1, 0xFF, sizeof(K), st));
        else S512_HIP(hipMemcpyAsync(band + 1, smp + (uint32_t)q1, sizeof(K), hipMemcpyDeviceToDevice, st));
        // a band from the smallest key has nothing below it: no second count
        k_count_band<T><<<grid, 256, 0, st>>>(a, n, band, bcnt, lo_min ? nullptr : bbelow);
        S512_CHECK_LAUNCH();
        k_sel_scan<<<1, 1024, 0, st>>>(bcnt, (uint32_t)nblk, bstart);
        S512_CHECK_LAUNCH();
        if (!lo_min) {
            k_sel_scan<<<1, 1024, 0, st>>>(bbelow, (uint32_t)nblk, sbelow);
            S512_CHECK_LAUNCH();
        }
        unsigned long long hc[2] = {0, 0};
        S512_HIP(hipMemcpyAsync(&hc[0], bstart + nblk, sizeof(hc[0]), hipMemcpyDeviceToHost, st));
        if (!lo_min) S512_HIP(hipMemcpyAsync(&hc[1], sbelow + nblk, sizeof(hc[1]), hipMemcpyDeviceToHost, st));
        S512_HIP(hipStreamSynchronize(st));
        c = (size_t)hc[0];
        nb = (size_t)hc[1];
        if ((nb <= r0 && nb + c >= r1) || (lo_min && hi_max)) break;
    }
    if (s512_status rc = sc.keys.reserve(std::max<size_t>(c, 1) * sizeof(T), st)) return rc;
    T* out = (T*)sc.keys.ptr;
    k_filter_band<T><<<grid, 256, 0, st>>>(a, n, band, bstart, out, c);
    S512_CHECK_LAUNCH();
    EngineParams fp = params_for(nullptr, sizeof(K) == 8 ? TUNE_F64 : TUNE_I32, c);
    if (s512_status rc = sort_engine<T, false>(out, nullptr, c, st, fp, nullptr, false)) return rc;
    *outp = out;
    *cout = c;
    *below = nb;
    return S512_OK;
}

template <class T>
s512_status topk_any(const T* keys, size_t n, size_t k, T* out, hipStream_t st) {
    if (k == 0) return S512_OK;
    if (!keys || !out) { set_error("null array"); return S512_ERR_INVALID; }
    if (k > n) { set_error("k = %zu exceeds n = %zu", k, n); return S512_ERR_INVALID; }
    const bool kdev = is_device_pointer(keys), odev = is_device_pointer(out);
    HostStage hs;
    void* dk = (void*)keys;
    if (!kdev)
        if (s512_status rc = hs.in(0, keys, n * sizeof(T), &dk)) return rc;
    T* pre = nullptr;
    size_t c = 0, nb = 0;  // the band of ranks [0, k) starts at the smallest key: nb = 0
    if (s512_status rc = select_band<T>((const T*)dk, n, 0, k, st, &pre, &c, &nb)) return rc;
    if (odev) {
        S512_HIP(hipMemcpyAsync(out, pre, k * sizeof(T), hipMemcpyDeviceToDevice, st));
        S512_HIP(hipStreamSynchronize(st));
    } else {
        S512_HIP(hipStreamSynchronize(st));
        S512_HIP(hipMemcpy(out, pre, k * sizeof(T), hipMemcpyDeviceToHost));
    }
    return S512_OK;
}

template <class T>
s512_status select_any(const T* keys, size_t n, size_t k, T* kth, hipStream_t st) {
    if (!keys || !kth) { set_error("null pointer"); return S512_ERR_INVALID; }
    if (k >= n) { set_error("k = %zu out of range for n = %zu", k, n); return S512_ERR_INVALID; }
    const bool kdev = is_device_pointer(keys);
    HostStage hs;
    void* dk = (void*)keys;
    if (!kdev)
        if (s512_status rc = hs.in(0, keys, n * sizeof(T), &dk)) return rc;
    T* pre = nullptr;
    size_t c = 0, nb = 0;
    if (s512_status rc = select_band<T>((const T*)dk, n, k, k + 1, st, &pre, &c, &nb)) return rc;
    S512_HIP(hipStreamSynchronize(st));
    S512_HIP(hipMemcpy(kth, pre + (k - nb), sizeof(T), hipMemcpyDeviceToHost));
    return S512_OK;
}

}  // namespace

void release_select_scratch(int dev) {
    if (dev < 0 || dev >= 64) return;
    release_sel<int32_t>(dev);
    release_sel<double>(dev);
    release_sel<int64_t>(dev);
    release_sel<uint64_t>(dev);
}

}  // namespace s512

using namespace s512;

#define S512_SELECT(NAME, T)                                                                              \
    extern "C" s512_status s512_topk_##NAME(const T* keys, size_t n, size_t k, T* out, void* stream) {    \
        return topk_any<T>(keys, n, k, out, (hipStream_t)stream);                                         \
    }                                                                                                     \
    extern "C" s512_status s512_select_##NAME(const T* keys, size_t n, size_t k, T* kth, void* stream) {  \
        return select_any<T>(keys, n, k, kth, (hipStream_t)stream);                                       \
    }
S512_SELECT(i32, int32_t)
S512_SELECT(f64, double)
S512_SELECT(i64, int64_t)
S512_SELECT(u64, uint64_t)
#undef S512_SELECT
