// s512_select.hip -- top-k (the k smallest keys, ascending) and selection
// (the k-th smallest key).  No reference counterpart: built from this
// library's sort for callers that need only a prefix of the order.
//
// 1. A threshold t from a sorted sample: a sample quantile a little above k/n,
//    so that #{key <= t} is likely >= k and not much larger.
// 2. k_count_le counts the keys <= t (encoded order, one pass over the keys).
//    Too few: t moves to a higher sample quantile (finally the largest key).
// 3. k_filter_le copies every key <= t into scratch (wave-aggregated slots,
//    in no particular order), and the library's sort orders those c keys; the
//    first k are the answer.  Equal keys make c larger than k, never wrong.
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

template <class T>
__global__ __launch_bounds__(256) void k_count_le(const T* a, uint64_t n, const typename KeyTraits<T>::K* t,
                                                  unsigned long long* cnt) {
    using K = typename KeyTraits<T>::K;
    const K th = *t;
    uint32_t c = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        c += KeyTraits<T>::enc(a[i]) <= th ? 1u : 0u;
    // wave sum, then one atomic per wave
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d);
    if ((threadIdx.x & 63u) == 0 && c) atomicAdd(cnt, (unsigned long long)c);
}

// Every key <= t to out[slot] (slots from one atomic per wave and pass).
template <class T>
__global__ __launch_bounds__(256) void k_filter_le(const T* a, uint64_t n, const typename KeyTraits<T>::K* t,
                                                   unsigned long long* cursor, T* out, uint64_t cap) {
    using K = typename KeyTraits<T>::K;
    const K th = *t;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    // every lane of a wave runs the same number of iterations (the ballot needs all of them)
    const uint64_t wave0 = blockIdx.x * (uint64_t)blockDim.x + (threadIdx.x & ~63u);
    for (uint64_t base = wave0; base < n; base += stride) {
        const uint64_t i = base + lane;
        const bool in = i < n;
        const T x = in ? a[i] : T(0);
        const bool keep = in && KeyTraits<T>::enc(x) <= th;
        const uint64_t m = __ballot(keep);
        if (!m) continue;
        unsigned long long w = 0;
        if (lane == 0) w = atomicAdd(cursor, (unsigned long long)__popcll(m));
        w = __shfl(w, 0);
        const uint64_t slot = w + __popcll(m & ((1ull << lane) - 1ull));
        if (keep && slot < cap) out[slot] = x;
    }
}

template <class T>
struct SelScratch {
    DevBuf smp, cnt, keys;
};

template <class T>
SelScratch<T>& sel_scratch(int dev) {
    static SelScratch<T> s[64];
    return s[dev];
}

// The c smallest-or-equal keys sorted into scratch; *outp points at them.
template <class T>
s512_status select_prefix(const T* a, size_t n, size_t k, hipStream_t st, T** outp, size_t* cout) {
    using K = typename KeyTraits<T>::K;
    int dev;
    if (s512_status rc = current_device(&dev)) return rc;
    static std::mutex mu[64];
    std::lock_guard<std::mutex> lk(mu[dev]);
    SelScratch<T>& sc = sel_scratch<T>(dev);
    const uint32_t S = (uint32_t)std::min<uint64_t>(n, SEL_SAMPLE);
    if (s512_status rc = sc.smp.reserve((size_t)S * sizeof(K) + 256, st)) return rc;
    if (s512_status rc = sc.cnt.reserve(4 * sizeof(unsigned long long), st)) return rc;
    K* smp = (K*)sc.smp.ptr;
    unsigned long long* cnt = (unsigned long long*)sc.cnt.ptr;
    k_sel_sample<T><<<(S + 255) / 256, 256, 0, st>>>(a, n, S, smp);
    S512_CHECK_LAUNCH();
    EngineParams sp = params_for(nullptr, sizeof(K) == 8 ? TUNE_F64 : TUNE_I32, S);
    if (s512_status rc = sort_engine<K, false>(smp, nullptr, S, st, sp, nullptr, false)) return rc;
    const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, 8192);
    // sample rank for the threshold: expected rank k/n*S plus 5% and 4 sigma,
    // widened 8x per retry; the last try takes the largest key (c = n)
    const double p = (double)k / (double)n;
    size_t c = 0;
    const K* th = nullptr;
    for (int attempt = 0;; ++attempt) {
        const double sig = std::sqrt(std::max(p * (1.0 - p) * S, 1.0));
        double q = p * S * 1.05 + (4.0 + 8.0 * attempt) * sig * (1 << (3 * attempt)) + 2.0;
        const uint32_t r = (attempt >= 3 || q >= S - 1) ? S - 1 : (uint32_t)q;
        th = smp + r;
        S512_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), st));
        if (r == S - 1) {
            // the largest sampled key may still miss larger keys: take all
            // keys up to the maximal encoded value (all ones) instead
            S512_HIP(hipMemsetAsync(smp + S, 0xFF, sizeof(K), st));
            th = smp + S;
        }
        k_count_le<T><<<grid, 256, 0, st>>>(a, n, th, cnt);
        S512_CHECK_LAUNCH();
        unsigned long long hc = 0;
        S512_HIP(hipMemcpyAsync(&hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, st));
        S512_HIP(hipStreamSynchronize(st));
        c = (size_t)hc;
        if (c >= k || th == smp + S) break;
    }
    if (s512_status rc = sc.keys.reserve(std::max<size_t>(c, 1) * sizeof(T), st)) return rc;
    T* out = (T*)sc.keys.ptr;
    S512_HIP(hipMemsetAsync(cnt + 1, 0, sizeof(unsigned long long), st));
    k_filter_le<T><<<grid, 256, 0, st>>>(a, n, th, cnt + 1, out, c);
    S512_CHECK_LAUNCH();
    EngineParams fp = params_for(nullptr, sizeof(K) == 8 ? TUNE_F64 : TUNE_I32, c);
    if (s512_status rc = sort_engine<T, false>(out, nullptr, c, st, fp, nullptr, false)) return rc;
    *outp = out;
    *cout = c;
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
    size_t c = 0;
    if (s512_status rc = select_prefix<T>((const T*)dk, n, k, st, &pre, &c)) return rc;
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
    size_t c = 0;
    if (s512_status rc = select_prefix<T>((const T*)dk, n, k + 1, st, &pre, &c)) return rc;
    S512_HIP(hipStreamSynchronize(st));
    S512_HIP(hipMemcpy(kth, pre + k, sizeof(T), hipMemcpyDeviceToHost));
    return S512_OK;
}

}  // namespace
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
