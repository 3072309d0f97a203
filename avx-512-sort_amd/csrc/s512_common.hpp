// s512_common.hpp -- shared device/host definitions of the MI355X sort library.
//
// Keys are compared as order-preserving unsigned integers ("encoded keys"):
//   int32  -> uint32  (flip the sign bit)
//   double -> uint64  (negative: invert all bits; otherwise set the sign bit)
// Every key is compared exactly; no precision is ever lowered.  For doubles the
// encoded order equals `<` for all non-NaN values except that -0.0 sorts before
// +0.0 (they compare equal in the reference, sort512.hpp:52-54, so either order
// is a valid ascending result).  Memory always holds the caller's raw type;
// keys are encoded on load and decoded on store.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace s512 {

template <class T> struct KeyTraits;

template <> struct KeyTraits<int32_t> {
    using K = uint32_t;
    static constexpr K KMAX = 0xFFFFFFFFu;
    __host__ __device__ static inline K enc(int32_t x) { return (uint32_t)x ^ 0x80000000u; }
    __host__ __device__ static inline K bits(int32_t x) { return (uint32_t)x; }
    __host__ __device__ static inline K enc_bits(K b) { return b ^ 0x80000000u; }
    __host__ __device__ static inline int32_t dec(K k) { return (int32_t)(k ^ 0x80000000u); }
};

template <> struct KeyTraits<double> {
    using K = uint64_t;
    static constexpr K KMAX = 0xFFFFFFFFFFFFFFFFull;
    // negative: all bits flipped; non-negative: sign bit set.  Branch-free:
    // xor with (sign-extended sign bit | sign bit).
    __host__ __device__ static inline K enc(double x) { return enc_bits(bits(x)); }
    __host__ __device__ static inline K bits(double x) { return __builtin_bit_cast(uint64_t, x); }
    __host__ __device__ static inline K enc_bits(K b) {
        const uint64_t m = (uint64_t)((int64_t)b >> 63);
        return b ^ (m | 0x8000000000000000ull);
    }
    __host__ __device__ static inline double dec(K k) {
        const uint64_t m = ~(uint64_t)((int64_t)k >> 63);
        return __builtin_bit_cast(double, k ^ (m | 0x8000000000000000ull));
    }
};

template <> struct KeyTraits<int64_t> {
    using K = uint64_t;
    static constexpr K KMAX = 0xFFFFFFFFFFFFFFFFull;
    __host__ __device__ static inline K enc(int64_t x) { return (uint64_t)x ^ 0x8000000000000000ull; }
    __host__ __device__ static inline int64_t dec(K k) { return (int64_t)(k ^ 0x8000000000000000ull); }
    __host__ __device__ static inline K bits(int64_t x) { return (uint64_t)x; }
    __host__ __device__ static inline K enc_bits(K b) { return b ^ 0x8000000000000000ull; }
};

// Identity traits: unsigned keys, and arrays that already hold encoded keys
// (splitter samples).
template <class KT> struct RawKeyTraits {
    using K = KT;
    static constexpr K KMAX = (K)~(K)0;
    __host__ __device__ static inline K enc(K x) { return x; }
    __host__ __device__ static inline K dec(K k) { return k; }
    __host__ __device__ static inline K bits(K x) { return x; }
    __host__ __device__ static inline K enc_bits(K b) { return b; }
};
template <> struct KeyTraits<uint32_t> : RawKeyTraits<uint32_t> {};
template <> struct KeyTraits<uint64_t> : RawKeyTraits<uint64_t> {};

// A contiguous run of keys that still has work to do.  `buf` says which of the
// two equally sized arrays (0 = caller's array, 1 = scratch) currently holds
// it; a run keeps the same [begin, begin+size) range in both arrays.
struct Seg {
    uint64_t begin;
    uint64_t size;
    uint32_t buf;
    uint32_t aux;   // planner-private (level of origin)
};

// Per-segment parameters of one multiway partition level.
struct PartSeg {
    uint64_t begin;
    uint64_t size;
    uint32_t buf;        // source array (destination is 1 - buf)
    uint32_t logf;       // fan-out F = 2^logf, F-1 splitters, 2F buckets
    uint32_t nchunks;    // chunks of `chunk` keys
    uint32_t nsample;    // samples drawn for the splitters
    uint64_t chunk;      // keys per chunk (last chunk may be short)
    uint64_t hist_base;  // offset of this segment's [nchunks][2F] histogram
    uint64_t tree_base;  // offset of this segment's sorted splitters ([F] keys)
    uint64_t bkt_base;   // offset of this segment's [2F] bucket sizes / starts
    uint64_t guide_base; // offset of this segment's guide table ([8F] cells)
    uint64_t kmin;       // guide table origin (encoded key), set on the device
    uint32_t shift;      // guide cell width 2^shift (encoded units), set on the device
    uint32_t mode;       // 0: sampled splitters, 2: the learned model's CDF table
    uint32_t lshift;     // learned: bucket = (fine bucket >> lshift) - lbase
    uint32_t lbase;
    uint32_t flags;      // classifier flags (GUIDE_SUBTABLES, GUIDE_EQBUCKETS), set on the device
    uint64_t sample_base;  // offset of this segment's samples in the sample buffer
    // one-pass levels only (s512_onepass.hpp)
    uint64_t dst;        // exact mode: position of the segment's first key in the destination;
                         // gapped mode: first slot of the segment's pair regions
    uint64_t fin;        // position of the segment's first key in sorted order (caller's array)
    uint64_t st_base;    // first look-back status word of the segment ([nchunks][2F])
};

struct ChunkRef {
    uint32_t seg;
    uint32_t idx;
};


__host__ __device__ constexpr int ilog2(unsigned v) { return v <= 1 ? 0 : 1 + ilog2(v >> 1); }

}  // namespace s512
