// s512_engine.hpp -- host interface of the multiway sort engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/s512.h"

namespace s512 {

struct EngineParams {
    uint32_t target_leaf;    // desired mean bucket size after the last level
    uint32_t oversample;     // samples per splitter
    uint32_t target_chunks;  // chunks per level (load balance vs. histogram size)
    uint32_t max_chunk;      // keys per chunk, upper bound
    int max_levels;          // safety bound on partition levels
    uint64_t seed;           // sample positions
};

struct EngineStats {
    int levels;
    uint64_t tiles;
};

EngineParams default_params(bool wide_keys);

template <class T, bool KV>
s512_status sort_engine(T* a, int32_t* va, size_t n, hipStream_t st, const EngineParams& prm, EngineStats* stats);

}  // namespace s512
