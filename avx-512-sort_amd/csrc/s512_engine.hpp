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
    int max_logf;            // cap on log2 fan-out per level
    bool f64_minmax;         // fp64 leaves by v_min/v_max when the input allows
    bool learned;            // learned CDF partitioning for the first two levels
    uint32_t sample_permille;// learned: sample size (per mille of n)
    uint32_t train_steps;    // learned: Adam steps
    float learn_rate;        // learned: Adam step size
    float imbalance;         // learned: fall back when a level-1 bucket exceeds this x mean
    float max_loss;          // learned: reject a network whose training loss (MSE on the CDF) is above
    uint64_t seed;           // sample positions
    int scatter_shape;       // k_scatter workgroup shape (0 .. SCATTER_SHAPES-1)
    bool ids;                // count stores each key's bucket; scatter reads it
    int count_shape;         // k_count shape (0 .. COUNT_SHAPES-1)
    int tile_plan;           // tile sort shapes (0 .. TILE_PLANS-1)
    int engine;              // S512_ENGINE_*: two-pass levels (auto) or one-pass levels
    uint64_t onepass_min;    // one-pass: arrays of at least this many keys take the one-pass levels
    float region_z;          // one-pass root level: region margin in standard deviations of its estimate
};

EngineParams default_params(bool wide_keys);
// Defaults overridden by the non-zero fields of `opt` (may be null).
EngineParams params_from(const s512_options* opt, bool wide_keys);

// Workload kinds of the autotuned defaults table (s512_tune_table.inc).
enum TuneKind { TUNE_I32 = 0, TUNE_F64 = 1, TUNE_KV = 2, TUNE_SEG_I32 = 3, TUNE_SEG_F64 = 4 };
// Defaults, then the autotuned entry for (kind, n keys), then the non-zero
// fields of `opt` (explicit options always win).
EngineParams params_for(const s512_options* opt, int kind, uint64_t n);

template <class T, bool KV>
s512_status sort_engine(T* a, int32_t* va, size_t n, hipStream_t st, const EngineParams& prm, s512_stats* stats,
                        bool profile);

struct DeviceState;
// Same, with the device workspace lock already held by the caller.
template <class T, bool KV>
s512_status sort_engine_locked(T* a, int32_t* va, size_t n, hipStream_t st, const EngineParams& prm,
                               s512_stats* stats, bool profile, DeviceState& ds);

// Sorted runs of width w of a[0, n) merged into one (merge path, through the
// device's scratch buffer); caller holds ds.mu.
template <class T>
s512_status merge_runs_locked(T* a, size_t n, uint64_t w, hipStream_t st, DeviceState& ds);

}  // namespace s512
