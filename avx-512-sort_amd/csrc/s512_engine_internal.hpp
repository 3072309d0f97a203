// s512_engine_internal.hpp -- pieces of the sort engine shared with the
// segmented sort: the partition level loop and the plain-key tile launches.
#pragma once
#include <vector>
#include "s512_kernels.hpp"
#include "s512_learned.hpp"
#include "s512_runtime.hpp"
#include "s512_engine.hpp"

namespace s512 {

// State of the learned partitioner across the level loop: the model's CDF
// tables place the keys of the root level (2^lf1 buckets) and of its buckets.
struct LearnedModel;
struct LevelState {
    bool use_learned = false;
    const LearnedModel* model = nullptr;  // the learned model (its rejection flags)
    bool check_root = false;               // the model's rejection test is still due (first level)
    uint32_t lf1 = 0;
    LearnedModelDev tabs{nullptr, nullptr};  // the root's CDF table, the root buckets' local tables
    hipEvent_t trained = nullptr;            // tabs.tab2 is ready once this event completes
    const float* loss_host = nullptr;      // training loss (pinned), valid after `trained`
    float max_loss = 0.0f;
    bool pending = false;                  // the second level still has to wait for `trained`
};

// The partition level loop over an initial list of segments of the arrays
// a (buf 0) and b (buf 1), both n keys long (key/value: va, vb).  Synchronises
// the stream between levels; returns with every segment sorted into a.
template <class T, bool KV>
s512_status run_levels(T* a, int32_t* va, T* b, int32_t* vb, size_t n, std::vector<Seg> segs, hipStream_t st,
                       const EngineParams& prm, s512_stats* stats, Profiler& pf, DeviceState& ds,
                       const LearnedModelDev& model, LevelState& lstate);

// Capacities (keys) of the tile classes for plain keys of type T.
template <class T>
const uint32_t* tile_caps(int* ncls);

// The engine's tile classes and launches (keys, or int32 key/value pairs):
// sort every listed segment of class cls from (buf ? b : a) into a.
template <class T, bool KV>
const uint32_t* tile_class_caps(int* ncls);
template <class T, bool KV>
void launch_tile_list(int cls, int plan, const TileSeg* segs, const uint32_t* dcount, uint32_t count, T* a,
                      int32_t* va, const T* b, const int32_t* vb, hipStream_t st, bool f64_fast);

// One-pass levels (s512_onepass.hip): the root level and the second level
// each read and write every key once; segments that need more levels are
// finished by run_levels.  *done = false (with nothing changed) when the
// root level cannot be done that way (a capacity region overflowed); the
// caller then runs the multi-level engine on the whole array.
template <class T, bool KV>
s512_status onepass_engine(T* a, int32_t* va, size_t n, hipStream_t st, const EngineParams& prm, s512_stats* stats,
                           Profiler& pf, DeviceState& ds, bool* done);

// Launch the tile sort of class cls over the device-counted list segs with a
// grid sized for grid_items segments (the kernels loop over the true count).
template <class T>
void launch_tile_fixed(int cls, int plan, const TileSeg* segs, const uint32_t* dcount, uint32_t grid_items, T* a,
                       const T* b, hipStream_t st);

}  // namespace s512
