// s512_learned_host.hpp -- host side of the learned partitioner.
#pragma once
#include "s512_learned.hpp"
#include "s512_runtime.hpp"
#include "s512_engine.hpp"

namespace s512 {

struct LearnedModel {
    LearnedModelDev tabs;   // CDF tables: the sample's (first level), the network's (second level)
    const float* params;    // device parameters (trained once `trained` completes)
    bool degenerate;        // all sampled keys equal: no model
    bool point_mass;        // one key value holds more than the imbalance
                            // limit's share of the sample: rejected
    uint32_t lf1;           // log2 of the first-level fan-out the model serves
    bool training;          // training launched on ds.lstream
    hipEvent_t trained;     // recorded on ds.lstream after the network's table
    const float* loss_host; // final training loss (pinned), valid after `trained`
};

// Samples ~prm.sample_permille/1000 of a (device array of n keys) and builds
// the sample's CDF table from histograms on stream st (synchronised); then,
// unless the sample shows a point mass, launches the network's training and
// its table on ds.lstream, which runs while the caller partitions the first
// level.  Caller holds ds.mu and must wait for `trained` before the call
// returns (or before using tabs.tab2, the root buckets' local tables).
template <class T>
s512_status build_learned_model(const T* a, size_t n, hipStream_t st, DeviceState& ds, const EngineParams& prm,
                                LearnedModel* out, s512_stats* stats);

}  // namespace s512
