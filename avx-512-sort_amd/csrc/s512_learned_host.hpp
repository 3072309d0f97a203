// s512_learned_host.hpp -- host side of the learned partitioner.
#pragma once
#include "s512_learned.hpp"
#include "s512_runtime.hpp"
#include "s512_engine.hpp"

namespace s512 {

struct LearnedModel {
    LearnedModelDev tabs;   // CDF tables: the sample's (first level), the network's (second level)
    const float* params;    // device parameters (trained once `trained` completes)
    // pinned, written on the caller's stream after the sample's table: the
    // smallest and largest sampled encoded key (equal: degenerate sample), then
    // (uint32 in the third word) the sampled keys on repeated values (over an
    // eighth of the sample: a table cannot split them).  Read once the stream
    // has passed that point.
    const uint64_t* flags_host;
    uint32_t ns;            // sampled keys
    uint32_t lf1;           // log2 of the first-level fan-out the model serves
    bool training;          // training launched on ds.lstream
    hipEvent_t trained;     // recorded on ds.lstream after the network's table
    const float* loss_host; // final training loss (pinned), valid after `trained`
    hipEvent_t ev0, ev1;    // (statistics only) around the sample and its table
};

// Rejection test of flags_host: a degenerate sample, or repeated values over
// an eighth of it.
inline bool learned_model_rejected(const LearnedModel& m) {
    const uint32_t repeated = *reinterpret_cast<const uint32_t*>(m.flags_host + 2);
    return m.flags_host[1] <= m.flags_host[0] || repeated > m.ns / 8;
}

// Samples ~prm.sample_permille/1000 of a (device array of n keys) and builds
// the sample's CDF table from histograms on stream st, then launches the
// network's training and its tables on ds.lstream, which runs while the caller
// partitions the first level.  Nothing waits on the host: the caller reads
// flags_host after a later synchronisation of st and drops the model for the
// second level if it is rejected (the first level's buckets are ordered key
// ranges either way).  Caller holds ds.mu and must wait for `trained` before
// the call returns (or before using tabs.tab2, the root buckets' local tables).
template <class T>
s512_status build_learned_model(const T* a, size_t n, hipStream_t st, DeviceState& ds, const EngineParams& prm,
                                LearnedModel* out, s512_stats* stats);

}  // namespace s512
