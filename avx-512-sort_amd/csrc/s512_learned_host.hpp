// s512_learned_host.hpp -- host side of the learned partitioner.
#pragma once
#include "s512_learned.hpp"
#include "s512_runtime.hpp"
#include "s512_engine.hpp"

namespace s512 {

struct LearnedModel {
    LearnedModelDev knots;  // device knots (encoded key, CDF value)
    const float* params;    // device parameters after training
    float loss;             // final full-batch training loss
    bool degenerate;        // all sampled keys equal: no model
    bool point_mass;        // one feature value holds more than the imbalance
                            // limit's share of the sample: rejected
    uint32_t lf1;           // log2 of the first-level fan-out the model serves
};

// Samples ~prm.sample_permille/1000 of a (device array of n keys), sorts the
// sample with the engine, trains the model and computes its knots.  Caller
// holds ds.mu.  Synchronises the stream.
template <class T>
s512_status build_learned_model(const T* a, size_t n, hipStream_t st, DeviceState& ds, const EngineParams& prm,
                                LearnedModel* out, s512_stats* stats);

}  // namespace s512
