// Flattens a Scenario into the plain arrays of the C ABI (include/pansim_hip.h)
// and computes the derived progression quantities: the geometric parameter p
// of every (matrix, state, variant) from BasicLengthAbstract::LengthOfState
// (src/tools/progressionMatrices/basicLengthAbstract.cpp:4-44).
#pragma once
#include <vector>

#include "../../../include/pansim_hip.h"
#include "scenario.h"

namespace ps {

struct ModelArrays {
    pansim_dims dims{};
    pansim_tables tables{};
    pansim_locations locations{};
    pansim_agents agents{};
    // storage behind the pointers above
    std::vector<uint32_t> edgeOff, badCount, edgeTo, eventOffset, locType32;
    std::vector<float> edgeChance, lenAvg, lenMax;
    std::vector<double> lenP;
    std::vector<pansim_event> events;
};

// Geometric parameter of a state length with mean avgLength truncated at
// maxLength (secant solve of the truncated mean).
double length_parameter(float avgLength, float maxLength);

// The returned object's pointers refer into itself and into `sc`; both must
// outlive the pansim_create call.
void build_model(const Scenario& sc, ModelArrays& m);

std::vector<pansim_event> to_c_events(const std::vector<Event>& ev);

}  // namespace ps
