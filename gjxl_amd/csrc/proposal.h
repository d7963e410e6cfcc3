// Internal launchers of the learned proposal (proposal.hip) and of the
// pruned candidate costs (ac_strategy.hip), for the plain path's learned run.
#pragma once

#include "common.h"

namespace gjxl_amd {

GJXLResult LaunchProposalFeatures(GJXLAmdDevice* device, const float* opsin, int width,
                                  int height, const float* quant_field, uint16_t* features);
GJXLResult LaunchProposalDecide(GJXLAmdDevice* device, const float* head, int bw, int bh,
                                float threshold, int top_k, float p_min, uint8_t* allow);
GJXLResult LaunchProposalWarmStart(GJXLAmdDevice* device, const float* head,
                                   const float* adjusted, size_t n, float limit, float offset,
                                   float lower, float upper, float* field);
// Partial sums of the norm's work groups (device doubles).
constexpr int kProposalNormPartials = 256;
// (mean of d^3)^(1/3) of n block distances into *out (device); partial
// holds kProposalNormPartials doubles.
GJXLResult LaunchProposalNorm3(GJXLAmdDevice* device, const float* distances, size_t n,
                               double* partial, float* out);
GJXLResult LaunchProposalSample(GJXLAmdDevice* device, const uint8_t* cells, const float* field,
                                const float* adjusted, size_t n, uint8_t* labels,
                                float* targets);

// The forward of gjxl_amd_proposal_forward on `stream` (the network's
// buffers are shared: one forward at a time per network); *head (may be
// null) receives the head's device address.
GJXLResult ProposalForward(GJXLAmdProposal* net, hipStream_t stream, const uint16_t* features,
                           uint32_t n, uint32_t h, uint32_t w, const float** head);

}  // namespace gjxl_amd
