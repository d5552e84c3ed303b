// Inputs of the learned plane proposal (the vacd_mlp model of include/vacd.h).
//
// A part is described to the model by a fixed-size set of its visibility
// edges and each candidate plane by its unit normal and offset, all in the
// part's own frame (bbox centre at the origin, longest bbox side 2), so the
// model sees parts of every size alike.  The reference has no counterpart:
// it scores every candidate plane exactly (lib/src/process.cpp:82-134).
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

#include "mesh.hpp"

namespace vhost {

constexpr int kProposalEdges = 512;  // edges per part seen by the set encoder
constexpr int kEdgeFeatures = 7;     // u (3), v (3), |v - u|

// x' = (x - c) / s
struct PartFrame {
  Vec3 c{0.0, 0.0, 0.0};
  double s = 1.0;
};

// Frame of a vertex set: c = bbox centre, s = half the longest bbox side.
PartFrame part_frame(const std::vector<Vec3> &v);

// kProposalEdges x 7 floats.  With E <= 512 edges every edge is taken and the
// list repeats cyclically (the encoder max-pools, so repeats change nothing);
// otherwise one edge from each of 512 equal strata of the list, drawn from
// seed.  Endpoints are ordered lexicographically in the frame.
void edge_features(const std::vector<Vec3> &v, const std::vector<std::pair<unsigned, unsigned>> &edges,
                   uint64_t seed, const PartFrame &fr, float *out);

// count planes -> (n, d') per plane in the frame, n of unit length:
// n.x + d = 0 with x = s x' + c gives n.x' + (n.c + d) / s = 0; the sign of
// (n, d') is chosen so that d' > 0 (or, for d' = 0, the first non-zero
// component of n is positive).
void plane_inputs(const std::vector<Plane> &planes, size_t count, const PartFrame &fr, float *out);

uint64_t mix64(uint64_t x);

}  // namespace vhost
