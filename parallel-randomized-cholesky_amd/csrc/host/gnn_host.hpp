// gnn_host.hpp -- learned elimination ordering: graph features and ordering
// from network scores (the network kernels are in csrc/device/gnn.hip).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "sparse.hpp"

namespace rc {

int64_t gnn_weight_count();

// Graph of the first nvert vertices of `a` (off-diagonal entries, |value| as
// weight): CSR, forward / transposed aggregation weights, centred inputs.
struct GnnGraph {
    int n = 0;
    std::vector<int> ptr, ind;
    std::vector<float> a, at, x0;   // x0: n x 2
};
void gnn_graph(const Csc &a, int nvert, GnnGraph &g);

// Elimination order from scores: a seeded draw from the Plackett-Luce
// distribution of the scores (descending score + Gumbel noise), so equal
// scores give a uniformly random order.
std::vector<int> order_from_scores(const std::vector<float> &scores, uint64_t seed);

}  // namespace rc
