// gnn_host.hpp -- learned elimination ordering (graph network) host entry.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "sparse.hpp"

namespace rc {

// Number of float32 parameters of the ordering network.
int64_t gnn_weight_count();

// Ordering of the first nvert vertices of `a` (ground `skip` excluded from
// the features) by the network's per-vertex scores.
int gnn_order(const Csc &a, int nvert, int skip, const std::vector<float> &w, uint64_t seed, std::vector<int> &perm,
              std::string &err);

}  // namespace rc
