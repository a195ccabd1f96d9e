// gnn_host.cpp -- learned ordering entry (network kernels land separately).
#include "gnn_host.hpp"

#include "../../../include/rchol.h"

namespace rc {

int64_t gnn_weight_count() { return 0; }

int gnn_order(const Csc &, int, int, const std::vector<float> &, uint64_t, std::vector<int> &, std::string &err)
{
    err = "learned ordering network not available";
    return RC_ERR_STATE;
}

}  // namespace rc
