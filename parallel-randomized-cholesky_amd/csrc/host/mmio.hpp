// mmio.hpp -- Matrix Market coordinate I/O.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace rc {

struct MtxTriplets {
    int nrows = 0, ncols = 0;
    bool symmetric = false;
    std::vector<int> rows, cols;   // 0-based, symmetric files expanded
    std::vector<double> vals;
};

// threads: host threads of the entry parser (0 = all hardware threads).
int read_mtx(const std::string &path, MtxTriplets &t, std::string &err, int threads = 0);
int write_mtx_csc(const std::string &path, int nrows, int ncols, const int64_t *colptr, const int *rowind,
                  const double *val, std::string &err);

}  // namespace rc
