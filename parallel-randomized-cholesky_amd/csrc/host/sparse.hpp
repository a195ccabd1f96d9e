// sparse.hpp -- host-side matrix preprocessing on the input boundary:
// Laplacian construction, SDDM ground append / trim, symmetric permutation,
// orderings.  All O(nnz) except the per-column sorts.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace rc {

struct Csc {                      // square, column-compressed, rows sorted
    int n = 0;
    std::vector<int64_t> ptr;     // n+1
    std::vector<int> ind;
    std::vector<double> val;
    int64_t nnz() const { return ptr.empty() ? 0 : ptr.back(); }
};

// pre_process.hpp:189-306, 338-357: off-diagonals -> -|v| (explicit zeros
// dropped), diagonals -> +|v|; if the input has no diagonal entry at all the
// diagonal is appended as the column sum of |off-diagonals|.
// kind_sddm: keep the diagonal as given (no append) -- the SDDM is not a
// Laplacian yet.  Returns 0 or a negative rc_status.
int build_laplacian(int n, int64_t m, const int *rows, const int *cols, const double *vals, Csc &out);

// Symmetric in pattern and values (exact).
bool is_symmetric(const Csc &a);

// write_graph.jl:367-383.  Returns 1 if a ground vertex was appended (out has
// n+1 vertices), 0 if the matrix already sums to zero (out = a), or
// RC_ERR_NOT_DOMINANT.
int ground_append(const Csc &a, Csc &out);

// Remove the last row and column (driver_physics.cu:1595-1734).
void trim_last(const Csc &a, Csc &out);

// out = a[perm, perm], perm[new] = old; rows of every column sorted.
void permute_sym(const Csc &a, const std::vector<int> &perm, Csc &out);

// Orderings (write_graph.jl:46-78, 333-365).  `count_skip` excludes entries
// with row == count_skip from the column counts (the ground vertex).
std::vector<int> random_perm(int n, uint64_t seed);
std::vector<int> nnz_sort_perm(const Csc &a, int nvert, uint64_t seed, int count_skip);
// Approximate minimum degree (Amestoy, Davis, Duff 1996) on the pattern of
// the first nvert vertices.
std::vector<int> amd_perm(const Csc &a, int nvert);
// Nested dissection by BFS level-structure separators (nd.cpp) on the first
// nvert vertices (ties broken by `seed`), `threads` host threads (0 = all).
std::vector<int> nd_perm(const Csc &a, int nvert, uint64_t seed, int threads);
// Recursive bisection order (nd_perm with each separator between its halves).
std::vector<int> rb_perm(const Csc &a, int nvert, uint64_t seed, int threads);

// Strict lower part as int32 CSC with |values| (cpu pre_process.hpp:763-850).
int lower_part(const Csc &a, std::vector<int> &ptr, std::vector<int> &row, std::vector<double> &w);

// Connected components of the pattern (for per-component RHS projection).
int components(const Csc &a, std::vector<int> &comp);

}  // namespace rc
