// gnn_host.cpp -- graph features and score ordering for the ordering network.
#include "gnn_host.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

#include "../../../include/rchol.h"
#include "../device/kernels.h"
#include "philox.h"

namespace rc {

int64_t gnn_weight_count() { return kGnnParams; }

void gnn_graph(const Csc &a, int nvert, GnnGraph &g)
{
    const int n = nvert;
    g.n = n;
    g.ptr.assign((size_t)n + 1, 0);
    std::vector<double> wdeg((size_t)n, 0.0);
    std::vector<int> deg((size_t)n, 0);
    for (int j = 0; j < n; ++j)
        for (int64_t p = a.ptr[(size_t)j]; p < a.ptr[(size_t)j + 1]; ++p) {
            int i = a.ind[(size_t)p];
            if (i == j || i >= n) continue;
            deg[(size_t)j]++;
            wdeg[(size_t)j] += std::fabs(a.val[(size_t)p]);
        }
    for (int j = 0; j < n; ++j) g.ptr[(size_t)j + 1] = g.ptr[(size_t)j] + deg[(size_t)j];
    g.ind.resize((size_t)g.ptr[(size_t)n]);
    g.a.resize(g.ind.size());
    g.at.resize(g.ind.size());
    for (int j = 0; j < n; ++j) {
        int q = g.ptr[(size_t)j];
        for (int64_t p = a.ptr[(size_t)j]; p < a.ptr[(size_t)j + 1]; ++p) {
            int i = a.ind[(size_t)p];
            if (i == j || i >= n) continue;
            double w = std::fabs(a.val[(size_t)p]);
            g.ind[(size_t)q] = i;
            g.a[(size_t)q] = (float)(w / wdeg[(size_t)j]);
            g.at[(size_t)q] = (float)(w / wdeg[(size_t)i]);
            ++q;
        }
    }
    // inputs: log degree and log weighted degree, centred over the graph
    double m0 = 0.0, m1 = 0.0;
    int cnt = 0;
    for (int j = 0; j < n; ++j)
        if (deg[(size_t)j] > 0) {
            m0 += std::log((double)deg[(size_t)j]);
            m1 += std::log(wdeg[(size_t)j]);
            ++cnt;
        }
    if (cnt) { m0 /= cnt; m1 /= cnt; }
    g.x0.assign((size_t)n * 2, 0.f);
    for (int j = 0; j < n; ++j)
        if (deg[(size_t)j] > 0) {
            g.x0[(size_t)j * 2] = (float)(std::log((double)deg[(size_t)j]) - m0);
            g.x0[(size_t)j * 2 + 1] = (float)(std::log(wdeg[(size_t)j]) - m1);
        }
}

std::vector<int> order_from_scores(const std::vector<float> &scores, uint64_t seed)
{
    // One draw from the Plackett-Luce distribution of the scores (the
    // distribution the network is trained under): sort score + Gumbel noise
    // in descending order.  Gumbel(0,1) = -log(-log u), u uniform in (0,1).
    const int n = (int)scores.size();
    std::vector<double> key((size_t)n);
    for (int i = 0; i < n; ++i) {
        const U4 o = philox_keyed(seed, (uint32_t)i, 0u, kTagGumbel);
        const double u = ((double)((((uint64_t)o.y << 32) | o.x) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        key[(size_t)i] = (double)scores[(size_t)i] - std::log(-std::log(u));
    }
    std::vector<int> p((size_t)n);
    std::iota(p.begin(), p.end(), 0);
    std::stable_sort(p.begin(), p.end(), [&](int x, int y) { return key[(size_t)x] > key[(size_t)y]; });
    return p;
}

}  // namespace rc
