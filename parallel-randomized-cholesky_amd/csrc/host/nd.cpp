// nd.cpp -- nested-dissection ordering (write_graph.jl:71-75 calls Metis for
// it; Metis is not available, so this is an own level-structure dissection in
// the manner of George's 1973 nested dissection of grids).
//
// A subgraph of more than kLeaf vertices is split by a vertex separator:
//   1. a pseudo-peripheral vertex u (two BFS sweeps: start anywhere, restart
//      from the last vertex reached);
//   2. the BFS level structure from u; the separator is the level m at which
//      the cumulative vertex count first reaches half the subgraph;
//   3. thinning: a vertex of level m with no neighbour in level m+1 separates
//      nothing and joins the first half.
// The first half (levels < m plus the thinned vertices) is ordered first, the
// second half (levels > m) next, the separator last; both halves recurse.  A
// subgraph the BFS does not cover is split into the reached part and the rest
// (no separator).  Leaves and separators are ordered by ascending degree, ties
// in a seeded random order.  Subgraphs are disjoint, so independent ones are
// processed by several host threads at once over shared per-vertex arrays.
//
// rb_perm is the same recursion with each separator placed between its two
// halves instead of after them: a recursive bisection order, in which the
// vertices of every subgraph (a compact region of the graph) are contiguous.
// The triangular sweeps order the rows of each level by it (locality of their
// dependency gathers).
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "sparse.hpp"

namespace rc {

namespace {

constexpr int kLeaf = 64;

struct NdTask {
    std::vector<int> verts;
    int64_t out;          // first slot of this subgraph in the ordering
};

struct NdState {
    const Csc *a;
    int nvert;
    std::vector<int> deg;          // degree among the first nvert vertices
    std::vector<int> mark;         // task id owning a vertex while it is processed
    std::vector<int> lev;          // BFS level (per task, reset before each BFS)
    std::vector<int> perm;         // perm[new] = old
    std::vector<uint32_t> key;     // seeded random tie-break per vertex
    std::atomic<int> next_id{1};
    bool sep_last = true;          // false: separator between the halves (bisection order)
};

// Owner of vertex w.  Neighbours of a subgraph's vertices may belong to a
// subgraph another thread is marking at the same time: its id is never ours,
// so a relaxed load is enough (levels are only read for our own vertices).
inline int mark_of(const NdState &S, int w)
{
    return __atomic_load_n(&S.mark[(size_t)w], __ATOMIC_RELAXED);
}

// BFS from `src` over the vertices marked `id`; fills `order` (visit order)
// and lev[]; returns the number of levels.
int bfs(NdState &S, int id, int src, std::vector<int> &order)
{
    const Csc &a = *S.a;
    order.clear();
    order.push_back(src);
    S.lev[(size_t)src] = 0;
    int nlev = 1;
    for (size_t h = 0; h < order.size(); ++h) {
        const int v = order[h];
        const int lv = S.lev[(size_t)v];
        for (int64_t p = a.ptr[(size_t)v]; p < a.ptr[(size_t)v + 1]; ++p) {
            const int w = a.ind[(size_t)p];
            if (w >= S.nvert || w == v || mark_of(S, w) != id || S.lev[(size_t)w] >= 0) continue;
            S.lev[(size_t)w] = lv + 1;
            if (lv + 2 > nlev) nlev = lv + 2;
            order.push_back(w);
        }
    }
    return nlev;
}

// Leaves and separators: ascending degree, ties in a seeded random order (an
// id order would chain neighbouring vertices into one long dependency path).
void sort_by_degree(NdState &S, std::vector<int> &v)
{
    std::sort(v.begin(), v.end(), [&](int x, int y) {
        const int dx = S.deg[(size_t)x], dy = S.deg[(size_t)y];
        if (dx != dy) return dx < dy;
        const uint32_t kx = S.key[(size_t)x], ky = S.key[(size_t)y];
        return kx != ky ? kx < ky : x < y;
    });
}

void write_leaf(NdState &S, std::vector<int> &v, int64_t out)
{
    sort_by_degree(S, v);
    std::copy(v.begin(), v.end(), S.perm.begin() + out);
}

// Processes one task; appends its child tasks to `kids`.
void split(NdState &S, NdTask &t, std::vector<NdTask> &kids)
{
    const int size = (int)t.verts.size();
    if (size <= kLeaf) {
        write_leaf(S, t.verts, t.out);
        return;
    }
    const int id = S.next_id.fetch_add(1);
    for (int v : t.verts) {
        __atomic_store_n(&S.mark[(size_t)v], id, __ATOMIC_RELAXED);
        S.lev[(size_t)v] = -1;
    }
    std::vector<int> order;
    order.reserve((size_t)size);
    int src = *std::min_element(t.verts.begin(), t.verts.end());
    int nlev = bfs(S, id, src, order);
    if ((int)order.size() < size) {
        // not connected: the reached part and the rest
        std::vector<int> rest;
        rest.reserve((size_t)size - order.size());
        for (int v : t.verts)
            if (S.lev[(size_t)v] < 0) rest.push_back(v);
        const int64_t n0 = (int64_t)order.size();
        kids.push_back({std::move(order), t.out});
        kids.push_back({std::move(rest), t.out + n0});
        return;
    }
    for (int sweep = 0; sweep < 2; ++sweep) {      // pseudo-peripheral start
        const int far = order.back();
        for (int v : t.verts) S.lev[(size_t)v] = -1;
        const int nl = bfs(S, id, far, order);
        if (nl <= nlev && sweep > 0) break;
        nlev = nl;
    }
    if (nlev <= 2) {
        write_leaf(S, t.verts, t.out);
        return;
    }
    std::vector<int> cnt((size_t)nlev + 1, 0);
    for (int v : t.verts) cnt[(size_t)S.lev[(size_t)v]]++;
    int m = 0;
    int64_t acc = 0;
    while (m < nlev - 1 && acc + cnt[(size_t)m] < size / 2) acc += cnt[(size_t)m++];
    if (m == 0) m = 1;
    if (m >= nlev - 1) m = nlev - 2;
    const Csc &a = *S.a;
    std::vector<int> A, B, Sep;
    A.reserve((size_t)size / 2 + 1);
    B.reserve((size_t)size / 2 + 1);
    for (int v : order) {            // BFS order: neighbouring vertices stay together
        const int lv = S.lev[(size_t)v];
        if (lv < m) {
            A.push_back(v);
        } else if (lv > m) {
            B.push_back(v);
        } else {
            bool touches_next = false;
            for (int64_t p = a.ptr[(size_t)v]; p < a.ptr[(size_t)v + 1] && !touches_next; ++p) {
                const int w = a.ind[(size_t)p];
                touches_next = w < S.nvert && mark_of(S, w) == id && S.lev[(size_t)w] == m + 1;
            }
            if (touches_next) Sep.push_back(v);
            else A.push_back(v);
        }
    }
    if (S.sep_last) sort_by_degree(S, Sep);     // the bisection order keeps the separator in BFS order (locality)
    const int64_t sa = (int64_t)A.size(), sb = (int64_t)B.size(), ss = (int64_t)Sep.size();
    const int64_t sep_at = S.sep_last ? t.out + sa + sb : t.out + sa;
    std::copy(Sep.begin(), Sep.end(), S.perm.begin() + sep_at);
    if (sa) kids.push_back({std::move(A), t.out});
    if (sb) kids.push_back({std::move(B), S.sep_last ? t.out + sa : t.out + sa + ss});
}

}  // namespace

static std::vector<int> dissect(const Csc &a, int nvert, uint64_t seed, int threads, bool sep_last)
{
    NdState S;
    S.a = &a;
    S.nvert = nvert;
    S.sep_last = sep_last;
    S.deg.assign((size_t)nvert, 0);
    for (int v = 0; v < nvert; ++v)
        for (int64_t p = a.ptr[(size_t)v]; p < a.ptr[(size_t)v + 1]; ++p) {
            const int w = a.ind[(size_t)p];
            if (w < nvert && w != v) S.deg[(size_t)v]++;
        }
    S.key.resize((size_t)nvert);
    for (int v = 0; v < nvert; ++v) {             // splitmix64 of (seed, v)
        uint64_t z = seed * 0x9E3779B97F4A7C15ULL + (uint64_t)v + 1;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        S.key[(size_t)v] = (uint32_t)((z ^ (z >> 31)) >> 32);
    }
    S.mark.assign((size_t)nvert, 0);
    S.lev.assign((size_t)nvert, -1);
    S.perm.assign((size_t)nvert, -1);
    if (nvert == 0) return {};
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = std::min(threads, 64);

    // work queue of disjoint subgraphs
    std::deque<NdTask> q;
    {
        NdTask root;
        root.verts.resize((size_t)nvert);
        for (int v = 0; v < nvert; ++v) root.verts[(size_t)v] = v;
        root.out = 0;
        q.push_back(std::move(root));
    }
    std::mutex mu;
    std::condition_variable cv;
    int busy = 0;
    auto worker = [&]() {
        std::vector<NdTask> kids;
        while (true) {
            NdTask t;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return !q.empty() || busy == 0; });
                if (q.empty()) return;            // nothing queued and nobody working: done
                t = std::move(q.front());
                q.pop_front();
                ++busy;
            }
            kids.clear();
            split(S, t, kids);
            {
                std::lock_guard<std::mutex> lk(mu);
                for (auto &k : kids) q.push_back(std::move(k));
                --busy;
            }
            cv.notify_all();
        }
    };
    if (threads == 1) {
        worker();
    } else {
        std::vector<std::thread> th;
        for (int i = 0; i < threads; ++i) th.emplace_back(worker);
        for (auto &x : th) x.join();
    }
    return std::move(S.perm);
}

std::vector<int> nd_perm(const Csc &a, int nvert, uint64_t seed, int threads)
{
    return dissect(a, nvert, seed, threads, true);
}

std::vector<int> rb_perm(const Csc &a, int nvert, uint64_t seed, int threads)
{
    return dissect(a, nvert, seed, threads, false);
}

}  // namespace rc
