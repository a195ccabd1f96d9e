// prep_device.cpp -- the prepared system built on the device from the loaded
// Laplacian and the ordering (kernels in csrc/device/prep.hip): permuted
// matrix, SDDM ground append or ground trim, strict-lower slice of the
// factored matrix, dependency counts, and the learned ordering's
// Plackett-Luce draw.  Same results as the host path of sparse.cpp /
// gnn_host.cpp, which host-only handles and rc_get_prepared_csc keep using.
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "handle.hpp"

namespace {

int bits_for(int64_t v)
{
    int b = 0;
    while (v > 0) { ++b; v >>= 1; }
    return b > 0 ? b : 1;
}

}  // namespace

int prepare_device(rc_handle_t h)
{
    hipStream_t s = h->stream;
    const Csc &A = h->base;
    const int n = A.n;
    if (A.nnz() > 2147483646LL) return fail(h, RC_ERR_ARG, "nnz exceeds int32 indexing");
    const int nnz = (int)A.nnz();

    // 1. the loaded matrix on the device, once per loaded matrix
    if (!h->base_dev) {
        std::vector<int> p32((size_t)n + 1);
        for (int i = 0; i <= n; ++i) p32[(size_t)i] = (int)A.ptr[(size_t)i];
        RC_HIP(h->b_ptr.ensure(p32.size() * sizeof(int)));
        RC_HIP(h->b_ind.ensure((size_t)std::max(nnz, 1) * sizeof(int)));
        RC_HIP(h->b_val.ensure((size_t)std::max(nnz, 1) * sizeof(double)));
        RC_HIP(hipMemcpyAsync(h->b_ptr.p, p32.data(), p32.size() * sizeof(int), hipMemcpyHostToDevice, s));
        if (nnz) {
            RC_HIP(hipMemcpyAsync(h->b_ind.p, A.ind.data(), (size_t)nnz * sizeof(int), hipMemcpyHostToDevice, s));
            RC_HIP(hipMemcpyAsync(h->b_val.p, A.val.data(), (size_t)nnz * sizeof(double), hipMemcpyHostToDevice, s));
        }
        RC_HIP(hipStreamSynchronize(s));
        h->base_dev = true;
    }

    // 2. the permutation over all n vertices (a grounded input keeps its
    //    ground last) and its inverse
    std::vector<int> full(h->perm);
    if (h->kind == RC_INPUT_GROUNDED) full.push_back(n - 1);
    RC_HIP(h->pfull.ensure((size_t)n * sizeof(int)));
    RC_HIP(h->pinv.ensure((size_t)n * sizeof(int)));
    RC_HIP(hipMemcpyAsync(h->pfull.p, full.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, s));
    launch_inverse_perm(n, h->pfull.as<int>(), h->pinv.as<int>(), s);

    // 3. permuted matrix P = A[perm, perm]: entries ordered by new row, then
    //    (stably) by new column, i.e. columns sorted by (row, input order)
    const size_t big = (size_t)nnz + 1;
    RC_HIP(h->keys0.ensure(big * sizeof(uint32_t)));
    RC_HIP(h->keys1.ensure(big * sizeof(uint32_t)));
    RC_HIP(h->vals0.ensure(big * sizeof(int)));
    RC_HIP(h->vals1.ensure(big * sizeof(int)));
    RC_HIP(h->rowk.ensure(big * sizeof(int)));
    RC_HIP(h->tmp_b.ensure(std::max(radix_tmp_ints((int)big), (size_t)n + 2) * sizeof(int)));
    uint32_t *k0 = h->keys0.as<uint32_t>(), *k1 = h->keys1.as<uint32_t>();
    int *v0 = h->vals0.as<int>(), *v1 = h->vals1.as<int>();
    const int kb = bits_for(n);
    launch_map_rows(nnz, h->b_ind.as<int>(), h->pinv.as<int>(), k0, s);
    launch_iota(v0, nnz, s);
    radix_sort_pairs(k0, v0, k1, v1, nnz, kb, h->tmp_b.as<int>(), s);
    launch_row_expand(n, h->b_ptr.as<int>(), h->pinv.as<int>(), h->rowk.as<int>(), s);   // new column of every entry
    launch_gather_key(h->rowk.as<int>(), v0, k0, nnz, s);
    radix_sort_pairs(k0, v0, k1, v1, nnz, kb, h->tmp_b.as<int>(), s);
    const bool grounded = h->kind == RC_INPUT_GROUNDED;
    DevVec &Pp = grounded ? h->f_ptr : h->a_ptr;
    DevVec &Pi = grounded ? h->f_ind : h->a_ind;
    DevVec &Pv = grounded ? h->f_val : h->a_val;
    RC_HIP(Pp.ensure(((size_t)n + 2) * sizeof(int)));
    RC_HIP(Pi.ensure((size_t)std::max(nnz, 1) * sizeof(int)));
    RC_HIP(Pv.ensure((size_t)std::max(nnz, 1) * sizeof(double)));
    launch_segment_bounds(nnz, k0, 0, n, Pp.as<int>(), s);
    launch_perm_finish(nnz, v0, h->b_ind.as<int>(), h->pinv.as<int>(), h->b_val.as<double>(), Pi.as<int>(),
                       Pv.as<double>(), s);

    // 4. the factored matrix and the PCG system
    RC_HIP(h->tmp_a.ensure(3 * ((size_t)n + 3) * sizeof(int)));
    RC_HIP(h->tmp_scan.ensure(scan_tmp_elems(n + 3) * sizeof(int) + 64));
    int *cnt = h->tmp_a.as<int>();
    const int *fp = nullptr, *fi = nullptr;
    const double *fv = nullptr;
    int64_t nnz_sys = nnz;
    if (h->kind == RC_INPUT_LAPLACIAN) {
        h->nf = h->ns = n;
        h->trim = false;
        fp = h->a_ptr.as<int>(); fi = h->a_ind.as<int>(); fv = h->a_val.as<double>();
    } else if (grounded) {
        // system = P without its last row and column (driver_physics.cu:1595-1734)
        if (n < 2) return fail(h, RC_ERR_ARG, "grounded input needs n >= 2");
        RC_HIP(h->a_ptr.ensure((size_t)n * sizeof(int)));
        launch_trim_count(n, h->f_ptr.as<int>(), h->f_ind.as<int>(), cnt, s);
        scan_exclusive(cnt, h->a_ptr.as<int>(), n - 1, h->tmp_scan.as<int>(), s);
        int t = 0;
        RC_HIP(hipMemcpyAsync(&t, h->a_ptr.as<int>() + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        RC_HIP(hipStreamSynchronize(s));
        nnz_sys = t;
        RC_HIP(h->a_ind.ensure((size_t)std::max(t, 1) * sizeof(int)));
        RC_HIP(h->a_val.ensure((size_t)std::max(t, 1) * sizeof(double)));
        launch_trim_fill(n, h->f_ptr.as<int>(), h->f_ind.as<int>(), h->f_val.as<double>(), h->a_ptr.as<int>(),
                         h->a_ind.as<int>(), h->a_val.as<double>(), s);
        h->nf = n;
        h->ns = n - 1;
        h->trim = true;
        fp = h->f_ptr.as<int>(); fi = h->f_ind.as<int>(); fv = h->f_val.as<double>();
    } else {
        // SDDM: system = P; ground appended after ordering (write_graph.jl:367-383)
        RC_HIP(h->p_cs.ensure((size_t)n * sizeof(double)));
        RC_HIP(h->p_app.ensure((size_t)n * sizeof(double)));
        RC_HIP(h->p_flag.ensure(((size_t)n + 1) * sizeof(int)));
        RC_HIP(h->p_pos.ensure(((size_t)n + 1) * sizeof(int)));
        launch_colsum(n, h->a_ptr.as<int>(), h->a_val.as<double>(), h->p_cs.as<double>(), h->p_app.as<double>(),
                      h->p_flag.as<int>(), s);
        std::vector<double> cs((size_t)n);
        RC_HIP(hipMemcpyAsync(cs.data(), h->p_cs.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s));
        RC_HIP(hipStreamSynchronize(s));
        // the decisions and the corner in the host path's sequential order
        double total = 0.0, corner = 0.0;
        int extra = 0;
        for (int j = 0; j < n; ++j) {
            total += cs[(size_t)j];
            const double v = -cs[(size_t)j];
            const double a = v > 0.0 ? 0.0 : v;
            corner += a;
            if (a != 0.0) ++extra;
        }
        corner = -corner;
        if (total < 0.0 && std::fabs(total) > 1e-4) return fail(h, RC_ERR_NOT_DOMINANT, "SDDM input is not diagonally dominant");
        h->ns = n;
        if (!(total > 1e-9)) {                   // already a Laplacian: factor the system itself
            h->nf = n;
            h->trim = false;
            fp = h->a_ptr.as<int>(); fi = h->a_ind.as<int>(); fv = h->a_val.as<double>();
        } else {
            if ((int64_t)nnz + 2 * (int64_t)extra + 1 > 2147483646LL)
                return fail(h, RC_ERR_ARG, "nnz exceeds int32 indexing");
            const int fnnz = nnz + 2 * extra + 1;
            scan_exclusive(h->p_flag.as<int>(), h->p_pos.as<int>(), n, h->tmp_scan.as<int>(), s);
            launch_ground_count(n, h->a_ptr.as<int>(), h->p_flag.as<int>(), extra, cnt, s);
            RC_HIP(h->f_ptr.ensure(((size_t)n + 2) * sizeof(int)));
            RC_HIP(h->f_ind.ensure((size_t)fnnz * sizeof(int)));
            RC_HIP(h->f_val.ensure((size_t)fnnz * sizeof(double)));
            scan_exclusive(cnt, h->f_ptr.as<int>(), n + 1, h->tmp_scan.as<int>(), s);
            launch_ground_fill(n, h->a_ptr.as<int>(), h->a_ind.as<int>(), h->a_val.as<double>(),
                               h->p_app.as<double>(), h->p_flag.as<int>(), h->p_pos.as<int>(), corner,
                               h->f_ptr.as<int>(), h->f_ind.as<int>(), h->f_val.as<double>(), s);
            h->nf = n + 1;
            h->trim = true;
            fp = h->f_ptr.as<int>(); fi = h->f_ind.as<int>(); fv = h->f_val.as<double>();
        }
    }
    h->nnz_sys = nnz_sys;
    h->fac_in_f = fp == h->f_ptr.as<int>();

    // 5. strict-lower slice of the factored matrix (|values|) and the
    //    dependency counts dep[k] = lower neighbours of k (pre_process.hpp:474-496)
    const int nf = h->nf;
    RC_HIP(h->dep_init.ensure((size_t)nf * sizeof(int)));
    RC_HIP(h->lo_ptr.ensure(((size_t)nf + 1) * sizeof(int)));
    launch_lower_count(nf, fp, fi, cnt, h->dep_init.as<int>(), s);
    scan_exclusive(cnt, h->lo_ptr.as<int>(), nf, h->tmp_scan.as<int>(), s);
    int nlow = 0;
    RC_HIP(hipMemcpyAsync(&nlow, h->lo_ptr.as<int>() + nf, sizeof(int), hipMemcpyDeviceToHost, s));
    RC_HIP(hipStreamSynchronize(s));
    h->nnz_low = nlow;
    RC_HIP(h->lo_row.ensure((size_t)std::max(nlow, 1) * sizeof(int)));
    RC_HIP(h->lo_w.ensure((size_t)std::max(nlow, 1) * sizeof(double)));
    launch_lower_fill(nf, fp, fi, fv, h->lo_ptr.as<int>(), h->lo_row.as<int>(), h->lo_w.as<double>(), s);

    // 6. the ordering in the user numbering (solve-side maps)
    const int ns = h->ns;
    RC_HIP(h->perm_d.ensure((size_t)std::max(ns, 1) * sizeof(int)));
    RC_HIP(hipMemcpyAsync(h->perm_d.p, h->perm.data(), (size_t)ns * sizeof(int), hipMemcpyHostToDevice, s));
    if (!h->loc_rank.empty()) {
        RC_HIP(h->loc_d.ensure((size_t)h->n_user * sizeof(int)));
        RC_HIP(hipMemcpyAsync(h->loc_d.p, h->loc_rank.data(), (size_t)h->n_user * sizeof(int), hipMemcpyHostToDevice, s));
    }
    RC_HIP(hipGetLastError());
    RC_HIP(hipStreamSynchronize(s));
    return RC_OK;
}

// One draw of the Plackett-Luce distribution of the scores already on the
// device (gnn output): keys score - log(-log u), u from Philox, sorted in
// descending order, ties in ascending vertex order (the host path's
// stable_sort).  perm[new] = old.
int learned_order_device(rc_handle_t h, const float *scores_dev, int n, uint64_t seed, std::vector<int> &perm)
{
    hipStream_t s = h->stream;
    const size_t big = (size_t)n + 1;
    RC_HIP(h->keys0.ensure(big * sizeof(uint32_t)));
    RC_HIP(h->keys1.ensure(big * sizeof(uint32_t)));
    RC_HIP(h->vals0.ensure(big * sizeof(int)));
    RC_HIP(h->vals1.ensure(big * sizeof(int)));
    RC_HIP(h->rowk.ensure(big * sizeof(int)));
    RC_HIP(h->tmp_b.ensure(radix_tmp_ints((int)big) * sizeof(int)));
    uint32_t *lo = h->keys0.as<uint32_t>(), *tmp = h->keys1.as<uint32_t>();
    uint32_t *hi = reinterpret_cast<uint32_t *>(h->rowk.as<int>());
    int *v0 = h->vals0.as<int>(), *v1 = h->vals1.as<int>();
    launch_gumbel_keys(n, scores_dev, seed, hi, lo, s);
    launch_iota(v0, n, s);
    radix_sort_pairs(lo, v0, tmp, v1, n, 32, h->tmp_b.as<int>(), s);       // low word first (LSD over 64 bits)
    launch_gather_key(reinterpret_cast<const int *>(hi), v0, lo, n, s);
    radix_sort_pairs(lo, v0, tmp, v1, n, 32, h->tmp_b.as<int>(), s);
    perm.resize((size_t)n);
    RC_HIP(hipMemcpyAsync(perm.data(), v0, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, s));
    RC_HIP(hipStreamSynchronize(s));
    return RC_OK;
}
