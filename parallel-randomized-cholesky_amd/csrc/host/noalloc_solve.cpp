// noalloc_solve.cpp -- rc_solve_noalloc: the PCG of rc_solve on right-hand
// side and solution vectors the caller keeps in device memory, with every
// vector, scalar block, partial sum and claim counter of the solve carved out
// of the fixed per-device workspace (noalloc.hip).  Nothing here allocates:
// no hipMalloc / hipFree / hipHostMalloc, no graph instantiation, no events
// (the PCG time is taken on the host clock).
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <string>

#include "handle.hpp"

namespace {

size_t up256(size_t b) { return (b + 255) / 256 * 256; }

// Workspace layout for n rows and value size s (bytes of each part, in order):
// scalars, partials (2 x reduce blocks), claim counters, then x, r, ap, p0,
// p1, b (n + 1 doubles each) and y, z (n + 1 values of the factor precision).
size_t workspace_bytes(size_t n, size_t s, int reduce_blocks)
{
    return up256(sizeof(PcgScalars)) + up256((size_t)reduce_blocks * 2 * sizeof(double)) +
           up256(kClaimWords * sizeof(unsigned int)) + 6 * up256((n + 1) * 8) + 2 * up256((n + 1) * s);
}

template <typename T>
void launch_init(const SolveArgs &a, hipStream_t s) { launch_pcg_init<T>(a, s); }

int solve_noalloc(rc_handle_t h, const double *b_dev, double *x_dev, double tol, rc_solve_stats *st)
{
    if (h->host_only) return fail(h, RC_ERR_NO_DEVICE, "host-only handle");
    if (!b_dev || !x_dev || !(tol >= 0.0)) return fail(h, RC_ERR_ARG, "bad solve arguments");
    if (!h->factored) return fail(h, RC_ERR_STATE, "rc_factor must precede rc_solve_noalloc");
    if (h->nsys > 0) return fail(h, RC_ERR_ARG, "rc_solve_noalloc solves one system (batches use rc_solve_batch)");
    if (h->na_err != hipSuccess || !h->na_ws)
        return fail(h, RC_ERR_HIP, std::string("fixed workspace unavailable: ") + hipGetErrorString(h->na_err));
    const size_t n = (size_t)h->ns;
    const size_t vs = h->prec == 32 ? 4 : 8;
    const size_t need = workspace_bytes(n, vs, h->reduce_blocks);
    if (need > kNoAllocScratchBytes)
        return fail(h, RC_ERR_ARG,
                    "system of " + std::to_string(n) + " rows needs " + std::to_string(need) +
                        " bytes of workspace; rc_solve_noalloc holds " + std::to_string(kNoAllocScratchBytes) +
                        " (use rc_solve)");
    RC_HIP(hipSetDevice(h->opt.device));
    hipStream_t s = h->stream;

    // carve the workspace
    char *w = static_cast<char *>(h->na_ws);
    auto take = [&](size_t bytes) {
        char *p = w;
        w += up256(bytes);
        return p;
    };
    SolveArgs a = solve_args(h);
    a.sc = reinterpret_cast<PcgScalars *>(take(sizeof(PcgScalars)));
    a.partials = reinterpret_cast<double *>(take((size_t)h->reduce_blocks * 2 * sizeof(double)));
    unsigned int *ctr = reinterpret_cast<unsigned int *>(take(kClaimWords * sizeof(unsigned int)));
    a.ctr_fwd = ctr;
    a.ctr_bwd = ctr + kClaimQueues * kClaimStride;
    a.x = reinterpret_cast<double *>(take((n + 1) * 8));
    a.r = reinterpret_cast<double *>(take((n + 1) * 8));
    a.ap = reinterpret_cast<double *>(take((n + 1) * 8));
    a.p0 = reinterpret_cast<double *>(take((n + 1) * 8));
    a.p1 = reinterpret_cast<double *>(take((n + 1) * 8));
    double *b = reinterpret_cast<double *>(take((n + 1) * 8));
    a.b = b;
    a.y = take((n + 1) * vs);
    a.z = take((n + 1) * vs);

    // scalars: every field (arrival counters included) starts from this copy;
    // the prepare kernel zeroes x, r, p, Ap and the dummy slots and resets the
    // claim counters; partials are written before they are read
    PcgScalars sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.tol = tol;
    sc.maxit = h->opt.maxit;
    sc.stop_rule = h->opt.stop_rule;
    const auto t0 = std::chrono::steady_clock::now();
    RC_HIP(hipMemcpyAsync(a.sc, &sc, sizeof(sc), hipMemcpyHostToDevice, s));
    launch_gather(b_dev, h->idx_user.as<int>(), b, (int)n, s);     // b in position order
    launch_zero_slot(b, (int)n, s);
    if (h->prec == 32) launch_init<float>(a, s);
    else launch_init<double>(a, s);
    RC_HIP(hipGetLastError());
    const auto t1 = std::chrono::steady_clock::now();
    int done = 0;
    for (int it = 0; it < h->opt.maxit + 2 && !done;) {
        for (int t = 0; t < 8; ++t, ++it) {
            if (h->prec == 32) launch_pcg_iteration<float>(a, it & 1, s);
            else launch_pcg_iteration<double>(a, it & 1, s);
        }
        RC_HIP(hipGetLastError());
        RC_HIP(hipMemcpyAsync(&done, reinterpret_cast<char *>(a.sc) + offsetof(PcgScalars, done), sizeof(int),
                              hipMemcpyDeviceToHost, s));
        RC_HIP(hipStreamSynchronize(s));
    }
    const auto t2 = std::chrono::steady_clock::now();
    launch_residual(a, s);
    launch_scatter(a.x, h->idx_user.as<int>(), x_dev, (int)n, s);   // x back in the loaded order
    RC_HIP(hipMemcpyAsync(&sc, a.sc, sizeof(sc), hipMemcpyDeviceToHost, s));
    RC_HIP(hipStreamSynchronize(s));
    if (st) {
        st->iterations = sc.iter;
        st->converged = (sc.status == 0 && sc.done && sc.iter < h->opt.maxit) ? 1 : 0;
        if (sc.status == 0 && sc.done && sc.iter == h->opt.maxit) {
            const bool met = h->opt.stop_rule == RC_STOP_REFERENCE ? (std::sqrt(sc.rz) / std::sqrt(sc.rz0) <= tol)
                                                                    : (std::sqrt(sc.rr) / std::sqrt(sc.bb) <= tol);
            st->converged = met ? 1 : 0;
        }
        st->rel_res_true = sc.bnorm2 > 0.0 ? std::sqrt(sc.res2) / std::sqrt(sc.bnorm2) : 0.0;
        st->rel_res_precond = sc.rz0 > 0.0 ? std::sqrt(std::max(sc.rz, 0.0)) / std::sqrt(sc.rz0) : 0.0;
        st->setup_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        st->pcg_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
        st->err_vs_oracle = -1.0;
    }
    if (sc.status & kPcgTimeout) return fail(h, RC_ERR_TIMEOUT, "triangular solve wait timed out");
    if (sc.status & kPcgBreakdown) return fail(h, RC_ERR_BREAKDOWN, "PCG breakdown (p'Ap <= 0)");
    return RC_OK;
}

}  // namespace

extern "C" {

int rc_solve_noalloc(rc_handle_t h, const double *b_dev, double *x_dev, double tol, rc_solve_stats *st)
{
    RC_CHECK_HANDLE();
    return guarded(h, [&]() { return solve_noalloc(h, b_dev, x_dev, tol, st); });
}

}  // extern "C"
