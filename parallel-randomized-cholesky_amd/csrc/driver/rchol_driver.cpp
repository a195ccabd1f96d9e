// rchol_driver -- command line over librchol.so with the reference's argv
// surface (no counterpart beyond the CLIs themselves):
//
//   GPU form (gpu_implementation/driver.cu:2148-2157, driver_physics.cu:2120-2129)
//     rchol_driver <matrix.mtx> <num_blocks> <check_solve 0|1> <tol> [options]
//   CPU form (cpu_implementation/driver_local.cpp:1325-1347)
//     rchol_driver <matrix.mtx> <threads> <out_prefix|""> [1] [options]
//
// The reference selects Laplacian vs SDDM-converted input by building
// driver.cu or driver_physics.cu; here --physics (or the CPU form's trailing
// 1) marks a Matrix Market file whose last vertex is the appended ground
// (write_graph.jl:367-383).  --sddm takes a raw SDDM and appends the ground
// after ordering.  The input is factored in file order unless --order is
// given (the reference orders offline in Julia).  In the CPU form the
// factorization still runs on the GPU; <threads> is accepted and ignored.
//
// options: --physics | --sddm | --order {as-is,nnz-sort,random,amd,learned}
//          --order-seed S  --seed S  --precision {32,64}  --stop {reference,true}
//          --rhs {zero-sum,normal}  --rhs-seed S  --maxit N  --dump PREFIX
//          --write-ordered PATH  --gpu-form | --cpu-form  --json
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/rchol.h"

namespace {

bool parse_double(const char *s, double &v)
{
    char *end = nullptr;
    v = std::strtod(s, &end);
    return end && *end == '\0' && end != s;
}

int die(rc_handle_t h, int st, const char *what)
{
    std::fprintf(stderr, "rchol_driver: %s failed: %s (%s)\n", what, rc_status_string(st),
                 h ? rc_last_error(h) : "");
    if (h) rc_destroy(h);
    return 1;
}

void usage()
{
    std::fprintf(stderr,
                 "usage: rchol_driver <matrix.mtx> <num_blocks> <check_solve 0|1> <tol> [options]\n"
                 "       rchol_driver <matrix.mtx> <threads> <out_prefix|\"\"> [1] [options]\n");
}

}  // namespace

int main(int argc, char **argv)
{
    std::vector<std::string> pos;
    int kind = RC_INPUT_LAPLACIAN, order = -1, precision = 64, stop = RC_STOP_REFERENCE, rhs_kind = -1, maxit = 300;
    unsigned long long seed = 1, order_seed = 1, rhs_seed = 0;
    std::string dump, write_ordered;
    int form = 0;  // 0 auto, 1 gpu, 2 cpu
    bool json = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&](const char *name) -> const char * {
            if (i + 1 >= argc) { std::fprintf(stderr, "missing value for %s\n", name); std::exit(2); }
            return argv[++i];
        };
        if (a == "--physics") kind = RC_INPUT_GROUNDED;
        else if (a == "--sddm") kind = RC_INPUT_SDDM;
        else if (a == "--order") {
            std::string o = next("--order");
            order = o == "as-is" ? RC_ORDER_AS_IS : o == "nnz-sort" ? RC_ORDER_NNZ_SORT : o == "random" ? RC_ORDER_RANDOM
                  : o == "amd" ? RC_ORDER_AMD : o == "learned" ? RC_ORDER_LEARNED : -2;
            if (order == -2) { std::fprintf(stderr, "unknown ordering %s\n", o.c_str()); return 2; }
        } else if (a == "--order-seed") order_seed = std::strtoull(next("--order-seed"), nullptr, 10);
        else if (a == "--seed") seed = std::strtoull(next("--seed"), nullptr, 10);
        else if (a == "--precision") precision = std::atoi(next("--precision"));
        else if (a == "--stop") stop = std::string(next("--stop")) == "true" ? RC_STOP_TRUE_RESIDUAL : RC_STOP_REFERENCE;
        else if (a == "--rhs") rhs_kind = std::string(next("--rhs")) == "normal" ? RC_RHS_NORMAL_PROJECTED : RC_RHS_ZERO_SUM_UNIFORM;
        else if (a == "--rhs-seed") rhs_seed = std::strtoull(next("--rhs-seed"), nullptr, 10);
        else if (a == "--maxit") maxit = std::atoi(next("--maxit"));
        else if (a == "--dump") dump = next("--dump");
        else if (a == "--write-ordered") write_ordered = next("--write-ordered");
        else if (a == "--gpu-form") form = 1;
        else if (a == "--cpu-form") form = 2;
        else if (a == "--json") json = true;
        else if (a == "-h" || a == "--help") { usage(); return 0; }
        else pos.push_back(a);
    }
    if (pos.size() < 3) { usage(); return 2; }
    double tol = 0.0;
    if (form == 0) {
        bool gpu_like = pos.size() >= 4 && (pos[2] == "0" || pos[2] == "1") && parse_double(pos[3].c_str(), tol) &&
                        pos[3] != "1";
        form = gpu_like ? 1 : 2;
    }
    const std::string path = pos[0];
    const int num_blocks = std::atoi(pos[1].c_str());
    bool check_solve = false;
    if (form == 1) {
        if (pos.size() < 4 || !parse_double(pos[3].c_str(), tol)) { usage(); return 2; }
        check_solve = pos[2] == "1";
    } else {
        if (dump.empty()) dump = pos[2];
        if (pos.size() >= 4 && pos[3] == "1") kind = RC_INPUT_GROUNDED;   // driver_local.cpp:1334-1346
    }

    rc_options opt;
    rc_default_options(&opt);
    opt.seed = seed;
    opt.num_blocks = form == 1 ? num_blocks : 0;
    opt.factor_precision = precision;
    opt.stop_rule = stop;
    opt.maxit = maxit;
    rc_handle_t h = nullptr;
    int st = rc_create(&h, &opt);
    if (st != RC_OK) return die(nullptr, st, "rc_create");

    std::printf("problem: %s\n", path.c_str());
    auto t0 = std::chrono::steady_clock::now();
    if ((st = rc_load_mtx(h, path.c_str(), kind)) != RC_OK) return die(h, st, "rc_load_mtx");
    int64_t n = 0, nnz = 0;
    rc_get_size(h, &n, &nnz);
    std::printf("number of nodes: %lld\nnumber of nonzeros: %lld\n", (long long)n, (long long)nnz);
    if (order >= 0 && (st = rc_order(h, order, order_seed)) != RC_OK) return die(h, st, "rc_order");
    if (!write_ordered.empty() && (st = rc_write_prepared_mtx(h, write_ordered.c_str())) != RC_OK)
        return die(h, st, "rc_write_prepared_mtx");
    double read_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    rc_factor_stats fs;
    if ((st = rc_factor(h, &fs)) != RC_OK) return die(h, st, "rc_factor");
    std::printf("Kernel execution time: %f ms\n", fs.factor_ms);
    std::printf("total conversion time: %f ms\n", fs.assemble_ms);
    std::printf("total nnz: %lld\n", (long long)fs.nnz_factor);
    std::printf("nnz ratio: %f, factor nnz: %lld, original lower triangular nnz (exclude diagonal): %lld\n",
                (double)fs.nnz_factor / (double)(fs.nnz_lower > 0 ? fs.nnz_lower : 1), (long long)fs.nnz_factor,
                (long long)fs.nnz_lower);
    std::printf("levels: %d, fills: %lld, max degree: %lld\n", fs.nlevels, (long long)fs.fills,
                (long long)fs.max_degree);
    if (!dump.empty() && (st = rc_dump_factor(h, dump.c_str())) != RC_OK) return die(h, st, "rc_dump_factor");

    rc_solve_stats ss;
    std::memset(&ss, 0, sizeof(ss));
    if (check_solve) {
        std::vector<double> b((size_t)n), x((size_t)n);
        int kind_rhs = rhs_kind >= 0 ? rhs_kind : RC_RHS_ZERO_SUM_UNIFORM;   // solver.hpp:427-431
        if ((st = rc_make_rhs(h, kind_rhs, rhs_seed, b.data())) != RC_OK) return die(h, st, "rc_make_rhs");
        st = rc_solve(h, b.data(), x.data(), tol, &ss);
        if (st != RC_OK && st != RC_ERR_BREAKDOWN) return die(h, st, "rc_solve");
        std::printf("solve preprocess time: %f ms\n", ss.setup_ms);
        std::printf("final iteration: %d, residual: %g\n", ss.iterations, ss.rel_res_precond);
        std::printf("pcg time: %f ms\n", ss.pcg_ms);
        std::printf("normalized diff norm: %.16f\n", ss.rel_res_true);
    }
    if (json) {
        std::printf("{\"problem\": \"%s\", \"n\": %lld, \"nnz\": %lld, \"read_order_s\": %.6f, \"factor_ms\": %.4f, "
                    "\"assemble_ms\": %.4f, \"nnz_factor\": %lld, \"fills\": %lld, \"levels\": %d, "
                    "\"iterations\": %d, \"pcg_ms\": %.4f, \"rel_res_true\": %.6e, \"converged\": %d}\n",
                    path.c_str(), (long long)n, (long long)nnz, read_s, fs.factor_ms, fs.assemble_ms,
                    (long long)fs.nnz_factor, (long long)fs.fills, fs.nlevels, ss.iterations, ss.pcg_ms,
                    ss.rel_res_true, ss.converged);
    }
    rc_destroy(h);
    return 0;
}
