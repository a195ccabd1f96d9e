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
// factorization still runs on the GPU; <threads> sets the host threads of the
// input path (Matrix Market parse, ordering).  --dry-run parses the command
// line, loads and orders the input on a host-only handle and prints the
// parsed settings as JSON (no device work).
//
// options: --physics | --sddm | --order {as-is,nnz-sort,random,amd,learned,nd}
//          --order-seed S  --seed S  --precision {32,64}  --stop {reference,true}
//          --rhs {zero-sum,normal}  --rhs-seed S  --maxit N  --dump PREFIX
//          --write-ordered PATH  --gpu-form | --cpu-form  --json  --dry-run
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

bool parse_int(const char *s, long long &v)
{
    char *end = nullptr;
    v = std::strtoll(s, &end, 10);
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

// Everything the command line selects.
struct CliArgs {
    std::string path;
    int form = 0;                 // 1 GPU form, 2 CPU form
    bool ambiguous = false;       // the positionals fit both forms (GPU form taken)
    int num_blocks = 0;           // GPU form argv[2]
    int threads = 0;              // CPU form argv[2]: host threads of the input path
    bool check_solve = false;
    double tol = 0.0;
    int kind = RC_INPUT_LAPLACIAN;
    int order = -1, precision = 64, stop = RC_STOP_REFERENCE, rhs_kind = -1, maxit = 300;
    unsigned long long seed = 1, order_seed = 1, rhs_seed = 0;
    std::string dump, write_ordered;
    bool json = false, dry_run = false;
};

// Parses argv; returns 0, or 2 with a message in err (usage error).
//
// The two reference forms are told apart by their positionals:
//   GPU  <mtx> <num_blocks:int> <check_solve:0|1> <tol:double>   (exactly four)
//   CPU  <mtx> <threads:int> <out_prefix> [1]                    (three, or four ending in 1)
// Four positionals whose third is 0/1 and whose fourth is "1" fit both (a GPU
// run with tol 1, or a CPU run with prefix "0"/"1" and the SDDM flag): the
// GPU form is taken and `ambiguous` is set; --cpu-form / --gpu-form decide.
int parse_cli(int argc, char **argv, CliArgs &a, std::string &err)
{
    std::vector<std::string> pos;
    int forced = 0;
    for (int i = 1; i < argc; ++i) {
        std::string s = argv[i];
        auto next = [&](const char *name, std::string &out) -> bool {
            if (i + 1 >= argc) { err = std::string("missing value for ") + name; return false; }
            out = argv[++i];
            return true;
        };
        std::string v;
        if (s == "--physics") a.kind = RC_INPUT_GROUNDED;
        else if (s == "--sddm") a.kind = RC_INPUT_SDDM;
        else if (s == "--order") {
            if (!next("--order", v)) return 2;
            a.order = v == "as-is" ? RC_ORDER_AS_IS : v == "nnz-sort" ? RC_ORDER_NNZ_SORT : v == "random" ? RC_ORDER_RANDOM
                    : v == "amd" ? RC_ORDER_AMD : v == "learned" ? RC_ORDER_LEARNED : v == "nd" ? RC_ORDER_ND : -2;
            if (a.order == -2) { err = "unknown ordering " + v; return 2; }
        } else if (s == "--order-seed") { if (!next(s.c_str(), v)) return 2; a.order_seed = std::strtoull(v.c_str(), nullptr, 10); }
        else if (s == "--seed") { if (!next(s.c_str(), v)) return 2; a.seed = std::strtoull(v.c_str(), nullptr, 10); }
        else if (s == "--precision") {
            if (!next(s.c_str(), v)) return 2;
            a.precision = std::atoi(v.c_str());
            if (a.precision != 32 && a.precision != 64) { err = "--precision takes 32 or 64"; return 2; }
        } else if (s == "--stop") { if (!next(s.c_str(), v)) return 2; a.stop = v == "true" ? RC_STOP_TRUE_RESIDUAL : RC_STOP_REFERENCE; }
        else if (s == "--rhs") { if (!next(s.c_str(), v)) return 2; a.rhs_kind = v == "normal" ? RC_RHS_NORMAL_PROJECTED : RC_RHS_ZERO_SUM_UNIFORM; }
        else if (s == "--rhs-seed") { if (!next(s.c_str(), v)) return 2; a.rhs_seed = std::strtoull(v.c_str(), nullptr, 10); }
        else if (s == "--maxit") { if (!next(s.c_str(), v)) return 2; a.maxit = std::atoi(v.c_str()); }
        else if (s == "--dump") { if (!next(s.c_str(), v)) return 2; a.dump = v; }
        else if (s == "--write-ordered") { if (!next(s.c_str(), v)) return 2; a.write_ordered = v; }
        else if (s == "--gpu-form") forced = 1;
        else if (s == "--cpu-form") forced = 2;
        else if (s == "--json") a.json = true;
        else if (s == "--dry-run") a.dry_run = true;
        else if (s.size() > 2 && s[0] == '-' && s[1] == '-') { err = "unknown option " + s; return 2; }
        else pos.push_back(s);
    }
    if (pos.size() < 3 || pos.size() > 4) { err = "expected 3 or 4 positional arguments"; return 2; }
    a.path = pos[0];
    long long count = 0;
    if (!parse_int(pos[1].c_str(), count) || count < 0 || count > 1 << 20) {
        err = "argument 2 (num_blocks / threads) must be a non-negative integer";
        return 2;
    }
    double tol = 0.0;
    const bool gpu_fits = pos.size() == 4 && (pos[2] == "0" || pos[2] == "1") && parse_double(pos[3].c_str(), tol);
    const bool cpu_fits = pos.size() == 3 || pos[3] == "1";
    a.form = forced ? forced : gpu_fits ? 1 : 2;
    a.ambiguous = !forced && gpu_fits && cpu_fits;
    if (a.form == 1) {
        if (!gpu_fits) { err = "GPU form: <matrix.mtx> <num_blocks> <check_solve 0|1> <tol>"; return 2; }
        a.num_blocks = (int)count;
        a.check_solve = pos[2] == "1";
        a.tol = tol;
        if (!(tol >= 0.0)) { err = "tolerance must be >= 0"; return 2; }
    } else {
        if (!cpu_fits) { err = "CPU form: <matrix.mtx> <threads> <out_prefix|\"\"> [1]"; return 2; }
        a.threads = (int)count;
        if (a.dump.empty()) a.dump = pos[2];
        if (pos.size() == 4) a.kind = RC_INPUT_GROUNDED;          // driver_local.cpp:1334-1346
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    CliArgs a;
    std::string err;
    if (argc >= 2 && (std::string(argv[1]) == "-h" || std::string(argv[1]) == "--help")) { usage(); return 0; }
    if (parse_cli(argc, argv, a, err) != 0) {
        std::fprintf(stderr, "rchol_driver: %s\n", err.c_str());
        usage();
        return 2;
    }
    if (a.ambiguous)
        std::fprintf(stderr, "rchol_driver: arguments fit both forms; taking the GPU form "
                             "(<num_blocks> <check_solve> <tol>); --cpu-form selects the other\n");

    rc_options opt;
    rc_default_options(&opt);
    opt.device = a.dry_run ? -1 : 0;
    opt.seed = a.seed;
    opt.num_blocks = a.form == 1 ? a.num_blocks : 0;
    opt.factor_precision = a.precision;
    opt.stop_rule = a.stop;
    opt.maxit = a.maxit;
    opt.host_threads = a.form == 2 ? a.threads : 0;
    rc_handle_t h = nullptr;
    int st = rc_create(&h, &opt);
    if (st != RC_OK) return die(nullptr, st, "rc_create");

    std::printf("problem: %s\n", a.path.c_str());
    auto t0 = std::chrono::steady_clock::now();
    if ((st = rc_load_mtx(h, a.path.c_str(), a.kind)) != RC_OK) return die(h, st, "rc_load_mtx");
    int64_t n = 0, nnz = 0;
    rc_get_size(h, &n, &nnz);
    std::printf("number of nodes: %lld\nnumber of nonzeros: %lld\n", (long long)n, (long long)nnz);
    if (a.order >= 0 && (st = rc_order(h, a.order, a.order_seed)) != RC_OK) return die(h, st, "rc_order");
    if (!a.write_ordered.empty() && (st = rc_write_prepared_mtx(h, a.write_ordered.c_str())) != RC_OK)
        return die(h, st, "rc_write_prepared_mtx");
    double read_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (a.dry_run) {
        // the parsed command line and the loaded input, no device work
        std::printf("{\"form\": \"%s\", \"ambiguous\": %d, \"path\": \"%s\", \"num_blocks\": %d, \"threads\": %d, "
                    "\"check_solve\": %d, \"tol\": %.17g, \"kind\": %d, \"order\": %d, \"dump\": \"%s\", "
                    "\"precision\": %d, \"n\": %lld, \"nnz\": %lld}\n",
                    a.form == 1 ? "gpu" : "cpu", a.ambiguous ? 1 : 0, a.path.c_str(), a.num_blocks, a.threads,
                    a.check_solve ? 1 : 0, a.tol, a.kind, a.order, a.dump.c_str(), a.precision, (long long)n,
                    (long long)nnz);
        rc_destroy(h);
        return 0;
    }

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
    if (!a.dump.empty() && (st = rc_dump_factor(h, a.dump.c_str())) != RC_OK) return die(h, st, "rc_dump_factor");

    rc_solve_stats ss;
    std::memset(&ss, 0, sizeof(ss));
    if (a.check_solve) {
        std::vector<double> b((size_t)n), x((size_t)n);
        int kind_rhs = a.rhs_kind >= 0 ? a.rhs_kind : RC_RHS_ZERO_SUM_UNIFORM;   // solver.hpp:427-431
        if ((st = rc_make_rhs(h, kind_rhs, a.rhs_seed, b.data())) != RC_OK) return die(h, st, "rc_make_rhs");
        st = rc_solve(h, b.data(), x.data(), a.tol, &ss);
        if (st != RC_OK && st != RC_ERR_BREAKDOWN) return die(h, st, "rc_solve");
        std::printf("solve preprocess time: %f ms\n", ss.setup_ms);
        std::printf("final iteration: %d, residual: %g\n", ss.iterations, ss.rel_res_precond);
        std::printf("pcg time: %f ms\n", ss.pcg_ms);
        std::printf("normalized diff norm: %.16f\n", ss.rel_res_true);
    }
    if (a.json) {
        std::printf("{\"problem\": \"%s\", \"n\": %lld, \"nnz\": %lld, \"read_order_s\": %.6f, \"factor_ms\": %.4f, "
                    "\"assemble_ms\": %.4f, \"nnz_factor\": %lld, \"fills\": %lld, \"levels\": %d, "
                    "\"iterations\": %d, \"pcg_ms\": %.4f, \"rel_res_true\": %.6e, \"converged\": %d}\n",
                    a.path.c_str(), (long long)n, (long long)nnz, read_s, fs.factor_ms, fs.assemble_ms,
                    (long long)fs.nnz_factor, (long long)fs.fills, fs.nlevels, ss.iterations, ss.pcg_ms,
                    ss.rel_res_true, ss.converged);
    }
    rc_destroy(h);
    return 0;
}
