// sb_cli.cpp -- `sb_solve <instance> [options]`, the drop-in for the
// reference's `./opt <instance>` (SB/CUDA_GPU/m2_opt.cu:93-103,206-210):
// same positional argument, same edge-list input, same output.txt with one
// energy H per replica (trial) in replica order.  The reference's compile-time
// defines become the options below, with the reference's values as defaults.
#include <cerrno>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sb.h"

static void usage()
{
    std::printf(
        "Usage:\n"
        "       sb_solve <instance> [options]\n"
        "  --variant asb|asb-m1|bsb|dsb  update rule (default asb, m2_opt.cu; asb-m1 is\n"
        "                            Bifurcation_m1.cu.c)\n"
        "  --precision fp32|split|ordered\n"
        "                            split (alias fp8bf16): J as FP8 E4M3 when integer with\n"
        "                            |J| <= 16, else BF16; X as BF16 hi + lo; FP32 sums\n"
        "  --steps N                 (default 100)\n"
        "  --dt X                    coupling step DeltaT (default 1.0)\n"
        "  --a0 X                    detune / a0 (default 1)\n"
        "  --c0 X                    coupling xi / c0 (default: auto from sigma_J)\n"
        "  --K X  --M N              aSB Kerr coefficient and sub-steps (1, 2)\n"
        "  --schedule linear         pump schedule p = k/steps\n"
        "  --replicas R              replicas (reference TIMES, default 1)\n"
        "  --init reference|uniform  initial state (default reference)\n"
        "  --seed S                  (default 0)\n"
        "  --one-based               edge list indices start at 1\n"
        "  --sparse                  CSR field (low-density instances; same sums as ordered)\n"
        "  --output FILE             energies (default output.txt)\n"
        "  --spins FILE              also write spins, one replica per line\n"
        "  --cut FILE                also write the cut value of the max-cut mapping W = -J,\n"
        "                            one replica per line\n"
        "  --device D                GPU ordinal (default 0)\n");
}

static bool parse_double(const char *s, double *out)
{
    char *e = nullptr;
    errno = 0;
    double v = std::strtod(s, &e);
    if (e == s || *e || errno) return false;
    *out = v;
    return true;
}

static bool parse_long(const char *s, long long *out)
{
    char *e = nullptr;
    errno = 0;
    long long v = std::strtoll(s, &e, 10);
    if (e == s || *e || errno) return false;
    *out = v;
    return true;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        usage();
        return 0;  // m2_opt.cu:93-98 prints usage and exits 0
    }
    sb_params p;
    sb_default_params(&p);
    const char *instance = nullptr;
    std::string out_path = "output.txt", spins_path, cut_path;
    int one_based = 0, device = 0, sparse = 0;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto need = [&](const char *name) -> const char * {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", name);
                std::exit(2);
            }
            return argv[++i];
        };
        double d;
        long long l;
        if (a == "--variant") {
            std::string v = need("--variant");
            if (v == "asb") p.variant = SB_ASB;
            else if (v == "bsb") p.variant = SB_BSB;
            else if (v == "dsb") p.variant = SB_DSB;
            else if (v == "asb-m1") p.variant = SB_ASB_M1;
            else { std::fprintf(stderr, "bad --variant %s\n", v.c_str()); return 2; }
        } else if (a == "--precision") {
            std::string v = need("--precision");
            if (v == "fp32") p.precision = SB_PREC_FP32;
            else if (v == "split" || v == "fp8bf16") p.precision = SB_PREC_SPLIT;
            else if (v == "ordered") p.precision = SB_PREC_ORDERED;
            else { std::fprintf(stderr, "bad --precision %s\n", v.c_str()); return 2; }
        } else if (a == "--init") {
            std::string v = need("--init");
            if (v == "reference") p.init = SB_INIT_REFERENCE;
            else if (v == "uniform") p.init = SB_INIT_UNIFORM;
            else { std::fprintf(stderr, "bad --init %s\n", v.c_str()); return 2; }
        } else if (a == "--schedule") {
            std::string v = need("--schedule");
            if (v != "linear") { std::fprintf(stderr, "bad --schedule %s\n", v.c_str()); return 2; }
            p.schedule = SB_SCHED_LINEAR;
        } else if (a == "--steps" && parse_long(need("--steps"), &l)) p.steps = (int32_t)l;
        else if (a == "--replicas" && parse_long(need("--replicas"), &l)) p.replicas = (int32_t)l;
        else if (a == "--M" && parse_long(need("--M"), &l)) p.M = (int32_t)l;
        else if (a == "--seed" && parse_long(need("--seed"), &l)) p.seed = (uint64_t)l;
        else if (a == "--device" && parse_long(need("--device"), &l)) device = (int)l;
        else if (a == "--dt" && parse_double(need("--dt"), &d)) p.dt = d;
        else if (a == "--a0" && parse_double(need("--a0"), &d)) p.a0 = d;
        else if (a == "--c0" && parse_double(need("--c0"), &d)) p.c0 = d;
        else if (a == "--K" && parse_double(need("--K"), &d)) p.K = d;
        else if (a == "--one-based") one_based = 1;
        else if (a == "--sparse") sparse = 1;
        else if (a == "--output") out_path = need("--output");
        else if (a == "--spins") spins_path = need("--spins");
        else if (a == "--cut") cut_path = need("--cut");
        else if (a == "--help" || a == "-h") { usage(); return 0; }
        else if (a.size() > 1 && a[0] == '-' && a[1] == '-') {
            std::fprintf(stderr, "bad option or value: %s\n", a.c_str());
            return 2;
        } else if (!instance) instance = argv[i];
        else { usage(); return 2; }
    }
    if (!instance) { usage(); return 0; }

    sb_ctx *ctx = nullptr;
    int rc = sb_ctx_create(device, &ctx);
    if (rc) {
        std::fprintf(stderr, "sb_ctx_create: %s (%s)\n", sb_status_string(rc), sb_last_error_message());
        return 1;
    }
    sb_problem *prob = nullptr;
    rc = sb_problem_from_edgelist_file(ctx, instance, one_based, &prob);
    if (rc) {
        std::fprintf(stderr, "%s: %s (%s)\n", instance, sb_status_string(rc), sb_last_error_message());
        sb_ctx_destroy(ctx);
        return 1;
    }
    if (sparse && (rc = sb_problem_set_sparse(prob, 1)) != SB_OK) {
        std::fprintf(stderr, "--sparse: %s (%s)\n", sb_status_string(rc), sb_last_error_message());
        sb_problem_destroy(prob);
        sb_ctx_destroy(ctx);
        return 1;
    }
    int32_t N = 0, integer = 0;
    sb_problem_info(prob, &N, &integer, nullptr, nullptr);
    const int R = p.replicas;
    std::vector<int64_t> Hi(R);
    std::vector<double> Hd(R);
    std::vector<int8_t> S(spins_path.empty() ? 0 : (size_t)R * N);
    std::vector<double> cut(R);
    sb_result res;
    std::memset(&res, 0, sizeof(res));
    res.energy_i64 = Hi.data();
    res.energy_f64 = Hd.data();
    res.spins = S.empty() ? nullptr : S.data();
    res.cut = cut.data();
    rc = sb_solve(ctx, prob, &p, &res);
    if (rc) {
        std::fprintf(stderr, "sb_solve: %s (%s)\n", sb_status_string(rc), sb_last_error_message());
        sb_problem_destroy(prob);
        sb_ctx_destroy(ctx);
        return 1;
    }
    std::printf("The elapsed time in gpu was %.2f ms\n", res.step_ms);
    std::printf("%lf\n", res.total_ms / 1000.0);
    FILE *f = std::fopen(out_path.c_str(), "w");
    if (!f) {
        std::fprintf(stderr, "cannot write %s\n", out_path.c_str());
        return 1;
    }
    // Integers when the energies are exact (integer J and field), else FP64.
    for (int r = 0; r < R; ++r) {
        if (res.energy_exact) std::fprintf(f, "%" PRId64 "\n", Hi[r]);
        else std::fprintf(f, "%.17g\n", Hd[r]);
    }
    std::fclose(f);
    if (!cut_path.empty()) {
        FILE *c = std::fopen(cut_path.c_str(), "w");
        if (!c) {
            std::fprintf(stderr, "cannot write %s\n", cut_path.c_str());
            return 1;
        }
        for (int r = 0; r < R; ++r) {
            if (res.energy_exact) std::fprintf(c, "%" PRId64 "\n", (int64_t)std::llround(cut[r]));
            else std::fprintf(c, "%.17g\n", cut[r]);
        }
        std::fclose(c);
    }
    if (!spins_path.empty()) {
        FILE *g = std::fopen(spins_path.c_str(), "w");
        if (!g) {
            std::fprintf(stderr, "cannot write %s\n", spins_path.c_str());
            return 1;
        }
        std::string line;
        for (int r = 0; r < R; ++r) {
            line.clear();
            for (int i = 0; i < N; ++i) {
                if (i) line.push_back(' ');
                line += S[(size_t)r * N + i] > 0 ? "1" : "-1";
            }
            std::fprintf(g, "%s\n", line.c_str());
        }
        std::fclose(g);
    }
    sb_problem_destroy(prob);
    sb_ctx_destroy(ctx);
    return 0;
}
