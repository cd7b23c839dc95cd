// qew_static: L-BFGS relaxation of a flat line, one minimum per pinning
// length. Drop-in for the reference's executables/qew_static/main.cpp:30-116:
// the same option names and defaults, the same one-line summary per pinning
// length and exit code 2 on a bad option; the output file holds the
// reference's dataset keys (HDF5, or a text key/value archive for a name
// ending in .kv). Added options: --m (L-BFGS history length) and --batched.
#include <cstdint>
#include <cstdio>
#include <exception>
#include <string>
#include <vector>

#include "qew/cli.hpp"
#include "qew/qew_io.hpp"
#include "qew/qew_runs.hpp"

using namespace qew;
using cli::Kind;

namespace {

const char *kTool = "qew_static";
const char *kSummary = "relaxes a flat line with L-BFGS for each pinning length";

std::vector<cli::OptionSpec> option_table() {
    return {
        {"model", Kind::Word, "arclength", "elastic energy", "linear|arclength"},
        {"amplitude", Kind::Real, "1", "disorder amplitude", ""},
        {"xi", Kind::Real, "0.1", "disorder correlation length", ""},
        {"Lx", Kind::Real, "32", "box length along the line", ""},
        {"Ly", Kind::Real, "1", "box height (disorder period)", ""},
        {"nx", Kind::Integer, "4096", "sites of the line = noise grid columns", ""},
        {"ny", Kind::Integer, "512", "noise grid rows", ""},
        {"driving_force", Kind::Real, "0", "uniform driving force", ""},
        {"pinning_lengths", Kind::Reals, "0.01,0.1,1,10,100", "pinning lengths over xi", ""},
        {"ftol", Kind::Real, "1e-6", "L-BFGS stop: max |grad|", ""},
        {"max_iter", Kind::Integer, "5000", "L-BFGS iteration cap", ""},
        {"m", Kind::Integer, "8", "L-BFGS history length", ""},
        {"batched", Kind::Integer, "1",
         "1: every pinning length in one batched launch; 0: one after the other", ""},
        {"seed", Kind::Seed, "1", "disorder seed", ""},
        {"output", Kind::Word, "", "output file (.kv: text archive) [static_<model>.h5]", ""},
    };
}

StaticConfig config_from(const cli::Options &o) {
    StaticConfig c;
    c.model = o.word("model") == "linear" ? Model::Linear : Model::Arclength;
    c.amplitude = o.real("amplitude");
    c.xi = o.real("xi");
    c.Lx = o.real("Lx");
    c.Ly = o.real("Ly");
    c.nx = o.integer("nx");
    c.ny = o.integer("ny");
    c.driving_force = o.real("driving_force");
    c.pinning_lengths_over_xi = o.reals("pinning_lengths");
    c.ftol = o.real("ftol");
    c.max_iter = o.integer("max_iter");
    c.m = o.integer("m");
    const int batched = o.integer("batched");
    if (batched != 0 && batched != 1) throw std::runtime_error("--batched is 0 or 1");
    c.batched = batched == 1;
    return c;
}

int run(const cli::Options &opt) {
    const StaticConfig cfg = config_from(opt);
    const std::string out =
        opt.given("output") ? opt.word("output") : "static_" + opt.word("model") + ".h5";
    initialize();
    FilteredNoise disorder(cfg.nx, cfg.ny, cfg.Lx, cfg.Ly, cfg.amplitude, cfg.xi, cfg.xi,
                           opt.seed("seed"));
    const std::vector<StaticProfile> minima = static_sweep(cfg, disorder.device_noise());
    for (const StaticProfile &pr : minima) {
        const LbfgsResult &r = pr.result;
        std::printf("lp/xi=%-6.3g lt=%-10.4g iters=%-7d max|F|=%.3e  %s\n",
                    pr.pinning_length_over_xi, (double)pr.line_tension, r.iterations,
                    (double)r.max_force, r.converged ? "converged" : "NOT converged");
    }
    write_output(out, [&](auto &w) { write_static(w, cfg, minima); });
    std::printf("wrote %s\n", out.c_str());
    return 0;
}

}  // namespace

int main(int argc, char *argv[]) {
    const std::vector<cli::OptionSpec> table = option_table();
    int status = 0;
    try {
        const cli::Options opt(table, argc, argv);
        if (opt.help()) std::fputs(opt.usage(kTool, kSummary).c_str(), stdout);
        else status = run(opt);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s: %s\n", kTool, e.what());
        std::fprintf(stderr, "(%s --help lists the options)\n", kTool);
        status = 2;
    }
    finalize();
    return status;
}
