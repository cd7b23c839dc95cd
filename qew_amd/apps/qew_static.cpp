// qew_static: L-BFGS relaxation of a flat line for a sweep of pinning lengths
// (the reference's executables/qew_static/main.cpp:30-116: same options, same
// defaults, same summary lines, exit code 2 on a bad option). Runs on the
// MI355X through libqew_amd. Output: the reference's dataset keys in an HDF5
// file (a qew-kv text archive when --output ends in .kv). Extra options: --m
// (L-BFGS history length, default 8 as in the reference's LbfgsParams) and
// --batched (0 minimises the pinning lengths one after the other as the
// reference does; 1, the default, all in one persistent batched launch).
#include <cstdint>
#include <cstdio>
#include <exception>
#include <string>
#include <vector>

#include "qew/cli.hpp"
#include "qew/qew_io.hpp"
#include "qew/qew_runs.hpp"

using namespace qew;

namespace {
const char *kUsage =
    "qew_static -- L-BFGS relaxation, sweep over pinning length.\n"
    "All options take a value; defaults match qew_static.py.\n\n"
    "  --model           linear|arclength      (default arclength)\n"
    "  --amplitude       disorder amplitude U0  (default 1)\n"
    "  --xi              correlation length     (default 0.1)\n"
    "  --Lx --Ly         system size            (default 32, 1)\n"
    "  --nx --ny         resolution             (default 4096, 512)\n"
    "  --driving_force   external force         (default 0)\n"
    "  --pinning_lengths comma list of lambda_p/xi (default 0.01,0.1,1,10,100)\n"
    "  --ftol            L-BFGS tolerance       (default 1e-6)\n"
    "  --max_iter        L-BFGS iteration cap   (default 5000)\n"
    "  --m               L-BFGS history length  (default 8)\n"
    "  --batched         1: all pinning lengths in one batched launch, 0: one\n"
    "                    host-driven minimisation per length (default 1)\n"
    "  --seed            noise RNG seed         (default 1)\n"
    "  --output          output file            (default static_<model>.h5; .kv: text)\n";

Model parse_model(const std::string &name) {
    if (name == "linear") return Model::Linear;
    if (name == "arclength") return Model::Arclength;
    throw std::runtime_error("--model must be 'linear' or 'arclength', got '" + name + "'");
}

int run(int argc, char **argv) {
    cli::Args args(argc, argv);
    if (args.help) {
        std::fputs(kUsage, stdout);
        return 0;
    }
    StaticConfig cfg;
    const std::string model_name = args.str("model", "arclength");
    cfg.model = parse_model(model_name);
    cfg.amplitude = args.dbl("amplitude", 1.0);
    cfg.xi = args.dbl("xi", 0.1);
    cfg.Lx = args.dbl("Lx", 32.0);
    cfg.Ly = args.dbl("Ly", 1.0);
    cfg.nx = args.integer("nx", 4096);
    cfg.ny = args.integer("ny", 512);
    cfg.driving_force = args.dbl("driving_force", 0.0);
    cfg.pinning_lengths_over_xi = args.dbl_list("pinning_lengths", {0.01, 0.1, 1.0, 10.0, 100.0});
    cfg.ftol = args.dbl("ftol", 1e-6);
    cfg.max_iter = args.integer("max_iter", 5000);
    cfg.m = args.integer("m", 8);
    const int batched = args.integer("batched", 1);
    if (batched != 0 && batched != 1) throw std::runtime_error("--batched must be 0 or 1");
    cfg.batched = batched == 1;
    const std::uint64_t seed = args.u64("seed", 1);
    const std::string fname = args.str("output", std::string("static_") + model_name + ".h5");
    args.require_all_used();

    initialize();
    FilteredNoise noise(cfg.nx, cfg.ny, cfg.Lx, cfg.Ly, cfg.amplitude, cfg.xi, cfg.xi, seed);
    const std::vector<StaticProfile> profiles = static_sweep(cfg, noise.device_noise());
    for (const StaticProfile &prof : profiles)
        std::printf("lp/xi=%-6.3g lt=%-10.4g iters=%-7d max|F|=%.3e  %s\n",
                    prof.pinning_length_over_xi, static_cast<double>(prof.line_tension),
                    prof.result.iterations, static_cast<double>(prof.result.max_force),
                    prof.result.converged ? "converged" : "NOT converged");
    write_output(fname, [&](auto &w) { write_static(w, cfg, profiles); });
    std::printf("wrote %s\n", fname.c_str());
    return 0;
}
}  // namespace

int main(int argc, char *argv[]) {
    int rc = 0;
    try {
        rc = run(argc, argv);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "qew_static: %s\n\n", e.what());
        std::fputs(kUsage, stderr);
        rc = 2;
    }
    finalize();
    return rc;
}
