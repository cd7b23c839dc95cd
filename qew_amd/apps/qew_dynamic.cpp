// qew_dynamic: annealed overdamped Langevin dynamics of a flat-started line
// (the reference's executables/qew_dynamic/main.cpp:29-116: same options,
// defaults and summary lines; exit code 2 on a bad option). Thermal kicks are
// Philox normals keyed by (seed + 1, step, site). Extra option: --dt, a fixed
// timestep replacing the reference's heuristic (0 keeps the heuristic).
#include <cstdint>
#include <cstdio>
#include <exception>
#include <string>

#include "qew/cli.hpp"
#include "qew/qew_io.hpp"
#include "qew/qew_runs.hpp"

using namespace qew;

namespace {
const char *kUsage =
    "qew_dynamic -- annealed overdamped Langevin dynamics.\n"
    "All options take a value; defaults match qew_dynamic.py.\n\n"
    "  --model           linear|arclength      (default arclength)\n"
    "  --amplitude       disorder amplitude U0  (default 1)\n"
    "  --xi              correlation length     (default 0.1)\n"
    "  --pinning_length  lambda_p/xi            (default 0.1)\n"
    "  --driving_force   external force         (default 0)\n"
    "  --Lx --Ly         system size            (default 2, 1)\n"
    "  --nx --ny         resolution             (default 2048, 1024)\n"
    "  --tau             damping time           (default 1)\n"
    "  --tl_start        anneal thermal length, start (default 0.01)\n"
    "  --tl_end          anneal thermal length, end   (default 0.0001)\n"
    "  --total_time      integration time       (default 1000; production ~1e6)\n"
    "  --dt              fixed timestep; 0 = heuristic (default 0)\n"
    "  --seed            noise/RNG seed          (default 1)\n"
    "  --output          output file            (default dynamic_<model>.h5; .kv: text)\n";

int run(int argc, char **argv) {
    cli::Args args(argc, argv);
    if (args.help) {
        std::fputs(kUsage, stdout);
        return 0;
    }
    DynamicConfig cfg;
    const std::string model_name = args.str("model", "arclength");
    if (model_name == "linear") cfg.model = Model::Linear;
    else if (model_name == "arclength") cfg.model = Model::Arclength;
    else throw std::runtime_error("--model must be 'linear' or 'arclength', got '" + model_name + "'");
    cfg.amplitude = args.dbl("amplitude", 1.0);
    cfg.xi = args.dbl("xi", 0.1);
    cfg.pinning_length_over_xi = args.dbl("pinning_length", 0.1);
    cfg.driving_force = args.dbl("driving_force", 0.0);
    cfg.Lx = args.dbl("Lx", 2.0);
    cfg.Ly = args.dbl("Ly", 1.0);
    cfg.nx = args.integer("nx", 2048);
    cfg.ny = args.integer("ny", 1024);
    cfg.seed = args.u64("seed", 1);
    cfg.tau = args.dbl("tau", 1.0);
    cfg.tl_start = args.dbl("tl_start", 0.01);
    cfg.tl_end = args.dbl("tl_end", 0.0001);
    cfg.total_time = args.dbl("total_time", 1000.0);
    cfg.dt = args.dbl("dt", 0.0);
    const std::string fname = args.str("output", std::string("dynamic_") + model_name + ".h5");
    args.require_all_used();

    initialize();
    FilteredNoise noise(cfg.nx, cfg.ny, cfg.Lx, cfg.Ly, cfg.amplitude, cfg.xi, cfg.xi, cfg.seed);
    const DynamicResult res = dynamic_anneal(cfg, noise.device_noise());
    std::printf("dt=%.3g  total_time=%.3g  steps=%ld  line_tension=%.4g\n",
                static_cast<double>(res.dt), static_cast<double>(res.total_time),
                static_cast<long>(res.total_time / res.dt), static_cast<double>(res.line_tension));
    for (const DynamicSnapshot &snap : res.snapshots)
        std::printf("  saved t=%.6g\n", static_cast<double>(snap.t));
    write_output(fname, [&](auto &w) { write_dynamic(w, cfg, res); });
    std::printf("wrote %s\n", fname.c_str());
    return 0;
}
}  // namespace

int main(int argc, char *argv[]) {
    int rc = 0;
    try {
        rc = run(argc, argv);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "qew_dynamic: %s\n\n", e.what());
        std::fputs(kUsage, stderr);
        rc = 2;
    }
    finalize();
    return rc;
}
