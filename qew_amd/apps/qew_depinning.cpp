// qew_depinning: Rosso-Krauth force ramp toward f_c (the reference's
// executables/qew_depinning/main.cpp:31-127: same options, RK knob defaults
// derived from Ly/ny, same summary lines; exit code 2 on a bad option).
// Extra option: --fc_rel_tol > 0 bisects the bracket the ramp found down to
// that relative width, warm-starting every probe from the highest blocked
// configuration (no reference counterpart; 0 keeps the reference's ramp only).
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
    "qew_depinning -- Rosso-Krauth force ramp toward the threshold f_c.\n"
    "All options take a value; defaults match qew_depinning.py.\n\n"
    "  --model           linear|arclength      (default arclength)\n"
    "  --amplitude       disorder amplitude U0  (default 1)\n"
    "  --xi              correlation length     (default 0.1)\n"
    "  --pinning_length  lambda_p/xi (stiff)    (default 5)\n"
    "  --Lx --Ly         system size            (default 8, 1)\n"
    "  --nx --ny         resolution             (default 1024, 256)\n"
    "  --f_step          force ramp increment   (default 1)\n"
    "  --f_max           force ramp cap         (default 40)\n"
    "  --fc_rel_tol      bisect f_c to this relative width; 0 = ramp only (default 0)\n"
    "  --seed            noise RNG seed          (default 1)\n"
    "  --output          output file            (default depinning_<model>.h5; .kv: text)\n"
    "  RK solver knobs (defaults derived from Ly/ny):\n"
    "  --dstep --max_advance --runaway --root_tol --sweep_tol --max_sweeps\n";

int run(int argc, char **argv) {
    cli::Args args(argc, argv);
    if (args.help) {
        std::fputs(kUsage, stdout);
        return 0;
    }
    DepinningConfig cfg;
    const std::string model_name = args.str("model", "arclength");
    if (model_name == "linear") cfg.model = Model::Linear;
    else if (model_name == "arclength") cfg.model = Model::Arclength;
    else throw std::runtime_error("--model must be 'linear' or 'arclength', got '" + model_name + "'");
    cfg.amplitude = args.dbl("amplitude", 1.0);
    cfg.xi = args.dbl("xi", 0.1);
    cfg.pinning_length_over_xi = args.dbl("pinning_length", 5.0);
    cfg.Lx = args.dbl("Lx", 8.0);
    cfg.Ly = args.dbl("Ly", 1.0);
    cfg.nx = args.integer("nx", 1024);
    cfg.ny = args.integer("ny", 256);
    cfg.f_step = args.dbl("f_step", 1.0);
    cfg.f_max = args.dbl("f_max", 40.0);
    cfg.fc_rel_tol = args.dbl("fc_rel_tol", 0.0);
    cfg.rp.dstep = args.dbl("dstep", 0.5 * (cfg.Ly / cfg.ny));
    cfg.rp.max_advance = args.dbl("max_advance", cfg.Ly);
    cfg.rp.runaway = args.dbl("runaway", 2.0 * cfg.Ly);
    cfg.rp.root_tol = args.dbl("root_tol", 1e-10);
    cfg.rp.sweep_tol = args.dbl("sweep_tol", 1e-8);
    cfg.rp.max_sweeps = args.integer("max_sweeps", 200000);
    const std::uint64_t seed = args.u64("seed", 1);
    const std::string fname = args.str("output", std::string("depinning_") + model_name + ".h5");
    args.require_all_used();

    initialize();
    FilteredNoise noise(cfg.nx, cfg.ny, cfg.Lx, cfg.Ly, cfg.amplitude, cfg.xi, cfg.xi, seed);
    const DepinningResult res = depinning_ramp(cfg, noise.device_noise());
    std::printf("model=%s  lambda_p/xi=%.3g  line_tension=%.4g\n", model_name.c_str(),
                static_cast<double>(cfg.pinning_length_over_xi),
                static_cast<double>(res.line_tension));
    for (const DepinningStep &st : res.steps)
        std::printf("f=%7.3f  sweeps=%-7d mean_h=%8.4f  %s\n", static_cast<double>(st.f),
                    st.result.sweeps, static_cast<double>(st.result.mean_h),
                    st.result.blocked ? "blocked" : (st.result.runaway ? "DEPINNED" : "timeout"));
    if (res.f_c_upper_bound >= 0)
        std::printf("Line depinned near f=%.3f; stopping ramp.\n",
                    static_cast<double>(res.f_c_upper_bound));
    if (res.f_c >= 0)
        std::printf("f_c = %.12g  in [%.12g, %.12g] after %d bisection probes\n",
                    static_cast<double>(res.f_c), static_cast<double>(res.f_c_lower),
                    static_cast<double>(res.f_c_upper), res.bisection_probes);
    write_output(fname, [&](auto &w) { write_depinning(w, cfg, res); });
    std::printf("wrote %s\n", fname.c_str());
    return 0;
}
}  // namespace

int main(int argc, char *argv[]) {
    int rc = 0;
    try {
        rc = run(argc, argv);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "qew_depinning: %s\n\n", e.what());
        std::fputs(kUsage, stderr);
        rc = 2;
    }
    finalize();
    return rc;
}
