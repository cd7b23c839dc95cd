// qew_depinning: Rosso-Krauth force ramp toward the depinning threshold f_c.
// Drop-in for the reference's executables/qew_depinning/main.cpp:31-127: the
// same option names, the same solver defaults derived from Ly / ny, the same
// summary lines and exit code 2 on a bad option. Added options: --fc_rel_tol
// > 0 narrows the ramp's bracket to that relative width, each probe started
// from the highest blocked configuration (0 keeps the ramp alone), with
// --fc_probes forces probed at once per round (1 bisects).
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

const char *kTool = "qew_depinning";
const char *kSummary = "Rosso-Krauth force ramp toward the depinning threshold f_c";

std::vector<cli::OptionSpec> option_table() {
    return {
        {"model", Kind::Word, "arclength", "elastic energy", "linear|arclength"},
        {"amplitude", Kind::Real, "1", "disorder amplitude", ""},
        {"xi", Kind::Real, "0.1", "disorder correlation length", ""},
        {"pinning_length", Kind::Real, "5", "pinning length over xi", ""},
        {"Lx", Kind::Real, "8", "box length along the line", ""},
        {"Ly", Kind::Real, "1", "box height (disorder period)", ""},
        {"nx", Kind::Integer, "1024", "sites of the line = noise grid columns", ""},
        {"ny", Kind::Integer, "256", "noise grid rows", ""},
        {"f_step", Kind::Real, "1", "force increment of the ramp", ""},
        {"f_max", Kind::Real, "40", "largest force of the ramp", ""},
        {"fc_rel_tol", Kind::Real, "0", "bisect f_c to this relative width; 0: ramp only", ""},
        {"fc_probes", Kind::Integer, "1", "forces probed per round of that search (1: bisection)", ""},
        {"seed", Kind::Seed, "1", "disorder seed", ""},
        {"output", Kind::Word, "", "output file (.kv: text archive) [depinning_<model>.h5]", ""},
        {"dstep", Kind::Real, "", "march step of a site [Ly / (2 ny)]", ""},
        {"max_advance", Kind::Real, "", "largest advance of a site per sweep [Ly]", ""},
        {"runaway", Kind::Real, "", "mean advance that counts as depinned [2 Ly]", ""},
        {"root_tol", Kind::Real, "1e-10", "bisection width of a site's forward zero", ""},
        {"sweep_tol", Kind::Real, "1e-8", "total advance of a blocked sweep", ""},
        {"max_sweeps", Kind::Integer, "200000", "sweep cap per force", ""},
    };
}

DepinningConfig config_from(const cli::Options &o) {
    DepinningConfig c;
    c.model = o.word("model") == "linear" ? Model::Linear : Model::Arclength;
    c.amplitude = o.real("amplitude");
    c.xi = o.real("xi");
    c.pinning_length_over_xi = o.real("pinning_length");
    c.Lx = o.real("Lx");
    c.Ly = o.real("Ly");
    c.nx = o.integer("nx");
    c.ny = o.integer("ny");
    c.f_step = o.real("f_step");
    c.f_max = o.real("f_max");
    c.fc_rel_tol = o.real("fc_rel_tol");
    c.fc_probes = o.integer("fc_probes");
    if (c.fc_probes < 1) throw std::runtime_error("--fc_probes is at least 1");
    // knobs whose defaults scale with the disorder cell
    const double cell = c.Ly / c.ny;
    c.rp.dstep = o.given("dstep") ? o.real("dstep") : 0.5 * cell;
    c.rp.max_advance = o.given("max_advance") ? o.real("max_advance") : c.Ly;
    c.rp.runaway = o.given("runaway") ? o.real("runaway") : 2.0 * c.Ly;
    c.rp.root_tol = o.real("root_tol");
    c.rp.sweep_tol = o.real("sweep_tol");
    c.rp.max_sweeps = o.integer("max_sweeps");
    return c;
}

const char *verdict(const RkResult &r) {
    if (r.blocked) return "blocked";
    return r.runaway ? "DEPINNED" : "timeout";
}

int run(const cli::Options &opt) {
    const DepinningConfig cfg = config_from(opt);
    const std::string model = opt.word("model");
    const std::string out = opt.given("output") ? opt.word("output") : "depinning_" + model + ".h5";
    initialize();
    FilteredNoise disorder(cfg.nx, cfg.ny, cfg.Lx, cfg.Ly, cfg.amplitude, cfg.xi, cfg.xi,
                           opt.seed("seed"));
    const DepinningResult res = depinning_ramp(cfg, disorder.device_noise());
    std::printf("model=%s  lambda_p/xi=%.3g  line_tension=%.4g\n", model.c_str(),
                (double)cfg.pinning_length_over_xi, (double)res.line_tension);
    for (const DepinningStep &s : res.steps)
        std::printf("f=%7.3f  sweeps=%-7d mean_h=%8.4f  %s\n", (double)s.f, s.result.sweeps,
                    (double)s.result.mean_h, verdict(s.result));
    if (res.f_c_upper_bound >= 0)
        std::printf("Line depinned near f=%.3f; stopping ramp.\n", (double)res.f_c_upper_bound);
    if (res.f_c >= 0)
        std::printf("f_c = %.12g  in [%.12g, %.12g] after %d bisection probes\n", (double)res.f_c,
                    (double)res.f_c_lower, (double)res.f_c_upper, res.bisection_probes);
    write_output(out, [&](auto &w) { write_depinning(w, cfg, res); });
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
