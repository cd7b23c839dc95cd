// qew_dynamic: annealed overdamped Langevin dynamics of a line started flat.
// Drop-in for the reference's executables/qew_dynamic/main.cpp:29-116: the
// same option names and defaults, the same summary lines and exit code 2 on a
// bad option. Thermal kicks are Philox normals keyed by (seed + 1, step,
// site). Added option: --dt, a fixed timestep (0 keeps the reference's
// heuristic).
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

const char *kTool = "qew_dynamic";
const char *kSummary = "overdamped Langevin dynamics with an annealed temperature";

std::vector<cli::OptionSpec> option_table() {
    return {
        {"model", Kind::Word, "arclength", "elastic energy", "linear|arclength"},
        {"amplitude", Kind::Real, "1", "disorder amplitude", ""},
        {"xi", Kind::Real, "0.1", "disorder correlation length", ""},
        {"pinning_length", Kind::Real, "0.1", "pinning length over xi", ""},
        {"driving_force", Kind::Real, "0", "uniform driving force", ""},
        {"Lx", Kind::Real, "2", "box length along the line", ""},
        {"Ly", Kind::Real, "1", "box height (disorder period)", ""},
        {"nx", Kind::Integer, "2048", "sites of the line = noise grid columns", ""},
        {"ny", Kind::Integer, "1024", "noise grid rows", ""},
        {"tau", Kind::Real, "1", "damping time", ""},
        {"tl_start", Kind::Real, "0.01", "thermal length at the start of the anneal", ""},
        {"tl_end", Kind::Real, "0.0001", "thermal length at its end", ""},
        {"total_time", Kind::Real, "1000", "integration time", ""},
        {"dt", Kind::Real, "0", "fixed timestep; 0 picks it from the stiffness", ""},
        {"seed", Kind::Seed, "1", "disorder seed (the thermal stream uses seed + 1)", ""},
        {"output", Kind::Word, "", "output file (.kv: text archive) [dynamic_<model>.h5]", ""},
    };
}

DynamicConfig config_from(const cli::Options &o) {
    DynamicConfig c;
    c.model = o.word("model") == "linear" ? Model::Linear : Model::Arclength;
    c.amplitude = o.real("amplitude");
    c.xi = o.real("xi");
    c.pinning_length_over_xi = o.real("pinning_length");
    c.driving_force = o.real("driving_force");
    c.Lx = o.real("Lx");
    c.Ly = o.real("Ly");
    c.nx = o.integer("nx");
    c.ny = o.integer("ny");
    c.tau = o.real("tau");
    c.tl_start = o.real("tl_start");
    c.tl_end = o.real("tl_end");
    c.total_time = o.real("total_time");
    c.dt = o.real("dt");
    c.seed = o.seed("seed");
    return c;
}

int run(const cli::Options &opt) {
    const DynamicConfig cfg = config_from(opt);
    const std::string out =
        opt.given("output") ? opt.word("output") : "dynamic_" + opt.word("model") + ".h5";
    initialize();
    FilteredNoise disorder(cfg.nx, cfg.ny, cfg.Lx, cfg.Ly, cfg.amplitude, cfg.xi, cfg.xi, cfg.seed);
    const DynamicResult res = dynamic_anneal(cfg, disorder.device_noise());
    std::printf("dt=%.3g  total_time=%.3g  steps=%ld  line_tension=%.4g\n", (double)res.dt,
                (double)res.total_time, (long)(res.total_time / res.dt), (double)res.line_tension);
    for (const DynamicSnapshot &s : res.snapshots) std::printf("  saved t=%.6g\n", (double)s.t);
    write_output(out, [&](auto &w) { write_dynamic(w, cfg, res); });
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
