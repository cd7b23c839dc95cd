// qew_benchmark: the reference's hot-loop benchmark
// (executables/benchmark/main.cpp:17-101) on the MI355X through libqew_amd.
// Times gradient(), objective() and one Langevin step on a 262144-site line
// (noise setup excluded, one warm-up gradient first), then an end-to-end
// preconditioned L-BFGS relaxation shaped like a qew_static run, and prints
// the reference's two summary lines. The reference's `threads=` field (the
// OpenMP concurrency) becomes `device=` (the HIP device ordinal). Extra
// options (no reference counterpart): --reps (repetitions per kernel,
// default 200, the reference's fixed count) and --device.
#include <chrono>
#include <cstdio>
#include <exception>
#include <string>
#include <vector>

#include "qew/cli.hpp"
#include "qew/qew.hpp"

using namespace qew;

namespace {

const char *kTool = "qew_benchmark";
const char *kSummary =
    "hot-loop timings (gradient, objective, Langevin step) on a 262144-site line,\n"
    "then one preconditioned L-BFGS relaxation";

std::vector<cli::OptionSpec> option_table() {
    return {
        {"reps", cli::Kind::Integer, "200", "repetitions per timed kernel", ""},
        {"device", cli::Kind::Integer, "0", "HIP device ordinal", ""},
    };
}

class Timer {
public:
    void reset() { t0_ = std::chrono::steady_clock::now(); }
    double seconds() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    }

private:
    std::chrono::steady_clock::time_point t0_ = std::chrono::steady_clock::now();
};

int run(const cli::Options &opt) {
    const int reps = opt.integer("reps");
    const int device = opt.integer("device");
    if (reps < 1) throw std::runtime_error("--reps is at least 1");
    initialize(device);

    const int nx = 262144;  // line length (the parallel dimension)
    const int ny = 64;      // noise y-resolution (setup only)
    const real_t Lx = 256.0, Ly = 1.0;

    Timer timer;
    FilteredNoise noise(nx, ny, Lx, Ly, 1.0, 0.1, 0.1, /*seed=*/1);
    fence();
    const double t_setup = timer.seconds();
    const DeviceNoise dn = noise.device_noise();

    Params p;
    p.physical_size = Lx;
    p.line_tension = 1.0;
    p.driving_force = 0.1;
    p.model = Model::Arclength;

    View1D h("h", nx), grad("grad", nx);
    deep_copy(h, 0.5 * Ly);

    gradient(h, p, dn, grad);  // warm-up: the first launch pays one-off costs
    fence();

    timer.reset();
    for (int k = 0; k < reps; ++k) gradient(h, p, dn, grad);
    fence();
    const double t_grad = timer.seconds();

    timer.reset();
    volatile real_t sink = 0;
    for (int k = 0; k < reps; ++k) sink = sink + objective(h, p, dn);
    fence();
    const double t_obj = timer.seconds();
    (void)sink;

    LangevinDynamics dyn(nx, /*seed=*/7);
    timer.reset();
    for (int k = 0; k < reps; ++k) dyn.step(h, p, dn, 0.01, 0.01);
    fence();
    const double t_dyn = timer.seconds();

    std::printf(
        "device=%2d  setup=%6.3fs  grad=%7.3f ms/ea  obj=%7.3f ms/ea  "
        "dyn=%7.3f ms/ea\n",
        device, t_setup, 1e3 * t_grad / reps, 1e3 * t_obj / reps, 1e3 * t_dyn / reps);

    // End-to-end L-BFGS relaxation shaped like a qew_static run
    // (preconditioned, flat start, lambda_p/xi = 1).
    const int snx = 4096, sny = 512;
    const real_t sLx = 32.0, sLy = 1.0, sxi = 0.1;
    FilteredNoise snoise(snx, sny, sLx, sLy, 1.0, sxi, sxi, /*seed=*/1);
    Params sp;
    sp.physical_size = sLx;
    sp.line_tension = 1.0;
    sp.driving_force = 0.0;
    sp.model = Model::Arclength;
    View1D hs("hs", snx);
    deep_copy(hs, 0.5 * sLy);
    LbfgsParams sopt;
    sopt.ftol = 1e-6;
    sopt.max_iter = 5000;
    sopt.precondition = true;
    sopt.precond_shift = 0.5 * (sLx / snx) / (sxi * sxi);
    fence();
    timer.reset();
    const LbfgsResult sres = lbfgs_minimize(hs, sp, snoise.device_noise(), sopt);
    const double t_lbfgs = timer.seconds();
    std::printf("lbfgs(n=%d, precond): %6.3fs  iters=%d  feval=%ld  %s\n", snx, t_lbfgs,
                sres.iterations, sres.n_func_evals, sres.converged ? "converged" : "NOT converged");
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
