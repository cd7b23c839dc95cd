// Writes the output files of the reference's writer tests
// (tests/test_executables.cpp:199-289: WriteStaticKeysAndValues,
// WriteDynamicKeysAndValues, WriteDepinningSkipsRunawayAndWritesFc) through
// include/qew/qew_io.hpp, once as HDF5 and once as a qew-kv archive, into the
// directory given as argv[1]. The keys and values are read back and checked
// by tests/test_io.py with two readers (HDF5 through ctypes, and h5dump).
// Host only: no device call is made.
#include <cstdio>
#include <exception>
#include <string>
#include <vector>

#include "qew/qew_io.hpp"

using namespace qew;

namespace {

template <class Writer>
void write_all(const std::string &dir, const std::string &ext) {
    {
        StaticConfig cfg;
        cfg.Lx = 4.0;
        cfg.driving_force = 0.0;
        std::vector<StaticProfile> profiles(1);
        profiles[0].pinning_length_over_xi = 0.1;
        profiles[0].line_tension = 0.5;
        profiles[0].h = {1.0, 2.0, 3.0, 4.0};
        Writer w(dir + "/test_write_static" + ext);
        write_static(w, cfg, profiles);
    }
    {
        DynamicConfig cfg;
        cfg.Lx = 2.0;
        cfg.driving_force = 0.0;
        cfg.pinning_length_over_xi = 0.1;
        DynamicResult res;
        res.line_tension = 2.0;
        res.snapshots.push_back(DynamicSnapshot{1.0, {5.0, 6.0}});
        Writer w(dir + "/test_write_dynamic" + ext);
        write_dynamic(w, cfg, res);
    }
    {
        DepinningConfig cfg;
        cfg.Lx = 8.0;
        cfg.pinning_length_over_xi = 5.0;
        DepinningResult res;
        res.f_c_upper_bound = 0.8;
        DepinningStep blocked;
        blocked.f = 0.0;
        blocked.result.blocked = true;
        blocked.h = {0.5, 0.5, 0.5};
        DepinningStep runaway;
        runaway.f = 0.8;
        runaway.result.blocked = false;
        runaway.result.runaway = true;  // h intentionally empty
        res.steps = {blocked, runaway};
        Writer w(dir + "/test_write_depinning" + ext);
        write_depinning(w, cfg, res);
    }
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s <output dir>\n", argv[0]);
        return 2;
    }
    try {
        write_all<KvWriter>(argv[1], ".kv");
        write_all<H5Writer>(argv[1], ".h5");
        // A second dataset under an existing key is an error, not a silent
        // overwrite (HighFive throws on it too).
        bool threw = false;
        try {
            H5Writer w(std::string(argv[1]) + "/duplicate.h5");
            w.put("a/b", 1.0);
            w.put("a/b", 2.0);
        } catch (const std::runtime_error &) {
            threw = true;
        }
        if (!threw) {
            std::printf("FAIL duplicate key accepted\n");
            return 1;
        }
        // The dispatcher: .kv selects the text archive, anything else HDF5.
        write_output(std::string(argv[1]) + "/dispatch.kv", [](auto &w) { w.put("x", 3.0); });
        write_output(std::string(argv[1]) + "/dispatch.h5", [](auto &w) { w.put("x", 4.0); });
    } catch (const std::exception &e) {
        std::printf("FAIL exception: %s\n", e.what());
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
