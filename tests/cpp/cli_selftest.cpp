// Host-only checks of include/qew/cli.hpp, restating the reference's parser
// tests (tests/test_cli.cpp:24-82). Prints one line per failed check and
// exits non-zero if any failed.
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "qew/cli.hpp"

using qew::cli::Args;

static int failures = 0;
#define CHECK(cond)                                                   \
    do {                                                              \
        if (!(cond)) {                                                \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                               \
        }                                                             \
    } while (0)

template <class F>
static bool throws(F f) {
    try {
        f();
    } catch (const std::runtime_error &) {
        return true;
    }
    return false;
}

static Args make(std::vector<const char *> v) {
    v.insert(v.begin(), "prog");
    return Args(static_cast<int>(v.size()), const_cast<char **>(v.data()));
}

int main() {
    {  // defaults kept when absent
        Args a = make({});
        CHECK(!a.help);
        CHECK(a.integer("nx", 4096) == 4096);
        CHECK(a.dbl("ftol", 1e-6) == 1e-6);
        CHECK(a.u64("seed", 1) == 1u);
        CHECK(a.str("model", "arclength") == "arclength");
        const std::vector<double> def = {0.01, 0.1, 1.0};
        CHECK(a.dbl_list("pinning_lengths", def) == def);
        CHECK(!throws([&] { a.require_all_used(); }));
    }
    {  // both spellings
        Args a = make({"--nx", "8192", "--ftol=1e-8", "--model", "linear"});
        CHECK(a.integer("nx", 4096) == 8192);
        CHECK(a.dbl("ftol", 1e-6) == 1e-8);
        CHECK(a.str("model", "arclength") == "linear");
        CHECK(!throws([&] { a.require_all_used(); }));
    }
    {  // negative values are values
        Args a = make({"--driving_force", "-1.5"});
        CHECK(a.dbl("driving_force", 0.0) == -1.5);
    }
    {  // comma list
        Args a = make({"--pinning_lengths", "0.1,1,10,100"});
        const std::vector<double> got = a.dbl_list("pinning_lengths", {});
        CHECK(got.size() == 4 && got[0] == 0.1 && got[3] == 100.0);
    }
    CHECK(make({"--help"}).help);
    CHECK(make({"-h"}).help);
    {  // unknown option rejected
        Args a = make({"--bogus", "3"});
        a.integer("nx", 1);
        CHECK(throws([&] { a.require_all_used(); }));
    }
    {  // malformed numbers
        Args a = make({"--nx", "abc"});
        CHECK(throws([&] { a.integer("nx", 1); }));
        Args b = make({"--ftol", "1.0x"});
        CHECK(throws([&] { b.dbl("ftol", 1.0); }));
    }
    CHECK(throws([] { make({"--nx"}); }));          // missing value
    CHECK(throws([] { make({"positional"}); }));    // non-option token
    {  // --kokkos-* options are consumed, like Kokkos::initialize does
        Args a = make({"--kokkos-num-threads=4", "--kokkos-disable-warnings", "--nx", "16"});
        CHECK(a.integer("nx", 1) == 16);
        CHECK(!throws([&] { a.require_all_used(); }));
    }
    std::printf("%d failures\n", failures);
    return failures ? 1 : 0;
}
