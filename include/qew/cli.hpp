// `--key value` / `--key=value` command-line parser for the qew tools, with
// the conventions of the reference's executables/cli.h:28-128: every option
// takes a value (values may be negative numbers), --help / -h is the only
// flag, defaults are given at each read site, comma-separated lists of
// doubles, malformed numbers and unknown keys are errors (std::runtime_error,
// exit code 2 in the tools). Options starting with --kokkos- are accepted and
// ignored, as Kokkos::initialize consumes them in the reference.
#ifndef QEW_CLI_HPP
#define QEW_CLI_HPP

#include <cstdint>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace qew::cli {

class Args {
public:
    Args(int argc, char **argv) {
        for (int i = 1; i < argc; ++i) {
            const std::string tok = argv[i];
            if (tok == "--help" || tok == "-h") {
                help = true;
                continue;
            }
            if (tok.rfind("--", 0) != 0)
                throw std::runtime_error("unexpected argument '" + tok +
                                         "' (options start with --)");
            std::string key = tok.substr(2), val;
            const auto eq = key.find('=');
            const bool kokkos = key.rfind("kokkos-", 0) == 0;
            if (eq != std::string::npos) {
                val = key.substr(eq + 1);
                key = key.substr(0, eq);
            } else if (kokkos) {
                // Kokkos flags may come without a value
                if (i + 1 < argc && std::string(argv[i + 1]).rfind("--", 0) != 0) ++i;
                continue;
            } else if (i + 1 < argc) {
                val = argv[++i];
            } else {
                throw std::runtime_error("option --" + key + " needs a value");
            }
            if (key.empty()) throw std::runtime_error("empty option name in '" + tok + "'");
            if (kokkos) continue;
            kv_[key] = val;
        }
    }

    bool help = false;

    bool has(const std::string &k) const { return kv_.count(k) != 0; }

    std::string str(const std::string &k, const std::string &def) {
        const auto it = kv_.find(k);
        if (it == kv_.end()) return def;
        used_.insert(k);
        return it->second;
    }
    double dbl(const std::string &k, double def) { return num<double>(k, def); }
    int integer(const std::string &k, int def) { return num<int>(k, def); }
    std::uint64_t u64(const std::string &k, std::uint64_t def) { return num<std::uint64_t>(k, def); }

    std::vector<double> dbl_list(const std::string &k, const std::vector<double> &def) {
        const auto it = kv_.find(k);
        if (it == kv_.end()) return def;
        used_.insert(k);
        std::vector<double> out;
        std::stringstream ss(it->second);
        std::string item;
        while (std::getline(ss, item, ','))
            if (!item.empty()) out.push_back(parse<double>(k, item));
        if (out.empty())
            throw std::runtime_error("option --" + k + ": empty list '" + it->second + "'");
        return out;
    }

    void require_all_used() const {
        std::string unknown;
        for (const auto &kv : kv_)
            if (!used_.count(kv.first)) unknown += (unknown.empty() ? "" : ", ") + ("--" + kv.first);
        if (!unknown.empty()) throw std::runtime_error("unknown option(s): " + unknown);
    }

private:
    template <class T>
    T num(const std::string &k, T def) {
        const auto it = kv_.find(k);
        if (it == kv_.end()) return def;
        used_.insert(k);
        return parse<T>(k, it->second);
    }
    template <class T>
    static T parse(const std::string &k, const std::string &v) {
        std::istringstream ss(v);
        T out{};
        ss >> out;
        if (ss.fail() || ss.peek() != std::char_traits<char>::eof())
            throw std::runtime_error("option --" + k + ": cannot parse '" + v + "' as a number");
        return out;
    }

    std::map<std::string, std::string> kv_;
    std::set<std::string> used_;
};

}  // namespace qew::cli

#endif  // QEW_CLI_HPP
