// Command-line options of the qew tools.
//
// Interface kept from the reference's tools (executables/cli.h:28-128 and the
// three main.cpp): options are spelled `--name value` or `--name=value`, every
// option takes a value (the token after the name is taken as it stands, so
// negative numbers are values), `--help` / `-h` is the only flag, lists are
// comma-separated numbers, and a malformed value or an unknown name is an
// error that the tools turn into exit code 2. Options starting with
// `--kokkos-` are dropped, as Kokkos::initialize consumes them there.
//
// The implementation is table driven: each tool declares one OptionSpec row
// per option (name, value kind, default, allowed words, help), and that table
// produces the parsing, the checks of every value, the rejection of names no
// row declares and the --help text. Args gives the same scanner and
// conversions with defaults supplied at each read, the reference's query
// interface, which its parser tests use.
#ifndef QEW_CLI_HPP
#define QEW_CLI_HPP

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace qew::cli {

// ---- scanning -------------------------------------------------------------

// argv split into (name, text) settings, in command-line order.
struct Scan {
    bool help = false;
    std::vector<std::pair<std::string, std::string>> settings;

    // The text of the last setting of `name`, or null.
    const std::string *last(const std::string &name) const {
        for (auto it = settings.rbegin(); it != settings.rend(); ++it)
            if (it->first == name) return &it->second;
        return nullptr;
    }
};

inline bool starts_with(const std::string &s, const char *prefix) {
    return s.compare(0, std::char_traits<char>::length(prefix), prefix) == 0;
}

inline Scan scan(int argc, char **argv) {
    Scan out;
    int k = 1;
    while (k < argc) {
        const std::string word = argv[k++];
        if (word == "-h" || word == "--help") {
            out.help = true;
            continue;
        }
        if (!starts_with(word, "--"))
            throw std::runtime_error("'" + word + "' is not an option (options are --name value)");
        const std::string body = word.substr(2);
        const size_t eq = body.find('=');
        std::string name = body.substr(0, eq);
        if (starts_with(name, "kokkos-")) {
            // --kokkos-x=v, --kokkos-x v or the bare flag --kokkos-x
            if (eq == std::string::npos && k < argc && !starts_with(argv[k], "--")) ++k;
            continue;
        }
        if (name.empty()) throw std::runtime_error("'" + word + "' has no option name after --");
        std::string text;
        if (eq != std::string::npos) {
            text = body.substr(eq + 1);
        } else {
            if (k == argc) throw std::runtime_error("--" + name + " is missing its value");
            text = argv[k++];
        }
        out.settings.emplace_back(std::move(name), std::move(text));
    }
    return out;
}

// ---- value conversions ------------------------------------------------------

[[noreturn]] inline void bad_value(const std::string &name, const std::string &text,
                                   const char *wanted) {
    throw std::runtime_error("--" + name + " wants " + wanted + ", not '" + text + "'");
}

inline double to_real(const std::string &name, const std::string &text) {
    const char *s = text.c_str();
    char *end = nullptr;
    errno = 0;
    const double v = std::strtod(s, &end);
    if (text.empty() || end != s + text.size() || errno == ERANGE || !std::isfinite(v))
        bad_value(name, text, "a finite real number");
    return v;
}

inline int to_int(const std::string &name, const std::string &text) {
    const char *s = text.c_str();
    char *end = nullptr;
    errno = 0;
    const long long v = std::strtoll(s, &end, 10);
    if (text.empty() || end != s + text.size() || errno == ERANGE || v < INT32_MIN ||
        v > INT32_MAX)
        bad_value(name, text, "an integer");
    return static_cast<int>(v);
}

inline std::uint64_t to_u64(const std::string &name, const std::string &text) {
    const char *s = text.c_str();
    char *end = nullptr;
    errno = 0;
    const unsigned long long v = std::strtoull(s, &end, 10);
    if (text.empty() || text[0] == '-' || end != s + text.size() || errno == ERANGE)
        bad_value(name, text, "a non-negative integer");
    return static_cast<std::uint64_t>(v);
}

// "a,b,c": empty items (",,", a trailing comma) are skipped; no item at all
// is an error.
inline std::vector<double> to_reals(const std::string &name, const std::string &text) {
    std::vector<double> out;
    size_t from = 0;
    while (from <= text.size()) {
        size_t to = text.find(',', from);
        if (to == std::string::npos) to = text.size();
        if (to > from) out.push_back(to_real(name, text.substr(from, to - from)));
        from = to + 1;
    }
    if (out.empty()) bad_value(name, text, "a comma-separated list of numbers");
    return out;
}

// ---- table-driven options (the tools) --------------------------------------

enum class Kind { Integer, Seed, Real, Reals, Word };

struct OptionSpec {
    std::string name;
    Kind kind;
    std::string fallback;  // default as text; empty: the tool derives it (see help)
    std::string help;
    std::string words;     // Kind::Word: the allowed values, '|'-separated; empty: any
};

class Options {
public:
    Options(std::vector<OptionSpec> table, int argc, char **argv)
        : table_(std::move(table)), scan_(scan(argc, argv)) {
        std::string unknown;
        for (const auto &s : scan_.settings)
            if (!find_row(s.first) && unknown.find("--" + s.first + " ") == std::string::npos)
                unknown += "--" + s.first + " ";
        if (!unknown.empty()) {
            unknown.pop_back();
            throw std::runtime_error("no such option: " + unknown + " (--help lists them)");
        }
        // every given value is checked now, read or not
        for (const auto &row : table_)
            if (scan_.last(row.name)) check(row);
    }

    bool help() const { return scan_.help; }
    bool given(const std::string &name) const { return scan_.last(row(name).name) != nullptr; }

    int integer(const std::string &name) const { return to_int(name, text(row(name), Kind::Integer)); }
    std::uint64_t seed(const std::string &name) const { return to_u64(name, text(row(name), Kind::Seed)); }
    double real(const std::string &name) const { return to_real(name, text(row(name), Kind::Real)); }
    std::vector<double> reals(const std::string &name) const {
        return to_reals(name, text(row(name), Kind::Reals));
    }
    std::string word(const std::string &name) const { return text(row(name), Kind::Word); }

    // --help text: one line per row, in table order.
    std::string usage(const std::string &tool, const std::string &summary) const {
        std::string u = tool + " -- " + summary + "\n\nOptions (--name value or --name=value):\n";
        for (const auto &r : table_) {
            std::string left = "  --" + r.name + " " + placeholder(r);
            if (left.size() < 30) left.resize(30, ' ');
            else left += "\n" + std::string(30, ' ');
            u += left + r.help;
            if (!r.fallback.empty()) u += " [" + r.fallback + "]";
            u += "\n";
        }
        return u;
    }

private:
    const OptionSpec *find_row(const std::string &name) const {
        for (const auto &r : table_)
            if (r.name == name) return &r;
        return nullptr;
    }
    const OptionSpec &row(const std::string &name) const {
        const OptionSpec *r = find_row(name);
        if (!r) throw std::logic_error("option table has no row --" + name);
        return *r;
    }
    std::string text(const OptionSpec &r, Kind want) const {
        if (r.kind != want) throw std::logic_error("--" + r.name + " is read as another kind");
        const std::string *t = scan_.last(r.name);
        if (t) return *t;
        if (r.fallback.empty())
            throw std::logic_error("--" + r.name + " has no default: check given() first");
        return r.fallback;
    }
    static std::string placeholder(const OptionSpec &r) {
        switch (r.kind) {
            case Kind::Integer: return "<int>";
            case Kind::Seed: return "<uint>";
            case Kind::Real: return "<x>";
            case Kind::Reals: return "<x,x,...>";
            case Kind::Word: return r.words.empty() ? "<text>" : r.words;
        }
        return "";
    }
    void check(const OptionSpec &r) const {
        const std::string &t = *scan_.last(r.name);
        switch (r.kind) {
            case Kind::Integer: (void)to_int(r.name, t); break;
            case Kind::Seed: (void)to_u64(r.name, t); break;
            case Kind::Real: (void)to_real(r.name, t); break;
            case Kind::Reals: (void)to_reals(r.name, t); break;
            case Kind::Word:
                if (!r.words.empty() && ("|" + r.words + "|").find("|" + t + "|") == std::string::npos)
                    bad_value(r.name, t, ("one of " + r.words).c_str());
                break;
        }
    }

    std::vector<OptionSpec> table_;
    Scan scan_;
};

// ---- query interface (defaults at each read) --------------------------------

class Args {
public:
    Args(int argc, char **argv) : scan_(scan(argc, argv)), read_(scan_.settings.size(), false) {
        help = scan_.help;
    }

    bool help = false;

    bool has(const std::string &k) const { return scan_.last(k) != nullptr; }
    std::string str(const std::string &k, const std::string &def) {
        const std::string *t = take(k);
        return t ? *t : def;
    }
    double dbl(const std::string &k, double def) {
        const std::string *t = take(k);
        return t ? to_real(k, *t) : def;
    }
    int integer(const std::string &k, int def) {
        const std::string *t = take(k);
        return t ? to_int(k, *t) : def;
    }
    std::uint64_t u64(const std::string &k, std::uint64_t def) {
        const std::string *t = take(k);
        return t ? to_u64(k, *t) : def;
    }
    std::vector<double> dbl_list(const std::string &k, const std::vector<double> &def) {
        const std::string *t = take(k);
        return t ? to_reals(k, *t) : def;
    }
    // Options given but never read are names the caller does not know.
    void require_all_used() const {
        std::string unread;
        for (size_t s = 0; s < read_.size(); ++s)
            if (!read_[s]) unread += " --" + scan_.settings[s].first;
        if (!unread.empty()) throw std::runtime_error("no such option:" + unread);
    }

private:
    const std::string *take(const std::string &k) {
        const std::string *t = nullptr;
        for (size_t s = 0; s < read_.size(); ++s)
            if (scan_.settings[s].first == k) {
                read_[s] = true;
                t = &scan_.settings[s].second;
            }
        return t;
    }

    Scan scan_;
    std::vector<bool> read_;
};

}  // namespace qew::cli

#endif  // QEW_CLI_HPP
