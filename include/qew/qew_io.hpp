// Output of the qew tools. The dataset keys are the reference's HDF5 key
// contract with its analysis scripts (src/qew_io.h:24-63, formatted like
// src/qew_format.h:13-17); they are written into a plain-text key/value
// archive ("qew-kv 1": one dataset per line, `key<TAB>count<TAB>values`,
// values as %.17g), which qew_amd.io.read_kv loads into a dict. HDF5 itself is
// not linked.
#ifndef QEW_IO_HPP
#define QEW_IO_HPP

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "qew_runs.hpp"

namespace qew {

inline std::string fmt(const char *spec, double v) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), spec, v);
    return std::string(buf);
}

class KvWriter {
public:
    explicit KvWriter(const std::string &path) : f_(std::fopen(path.c_str(), "w")) {
        if (!f_) throw std::runtime_error("cannot open " + path + " for writing");
        std::fputs("# qew-kv 1\n", f_);
    }
    ~KvWriter() {
        if (f_) std::fclose(f_);
    }
    KvWriter(const KvWriter &) = delete;
    KvWriter &operator=(const KvWriter &) = delete;

    void put(const std::string &key, const std::vector<double> &v) {
        std::fprintf(f_, "%s\t%zu\t", key.c_str(), v.size());
        for (size_t i = 0; i < v.size(); ++i) std::fprintf(f_, i ? " %.17g" : "%.17g", v[i]);
        std::fputc('\n', f_);
    }
    void put(const std::string &key, double v) { put(key, std::vector<double>{v}); }

private:
    std::FILE *f_;
};

inline void write_static(KvWriter &w, const StaticConfig &c,
                         const std::vector<StaticProfile> &profiles) {
    w.put("physical_size", c.Lx);
    for (const StaticProfile &prof : profiles) {
        const std::string prefix = "driving_force=" + fmt("%.1g", c.driving_force) +
                                   "/pinning_length=" + fmt("%.3g", prof.pinning_length_over_xi);
        w.put(prefix + "/pinning_length", prof.pinning_length_over_xi);
        w.put(prefix + "/h", prof.h);
    }
}

inline void write_dynamic(KvWriter &w, const DynamicConfig &c, const DynamicResult &res) {
    w.put("physical_size", c.Lx);
    for (const DynamicSnapshot &snap : res.snapshots) {
        const std::string prefix = "timestep=" + fmt("%.6g", snap.t) + "/driving_force=" +
                                   fmt("%.1g", c.driving_force) + "/line_tension=" +
                                   fmt("%.3g", res.line_tension);
        w.put(prefix + "/pinning_length", c.pinning_length_over_xi);
        w.put(prefix + "/h", snap.h);
    }
}

inline void write_depinning(KvWriter &w, const DepinningConfig &c, const DepinningResult &res) {
    w.put("physical_size", c.Lx);
    for (const DepinningStep &st : res.steps) {
        if (!st.result.blocked) continue;
        const std::string prefix = "driving_force=" + fmt("%.4g", st.f);
        w.put(prefix + "/h", st.h);
        w.put(prefix + "/pinning_length", c.pinning_length_over_xi);
    }
    if (res.f_c_upper_bound >= 0) w.put("f_c_upper_bound", res.f_c_upper_bound);
    if (res.f_c >= 0) {
        // bisection results (no reference counterpart)
        w.put("f_c", res.f_c);
        w.put("f_c_lower", res.f_c_lower);
        w.put("f_c_upper", res.f_c_upper);
        w.put("critical/h", res.h_critical);
    }
}

}  // namespace qew

#endif  // QEW_IO_HPP
