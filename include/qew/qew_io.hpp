// Output of the qew tools. The dataset keys are the reference's HDF5 key
// contract with its analysis scripts (src/qew_io.h:24-63, formatted like
// src/qew_format.h:13-17). Two containers take the same keys:
//  * H5Writer: an HDF5 file, laid out as the reference's HighFive writer lays
//    it out (one FP64 dataset per key, intermediate groups created from the
//    '/'-separated key, a scalar dataspace for single values and a 1-D one for
//    profiles). The HDF5 C library is loaded at run time with dlopen (path:
//    $QEW_HDF5_LIBRARY, else the one the tools were configured with), so the
//    tools start and write text archives on hosts without it.
//  * KvWriter: a plain-text key/value archive ("qew-kv 1": one dataset per
//    line, `key<TAB>count<TAB>values`, values as %.17g).
// qew_amd.io.read_output loads either into a dict.
#ifndef QEW_IO_HPP
#define QEW_IO_HPP

#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#if __has_include(<hdf5.h>)
#include <hdf5.h>
#define QEW_IO_HAVE_HDF5_HEADER 1
#endif

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

#ifdef QEW_IO_HAVE_HDF5_HEADER
#ifndef QEW_HDF5_LIBRARY
#define QEW_HDF5_LIBRARY "libhdf5.so"
#endif

// The HDF5 C entry points the writer uses, resolved once per process.
class Hdf5Api {
public:
    decltype(&H5open) open = nullptr;
    decltype(&H5Eset_auto2) eset_auto = nullptr;
    decltype(&H5Fcreate) fcreate = nullptr;
    decltype(&H5Fclose) fclose = nullptr;
    decltype(&H5Screate) screate = nullptr;
    decltype(&H5Screate_simple) screate_simple = nullptr;
    decltype(&H5Sclose) sclose = nullptr;
    decltype(&H5Pcreate) pcreate = nullptr;
    decltype(&H5Pset_create_intermediate_group) pset_intermediate = nullptr;
    decltype(&H5Pclose) pclose = nullptr;
    decltype(&H5Dcreate2) dcreate = nullptr;
    decltype(&H5Dwrite) dwrite = nullptr;
    decltype(&H5Dclose) dclose = nullptr;
    hid_t native_double = -1;  // H5T_NATIVE_DOUBLE
    hid_t link_create = -1;    // H5P_LINK_CREATE

    // The loaded API, or a runtime_error naming the library that could not be
    // loaded. The handle stays open until the process ends.
    static const Hdf5Api &get() {
        static const std::unique_ptr<Hdf5Api> api = load();
        if (!api->fcreate)
            throw std::runtime_error("HDF5 output needs the HDF5 C library (" + api->error +
                                     "); set QEW_HDF5_LIBRARY or write a .kv archive");
        return *api;
    }

private:
    std::string error;

    static std::unique_ptr<Hdf5Api> load() {
        std::unique_ptr<Hdf5Api> a(new Hdf5Api);
        const char *env = std::getenv("QEW_HDF5_LIBRARY");
        const std::string path = env && *env ? env : QEW_HDF5_LIBRARY;
        void *lib = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!lib) {
            a->error = "dlopen " + path + ": " + dlerror();
            return a;
        }
        bool ok = true;
        auto sym = [&](auto &fn, const char *name) {
            fn = reinterpret_cast<std::remove_reference_t<decltype(fn)>>(dlsym(lib, name));
            if (!fn) {
                ok = false;
                a->error = path + " lacks " + name;
            }
        };
        Hdf5Api t;
        sym(t.open, "H5open");
        sym(t.eset_auto, "H5Eset_auto2");
        sym(t.fcreate, "H5Fcreate");
        sym(t.fclose, "H5Fclose");
        sym(t.screate, "H5Screate");
        sym(t.screate_simple, "H5Screate_simple");
        sym(t.sclose, "H5Sclose");
        sym(t.pcreate, "H5Pcreate");
        sym(t.pset_intermediate, "H5Pset_create_intermediate_group");
        sym(t.pclose, "H5Pclose");
        sym(t.dcreate, "H5Dcreate2");
        sym(t.dwrite, "H5Dwrite");
        sym(t.dclose, "H5Dclose");
        const hid_t *nd = static_cast<const hid_t *>(dlsym(lib, "H5T_NATIVE_DOUBLE_g"));
        const hid_t *lc = static_cast<const hid_t *>(dlsym(lib, "H5P_CLS_LINK_CREATE_ID_g"));
        if (!ok || !nd || !lc || t.open() < 0) {
            if (ok) a->error = path + ": cannot initialise the library";
            return a;
        }
        t.eset_auto(H5E_DEFAULT, nullptr, nullptr);  // errors become exceptions here
        t.native_double = *nd;                       // valid once H5open has run
        t.link_create = *lc;
        *a = t;
        return a;
    }
};

class H5Writer {
public:
    // Creates (truncates) the file, like HighFive::File::Truncate in the
    // reference tools (executables/qew_static/main.cpp:94).
    explicit H5Writer(const std::string &path) : api_(Hdf5Api::get()), path_(path) {
        // 0x0002 is H5F_ACC_TRUNC, whose header macro would also call into the
        // library at link time.
        file_ = api_.fcreate(path.c_str(), 0x0002u, H5P_DEFAULT, H5P_DEFAULT);
        if (file_ < 0) throw std::runtime_error("cannot create HDF5 file " + path);
        lcpl_ = api_.pcreate(api_.link_create);
        if (lcpl_ < 0 || api_.pset_intermediate(lcpl_, 1) < 0) {
            api_.fclose(file_);
            throw std::runtime_error("HDF5: cannot set up link creation for " + path);
        }
    }
    ~H5Writer() {
        api_.pclose(lcpl_);
        api_.fclose(file_);
    }
    H5Writer(const H5Writer &) = delete;
    H5Writer &operator=(const H5Writer &) = delete;

    // A 1-D FP64 dataset of v.size() elements.
    void put(const std::string &key, const std::vector<double> &v) {
        const hsize_t dims[1] = {static_cast<hsize_t>(v.size())};
        write(key, api_.screate_simple(1, dims, nullptr), v.data());
    }
    // A scalar FP64 dataset.
    void put(const std::string &key, double v) { write(key, api_.screate(H5S_SCALAR), &v); }

private:
    const Hdf5Api &api_;
    std::string path_;
    hid_t file_ = -1;
    hid_t lcpl_ = -1;

    void write(const std::string &key, hid_t space, const double *data) {
        if (space < 0) throw std::runtime_error("HDF5: cannot create a dataspace for " + key);
        const hid_t ds = api_.dcreate(file_, key.c_str(), api_.native_double, space, lcpl_,
                                      H5P_DEFAULT, H5P_DEFAULT);
        herr_t rc = -1;
        if (ds >= 0) {
            rc = api_.dwrite(ds, api_.native_double, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
            api_.dclose(ds);
        }
        api_.sclose(space);
        if (rc < 0)
            throw std::runtime_error("HDF5: cannot write dataset '" + key + "' to " + path_);
    }
};
#endif  // QEW_IO_HAVE_HDF5_HEADER

// True when the tools write `path` as a qew-kv text archive (a .kv suffix)
// rather than HDF5.
inline bool is_kv_path(const std::string &path) {
    return path.size() >= 3 && path.compare(path.size() - 3, 3, ".kv") == 0;
}

// The writers below take either container (KvWriter or H5Writer).
template <class Writer>
inline void write_static(Writer &w, const StaticConfig &c,
                         const std::vector<StaticProfile> &profiles) {
    w.put("physical_size", c.Lx);
    for (const StaticProfile &prof : profiles) {
        const std::string prefix = "driving_force=" + fmt("%.1g", c.driving_force) +
                                   "/pinning_length=" + fmt("%.3g", prof.pinning_length_over_xi);
        w.put(prefix + "/pinning_length", prof.pinning_length_over_xi);
        w.put(prefix + "/h", prof.h);
    }
}

template <class Writer>
inline void write_dynamic(Writer &w, const DynamicConfig &c, const DynamicResult &res) {
    w.put("physical_size", c.Lx);
    for (const DynamicSnapshot &snap : res.snapshots) {
        const std::string prefix = "timestep=" + fmt("%.6g", snap.t) + "/driving_force=" +
                                   fmt("%.1g", c.driving_force) + "/line_tension=" +
                                   fmt("%.3g", res.line_tension);
        w.put(prefix + "/pinning_length", c.pinning_length_over_xi);
        w.put(prefix + "/h", snap.h);
    }
}

template <class Writer>
inline void write_depinning(Writer &w, const DepinningConfig &c, const DepinningResult &res) {
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

// Writes with `write(writer)` into `path`: a qew-kv archive for a .kv path,
// HDF5 otherwise (a runtime_error when HDF5 is unavailable).
template <class Fn>
inline void write_output(const std::string &path, Fn &&write) {
    if (is_kv_path(path)) {
        KvWriter w(path);
        write(w);
        return;
    }
#ifdef QEW_IO_HAVE_HDF5_HEADER
    H5Writer w(path);
    write(w);
#else
    throw std::runtime_error("this build has no HDF5 support; write a .kv archive instead of " +
                             path);
#endif
}

}  // namespace qew

#endif  // QEW_IO_HPP
