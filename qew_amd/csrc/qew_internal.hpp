// Host-side internals shared by the library's translation units: the opaque
// handle types behind the C ABI, error plumbing and the stream-ordered scratch
// buffers a context keeps for single-line calls.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/qew_amd.h"
#include "qew_core.hpp"

namespace qew_host {

// Thrown inside the library, turned into a qew_status at the ABI boundary.
struct Error : std::runtime_error {
    qew_status code;
    Error(qew_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

#define QEW_HIP(expr)                                                                  \
    do {                                                                               \
        hipError_t qew_e_ = (expr);                                                    \
        if (qew_e_ != hipSuccess)                                                      \
            throw ::qew_host::Error(                                                   \
                qew_e_ == hipErrorOutOfMemory ? QEW_ERR_OUT_OF_MEMORY : QEW_ERR_HIP,   \
                std::string(#expr) + ": " + hipGetErrorString(qew_e_));               \
    } while (0)

#define QEW_REQUIRE(cond, msg)                                                         \
    do {                                                                               \
        if (!(cond)) throw ::qew_host::Error(QEW_ERR_INVALID_ARGUMENT, (msg));        \
    } while (0)

// Runs f and maps exceptions to a status.
template <class F>
qew_status guard(F &&f) {
    try {
        f();
        return QEW_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return QEW_ERR_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return QEW_ERR_HIP;
    }
}

// A device buffer that grows on demand (never shrinks) and frees on destroy.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    void ensure(size_t need) {
        if (need <= bytes) return;
        if (ptr) QEW_HIP(hipFree(ptr));
        ptr = nullptr;
        bytes = 0;
        QEW_HIP(hipMalloc(&ptr, need));
        bytes = need;
    }
    template <class T>
    T *as() const { return static_cast<T *>(ptr); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

// Two timing events, destroyed on every exit path (launch-plan timing).
struct EventPair {
    hipEvent_t start = nullptr, stop = nullptr;
    EventPair() {
        QEW_HIP(hipEventCreate(&start));
        hipError_t e = hipEventCreate(&stop);
        if (e != hipSuccess) {
            (void)hipEventDestroy(start);
            QEW_HIP(e);
        }
    }
    ~EventPair() {
        (void)hipEventDestroy(start);
        (void)hipEventDestroy(stop);
    }
    EventPair(const EventPair &) = delete;
    EventPair &operator=(const EventPair &) = delete;
    float elapsed_ms() const {
        float ms = 0;
        QEW_HIP(hipEventElapsedTime(&ms, start, stop));
        return ms;
    }
};

// Pinned host staging for small results (host scalars returned by
// objective(), dot products in the host-driven L-BFGS, ...).
struct PinnedBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    void ensure(size_t need) {
        if (need <= bytes) return;
        if (ptr) QEW_HIP(hipHostFree(ptr));
        ptr = nullptr;
        QEW_HIP(hipHostMalloc(&ptr, need, hipHostMallocDefault));
        bytes = need;
    }
    template <class T>
    T *as() const { return static_cast<T *>(ptr); }
    void release() {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

}  // namespace qew_host

struct qew_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // scratch for single-line calls: per-tile energy partials, the energy
    // result slot and solver work buffers
    qew_host::DevBuf partial, result, work;
    qew_host::PinnedBuf pinned;
    int num_cus = 0;
};

struct qew_noise_s {
    qew_ctx ctx = nullptr;
    qew_dev::NoiseDesc desc{};
    void *table = nullptr;  // device table, see NoiseDesc for the layout
    std::vector<double> field;         // noise_on_grid (generated / from-field tables)
};

struct qew_batch_s {
    qew_ctx ctx = nullptr;
    int nlines = 0;
    int64_t total = 0;
    int ntiles = 0;
    int max_n = 0;
    std::vector<qew_dev::LineDesc> lines;
    std::vector<qew_dev::NoiseDesc> noises;
    std::vector<int64_t> offsets;
    qew_host::DevBuf d_lines, d_noises, d_tile_line, d_partial;
    qew_host::DevBuf d_scratch;  // solver workspaces (grown on demand)
    // multi-step Langevin tiling: (line, first output site) per tile of
    // ltile_out output sites, built on first use; lv_k = launch plan picked by
    // timing (K > 0: K steps per launch on workgroup tiles, -K: on wave
    // tiles; 0: not yet tuned); d_tune = line-state copies the timing runs on
    int ltiles = 0, ltile_out = 0, lv_k = 0;
    // Rosso-Krauth path picked by timing: 0 not yet, 1 LDS line kernel,
    // 2 per-colour launches
    int rk_path = 0;
    qew_host::DevBuf d_ltile, d_tune;
    // persistent workgroups of the last qew_batch_lbfgs launch and the
    // workgroups it gave each line
    int lbfgs_workgroups = 0;
    std::vector<int> lbfgs_last_team;
    // workgroups per line picked by timing, per (FP32 state, preconditioner)
    struct TunedTeams {
        bool tuned = false;
        std::vector<int> team;
    } lbfgs_teams[4];
};

namespace qew_host {

inline void ctx_check(qew_ctx ctx) { QEW_REQUIRE(ctx != nullptr, "null context"); }

inline void use_device(qew_ctx ctx) { QEW_HIP(hipSetDevice(ctx->device)); }

// BF16 tables and midpoint-contracted FP32 tables are read only by the FP32
// batch evaluation and dynamics; every entry point that reads tables as
// per-cell FP64 or FP32 coefficients rejects them before launching anything.
inline void require_wide_table(const qew_dev::NoiseDesc &nd) {
    QEW_REQUIRE(nd.kind != qew_dev::TABLE ||
                    (nd.dtype != QEW_BF16 && nd.dtype != QEW_F32_MID && nd.dtype != QEW_F64_MID),
                "BF16 and midpoint-contracted tables are read only by the batch entry points "
                "that document them");
}
inline void require_wide_tables(const std::vector<qew_dev::NoiseDesc> &v) {
    for (const auto &nd : v) require_wide_table(nd);
}

// The tables an FP32-state batch call reads: FP32, BF16 or QEW_F32_MID, the
// last only for lines whose edge midpoints all lie at x-fraction 1/2 (site
// spacing L / n equal to the table's dx, so the FP32 locate's x scale is
// exactly 1/2, and n even, so the wrap edge's midpoint (n - 1) / 2 is a
// half-integer too).
inline void require_f32_tables(const std::vector<qew_dev::LineDesc> &lines,
                               const std::vector<qew_dev::NoiseDesc> &noises, const char *what) {
    for (const auto &nd : noises)
        QEW_REQUIRE(nd.kind == qew_dev::TABLE &&
                        (nd.dtype == QEW_F32 || nd.dtype == QEW_BF16 || nd.dtype == QEW_F32_MID),
                    std::string(what) + " needs FP32, BF16 or QEW_F32_MID coefficient tables");
    for (const auto &l : lines) {
        const qew_dev::NoiseDesc &nd = noises[l.noise];
        if (nd.dtype == QEW_F32_MID)
            QEW_REQUIRE(l.n % 2 == 0 && l.L / (double)l.n == nd.dx,
                        "a QEW_F32_MID table needs lines of even length whose site spacing "
                        "L / n equals the table's x spacing");
    }
}

// Makes sure the context's scratch covers the energy partials of `ntiles`
// tiles and the results of `nlines` lines.
void ctx_reserve_eval(qew_ctx ctx, int ntiles, int nlines);

// Resolves, once per device and with that device current, the address of the
// scratch array qew_eval_noalloc uses (called by qew_ctx_create).
void noalloc_scratch_init(int device);

// Single-line descriptor helper.
qew_dev::LineDesc single_line(int n, const qew_params *p);

// ---- kernels exposed to other translation units ----
// Fused gradient/energy over `nlines` lines. lines/noises: device arrays
// (batch) or, when single != nullptr, the one line passed by value.
struct EvalArgs {
    const qew_dev::LineDesc *lines;
    const qew_dev::NoiseDesc *noises;
    const int *tile_line;
    int ntiles;
    const void *h;
    const void *d;      // trial direction (objective at h + alpha d), or null
    double alpha;
    const double *anchor;
    void *grad;         // null: energy only
    double *energy;     // device, per line; null: gradient only
    void *partial;
    // FP64 single-line gradient only: per-edge dV/dy and (arclength) slope
    // written out when non-null
    double *edge_dvdy;
    double *edge_slope;
};

void launch_eval_f64(hipStream_t s, const EvalArgs &a, const qew_dev::LineDesc *one_line,
                     const qew_dev::NoiseDesc *one_noise);
void launch_eval_f32(hipStream_t s, const EvalArgs &a, int nlines);
// Sums the tile partials of an energy evaluation into energy[] (one line
// passed by value, or nlines batch lines).
void finish_energy_f64(hipStream_t s, const EvalArgs &a, const qew_dev::LineDesc *one_line,
                       int nlines);

// Global-memory Stockham FFT (batched 1-D complex transforms of FP64 pairs).
// data[b * bstride + k * stride] for b < batch, k < n; sign -1 forward,
// +1 backward (unnormalised). tmp must hold as many elements as data spans.
void fft_c2c(hipStream_t s, double2 *data, double2 *tmp, int n, int batch, int64_t stride,
             int64_t bstride, int sign);

}  // namespace qew_host
