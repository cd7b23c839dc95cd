/*
 * qew_amd.h -- C ABI of the MI355X (gfx950) qEW engine.
 *
 * Plain C types only: device buffers are raw pointers the caller allocated
 * (qew_device_alloc, hipMalloc, or a torch tensor's data_ptr), sizes are ints.
 * Every entry point returns a qew_status; on failure qew_last_error() holds a
 * message for the calling thread. No C++ exception crosses this boundary.
 *
 * The reference has no FFI layer: its API is a set of C++ header templates on
 * Kokkos views, templated on a `Noise` functor. Each entry point below names
 * the reference function it replaces (file:line in IMTEK-Simulation/qEW).
 * The C++ facade in include/qew/ restores the reference's names and types on
 * top of this ABI.
 *
 * Stream semantics: calls are ordered on the context's stream. Entry points
 * that return a host scalar or a host result struct synchronise that stream;
 * the batched entry points that write device buffers do not.
 */
#ifndef QEW_AMD_H
#define QEW_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum qew_status {
    QEW_OK = 0,
    QEW_ERR_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference */
    QEW_ERR_HIP = 2,              /* a HIP runtime call failed */
    QEW_ERR_OUT_OF_MEMORY = 3,
    QEW_ERR_UNSUPPORTED = 4
} qew_status;

/* Arithmetic / storage type of line state and coefficient tables. QEW_BF16
 * and QEW_F32_MID are table storage types only:
 *  - QEW_BF16: bfloat16 coefficients, rounded to nearest even from FP64 via
 *    FP32;
 *  - QEW_F32_MID: one FP32 value per cell, T(ix, iy) = sum_a w_a c(ix-1+a, iy)
 *    with the cubic B-spline weights at x-fraction 1/2 (1, 23, 23, 1)/48,
 *    contracted in FP64 and rounded once. Every edge midpoint of a line whose
 *    site spacing equals the table's x spacing (and whose length is even)
 *    lies at x-fraction 1/2, so its 4 x 4 gather reduces to 4 values of T.
 * Such tables are read by qew_batch_eval and qew_batch_langevin with QEW_F32
 * line state, and QEW_F32_MID tables also by qew_batch_lbfgs with QEW_F32
 * and by qew_batch_rk_block (FP64 arithmetic on the FP32 values, as on FP32
 * tables) (QEW_F32_MID only for lines with that geometry). QEW_F64_MID is the
 * same contraction kept in FP64, read only by qew_batch_lbfgs with QEW_F64
 * line state (the reference's arithmetic up to the order of the x and y
 * sums). Every other entry point rejects these tables with
 * QEW_ERR_INVALID_ARGUMENT. */
typedef enum qew_dtype {
    QEW_F64 = 0,
    QEW_F32 = 1,
    QEW_BF16 = 2,
    QEW_F32_MID = 3,
    QEW_F64_MID = 4
} qew_dtype;

/* qew::Model (src/qew_model.h:18). */
typedef enum qew_model { QEW_LINEAR = 0, QEW_ARCLENGTH = 1 } qew_model;

/* Potential kinds. TABLE is FilteredNoise/DeviceNoise
 * (src/filtered_noise.h:64-176); the others restate the analytic functors the
 * reference tests inject through the Noise concept (src/qew_model.h:31-48):
 *   SINCOS    p = {amp, kx, ky}: V = amp sin(kx x) cos(ky y)
 *             (tests/test_qew_model.cpp:22-35, AnalyticNoise)
 *   CONST     p = {v0}          (tests/test_qew_model.cpp:37-42, ConstNoise)
 *   ZERO                        (tests/test_lbfgs.cpp:24-27, ZeroNoise)
 *   HARMONIC  p = {k, y0}: V = k (y - y0)^2 / 2
 *             (tests/test_lbfgs.cpp:29-38, HarmonicWell)
 *   WASHBOARD p = {A, k}:  V = A cos(k y)
 *             (tests/test_rosso_krauth.cpp:24-31, SinusoidNoise) */
typedef enum qew_noise_kind {
    QEW_NOISE_TABLE = 0,
    QEW_NOISE_SINCOS = 1,
    QEW_NOISE_CONST = 2,
    QEW_NOISE_ZERO = 3,
    QEW_NOISE_HARMONIC = 4,
    QEW_NOISE_WASHBOARD = 5
} qew_noise_kind;

typedef struct qew_ctx_s *qew_ctx;
typedef struct qew_noise_s *qew_noise;
typedef struct qew_batch_s *qew_batch;

/* qew::Params (src/qew_model.h:20-27). */
typedef struct qew_params {
    double physical_size; /* box length L_x; dx = L_x / n */
    double line_tension;
    double driving_force;
    int model;            /* qew_model */
} qew_params;

/* qew::LbfgsParams (src/lbfgs.h:32-52); same defaults via
 * qew_lbfgs_default_params. */
typedef struct qew_lbfgs_params {
    int m;
    double ftol;
    int max_iter;
    double c1;
    double ls_shrink;
    int max_ls;
    double curvature_eps;
    int precondition;
    double precond_shift;
} qew_lbfgs_params;

/* qew::LbfgsResult (src/lbfgs.h:54-61). */
typedef struct qew_lbfgs_result {
    int iterations;
    double max_force;
    int converged;
    long n_func_evals;
    long n_curvature_skips;
    long n_ls_restarts;
} qew_lbfgs_result;

/* qew::FireParams (src/fire.h:23-33); same defaults via
 * qew_fire_default_params, except that dt_max has none and must be > 0. */
typedef struct qew_fire_params {
    double dt_max;
    double dt_init;
    double ftol;
    int max_iter;
    int n_min;
    double f_inc;
    double f_dec;
    double alpha_start;
    double f_alpha;
} qew_fire_params;

/* qew::FireResult (src/fire.h:35-39). */
typedef struct qew_fire_result {
    int iterations;
    double max_force;
    int converged;
} qew_fire_result;

/* qew::RkParams / qew::RkResult (src/rosso_krauth.h:34-49). */
typedef struct qew_rk_params {
    double dstep;
    double max_advance;
    double root_tol;
    double sweep_tol;
    double runaway;
    int max_sweeps;
} qew_rk_params;

typedef struct qew_rk_result {
    int blocked;
    int runaway;
    int sweeps;
    double total_advance;
    double mean_h;
    /* rk_site_velocity calls the reference algorithm makes (no reference
     * field; sites the device skips because their height and both neighbours
     * are unchanged since they were last found blocked are counted) */
    long n_velocity_evals;
    /* site-velocity evaluations the device performed (skipped sites not
     * counted; every lane of a cooperative march or bisection counted) */
    long n_velocity_evals_done;
} qew_rk_result;

/* ---- errors, context, memory ------------------------------------------- */

const char *qew_last_error(void);
/* Version string of the library build (arch, build flags). */
const char *qew_version(void);

/* Context = device + HIP stream (NULL: the default stream) plus reusable
 * scratch. Replaces Kokkos::initialize / the default execution space
 * instance (README.md:145-147). */
qew_status qew_ctx_create(int device, void *hip_stream, qew_ctx *out);
qew_status qew_ctx_destroy(qew_ctx ctx);
qew_status qew_ctx_synchronize(qew_ctx ctx);

/* Device memory helpers for callers without their own allocator
 * (Kokkos::View allocation / deep_copy, src/qew_types.h:17-22). */
qew_status qew_device_alloc(qew_ctx ctx, size_t bytes, void **out);
qew_status qew_device_free(qew_ctx ctx, void *ptr);
qew_status qew_copy_to_device(qew_ctx ctx, void *dst, const void *src, size_t bytes);
qew_status qew_copy_to_host(qew_ctx ctx, void *dst, const void *src, size_t bytes);

/* ---- potentials (src/filtered_noise.h:183-191, filtered_noise.cpp:25-136) */

/* FilteredNoise(nx, ny, Lx, Ly, amplitude, xi_x, xi_y, seed)
 * (src/filtered_noise.cpp:25-82): random phases from mt19937_64 on the host,
 * band mask, inverse 2-D FFT, normalisation and the periodic cubic B-spline
 * prefilter on the device. `dtype` is the device table's storage type. */
qew_status qew_noise_generate(qew_ctx ctx, int nx, int ny, double Lx, double Ly,
                              double amplitude, double xi_x, double xi_y,
                              uint64_t seed, int dtype, qew_noise *out);
/* `count` FilteredNoise tables in one call (arrays of count entries): the
 * independent mt19937_64 phase streams are drawn on parallel host threads.
 * Each out[k] equals qew_noise_generate with the k-th arguments. */
qew_status qew_noise_generate_batch(qew_ctx ctx, int count, const int *nx, const int *ny,
                                    const double *Lx, const double *Ly,
                                    const double *amplitude, const double *xi_x,
                                    const double *xi_y, const uint64_t *seed, int dtype,
                                    qew_noise *out);
/* A copy of a potential with its table stored as `dtype` (FP64 -> FP32
 * rounds each coefficient once; QEW_F32_MID contracts the x-weights at
 * fraction 1/2 in FP64 and rounds each cell's value once; BF16 and
 * QEW_F32_MID tables are not converted further). No reference counterpart. */
qew_status qew_noise_convert(qew_noise noise, int dtype, qew_noise *out);
/* FilteredNoise(field, nx, ny, Lx, Ly) (src/filtered_noise.cpp:84-90): field
 * is host, row-major field[i*ny + j]. */
qew_status qew_noise_from_field(qew_ctx ctx, const double *field, int nx, int ny,
                                double Lx, double Ly, int dtype, qew_noise *out);
/* A table of already prefiltered coefficients, host row-major coef[i*ny + j]
 * (the DeviceNoise.coef view, src/filtered_noise.h:101-106). */
qew_status qew_noise_from_coefficients(qew_ctx ctx, const double *coef, int nx,
                                       int ny, double dx, double dy, int dtype,
                                       qew_noise *out);
/* Analytic potential (qew_noise_kind other than TABLE); p has 4 entries. */
qew_status qew_noise_analytic(qew_ctx ctx, int kind, const double *p, qew_noise *out);
qew_status qew_noise_destroy(qew_noise noise);
/* Table geometry; any pointer may be NULL. */
qew_status qew_noise_info(qew_noise noise, int *kind, int *dtype, int *nx, int *ny,
                          double *dx, double *dy);
/* Host copy of noise_on_grid (FilteredNoise::field(), src/filtered_noise.h:195)
 * for generated tables, row-major nx*ny. */
qew_status qew_noise_field(qew_noise noise, double *field_host);
/* Host copy of the coefficients, row-major coef[i*ny + j], as FP64 (for a
 * QEW_F32_MID table: its per-cell contracted values T(i, j)). */
qew_status qew_noise_coefficients(qew_noise noise, double *coef_host);
/* DeviceNoise::sample at m device points (src/filtered_noise.h:103-134);
 * x, y, v, dvdx, dvdy are device FP64 arrays. */
qew_status qew_noise_sample(qew_ctx ctx, qew_noise noise, int m, const double *x,
                            const double *y, double *v, double *dvdx, double *dvdy);

/* DeviceNoise::sample_value and DeviceNoise::sample_dy at m device points
 * (src/filtered_noise.h:138-176): the value-only and dy-only samplers, whose
 * arithmetic the objective and the gradient use. */
qew_status qew_noise_sample_value(qew_ctx ctx, qew_noise noise, int m, const double *x,
                                  const double *y, double *v);
qew_status qew_noise_sample_dy(qew_ctx ctx, qew_noise noise, int m, const double *x,
                               const double *y, double *dvdy);

/* ---- one line, FP64 (src/qew_model.h) ----------------------------------- */

/* objective(h, p, noise) (src/qew_model.h:101-106). h: device FP64[n]. */
qew_status qew_objective(qew_ctx ctx, qew_noise noise, const qew_params *p, int n,
                         const double *h, double *energy);
/* objective(x0, d, alpha, p, noise) (src/qew_model.h:108-119). */
qew_status qew_objective_trial(qew_ctx ctx, qew_noise noise, const qew_params *p,
                               int n, const double *x0, const double *d,
                               double alpha, double *energy);
/* gradient(h, p, noise, grad) (src/qew_model.h:149-201). grad: device FP64[n]. */
qew_status qew_gradient(qew_ctx ctx, qew_noise noise, const qew_params *p, int n,
                        const double *h, double *grad);
/* gradient(h, p, noise, grad, ws) (src/qew_model.h:126-201) with the
 * GradientWorkspace's per-edge quantities written out: edge_dvdy[i] = dV/dy at
 * the midpoint of edge (i, i+1 mod n), and for the arclength model
 * edge_slope[i] = s / sqrt(1 + s^2), s = (h[i+1] - h[i]) / dx (edge_slope may
 * be NULL; it is not written for the linear model). Device FP64[n] arrays. */
qew_status qew_gradient_edges(qew_ctx ctx, qew_noise noise, const qew_params *p, int n,
                              const double *h, double *grad, double *edge_dvdy,
                              double *edge_slope);
/* gradient and objective fused in one pass (no reference counterpart; both
 * results equal the two calls above). energy may be NULL. */
qew_status qew_eval(qew_ctx ctx, qew_noise noise, const qew_params *p, int n,
                    const double *h, double *grad, double *energy);

/* qew_eval without any allocation: no reference counterpart. For callers that
 * cannot allocate device or pinned host memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Gradient and energy
 * equal qew_eval's bit for bit (the same kernels); the energy is written to
 * the DEVICE pointer `energy` (FP64[1], or NULL for the gradient only) and the
 * call does not synchronise. The per-tile energy partials live in a
 * file-scope device array of 16384 doubles (128 KiB, held by every device
 * the library runs on until the process ends; its address is resolved per
 * device by qew_ctx_create), so n is limited to 16384 * 255 = 4177920 sites;
 * longer lines are rejected with QEW_ERR_INVALID_ARGUMENT. All calls on one
 * device share that array: calls that overlap on one device (other streams,
 * other threads) race and must be serialised by the caller. */
qew_status qew_eval_noalloc(qew_ctx ctx, qew_noise noise, const qew_params *p, int n,
                            const double *h, double *grad, double *energy);

/* ---- batches of independent lines --------------------------------------- */

/* A ragged batch: line l has lengths[l] sites, potential noises[l] and model
 * parameters params[l] (host arrays). Lines are stored back to back in device
 * buffers; qew_batch_offsets returns the nlines+1 prefix offsets. */
qew_status qew_batch_create(qew_ctx ctx, int nlines, const int *lengths,
                            const qew_noise *noises, const qew_params *params,
                            qew_batch *out);
qew_status qew_batch_destroy(qew_batch batch);
qew_status qew_batch_offsets(qew_batch batch, int64_t *offsets_host);
/* Per-line L-BFGS preconditioner shifts for qew_batch_lbfgs (host array of
 * nlines; a value > 0 replaces qew_lbfgs_params.precond_shift for its line,
 * e.g. the driver's 0.5 dx A / xi^2 when xi differs between lines). */
qew_status qew_batch_set_precond_shifts(qew_batch batch, const double *shifts);
/* Changes the driving force of every line (host array of nlines doubles). */
qew_status qew_batch_set_forces(qew_batch batch, const double *forces);

/* Gradient + energy of every line in one launch. dtype QEW_F64: h, grad are
 * FP64 and the result equals qew_eval line by line. dtype QEW_F32: h holds
 * FP32 offsets from a per-line anchor (device FP64[nlines], NULL = 0), the
 * tables must be FP32, grad is FP32, and the energy is accumulated with
 * compensated (Kahan/two-sum) FP32 sums. energy: device FP64[nlines] or NULL. */
qew_status qew_batch_eval(qew_batch batch, int dtype, const void *h,
                          const double *anchor, void *grad, double *energy);

/* ---- L-BFGS (src/lbfgs.h:94-287) ---------------------------------------- */

/* Reference defaults: m=8, ftol=1e-6, max_iter=10000, c1=1e-4, shrink 0.5,
 * max_ls=30, curvature_eps=1e-12, no preconditioner, shift 0 (auto). */
qew_lbfgs_params qew_lbfgs_default_params(void);
/* lbfgs_minimize(h, p, noise, lp): one line, FP64, iterations driven from the
 * host exactly as the reference does (the plain single-line GPU path). */
qew_status qew_lbfgs_minimize(qew_ctx ctx, qew_noise noise, const qew_params *p,
                              const qew_lbfgs_params *lp, int n, double *h,
                              qew_lbfgs_result *result);
/* Every line of the batch minimised by one persistent kernel (one workgroup
 * per line, device-side control flow, no host round trip per iteration).
 * h: device FP64 line state in/out. dtype QEW_F32 runs the line state and
 * gathers in FP32 (FP32 tables) with Kahan-compensated reductions and keeps
 * the Armijo test and the curvature pair update in FP64. Results: host array
 * of nlines. With lp->precondition, line lengths must be even, at most 16384,
 * and their half must have no prime factor above 7. */
qew_status qew_batch_lbfgs(qew_batch batch, int dtype, const qew_lbfgs_params *lp,
                           double *h, qew_lbfgs_result *results);
/* How the batch's last qew_batch_lbfgs launch ran: its persistent
 * workgroups, and (team_per_line: host int[nlines], or NULL) how many
 * workgroups solved each line. The first call for a batch (per precision and
 * preconditioner setting) times candidates on a copy of the line state and
 * keeps the fastest: one workgroup per line; teams of 2 or 4 on the longest
 * length and of 2 on the next; and, from the iteration counts of the first
 * candidate, teams of 4 and 2 on the lines with the most site-iterations.
 * QEW_LBFGS_TEAM=k (k workgroups on every line that k divides into segments
 * of at least 64 sites) or QEW_LBFGS_TEAMS="n1:k1,n2:k2" (k1 on lines of n1
 * sites and more, else k2 on lines of n2 and more, else 1) set them instead.
 * No reference counterpart. */
qew_status qew_batch_lbfgs_plan(qew_batch batch, int *workgroups, int *team_per_line);

/* ---- FIRE (src/fire.h:41-131) ------------------------------------------- */

qew_fire_params qew_fire_default_params(void);
/* fire_minimize(h, p, noise, fp): one line, FP64, h (device FP64[n]) relaxed
 * in place; the whole relaxation runs in one launch. */
qew_status qew_fire_minimize(qew_ctx ctx, qew_noise noise, const qew_params *p,
                             const qew_fire_params *fp, int n, double *h,
                             qew_fire_result *result);
/* fire_minimize on every line of the batch in one persistent launch (one
 * workgroup per line). h: device FP64 line state in/out; results: host array
 * of nlines. Tables may be FP64 or FP32 (read as FP64 values). */
qew_status qew_batch_fire(qew_batch batch, const qew_fire_params *fp, double *h,
                          qew_fire_result *results);

/* ---- Langevin (src/dynamics.h:27-70) ------------------------------------ */

/* LangevinDynamics::step(h, p, noise, drift, diff) for integrator step number
 * `step` of the stream keyed by `seed`: h += -drift*grad + diff*N(0,1), the
 * normal drawn from Philox4x32-10 keyed by (seed, step, site). diff == 0 draws
 * nothing (src/dynamics.h:59-63). h: device FP64[n]. */
qew_status qew_langevin_step(qew_ctx ctx, qew_noise noise, const qew_params *p,
                             int n, double *h, double drift, double diff,
                             uint64_t seed, uint64_t step);
/* nsteps Langevin steps of every line of the batch with per-step coefficients
 * drift[k], diff[k] (host arrays of nsteps) and step numbers step0 + k. dtype
 * as for qew_batch_eval (QEW_F32: h is FP32 offsets from anchor). */
qew_status qew_batch_langevin(qew_batch batch, int dtype, void *h, const double *anchor,
                              int nsteps, const double *drift, const double *diff,
                              uint64_t seed, uint64_t step0);

/* ---- Rosso-Krauth (src/rosso_krauth.h:52-194) --------------------------- */

/* rk_block(h, p, noise, rp): QEW_ERR_INVALID_ARGUMENT for odd n
 * (src/rosso_krauth.h:132-136). h: device FP64[n], in place. */
qew_status qew_rk_block(qew_ctx ctx, qew_noise noise, const qew_params *p,
                        const qew_rk_params *rp, int n, double *h,
                        qew_rk_result *result);
/* rk_block for every line of the batch at the batch's current forces; lines
 * finish independently. The sweep runs on the tables' storage type with the
 * line state, thresholds and bisection in FP64. results: host nlines. */
qew_status qew_batch_rk_block(qew_batch batch, const qew_rk_params *rp, double *h,
                              qew_rk_result *results);
/* rk_site_velocity at m device sites (src/rosso_krauth.h:52-82): v[k] for
 * site i[k] of a line of length n with trial height hi[k] and neighbours
 * hl[k], hr[k]; device arrays. */
/* rk_forward_zero at m device sites (src/rosso_krauth.h:88-124): out[k] is
 * the nearest forward zero of the velocity of site i[k] of a line of length n
 * with neighbours hl[k], hr[k] fixed, starting from hi[k] (hi[k] itself when
 * its velocity is <= 0); found[k] is 0 when the march reached
 * rp->max_advance without a zero (out[k] is then the marched-to height, the
 * reference's `found = false`); evals[k] counts the site-velocity
 * evaluations. found and evals may be NULL. Device arrays. March lengths
 * max_advance / dstep beyond 2^26 steps are rejected. */
qew_status qew_rk_forward_zero(qew_ctx ctx, qew_noise noise, const qew_params *p,
                               const qew_rk_params *rp, int n, int m, const int *i,
                               const double *hi, const double *hl, const double *hr,
                               double *out, int *found, int64_t *evals);
qew_status qew_rk_site_velocity(qew_ctx ctx, qew_noise noise, const qew_params *p,
                                int n, int m, const int *i, const double *hi,
                                const double *hl, const double *hr, double *v);

/* ---- roughness analysis -------------------------------------------------- */

/* Roughness of m periodic profiles of n heights on a box of length L (h:
 * device FP64[m * n], profile p at p * n). Replaces the height_difference_acf
 * / psd calls of the reference's post-processing
 * (tools/postprocess_static.py:20-47; the module they come from is not in the
 * reference). Outputs, each device FP64 or NULL to skip: acf[r - 1] =
 * A(r L / n) = 1/2 <(h(x + r) - h(x))^2>, r = 1 .. n/2, and psd[k - 1] =
 * C(2 pi k / L) = (L / n^2) |sum_x h(x) exp(-2 pi i k x / n)|^2, k = 1 .. n/2,
 * both averaged over the m profiles; w2[p] = <(h_p - <h_p>)^2>. */
qew_status qew_roughness(qew_ctx ctx, int m, int n, double L, const double *h,
                         double *acf, double *psd, double *w2);

#ifdef __cplusplus
}
#endif

#endif /* QEW_AMD_H */
