// Context, memory, potentials and batches behind the C ABI.
//
// Potentials: FilteredNoise (src/filtered_noise.cpp:25-136) is built the
// reference's way -- mt19937_64 phases drawn on the host for every mode in
// loop order, then the band mask -- but everything after the draw runs on the
// device: scatter of the in-band modes, inverse 2-D FFT, mean/variance
// normalisation, the periodic cubic B-spline prefilter (forward FFT, divide by
// Kx(i) Ky(j), inverse FFT) and the transpose into the x-contiguous
// coef[iy * nx + ix] layout of the requested storage type.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <future>
#include <string>
#include <thread>

#include "qew_internal.hpp"

using namespace qew_dev;

namespace qew_host {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string &msg) { g_last_error = msg; }

}  // namespace qew_host

using namespace qew_host;

namespace {

// in-band modes: (flat index, cos, sin)
struct Mode {
    int64_t idx;
    double c, s;
};

__global__ void scatter_modes(double2 *A, const Mode *modes, int64_t m) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) A[modes[k].idx] = make_double2(modes[k].c, modes[k].s);
}

__global__ void scale_complex(double2 *A, int64_t total, double s) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < total) A[k] = make_double2(A[k].x * s, A[k].y * s);
}

// per-block sums of Re(A) (pass 0) or of (Re(A) - mean)^2 (pass 1)
__global__ __launch_bounds__(256) void moment_partial(const double2 *A, int64_t total, int pass,
                                                      const double *mean, double *part) {
    __shared__ double red[16];
    double acc = 0;
    const double mu = pass ? *mean : 0.0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (int64_t)gridDim.x * blockDim.x) {
        const double v = A[k].x - mu;
        acc += pass ? v * v : v;
    }
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void moment_final(const double *part, int np, double *out,
                                                    double inv_n) {
    __shared__ double red[16];
    double acc = 0;
    for (int k = threadIdx.x; k < np; k += blockDim.x) acc += part[k];
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) *out = acc * inv_n;
}

// field = Re(A) * scale, written both as a complex (for the prefilter FFT)
// and as the real field.
__global__ void field_from_spectrum(double2 *A, double *field, int64_t total, double amp,
                                    const double *var) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    const double scale = amp / sqrt(*var);
    const double v = A[k].x * scale;
    field[k] = v;
    A[k] = make_double2(v, 0.0);
}

__global__ void real_to_complex(const double *f, double2 *A, int64_t total) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < total) A[k] = make_double2(f[k], 0.0);
}

// Shat(i, j) /= Kx(i) Ky(j), K(m) = 2/3 + cos(2 pi m / N) / 3
// (src/filtered_noise.cpp:110-119).
__global__ void prefilter_divide(double2 *A, int nx, int ny) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)nx * ny) return;
    const int i = (int)(k / ny), j = (int)(k % ny);
    const double two_pi = 6.283185307179586;
    const double kx = 2.0 / 3 + cos(two_pi * i / nx) / 3.0;
    const double ky = 2.0 / 3 + cos(two_pi * j / ny) / 3.0;
    const double den = kx * ky;
    A[k] = make_double2(A[k].x / den, A[k].y / den);
}

// table[j * pitch + i + 1] = (T) src[i * ny + j]; src is the real part of a
// complex array (stride 2) or a real array (stride 1).
template <typename T>
__global__ void transpose_to_table(const double *src, int sstride, T *table, int nx, int ny,
                                   int pitch) {
    __shared__ double tile[32][33];
    const int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        const int i = i0 + r, j = j0 + threadIdx.x;
        if (i < nx && j < ny) tile[r][threadIdx.x] = src[((int64_t)i * ny + j) * sstride];
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        const int j = j0 + r, i = i0 + threadIdx.x;
        if (i < nx && j < ny) table[(int64_t)j * pitch + i + 1] = (T)tile[threadIdx.x][r];
    }
}

// periodic x-halo: columns 0, nx + 1, nx + 2 repeat coef(nx - 1), coef(0),
// coef(1 mod nx)
template <typename T>
__global__ void fill_x_halo(T *table, int nx, int ny, int pitch) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ny) return;
    T *row = table + (int64_t)j * pitch;
    row[0] = row[nx];
    row[nx + 1] = row[1];
    row[nx + 2] = row[(1 % nx) + 1];
}

template <typename T>
__global__ void table_to_rowmajor(const T *table, double *dst, int nx, int ny, int pitch) {
    __shared__ double tile[32][33];
    const int j0 = blockIdx.y * 32, i0 = blockIdx.x * 32;
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        const int j = j0 + r, i = i0 + threadIdx.x;
        if (i < nx && j < ny) tile[r][threadIdx.x] = (double)table[(int64_t)j * pitch + i + 1];
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        const int i = i0 + r, j = j0 + threadIdx.x;
        if (i < nx && j < ny) dst[(int64_t)i * ny + j] = tile[threadIdx.x][r];
    }
}

dim3 grid1(int64_t total) { return dim3((unsigned)((total + 255) / 256)); }

void fft2(hipStream_t s, double2 *A, double2 *tmp, int nx, int ny, int sign) {
    fft_c2c(s, A, tmp, ny, nx, 1, ny, sign);  // rows
    fft_c2c(s, A, tmp, nx, ny, ny, 1, sign);  // columns
}

// Builds the device table from a device row-major FP64 source (complex real
// part when sstride == 2).
void make_table(qew_noise nz, const double *src, int sstride) {
    qew_ctx ctx = nz->ctx;
    const int nx = nz->desc.nx, ny = nz->desc.ny;
    const size_t es = nz->desc.dtype == 1 ? sizeof(float) : sizeof(double);
    const int pitch = ((nx + 3 + 3) / 4) * 4;
    nz->desc.pitch = pitch;
    QEW_HIP(hipMalloc(&nz->table, es * (size_t)pitch * ny));
    const dim3 g((ny + 31) / 32, (nx + 31) / 32), b(32, 8);
    if (nz->desc.dtype == 1) {
        hipLaunchKernelGGL(transpose_to_table<float>, g, b, 0, ctx->stream, src, sstride,
                           static_cast<float *>(nz->table), nx, ny, pitch);
        hipLaunchKernelGGL(fill_x_halo<float>, dim3((ny + 255) / 256), dim3(256), 0, ctx->stream,
                           static_cast<float *>(nz->table), nx, ny, pitch);
    } else {
        hipLaunchKernelGGL(transpose_to_table<double>, g, b, 0, ctx->stream, src, sstride,
                           static_cast<double *>(nz->table), nx, ny, pitch);
        hipLaunchKernelGGL(fill_x_halo<double>, dim3((ny + 255) / 256), dim3(256), 0, ctx->stream,
                           static_cast<double *>(nz->table), nx, ny, pitch);
    }
    QEW_HIP(hipGetLastError());
    nz->desc.coef = nz->table;
}

// Device prefilter of the complex array A (holding the field) into a table.
void prefilter_to_table(qew_noise nz, double2 *A, double2 *tmp) {
    qew_ctx ctx = nz->ctx;
    const int nx = nz->desc.nx, ny = nz->desc.ny;
    const int64_t total = (int64_t)nx * ny;
    fft2(ctx->stream, A, tmp, nx, ny, -1);
    hipLaunchKernelGGL(prefilter_divide, grid1(total), dim3(256), 0, ctx->stream, A, nx, ny);
    QEW_HIP(hipGetLastError());
    fft2(ctx->stream, A, tmp, nx, ny, +1);
    hipLaunchKernelGGL(scale_complex, grid1(total), dim3(256), 0, ctx->stream, A, total,
                       1.0 / ((double)nx * (double)ny));
    QEW_HIP(hipGetLastError());
    make_table(nz, reinterpret_cast<const double *>(A), 2);
}

// field_ is kept on the host (FilteredNoise::field()) up to this many points.
constexpr int64_t kKeepFieldMax = int64_t(1) << 26;

qew_noise new_noise(qew_ctx ctx, int kind, int dtype, int nx, int ny, double dx, double dy) {
    qew_noise nz = new qew_noise_s();
    nz->ctx = ctx;
    nz->desc.kind = kind;
    nz->desc.dtype = dtype;
    nz->desc.nx = nx;
    nz->desc.ny = ny;
    nz->desc.dx = dx;
    nz->desc.dy = dy;
    nz->desc.rdx = 1.0 / dx;
    nz->desc.rdy = 1.0 / dy;
    return nz;
}

void check_table_args(int nx, int ny, int dtype) {
    QEW_REQUIRE(nx >= 1 && ny >= 1, "noise grid must be at least 1x1");
    QEW_REQUIRE(dtype == QEW_F64 || dtype == QEW_F32, "dtype must be QEW_F64 or QEW_F32");
}

template <typename S, typename D>
__global__ void convert_table(const S *src, D *dst, int64_t total) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < total) dst[k] = (D)src[k];
}

// float -> bfloat16 bits, round to nearest even (NaN stays a quiet NaN)
__device__ __forceinline__ unsigned short bf16_bits(float f) {
    const unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x0040u);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// Padded BF16 rows (pitch >= nx + 7): element j of row iy holds
// coef((j - 1) mod nx, iy), rounded FP64 -> FP32 -> BF16. The source rows
// have the standard halo (j = 0 .. nx + 2).
template <typename S>
__global__ void bf16_table(const S *src, int s_pitch, unsigned short *dst, int d_pitch, int nx,
                           int ny) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)d_pitch * ny) return;
    const int j = (int)(k % d_pitch);
    const int iy = (int)(k / d_pitch);
    const int cell = ((j - 1) % nx + nx) % nx;
    dst[(int64_t)iy * d_pitch + j] = bf16_bits((float)src[(int64_t)iy * s_pitch + cell + 1]);
}

// Midpoint-contracted FP32 table (pitch >= nx, no halo): element ix of row
// iy holds sum_a w_a coef(ix - 1 + a mod nx, iy), w = the cubic B-spline
// weights at fraction 1/2, contracted in FP64 (a = 0..3 in order) and
// rounded to FP32 once; the padding elements are 0. The source rows have the
// standard halo (element j = coef(j - 1 mod nx)), so the four neighbours of
// ix are source elements ix .. ix + 3.
template <typename S, typename D>
__global__ void mid_table(const S *src, int s_pitch, D *dst, int d_pitch, int nx, int ny) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)d_pitch * ny) return;
    const int ix = (int)(k % d_pitch);
    const int iy = (int)(k / d_pitch);
    if (ix >= nx) {
        dst[k] = (D)0;
        return;
    }
    double w[4];
    qew_dev::bspline_w(0.5, w);
    const S *r = src + (int64_t)iy * s_pitch + ix;
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) acc += w[a] * (double)r[a];
    dst[k] = (D)acc;
}

template <typename D>
__global__ void mid_table_to_rowmajor(const D *table, double *dst, int nx, int ny, int pitch) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)nx * ny) return;
    const int i = (int)(k / ny), j = (int)(k % ny);
    dst[k] = (double)table[(int64_t)j * pitch + i];
}

__global__ void bf16_table_to_rowmajor(const unsigned short *table, double *dst, int nx, int ny,
                                       int pitch) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)nx * ny) return;
    const int i = (int)(k / ny), j = (int)(k % ny);
    dst[k] = (double)__uint_as_float((unsigned)table[(int64_t)j * pitch + i + 1] << 16);
}

// Host part of FilteredNoise (src/filtered_noise.cpp:35-51): the mt19937_64
// stream, one phase per mode in (i, j) loop order, drawn before the mask
// test; the in-band modes as (flat index, cos, sin).
std::vector<Mode> draw_modes(int nx, int ny, double dx, double dy, double xi_x, double xi_y,
                             uint64_t seed) {
    const double two_pi = 2.0 * std::acos(-1.0);
    std::vector<Mode> modes;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> phase(0, two_pi);
    auto fftfreq = [](int k, int n) {
        const int ks = (k < (n + 1) / 2) ? k : k - n;
        return (double)ks / (double)n;
    };
    for (int i = 0; i < nx; ++i) {
        const double fx = fftfreq(i, nx) / dx;
        for (int j = 0; j < ny; ++j) {
            const double fy = fftfreq(j, ny) / dy;
            const double ph = phase(rng);
            if (!((fx * xi_x) * (fx * xi_x) + (fy * xi_y) * (fy * xi_y) > 1.0))
                modes.push_back(Mode{(int64_t)i * ny + j, std::cos(ph), std::sin(ph)});
        }
    }
    return modes;
}

// Device scratch of the generator, reused across the tables of a batch.
struct GenScratch {
    qew_host::DevBuf A, tmp, f, part, modes;
    void release() {
        A.release();
        tmp.release();
        f.release();
        part.release();
        modes.release();
    }
};

// Device part: scatter, inverse 2-D FFT, normalisation, prefilter, table.
qew_noise build_from_modes(qew_ctx ctx, int nx, int ny, double dx, double dy, double amplitude,
                           int dtype, const std::vector<Mode> &modes, GenScratch &g) {
    qew_noise nz = new_noise(ctx, TABLE, dtype, nx, ny, dx, dy);
    const int64_t total = (int64_t)nx * ny;
    const int nparts = 1024;
    try {
        g.A.ensure(sizeof(double2) * total);
        g.tmp.ensure(sizeof(double2) * total);
        g.f.ensure(sizeof(double) * total);
        g.part.ensure(sizeof(double) * (nparts + 2));
        double2 *A = g.A.as<double2>(), *tmp = g.tmp.as<double2>();
        double *f = g.f.as<double>(), *part = g.part.as<double>();
        QEW_HIP(hipMemsetAsync(A, 0, sizeof(double2) * total, ctx->stream));
        if (!modes.empty()) {
            g.modes.ensure(sizeof(Mode) * modes.size());
            QEW_HIP(hipMemcpyAsync(g.modes.ptr, modes.data(), sizeof(Mode) * modes.size(),
                                   hipMemcpyHostToDevice, ctx->stream));
            hipLaunchKernelGGL(scatter_modes, grid1((int64_t)modes.size()), dim3(256), 0,
                               ctx->stream, A, g.modes.as<Mode>(), (int64_t)modes.size());
            QEW_HIP(hipGetLastError());
        }
        // ifft2 with the 1/(nx ny) backward normalisation
        fft2(ctx->stream, A, tmp, nx, ny, +1);
        hipLaunchKernelGGL(scale_complex, grid1(total), dim3(256), 0, ctx->stream, A, total,
                           1.0 / ((double)nx * (double)ny));
        // mean, then population variance about it (src/filtered_noise.cpp:61-76)
        double *mean = part + nparts, *var = part + nparts + 1;
        hipLaunchKernelGGL(moment_partial, dim3(nparts), dim3(256), 0, ctx->stream, A, total, 0,
                           mean, part);
        hipLaunchKernelGGL(moment_final, dim3(1), dim3(256), 0, ctx->stream, part, nparts, mean,
                           1.0 / (double)total);
        hipLaunchKernelGGL(moment_partial, dim3(nparts), dim3(256), 0, ctx->stream, A, total, 1,
                           mean, part);
        hipLaunchKernelGGL(moment_final, dim3(1), dim3(256), 0, ctx->stream, part, nparts, var,
                           1.0 / (double)total);
        hipLaunchKernelGGL(field_from_spectrum, grid1(total), dim3(256), 0, ctx->stream, A, f,
                           total, amplitude, var);
        QEW_HIP(hipGetLastError());
        prefilter_to_table(nz, A, tmp);
        if (total <= kKeepFieldMax) {
            nz->field.resize(total);
            QEW_HIP(hipMemcpyAsync(nz->field.data(), f, sizeof(double) * total,
                                   hipMemcpyDeviceToHost, ctx->stream));
        }
        QEW_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        if (nz->table) (void)hipFree(nz->table);
        delete nz;
        throw;
    }
    return nz;
}

}  // namespace

extern "C" {

const char *qew_last_error(void) { return qew_host::g_last_error.c_str(); }

const char *qew_version(void) {
    return "qew_amd 0.1 (gfx950, HIP, -ffp-contract=off, wave64)";
}

qew_status qew_ctx_create(int device, void *hip_stream, qew_ctx *out) {
    return guard([&] {
        QEW_REQUIRE(out != nullptr, "null output");
        int count = 0;
        QEW_HIP(hipGetDeviceCount(&count));
        QEW_REQUIRE(device >= 0 && device < count, "no such device");
        QEW_HIP(hipSetDevice(device));
        noalloc_scratch_init(device);
        qew_ctx c = new qew_ctx_s();
        c->device = device;
        // NULL selects the device's default (null) stream, which orders with
        // every blocking stream -- torch's default stream included.
        c->stream = static_cast<hipStream_t>(hip_stream);
        c->own_stream = false;
        hipDeviceProp_t prop;
        QEW_HIP(hipGetDeviceProperties(&prop, device));
        c->num_cus = prop.multiProcessorCount;
        *out = c;
    });
}

qew_status qew_ctx_destroy(qew_ctx ctx) {
    return guard([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        ctx->partial.release();
        ctx->result.release();
        ctx->work.release();
        ctx->pinned.release();
        if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

qew_status qew_ctx_synchronize(qew_ctx ctx) {
    return guard([&] {
        ctx_check(ctx);
        QEW_HIP(hipStreamSynchronize(ctx->stream));
    });
}

qew_status qew_device_alloc(qew_ctx ctx, size_t bytes, void **out) {
    return guard([&] {
        ctx_check(ctx);
        QEW_REQUIRE(out != nullptr, "null output");
        use_device(ctx);
        QEW_HIP(hipMalloc(out, std::max<size_t>(bytes, 1)));
    });
}

qew_status qew_device_free(qew_ctx ctx, void *ptr) {
    return guard([&] {
        ctx_check(ctx);
        if (!ptr) return;
        use_device(ctx);
        QEW_HIP(hipStreamSynchronize(ctx->stream));
        QEW_HIP(hipFree(ptr));
    });
}

qew_status qew_copy_to_device(qew_ctx ctx, void *dst, const void *src, size_t bytes) {
    return guard([&] {
        ctx_check(ctx);
        use_device(ctx);
        QEW_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        QEW_HIP(hipStreamSynchronize(ctx->stream));
    });
}

qew_status qew_copy_to_host(qew_ctx ctx, void *dst, const void *src, size_t bytes) {
    return guard([&] {
        ctx_check(ctx);
        use_device(ctx);
        QEW_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        QEW_HIP(hipStreamSynchronize(ctx->stream));
    });
}

qew_status qew_noise_analytic(qew_ctx ctx, int kind, const double *p, qew_noise *out) {
    return guard([&] {
        ctx_check(ctx);
        QEW_REQUIRE(out != nullptr, "null output");
        QEW_REQUIRE(kind >= QEW_NOISE_SINCOS && kind <= QEW_NOISE_WASHBOARD,
                    "analytic kind must be SINCOS, CONST, ZERO, HARMONIC or WASHBOARD");
        qew_noise nz = new_noise(ctx, kind, 0, 1, 1, 1.0, 1.0);
        for (int k = 0; k < 4; ++k) nz->desc.p[k] = p ? p[k] : 0.0;
        *out = nz;
    });
}

qew_status qew_noise_from_coefficients(qew_ctx ctx, const double *coef, int nx, int ny,
                                       double dx, double dy, int dtype, qew_noise *out) {
    return guard([&] {
        ctx_check(ctx);
        QEW_REQUIRE(out && coef, "null pointer");
        check_table_args(nx, ny, dtype);
        QEW_REQUIRE(dx > 0 && dy > 0, "grid spacings must be positive");
        use_device(ctx);
        qew_noise nz = new_noise(ctx, TABLE, dtype, nx, ny, dx, dy);
        double *stage = nullptr;
        try {
            const size_t bytes = sizeof(double) * (size_t)nx * ny;
            QEW_HIP(hipMalloc(&stage, bytes));
            QEW_HIP(hipMemcpyAsync(stage, coef, bytes, hipMemcpyHostToDevice, ctx->stream));
            make_table(nz, stage, 1);
            QEW_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            if (stage) (void)hipFree(stage);
            if (nz->table) (void)hipFree(nz->table);
            delete nz;
            throw;
        }
        QEW_HIP(hipFree(stage));
        *out = nz;
    });
}

qew_status qew_noise_from_field(qew_ctx ctx, const double *field, int nx, int ny, double Lx,
                                double Ly, int dtype, qew_noise *out) {
    return guard([&] {
        ctx_check(ctx);
        QEW_REQUIRE(out && field, "null pointer");
        check_table_args(nx, ny, dtype);
        QEW_REQUIRE(Lx > 0 && Ly > 0, "box lengths must be positive");
        use_device(ctx);
        qew_noise nz = new_noise(ctx, TABLE, dtype, nx, ny, Lx / nx, Ly / ny);
        const int64_t total = (int64_t)nx * ny;
        double2 *A = nullptr, *tmp = nullptr;
        double *f = nullptr;
        try {
            QEW_HIP(hipMalloc(&A, sizeof(double2) * total));
            QEW_HIP(hipMalloc(&tmp, sizeof(double2) * total));
            QEW_HIP(hipMalloc(&f, sizeof(double) * total));
            QEW_HIP(hipMemcpyAsync(f, field, sizeof(double) * total, hipMemcpyHostToDevice,
                                   ctx->stream));
            hipLaunchKernelGGL(real_to_complex, grid1(total), dim3(256), 0, ctx->stream, f, A, total);
            QEW_HIP(hipGetLastError());
            prefilter_to_table(nz, A, tmp);
            QEW_HIP(hipStreamSynchronize(ctx->stream));
            if (total <= kKeepFieldMax) nz->field.assign(field, field + total);
        } catch (...) {
            if (A) (void)hipFree(A);
            if (tmp) (void)hipFree(tmp);
            if (f) (void)hipFree(f);
            if (nz->table) (void)hipFree(nz->table);
            delete nz;
            throw;
        }
        (void)hipFree(A);
        (void)hipFree(tmp);
        (void)hipFree(f);
        *out = nz;
    });
}

qew_status qew_noise_generate(qew_ctx ctx, int nx, int ny, double Lx, double Ly,
                              double amplitude, double xi_x, double xi_y, uint64_t seed,
                              int dtype, qew_noise *out) {
    return guard([&] {
        ctx_check(ctx);
        QEW_REQUIRE(out != nullptr, "null output");
        check_table_args(nx, ny, dtype);
        QEW_REQUIRE(Lx > 0 && Ly > 0, "box lengths must be positive");
        use_device(ctx);
        const std::vector<Mode> modes = draw_modes(nx, ny, Lx / nx, Ly / ny, xi_x, xi_y, seed);
        GenScratch g;
        try {
            *out = build_from_modes(ctx, nx, ny, Lx / nx, Ly / ny, amplitude, dtype, modes, g);
        } catch (...) {
            g.release();
            throw;
        }
        g.release();
    });
}

qew_status qew_noise_generate_batch(qew_ctx ctx, int count, const int *nx, const int *ny,
                                    const double *Lx, const double *Ly, const double *amplitude,
                                    const double *xi_x, const double *xi_y, const uint64_t *seed,
                                    int dtype, qew_noise *out) {
    return guard([&] {
        ctx_check(ctx);
        QEW_REQUIRE(count >= 0, "negative count");
        QEW_REQUIRE(!count || (nx && ny && Lx && Ly && amplitude && xi_x && xi_y && seed && out),
                    "null pointer");
        for (int k = 0; k < count; ++k) {
            check_table_args(nx[k], ny[k], dtype);
            QEW_REQUIRE(Lx[k] > 0 && Ly[k] > 0, "box lengths must be positive");
            out[k] = nullptr;
        }
        use_device(ctx);
        // each table's phase stream is independent: host threads draw the
        // next chunk of `width` tables while the device builds the current one
        const int width = std::max(1, std::min(8, (int)std::thread::hardware_concurrency()));
        GenScratch g;
        using Chunk = std::vector<std::future<std::vector<Mode>>>;
        auto launch_chunk = [&](int k0) {
            Chunk c;
            for (int k = k0; k < std::min(count, k0 + width); ++k)
                c.push_back(std::async(std::launch::async, [=] {
                    return draw_modes(nx[k], ny[k], Lx[k] / nx[k], Ly[k] / ny[k], xi_x[k], xi_y[k],
                                      seed[k]);
                }));
            return c;
        };
        try {
            Chunk cur = launch_chunk(0);
            for (int k0 = 0; k0 < count; k0 += width) {
                Chunk next = (k0 + width < count) ? launch_chunk(k0 + width) : Chunk();
                for (int k = k0; k < std::min(count, k0 + width); ++k) {
                    const std::vector<Mode> modes = cur[k - k0].get();
                    out[k] = build_from_modes(ctx, nx[k], ny[k], Lx[k] / nx[k], Ly[k] / ny[k],
                                              amplitude[k], dtype, modes, g);
                }
                cur = std::move(next);
            }
        } catch (...) {
            g.release();
            for (int k = 0; k < count; ++k)
                if (out[k]) {
                    qew_noise_destroy(out[k]);
                    out[k] = nullptr;
                }
            throw;
        }
        g.release();
    });
}

qew_status qew_noise_convert(qew_noise noise, int dtype, qew_noise *out) {
    return guard([&] {
        QEW_REQUIRE(noise && out, "null pointer");
        QEW_REQUIRE(dtype == QEW_F64 || dtype == QEW_F32 || dtype == QEW_BF16 ||
                        dtype == QEW_F32_MID || dtype == QEW_F64_MID,
                    "dtype must be QEW_F64, QEW_F32, QEW_BF16, QEW_F32_MID or QEW_F64_MID");
        QEW_REQUIRE(noise->desc.kind != TABLE || noise->desc.dtype == QEW_F64 ||
                        noise->desc.dtype == QEW_F32,
                    "BF16 and midpoint-contracted tables are not converted further: convert the "
                    "FP64 or FP32 table");
        qew_ctx ctx = noise->ctx;
        use_device(ctx);
        qew_noise nz = new_noise(ctx, noise->desc.kind, dtype, noise->desc.nx, noise->desc.ny,
                                 noise->desc.dx, noise->desc.dy);
        for (int k = 0; k < 4; ++k) nz->desc.p[k] = noise->desc.p[k];
        nz->field = noise->field;
        if (noise->desc.kind == TABLE && (dtype == QEW_F32_MID || dtype == QEW_F64_MID)) {
            const int nx = noise->desc.nx, ny = noise->desc.ny;
            const int pitch = ((nx + 31) / 32) * 32;
            const bool wide = dtype == QEW_F64_MID;
            hipError_t e = hipMalloc(&nz->table, (wide ? sizeof(double) : sizeof(float)) *
                                                     (size_t)pitch * ny);
            if (e != hipSuccess) {
                delete nz;
                QEW_HIP(e);
            }
            nz->desc.pitch = pitch;
            nz->desc.coef = nz->table;
            const dim3 g = grid1((int64_t)pitch * ny);
            if (noise->desc.dtype == 1 && wide)
                hipLaunchKernelGGL((mid_table<float, double>), g, dim3(256), 0, ctx->stream,
                                   static_cast<const float *>(noise->table), noise->desc.pitch,
                                   static_cast<double *>(nz->table), pitch, nx, ny);
            else if (noise->desc.dtype == 1)
                hipLaunchKernelGGL((mid_table<float, float>), g, dim3(256), 0, ctx->stream,
                                   static_cast<const float *>(noise->table), noise->desc.pitch,
                                   static_cast<float *>(nz->table), pitch, nx, ny);
            else if (wide)
                hipLaunchKernelGGL((mid_table<double, double>), g, dim3(256), 0, ctx->stream,
                                   static_cast<const double *>(noise->table), noise->desc.pitch,
                                   static_cast<double *>(nz->table), pitch, nx, ny);
            else
                hipLaunchKernelGGL((mid_table<double, float>), g, dim3(256), 0, ctx->stream,
                                   static_cast<const double *>(noise->table), noise->desc.pitch,
                                   static_cast<float *>(nz->table), pitch, nx, ny);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) {
                (void)hipFree(nz->table);
                delete nz;
                QEW_HIP(e);
            }
        } else if (noise->desc.kind == TABLE && dtype == QEW_BF16) {
            const int nx = noise->desc.nx, ny = noise->desc.ny;
            const int pitch = ((nx + 7 + 7) / 8) * 8;
            hipError_t e = hipMalloc(&nz->table, sizeof(unsigned short) * (size_t)pitch * ny);
            if (e != hipSuccess) {
                delete nz;
                QEW_HIP(e);
            }
            nz->desc.pitch = pitch;
            nz->desc.coef = nz->table;
            const dim3 g = grid1((int64_t)pitch * ny);
            if (noise->desc.dtype == 1)
                hipLaunchKernelGGL(bf16_table<float>, g, dim3(256), 0, ctx->stream,
                                   static_cast<const float *>(noise->table), noise->desc.pitch,
                                   static_cast<unsigned short *>(nz->table), pitch, nx, ny);
            else
                hipLaunchKernelGGL(bf16_table<double>, g, dim3(256), 0, ctx->stream,
                                   static_cast<const double *>(noise->table), noise->desc.pitch,
                                   static_cast<unsigned short *>(nz->table), pitch, nx, ny);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) {
                (void)hipFree(nz->table);
                delete nz;
                QEW_HIP(e);
            }
        } else if (noise->desc.kind == TABLE) {
            const int pitch = noise->desc.pitch;
            const int64_t total = (int64_t)pitch * noise->desc.ny;
            const size_t es = dtype == QEW_F32 ? sizeof(float) : sizeof(double);
            hipError_t e = hipMalloc(&nz->table, es * (size_t)total);
            if (e != hipSuccess) {
                delete nz;
                QEW_HIP(e);
            }
            nz->desc.pitch = pitch;
            nz->desc.coef = nz->table;
            const bool from32 = noise->desc.dtype == 1;
            if (dtype == QEW_F32) {
                if (from32)
                    hipLaunchKernelGGL((convert_table<float, float>), grid1(total), dim3(256), 0,
                                       ctx->stream, static_cast<const float *>(noise->table),
                                       static_cast<float *>(nz->table), total);
                else
                    hipLaunchKernelGGL((convert_table<double, float>), grid1(total), dim3(256), 0,
                                       ctx->stream, static_cast<const double *>(noise->table),
                                       static_cast<float *>(nz->table), total);
            } else {
                if (from32)
                    hipLaunchKernelGGL((convert_table<float, double>), grid1(total), dim3(256), 0,
                                       ctx->stream, static_cast<const float *>(noise->table),
                                       static_cast<double *>(nz->table), total);
                else
                    hipLaunchKernelGGL((convert_table<double, double>), grid1(total), dim3(256), 0,
                                       ctx->stream, static_cast<const double *>(noise->table),
                                       static_cast<double *>(nz->table), total);
            }
            e = hipGetLastError();
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) {
                (void)hipFree(nz->table);
                delete nz;
                QEW_HIP(e);
            }
        }
        *out = nz;
    });
}

qew_status qew_noise_destroy(qew_noise noise) {
    return guard([&] {
        if (!noise) return;
        if (noise->table) {
            (void)hipSetDevice(noise->ctx->device);
            (void)hipStreamSynchronize(noise->ctx->stream);
            QEW_HIP(hipFree(noise->table));
        }
        delete noise;
    });
}

qew_status qew_noise_info(qew_noise noise, int *kind, int *dtype, int *nx, int *ny, double *dx,
                          double *dy) {
    return guard([&] {
        QEW_REQUIRE(noise != nullptr, "null noise");
        if (kind) *kind = noise->desc.kind;
        if (dtype) *dtype = noise->desc.dtype;
        if (nx) *nx = noise->desc.nx;
        if (ny) *ny = noise->desc.ny;
        if (dx) *dx = noise->desc.dx;
        if (dy) *dy = noise->desc.dy;
    });
}

qew_status qew_noise_field(qew_noise noise, double *field_host) {
    return guard([&] {
        QEW_REQUIRE(noise && field_host, "null pointer");
        if (noise->field.empty())
            throw Error(QEW_ERR_UNSUPPORTED,
                        "no host field kept for this potential (coefficient table, analytic, "
                        "or larger than 2^26 points)");
        std::memcpy(field_host, noise->field.data(), sizeof(double) * noise->field.size());
    });
}

qew_status qew_noise_coefficients(qew_noise noise, double *coef_host) {
    return guard([&] {
        QEW_REQUIRE(noise && coef_host, "null pointer");
        QEW_REQUIRE(noise->desc.kind == TABLE, "analytic potentials have no table");
        qew_ctx ctx = noise->ctx;
        use_device(ctx);
        const int nx = noise->desc.nx, ny = noise->desc.ny;
        double *d = nullptr;
        QEW_HIP(hipMalloc(&d, sizeof(double) * (size_t)nx * ny));
        const dim3 g((nx + 31) / 32, (ny + 31) / 32), b(32, 8);
        if (noise->desc.dtype == QEW_F32_MID)
            hipLaunchKernelGGL(mid_table_to_rowmajor<float>, grid1((int64_t)nx * ny), dim3(256), 0,
                               ctx->stream, static_cast<const float *>(noise->table), d, nx, ny,
                               noise->desc.pitch);
        else if (noise->desc.dtype == QEW_F64_MID)
            hipLaunchKernelGGL(mid_table_to_rowmajor<double>, grid1((int64_t)nx * ny), dim3(256),
                               0, ctx->stream, static_cast<const double *>(noise->table), d, nx,
                               ny, noise->desc.pitch);
        else if (noise->desc.dtype == 2)
            hipLaunchKernelGGL(bf16_table_to_rowmajor, grid1((int64_t)nx * ny), dim3(256), 0,
                               ctx->stream, static_cast<const unsigned short *>(noise->table), d,
                               nx, ny, noise->desc.pitch);
        else if (noise->desc.dtype == 1)
            hipLaunchKernelGGL(table_to_rowmajor<float>, g, b, 0, ctx->stream,
                               static_cast<const float *>(noise->table), d, nx, ny,
                               noise->desc.pitch);
        else
            hipLaunchKernelGGL(table_to_rowmajor<double>, g, b, 0, ctx->stream,
                               static_cast<const double *>(noise->table), d, nx, ny,
                               noise->desc.pitch);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(coef_host, d, sizeof(double) * (size_t)nx * ny,
                               hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        (void)hipFree(d);
        QEW_HIP(e);
    });
}

qew_status qew_batch_create(qew_ctx ctx, int nlines, const int *lengths,
                            const qew_noise *noises, const qew_params *params, qew_batch *out) {
    return guard([&] {
        ctx_check(ctx);
        QEW_REQUIRE(out && lengths && noises && params, "null pointer");
        QEW_REQUIRE(nlines >= 1, "a batch needs at least one line");
        use_device(ctx);
        qew_batch b = new qew_batch_s();
        b->ctx = ctx;
        b->nlines = nlines;
        b->offsets.resize(nlines + 1);
        std::vector<int> tile_line;
        int64_t off = 0;
        // distinct potentials get one descriptor each
        std::vector<qew_noise> uniq;
        try {
            for (int l = 0; l < nlines; ++l) {
                QEW_REQUIRE(lengths[l] >= 1, "line length must be >= 1");
                QEW_REQUIRE(noises[l] != nullptr, "null noise in batch");
                QEW_REQUIRE(params[l].model == 0 || params[l].model == 1, "bad model");
                int idx = -1;
                for (size_t u = 0; u < uniq.size(); ++u)
                    if (uniq[u] == noises[l]) idx = (int)u;
                if (idx < 0) {
                    idx = (int)uniq.size();
                    uniq.push_back(noises[l]);
                    b->noises.push_back(noises[l]->desc);
                }
                LineDesc ld = single_line(lengths[l], &params[l]);
                ld.off = off;
                ld.noise = idx;
                ld.tile0 = (int)tile_line.size();
                for (int t = 0; t < ld.ntiles; ++t) tile_line.push_back(l);
                b->lines.push_back(ld);
                b->offsets[l] = off;
                off += lengths[l];
                b->max_n = std::max(b->max_n, lengths[l]);
            }
            b->offsets[nlines] = off;
            b->total = off;
            b->ntiles = (int)tile_line.size();
            b->d_lines.ensure(sizeof(LineDesc) * nlines);
            b->d_noises.ensure(sizeof(NoiseDesc) * b->noises.size());
            b->d_tile_line.ensure(sizeof(int) * tile_line.size());
            b->d_partial.ensure(sizeof(double) * tile_line.size());
            QEW_HIP(hipMemcpyAsync(b->d_lines.ptr, b->lines.data(), sizeof(LineDesc) * nlines,
                                   hipMemcpyHostToDevice, ctx->stream));
            QEW_HIP(hipMemcpyAsync(b->d_noises.ptr, b->noises.data(),
                                   sizeof(NoiseDesc) * b->noises.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
            QEW_HIP(hipMemcpyAsync(b->d_tile_line.ptr, tile_line.data(),
                                   sizeof(int) * tile_line.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
            QEW_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            b->d_lines.release();
            b->d_noises.release();
            b->d_tile_line.release();
            b->d_partial.release();
            delete b;
            throw;
        }
        *out = b;
    });
}

qew_status qew_batch_destroy(qew_batch b) {
    return guard([&] {
        if (!b) return;
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
        b->d_lines.release();
        b->d_noises.release();
        b->d_tile_line.release();
        b->d_partial.release();
        b->d_scratch.release();
        b->d_ltile.release();
        b->d_tune.release();
        delete b;
    });
}

qew_status qew_batch_offsets(qew_batch b, int64_t *offsets_host) {
    return guard([&] {
        QEW_REQUIRE(b && offsets_host, "null pointer");
        std::memcpy(offsets_host, b->offsets.data(), sizeof(int64_t) * (b->nlines + 1));
    });
}

qew_status qew_batch_set_precond_shifts(qew_batch b, const double *shifts) {
    return guard([&] {
        QEW_REQUIRE(b && shifts, "null pointer");
        for (int l = 0; l < b->nlines; ++l) b->lines[l].shift = shifts[l];
        use_device(b->ctx);
        QEW_HIP(hipMemcpyAsync(b->d_lines.ptr, b->lines.data(), sizeof(LineDesc) * b->nlines,
                               hipMemcpyHostToDevice, b->ctx->stream));
        QEW_HIP(hipStreamSynchronize(b->ctx->stream));
    });
}

qew_status qew_batch_set_forces(qew_batch b, const double *forces) {
    return guard([&] {
        QEW_REQUIRE(b && forces, "null pointer");
        for (int l = 0; l < b->nlines; ++l) b->lines[l].f = forces[l];
        use_device(b->ctx);
        QEW_HIP(hipMemcpyAsync(b->d_lines.ptr, b->lines.data(), sizeof(LineDesc) * b->nlines,
                               hipMemcpyHostToDevice, b->ctx->stream));
        QEW_HIP(hipStreamSynchronize(b->ctx->stream));
    });
}

}  // extern "C"
