// Device-side core of the qEW engine for gfx950: descriptors, the bicubic
// B-spline sampler (exact FP64 restatement and an FP32 fast path), elastic
// stencils and wave64 reductions. Included by every .hip translation unit.
//
// All HIP code is compiled with -ffp-contract=off: the FP64 path reproduces
// the reference's operation order (src/filtered_noise.h:33-176,
// src/qew_model.h:53-194) without fused multiply-adds, so per-site results are
// bit-identical to a sequential FP64 CPU evaluation. The FP32 path asks for
// fmaf explicitly where it wants one.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qew_div.hpp"

namespace qew_dev {

constexpr int kWave = 64;
// Eval and one-step Langevin tiles: kTile threads per workgroup cover
// kTileSites = kTile - 1 consecutive sites. Thread t evaluates the edge to
// the right of site first - 1 + t, so thread 0's edge is the tile's left halo
// (the left edge of its first site) and is gathered in the same round as the
// others: no wave of the tile waits on a second, dependent gather.
constexpr int kTile = 256;
constexpr int kTileSites = kTile - 1;

enum Kind { TABLE = 0, SINCOS = 1, CONST = 2, ZERO = 3, HARMONIC = 4, WASHBOARD = 5 };

// A potential as seen by a kernel. The table is stored x-contiguous with a
// periodic x-halo: row iy holds pitch >= nx + 3 entries,
//   table[iy * pitch + j] = coef((j - 1) mod nx, iy),  j = 0 .. nx + 2,
// so the 4 x-neighbours ix-1 .. ix+2 of any cell ix are the contiguous
// entries j = ix .. ix+3 -- one 16-byte (FP32) or two (FP64) vector loads per
// y level, no wrap test -- and neighbouring edges (adjacent lanes) read
// adjacent x, so a wavefront's gathers coalesce along x. pitch is a multiple
// of 4 so rows start 16-byte aligned.
struct NoiseDesc {
    int kind;
    int dtype;  // 0 FP64 table, 1 FP32 table
    int nx, ny;
    int pitch;
    double dx, dy;
    const void *coef;
    double p[4];
    double rdx, rdy;  // RN(1 / dx), RN(1 / dy)
};

// One line of a batch: sites [off, off + n) of the batch's line buffers.
struct LineDesc {
    int64_t off;
    int n;
    int noise;  // index into the batch's NoiseDesc array
    int model;  // 0 linear, 1 arclength
    int tile0;  // first eval tile of this line
    int ntiles;
    int pad;
    double shift;  // L-BFGS preconditioner shift for this line (> 0 overrides)
    double L;   // physical_size
    double lt;  // line_tension
    double f;   // driving_force
};

// ---------------------------------------------------------------------------
// FP64 restatement of the reference sampler (src/filtered_noise.h:33-176).
// ---------------------------------------------------------------------------
// The weights of src/filtered_noise.h:33-61, bit for bit (the divisions by
// 6 through div6: the numerators 4 - 6 t^2 + 3 t^3 and 1 + 3 t + ... are at
// least 1, (1 - t)^3 >= 2^-159 and 3 + 6 t - 9 t^2 is 0 or a normal number,
// so their quotients are normal; t^3, -12 t + 9 t^2 and 3 t^2 can be tiny).
__device__ __forceinline__ void bspline_w(double t, double w[4]) {
    const double t2 = t * t;
    const double t3 = t2 * t;
    const double omt = 1.0 - t;
    w[0] = qew_div::div6(omt * omt * omt);
    w[1] = qew_div::div6(4.0 - 6.0 * t2 + 3.0 * t3);
    w[2] = qew_div::div6(1.0 + 3.0 * t + 3.0 * t2 - 3.0 * t3);
    w[3] = qew_div::div6_sub(t3);
}

__device__ __forceinline__ void bspline_dw(double t, double dw[4]) {
    const double t2 = t * t;
    const double omt = 1.0 - t;
    dw[0] = -omt * omt / 2.0;
    dw[1] = qew_div::div6_sub(-12.0 * t + 9.0 * t2);
    dw[2] = qew_div::div6(3.0 + 6.0 * t - 9.0 * t2);
    dw[3] = qew_div::div6_sub(3.0 * t2);
}

// Workgroups are dispatched round-robin over the part's 8 XCDs (block b on
// XCD b % 8), each with its own L2. This bijection of [0, nb) gives every XCD
// a contiguous range of tiles, so neighbouring tiles of a line -- which share
// the 128-B lines at their edges -- meet in the same L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int x = b % kXcds, s = b / kXcds;
    const int q = nb / kXcds, r = nb % kXcds;
    return x * q + (x < r ? x : r) + s;
}

// a / b, skipping the division when b is exactly 1 (a / 1.0 == a bit for bit)
__device__ __forceinline__ double div1(double a, double b) { return b == 1.0 ? a : a / b; }

// a / b for a spacing b > 0 with rb = RN(1 / b) at hand: a itself when b is
// exactly 1, else the division-free quotient (correctly rounded whenever it
// is normal or zero). Unlike div1 no division is evaluated on either side.
__device__ __forceinline__ double divr(double a, double b, double rb) {
    return b == 1.0 ? a : qew_div::divp(a, b, rb);
}

// A value the same on every lane, moved to scalar registers (the line
// constants computed on the vector unit would otherwise hold VGPRs).
__device__ __forceinline__ double uniform_f64(double v) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// g - floor(g / n) * n. For 0 <= g <= fl(n (1 - 2^-52)) the correctly rounded
// g / n is at most 1 - 2^-53, so the floor is 0 and g comes back unchanged:
// the division is skipped there with the same result bit for bit.
__device__ __forceinline__ double wrap_cells(double g, int n) {
    const double nd = (double)n;
    if (g >= 0.0 && g <= nd * (1.0 - 0x1p-52)) return g;
    return g - floor(g / nd) * nd;
}

// DeviceNoise::locate (src/filtered_noise.h:77-101). Division by a spacing of
// exactly 1 is skipped: x / 1.0 == x bit for bit.
__device__ __forceinline__ void locate64(const NoiseDesc &nd, double x, double y, int ixn[4],
                                         int iyn[4], double &fx, double &fy) {
    const int nx = nd.nx, ny = nd.ny;
    double gx = (nd.dx == 1.0) ? x : x / nd.dx;
    double gy = divr(y, nd.dy, nd.rdy);
    gx = wrap_cells(gx, nx);
    gy = wrap_cells(gy, ny);
    int ix = (int)floor(gx);
    int iy = (int)floor(gy);
    if (ix >= nx) { ix -= nx; gx -= (double)nx; }
    if (iy >= ny) { iy -= ny; gy -= (double)ny; }
    fx = gx - (double)ix;
    fy = gy - (double)iy;
    ixn[0] = (ix == 0) ? nx - 1 : ix - 1;
    ixn[1] = ix;
    ixn[2] = (ix + 1 == nx) ? 0 : ix + 1;
    ixn[3] = (ixn[2] + 1 == nx) ? 0 : ixn[2] + 1;
    iyn[0] = (iy == 0) ? ny - 1 : iy - 1;
    iyn[1] = iy;
    iyn[2] = (iy + 1 == ny) ? 0 : iy + 1;
    iyn[3] = (iyn[2] + 1 == ny) ? 0 : iyn[2] + 1;
}

// Vector types with element alignment only: a window starts at any element.
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));
#define QEW_GLOBAL __attribute__((address_space(1)))

// The 4 coefficients (ix-1 .. ix+2 mod nx, iy) as one global vector load
// (the table pointer travels in a descriptor, so the address space is stated
// explicitly; a generic pointer would compile to flat loads).
__device__ __forceinline__ void row4(const float *t, int pitch, int ix, int iy, float out[4]) {
    const f4u r = *(const QEW_GLOBAL f4u *)(t + (size_t)iy * pitch + ix);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
}

// The same 4 coefficients from a BF16 table (dtype 2): one 4-byte-aligned
// 16-byte load from the even element at or below the window (BF16 rows are
// padded to roundup(nx + 7, 8) elements, so the load stays in the row), then
// each bfloat16 widened to float by a 16-bit shift (little-endian: the even
// element is the low half of a dword).
typedef unsigned u4u __attribute__((ext_vector_type(4), aligned(4)));
__device__ __forceinline__ void row4_bf16(const unsigned short *t, int pitch, int ix, int iy,
                                          float out[4]) {
    const u4u r = *(const QEW_GLOBAL u4u *)(t + (size_t)iy * pitch + (ix & ~1));
    const bool odd = (ix & 1) != 0;
    const unsigned lo0 = r.x << 16, hi0 = r.x & 0xffff0000u;
    const unsigned lo1 = r.y << 16, hi1 = r.y & 0xffff0000u;
    const unsigned lo2 = r.z << 16;
    out[0] = __uint_as_float(odd ? hi0 : lo0);
    out[1] = __uint_as_float(odd ? lo1 : hi0);
    out[2] = __uint_as_float(odd ? hi1 : lo1);
    out[3] = __uint_as_float(odd ? lo2 : hi1);
}

template <typename C>
__device__ __forceinline__ void row4_64(const NoiseDesc &nd, int ix, int iy, double out[4]) {
    if constexpr (sizeof(C) == 8) {
        const double *p = static_cast<const double *>(nd.coef) + (size_t)iy * nd.pitch + ix;
        const d2u a = *(const QEW_GLOBAL d2u *)p;
        const d2u b = *(const QEW_GLOBAL d2u *)(p + 2);
        out[0] = a.x;
        out[1] = a.y;
        out[2] = b.x;
        out[3] = b.y;
    } else {
        float f[4];
        row4(static_cast<const float *>(nd.coef), nd.pitch, ix, iy, f);
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = (double)f[k];
    }
}

// The 4 x 4 coefficient block of a cell, c[a][b] = coef(x-row a, y level b),
// held in the table's storage type.
template <typename C>
__device__ __forceinline__ void gather_block64(const NoiseDesc &nd, int ix, const int iyn[4],
                                               C c[4][4]) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        double r[4];
        row4_64<C>(nd, ix, iyn[b], r);
#pragma unroll
        for (int a = 0; a < 4; ++a) c[a][b] = (C)r[a];
    }
}

// V from a held block and the weights (the reference's summation order,
// sample_value, src/filtered_noise.h:138-155): rows over y first, then over x.
template <typename C>
__device__ __forceinline__ double v_from_block(const C (&c)[4][4], const double wx[4],
                                               const double wy[4]) {
    double vv = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        double row = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) row += (double)c[a][b] * wy[b];
        vv += wx[a] * row;
    }
    return vv;
}

// Analytic potentials: the reference tests' functors, evaluated with the
// device math library (ocml).
__device__ inline void analytic_sample(const NoiseDesc &nd, double x, double y, double &v,
                                       double &dvdx, double &dvdy) {
    const double *p = nd.p;
    switch (nd.kind) {
    case SINCOS:
        v = p[0] * sin(p[1] * x) * cos(p[2] * y);
        dvdx = p[0] * p[1] * cos(p[1] * x) * cos(p[2] * y);
        dvdy = -p[0] * p[2] * sin(p[1] * x) * sin(p[2] * y);
        return;
    case CONST:
        v = p[0]; dvdx = 0; dvdy = 0;
        return;
    case HARMONIC: {
        const double d = y - p[1];
        v = 0.5 * p[0] * d * d; dvdx = 0; dvdy = p[0] * d;
        return;
    }
    case WASHBOARD:
        v = p[0] * cos(p[1] * y); dvdx = 0; dvdy = -p[0] * p[1] * sin(p[1] * y);
        return;
    default:
        v = 0; dvdx = 0; dvdy = 0;
        return;
    }
}

// V and dV/dy at one point from one 4x4 gather: the same arithmetic as
// sample_value (src/filtered_noise.h:138-155) and sample_dy (:159-176).
template <typename C>
__device__ __forceinline__ void sample_v_dy64(const NoiseDesc &nd, double x, double y,
                                              double &v, double &dvdy) {
    if (nd.kind != TABLE) {
        double dvdx;
        analytic_sample(nd, x, y, v, dvdx, dvdy);
        return;
    }
    int ixn[4], iyn[4];
    double fx, fy;
    locate64(nd, x, y, ixn, iyn, fx, fy);
    double wx[4], wy[4], dwy[4];
    bspline_w(fx, wx);
    bspline_w(fy, wy);
    bspline_dw(fy, dwy);
    double c[4][4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        double r[4];
        row4_64<C>(nd, ixn[1], iyn[b], r);
#pragma unroll
        for (int a = 0; a < 4; ++a) c[a][b] = r[a];
    }
    double vv = 0, dd = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        double row = 0, row_dy = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            row += c[a][b] * wy[b];
            row_dy += c[a][b] * dwy[b];
        }
        vv += wx[a] * row;
        dd += wx[a] * row_dy;
    }
    v = vv;
    dvdy = divr(dd, nd.dy, nd.rdy);
}

template <typename C>
__device__ __forceinline__ double sample_dy64(const NoiseDesc &nd, double x, double y) {
    if (nd.kind != TABLE) {
        double v, dvdx, dvdy;
        analytic_sample(nd, x, y, v, dvdx, dvdy);
        return dvdy;
    }
    int ixn[4], iyn[4];
    double fx, fy;
    locate64(nd, x, y, ixn, iyn, fx, fy);
    double wx[4], dwy[4];
    bspline_w(fx, wx);
    bspline_dw(fy, dwy);
    double c[4][4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        double r[4];
        row4_64<C>(nd, ixn[1], iyn[b], r);
#pragma unroll
        for (int a = 0; a < 4; ++a) c[a][b] = r[a];
    }
    double d = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        double row_dy = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) row_dy += c[a][b] * dwy[b];
        d += wx[a] * row_dy;
    }
    return divr(d, nd.dy, nd.rdy);
}

template <typename C>
__device__ __forceinline__ double sample_v64(const NoiseDesc &nd, double x, double y) {
    double v, dvdy;
    if (nd.kind != TABLE) {
        double dvdx;
        analytic_sample(nd, x, y, v, dvdx, dvdy);
        return v;
    }
    int ixn[4], iyn[4];
    double fx, fy;
    locate64(nd, x, y, ixn, iyn, fx, fy);
    double wx[4], wy[4];
    bspline_w(fx, wx);
    bspline_w(fy, wy);
    C c[4][4];
    gather_block64<C>(nd, ixn[1], iyn, c);
    return v_from_block(c, wx, wy);
}

// Full sample with both derivatives (src/filtered_noise.h:103-134).
template <typename C>
__device__ inline void sample_full64(const NoiseDesc &nd, double x, double y, double &v,
                                     double &dvdx, double &dvdy) {
    if (nd.kind != TABLE) {
        analytic_sample(nd, x, y, v, dvdx, dvdy);
        return;
    }
    int ixn[4], iyn[4];
    double fx, fy;
    locate64(nd, x, y, ixn, iyn, fx, fy);
    double wx[4], dwx[4], wy[4], dwy[4];
    bspline_w(fx, wx);
    bspline_dw(fx, dwx);
    bspline_w(fy, wy);
    bspline_dw(fy, dwy);
    double cc[4][4];
    for (int b = 0; b < 4; ++b) {
        double r[4];
        row4_64<C>(nd, ixn[1], iyn[b], r);
        for (int a = 0; a < 4; ++a) cc[a][b] = r[a];
    }
    double vv = 0, dgx = 0, dgy = 0;
    for (int a = 0; a < 4; ++a) {
        double row = 0, row_dy = 0;
        for (int b = 0; b < 4; ++b) {
            const double c = cc[a][b];
            row += c * wy[b];
            row_dy += c * dwy[b];
        }
        vv += wx[a] * row;
        dgx += dwx[a] * row;
        dgy += wx[a] * row_dy;
    }
    v = vv;
    dvdx = dgx / nd.dx;
    dvdy = dgy / nd.dy;
}

// ---------------------------------------------------------------------------
// FP32 fast sampler. The line height enters as an FP32 offset from a per-line
// anchor; the anchor's cell index is exact (integer) and only the small
// in-band offset is FP32, so the fraction keeps ~1e-7 cell resolution
// regardless of where the line sits in the table.
// ---------------------------------------------------------------------------
struct YAnchor {
    int cell;    // floor(anchor / dy) reduced into [0, ny)
    float frac;  // anchor / dy - floor(anchor / dy)
};

__device__ __forceinline__ YAnchor make_anchor(double anchor, double dy, int ny) {
    const double g = anchor / dy;
    const double fl = floor(g);
    YAnchor a;
    double c = fl - floor(fl / (double)ny) * (double)ny;
    int ci = (int)c;
    if (ci >= ny) ci -= ny;
    a.cell = ci;
    a.frac = (float)(g - fl);
    return a;
}

__device__ __forceinline__ void bspline_wf(float t, float w[4]) {
    const float t2 = t * t, t3 = t2 * t, omt = 1.0f - t;
    const float s = 1.0f / 6.0f;
    w[0] = omt * omt * omt * s;
    w[1] = __builtin_fmaf(3.0f, t3, __builtin_fmaf(-6.0f, t2, 4.0f)) * s;
    w[2] = __builtin_fmaf(-3.0f, t3, __builtin_fmaf(3.0f, t2, __builtin_fmaf(3.0f, t, 1.0f))) * s;
    w[3] = t3 * s;
}

__device__ __forceinline__ void bspline_dwf(float t, float dw[4]) {
    const float t2 = t * t, omt = 1.0f - t;
    dw[0] = -0.5f * omt * omt;
    dw[1] = __builtin_fmaf(1.5f, t2, -2.0f * t);
    dw[2] = __builtin_fmaf(-1.5f, t2, __builtin_fmaf(1.0f, t, 0.5f));
    dw[3] = 0.5f * t2;
}

// ---------------------------------------------------------------------------
// wave64 / workgroup reductions
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const T u = __shfl_xor(v, o, kWave);
        v = (u > v) ? u : v;
    }
    return v;
}

// Error-free sum of two floats (Knuth TwoSum): s + e == a + b exactly.
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Compensated pair (sum, correction) combine.
struct KSum {
    float s, c;
};

__device__ __forceinline__ KSum ksum_add(KSum x, KSum y) {
    float s, e;
    two_sum(x.s, y.s, s, e);
    KSum r;
    const float c = x.c + y.c + e;
    two_sum(s, c, r.s, r.c);
    return r;
}

__device__ __forceinline__ KSum wave_ksum(KSum v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        KSum u;
        u.s = __shfl_xor(v.s, o, kWave);
        u.c = __shfl_xor(v.c, o, kWave);
        v = ksum_add(v, u);
    }
    return v;
}

// Sum over a workgroup of blockDim.x threads (a multiple of 64, <= 1024).
// Every thread gets the result. `red` is LDS scratch of >= 16 entries.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *red) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    T r = 0;
    for (int w = 0; w < nw; ++w) r += red[w];
    return r;
}

template <typename T>
__device__ __forceinline__ T block_max(T v, T *red) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    v = wave_max(v);
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    T r = red[0];
    for (int w = 1; w < nw; ++w) r = (red[w] > r) ? red[w] : r;
    return r;
}

__device__ __forceinline__ KSum block_ksum(KSum v, KSum *red) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    v = wave_ksum(v);
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    KSum r = red[0];
    for (int w = 1; w < nw; ++w) r = ksum_add(r, red[w]);
    return r;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and the per-(seed, step,
// site) standard normal: counter = (site/2, step_lo, step_hi, 0), key = seed,
// two 53-bit uniforms, Box-Muller (cos for even sites, sin for odd sites).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32 x 32 -> 64-bit product per multiplier gives both halves
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0;
        const uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Both normals of the Philox block of sites 2 pair and 2 pair + 1: zc (cos,
// the even site) and zs (sin, the odd site), from one sincos.
__device__ __forceinline__ void philox_normal_pair(uint64_t seed, uint64_t step, uint64_t pair,
                                                   double &zc, double &zs) {
    uint32_t c[4] = {(uint32_t)pair, (uint32_t)step, (uint32_t)(step >> 32), 0u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t k1 = ((((uint64_t)c[0]) << 32) | c[1]) >> 11;
    const uint64_t k2 = ((((uint64_t)c[2]) << 32) | c[3]) >> 11;
    const double u1 = (double)(k1 + 1) * 0x1.0p-53;
    const double u2 = (double)k2 * 0x1.0p-53;
    const double rad = sqrt(-2.0 * log(u1));
    const double th = 6.283185307179586 * u2;
    double sn, cs;
    sincos(th, &sn, &cs);
    zc = rad * cs;
    zs = rad * sn;
}

__device__ __forceinline__ double philox_normal(uint64_t seed, uint64_t step, uint64_t site) {
    double zc, zs;
    philox_normal_pair(seed, step, site >> 1, zc, zs);
    return (site & 1) ? zs : zc;
}

// The same two normals with the Box-Muller transform in FP32, for kernels
// whose line state is FP32: the same Philox block and the same 53-bit
// uniforms, so each normal is the FP64 one to within a few FP32 ulps (the
// FP32 state rounds every update to that resolution anyway). ln u1 keeps
// full relative accuracy: log of u1 itself below 1/2, log1p of the exact
// u1 - 1 above (where u1 rounded to FP32 would lose it); the angle goes to
// sincospi as 2 u2 half-turns, which needs no range reduction.
__device__ __forceinline__ void philox_normal_pair32(uint64_t seed, uint64_t step, uint64_t pair,
                                                     float &zc, float &zs) {
    uint32_t c[4] = {(uint32_t)pair, (uint32_t)step, (uint32_t)(step >> 32), 0u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t k1 = ((((uint64_t)c[0]) << 32) | c[1]) >> 11;
    const uint64_t k2 = ((((uint64_t)c[2]) << 32) | c[3]) >> 11;
    const double u1 = (double)(k1 + 1) * 0x1.0p-53;
    const double u2 = (double)k2 * 0x1.0p-53;
    const float lu = u1 < 0.5 ? logf((float)u1) : log1pf((float)(u1 - 1.0));
    const float rad = sqrtf(-2.0f * lu);
    float sn, cs;
    sincospif(2.0f * (float)u2, &sn, &cs);
    zc = rad * cs;
    zs = rad * sn;
}

__device__ __forceinline__ float philox_normal32(uint64_t seed, uint64_t step, uint64_t site) {
    float zc, zs;
    philox_normal_pair32(seed, step, site >> 1, zc, zs);
    return (site & 1) ? zs : zc;
}

// The normal pair in the line state's precision (FP64 or FP32 Box-Muller),
// widened to double: every kernel of one state precision draws its kicks
// through this, so they agree bit for bit.
template <typename S>
__device__ __forceinline__ void state_normal_pair(uint64_t seed, uint64_t step, uint64_t pair,
                                                  double &zc, double &zs) {
    if constexpr (sizeof(S) == 8) {
        philox_normal_pair(seed, step, pair, zc, zs);
    } else {
        float c, s;
        philox_normal_pair32(seed, step, pair, c, s);
        zc = c;
        zs = s;
    }
}

template <typename S>
__device__ __forceinline__ double state_normal(uint64_t seed, uint64_t step, uint64_t site) {
    if constexpr (sizeof(S) == 8) return philox_normal(seed, step, site);
    else return philox_normal32(seed, step, site);
}

}  // namespace qew_dev

namespace qew_dev {

// ---------------------------------------------------------------------------
// Per-line geometry for kernels that work on one whole line per workgroup.
// ---------------------------------------------------------------------------
struct LineGeom {
    int n;
    int model;
    double dx, lt, f, L;
    double rdx, rdx2;  // RN(1 / dx), RN(1 / dx^2), dx^2 rounded
    NoiseDesc nd;
    // FP32 path
    YAnchor ya;
    double anchor;
    double cx;
    float inv_dy, inv_dx, dxf;
};

__device__ __forceinline__ LineGeom make_geom(const LineDesc &L, const NoiseDesc &nd,
                                              double anchor) {
    LineGeom g;
    g.n = L.n;
    g.model = L.model;
    g.dx = L.L / (double)L.n;
    g.rdx = uniform_f64(1.0 / g.dx);
    g.rdx2 = uniform_f64(1.0 / (g.dx * g.dx));
    g.lt = L.lt;
    g.f = L.f;
    g.L = L.L;
    g.nd = nd;
    g.anchor = anchor;
    g.ya = make_anchor(anchor, nd.dy, nd.ny);
    g.cx = 0.5 * g.dx / nd.dx;
    g.inv_dy = (float)(1.0 / nd.dy);
    g.inv_dx = (float)(1.0 / g.dx);
    g.dxf = (float)g.dx;
    return g;
}

// Cell and fractions of the FP32 edge (e, ep) of a line whose heights are
// FP32 offsets from the anchor (iy already reduced into [0, ny)).
struct Edge32Loc {
    int ix, iy;
    float fx, fy;
};
__device__ __forceinline__ Edge32Loc edge32_locate(const LineGeom &G, int e, int ep, float he,
                                                   float hep) {
    const int nx = G.nd.nx, ny = G.nd.ny;
    double gx = (double)(e + ep) * G.cx;
    if (gx >= (double)nx) gx -= floor(gx / (double)nx) * (double)nx;
    Edge32Loc l;
    l.ix = (int)gx;
    l.fx = (float)(gx - (double)l.ix);
    if (l.ix >= nx) l.ix -= nx;
    const float t = __builtin_fmaf((he + hep) * 0.5f, G.inv_dy, G.ya.frac);
    const float kf = floorf(t);
    l.fy = t - kf;
    l.iy = G.ya.cell + (int)kf;
    if (l.iy < 0 || l.iy >= ny) l.iy = ((l.iy % ny) + ny) % ny;
    return l;
}

// V and/or dV/dy from the window's coefficients cf[q][b] (x-row q, y level b)
template <bool WANT_V, bool WANT_DY>
__device__ __forceinline__ void edge32_combine(const LineGeom &G, const Edge32Loc &l,
                                               const float (&cf)[4][4], float &v, float &dvdy) {
    float wx[4], wy[4], dwy[4];
    bspline_wf(l.fx, wx);
    if (WANT_V) bspline_wf(l.fy, wy);
    if (WANT_DY) bspline_dwf(l.fy, dwy);
    float vv = 0.f, dd = 0.f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float r = 0.f, rd = 0.f;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const float c = cf[q][b];
            if (WANT_V) r = __builtin_fmaf(c, wy[b], r);
            if (WANT_DY) rd = __builtin_fmaf(c, dwy[b], rd);
        }
        if (WANT_V) vv = __builtin_fmaf(wx[q], r, vv);
        if (WANT_DY) dd = __builtin_fmaf(wx[q], rd, dd);
    }
    v = vv;
    dvdy = dd * G.inv_dy;
}

// FP32 edge (e, ep) of a line whose heights are FP32 offsets from the
// anchor: value and/or y-derivative of the potential at the edge midpoint.
// Midpoint-contracted table (dtype 3): T(ix, iy) already holds the x-sum
// with the weights at x-fraction 1/2, which is every edge's fraction on the
// lines such tables are accepted for (site spacing = table x spacing, even
// n). The edge (e, ep) lies in x cell (e + ep) / 2 (integer division: e + ep
// is odd, or n - 1 for the wrap edge), and V and dV/dy need only the 4
// values of that column at the edge's y levels.
template <bool WANT_V, bool WANT_DY>
__device__ __forceinline__ void edge32_mid(const LineGeom &G, int e, int ep, float he, float hep,
                                           float &v, float &dvdy) {
    const int nx = G.nd.nx, ny = G.nd.ny;
    int ix = (e + ep) >> 1;
    if (ix >= nx) ix -= (ix / nx) * nx;
    const float t = __builtin_fmaf((he + hep) * 0.5f, G.inv_dy, G.ya.frac);
    const float kf = floorf(t);
    const float fy = t - kf;
    int iy = G.ya.cell + (int)kf;
    if (iy < 0 || iy >= ny) iy = ((iy % ny) + ny) % ny;
    const int y2 = (iy + 1 == ny) ? 0 : iy + 1;
    const int ys[4] = {(iy == 0) ? ny - 1 : iy - 1, iy, y2, (y2 + 1 == ny) ? 0 : y2 + 1};
    const float *tab = static_cast<const float *>(G.nd.coef);
    float c[4];
#pragma unroll
    for (int b = 0; b < 4; ++b)
        c[b] = *(const QEW_GLOBAL float *)(tab + (size_t)ys[b] * G.nd.pitch + ix);
    float wy[4], dwy[4];
    if (WANT_V) bspline_wf(fy, wy);
    if (WANT_DY) bspline_dwf(fy, dwy);
    float vv = 0.f, dd = 0.f;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        if (WANT_V) vv = __builtin_fmaf(c[b], wy[b], vv);
        if (WANT_DY) dd = __builtin_fmaf(c[b], dwy[b], dd);
    }
    v = vv;
    dvdy = dd * G.inv_dy;
}

// Column of a midpoint-contracted table that edge (e, ep) reads (see
// edge32_mid).
__device__ __forceinline__ int mid_column(const NoiseDesc &nd, int e, int ep) {
    int ix = (e + ep) >> 1;
    if (ix >= nd.nx) ix -= (ix / nd.nx) * nd.nx;
    return ix;
}

// V at height y of column ix of a midpoint-contracted table, in FP64: the y
// locate and weights of locate64 / bspline_w, the 4 column values widened.
__device__ __forceinline__ void mid_locate_y64(const NoiseDesc &nd, double y, int iyn[4],
                                               double &fy) {
    const int ny = nd.ny;
    double gy = divr(y, nd.dy, nd.rdy);
    gy = wrap_cells(gy, ny);
    int iy = (int)floor(gy);
    if (iy >= ny) { iy -= ny; gy -= (double)ny; }
    fy = gy - (double)iy;
    iyn[0] = (iy == 0) ? ny - 1 : iy - 1;
    iyn[1] = iy;
    iyn[2] = (iy + 1 == ny) ? 0 : iy + 1;
    iyn[3] = (iyn[2] + 1 == ny) ? 0 : iyn[2] + 1;
}
template <typename V>
__device__ __forceinline__ void mid_column4(const NoiseDesc &nd, int ix, const int iyn[4],
                                            V c[4]) {
    const V *t = static_cast<const V *>(nd.coef);
#pragma unroll
    for (int b = 0; b < 4; ++b) c[b] = *(const QEW_GLOBAL V *)(t + (size_t)iyn[b] * nd.pitch + ix);
}
template <typename V>
__device__ __forceinline__ double mid_v_from(const V c[4], double fy) {
    double wy[4];
    bspline_w(fy, wy);
    double v = 0.0;
#pragma unroll
    for (int b = 0; b < 4; ++b) v += (double)c[b] * wy[b];
    return v;
}
template <typename V = float>
__device__ __forceinline__ double mid_v64(const NoiseDesc &nd, int ix, double y) {
    int iyn[4];
    double fy;
    mid_locate_y64(nd, y, iyn, fy);
    V c[4];
    mid_column4<V>(nd, ix, iyn, c);
    return mid_v_from<V>(c, fy);
}

// dV/dy at height y of column ix, in FP64 (stored values of type V).
template <typename V>
__device__ __forceinline__ double mid_dy64(const NoiseDesc &nd, int ix, double y) {
    int iyn[4];
    double fy, dwy[4];
    mid_locate_y64(nd, y, iyn, fy);
    bspline_dw(fy, dwy);
    V c[4];
    mid_column4<V>(nd, ix, iyn, c);
    double d = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) d += (double)c[b] * dwy[b];
    return divr(d, nd.dy, nd.rdy);
}

// The 4 x 4 window path of an FP32 or BF16 table.
template <bool WANT_V, bool WANT_DY>
__device__ __forceinline__ void edge32_grid(const LineGeom &G, int e, int ep, float he, float hep,
                                            float &v, float &dvdy) {
    const Edge32Loc l = edge32_locate(G, e, ep, he, hep);
    const int ny = G.nd.ny;
    const int y2 = (l.iy + 1 == ny) ? 0 : l.iy + 1;
    const int ys[4] = {(l.iy == 0) ? ny - 1 : l.iy - 1, l.iy, y2, (y2 + 1 == ny) ? 0 : y2 + 1};
    float cf[4][4];
    // one storage-type branch around the four row loads
    if (G.nd.dtype == 2) {
        const unsigned short *coef = static_cast<const unsigned short *>(G.nd.coef);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            float r[4];
            row4_bf16(coef, G.nd.pitch, l.ix, ys[b], r);
#pragma unroll
            for (int q = 0; q < 4; ++q) cf[q][b] = r[q];
        }
    } else {
        const float *coef = static_cast<const float *>(G.nd.coef);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            float r[4];
            row4(coef, G.nd.pitch, l.ix, ys[b], r);
#pragma unroll
            for (int q = 0; q < 4; ++q) cf[q][b] = r[q];
        }
    }
    edge32_combine<WANT_V, WANT_DY>(G, l, cf, v, dvdy);
}

// FP32 edge (e, ep) on any FP32-state table: FP32, BF16 or midpoint-contracted.
template <bool WANT_V, bool WANT_DY>
__device__ __forceinline__ void edge32(const LineGeom &G, int e, int ep, float he, float hep,
                                       float &v, float &dvdy) {
    if (G.nd.dtype == 3) edge32_mid<WANT_V, WANT_DY>(G, e, ep, he, hep, v, dvdy);
    else edge32_grid<WANT_V, WANT_DY>(G, e, ep, he, hep, v, dvdy);
}

}  // namespace qew_dev
