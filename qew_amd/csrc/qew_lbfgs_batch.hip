// Batched L-BFGS in one persistent launch (src/lbfgs.h:94-287, restated per
// line). One workgroup owns one line at a time and runs the whole
// minimisation -- two-loop recursion, FFT preconditioner, Armijo line search,
// curvature-pair update -- with every scalar decision taken on the device:
// reductions are workgroup-wide and return the same value to every thread, so
// control flow stays uniform and no host round trip happens per iteration.
// Workgroups pull lines from a queue ordered longest first; a finished line
// retires and its workgroup takes the next, which balances ragged batches.
//
// A workgroup is one CU's worth of latency hiding for a whole line, so the
// element loops issue the loads of kU elements per thread before using them
// (stream()), and the passes over the line are fused where the algorithm
// allows: each two-loop step applies its axpy and forms the next step's dot
// product in the same pass, the last one also negates d and forms g.d.
// Per-element arithmetic and each thread's summation order are those of the
// plain loops, so the results do not depend on the fusion.
//
// The preconditioner d = irfft(rfft(q) / eig) runs in LDS: the real line of
// length n is packed as n/2 complex values, transformed by an in-place
// Stockham FFT (radix 4, one radix-2 stage for odd log2), unpacked to the
// half spectrum, divided by eig, repacked and transformed back. Twiddles and
// the eig sines come from one table of exp(2 pi i j / M), M = 4 * the
// largest transform length, filled once per launch.
//
// The same file holds the batched FIRE relaxation (src/fire.h), which reuses
// the per-line gradient, element loops and reductions.
//
// Precision modes: FP64 (the reference's arithmetic; sums are workgroup tree
// sums) and FP32 (line offsets from a per-line anchor, FP32 tables, FP32
// vectors, Kahan-compensated FP32 dot products and energies) in which the
// Armijo test, the energies it compares and the curvature pair s.y, s.s, y.y
// stay FP64.
#include <algorithm>
#include <numeric>
#include <type_traits>
#include <vector>

#include "qew_internal.hpp"

using namespace qew_dev;

namespace {

constexpr int kLB = 512;        // threads per workgroup
constexpr int kU = 8;           // elements per thread with loads in flight (vector passes)
// line-search trial steps evaluated per pass (each holds a 16-coefficient
// gather in flight); the reference's backtracking here takes ~3.8 per iteration
constexpr int kLS = 4;
constexpr int kMaxM = 32;       // history length limit
constexpr int kMaxFftN = 8192;  // complex LDS transform length (line length 16384)

template <typename T>
struct C2 {
    T x, y;
};

struct BatchArgs {
    const LineDesc *lines;
    const NoiseDesc *noises;
    const int *order;
    int nlines;
    unsigned *next;
    double *h;
    void *work;
    qew_lbfgs_result *results;
    // twiddles: power-of-two lengths share C2<T> tw[j] = exp(2 pi i j / tw_m),
    // tw_m = 4 * the largest N = n / 2; in the MIXED kernels each length has
    // its own table, line l reading tw[tw_off[l] + j] = exp(2 pi i j / (2 n))
    const void *tw;
    int tw_m;
    qew_lbfgs_params lp;
    const int64_t *tw_off;
};

template <typename T>
__device__ __forceinline__ C2<T> cm(C2<T> a, C2<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// for (i = tid; i < n; i += kLB) use(i, ld(i)), with the loads of kU
// consecutive iterations issued before any of their uses. Full chunks run
// without per-element guards, so the uses (and any loads inside them, such
// as table gathers) share one basic block and can overlap.
template <int U = kU, typename Ld, typename Use>
__device__ __forceinline__ void stream(int n, Ld ld, Use use) {
    for (int b = threadIdx.x; b < n; b += U * kLB) {
        decltype(ld(0)) v[U];
        if (b + (U - 1) * kLB < n) {
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = ld(b + u * kLB);
#pragma unroll
            for (int u = 0; u < U; ++u) use(b + u * kLB, v[u]);
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = b + u * kLB;
                if (i < n) v[u] = ld(i);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = b + u * kLB;
                if (i < n) use(i, v[u]);
            }
        }
    }
}

template <typename T>
struct P2 {
    T a, b;
};
template <typename T>
struct P3 {
    T a, b, c;
};
template <typename T>
struct P4 {
    T a, b, c, d;
};

// ---- workgroup reductions (uniform results) ----
struct Red {
    double *d;
    KSum *k;
};

// dot-product accumulator: FP64 sum, or a Kahan pair for FP32 vectors
template <typename T>
struct Acc;
template <>
struct Acc<double> {
    double s = 0;
    __device__ void add(double p) { s += p; }
    __device__ __forceinline__ double reduce(Red r) { return block_sum(s, r.d); }
};
template <>
struct Acc<float> {
    KSum s{0.f, 0.f};
    __device__ void add(float p) { s = ksum_add(s, KSum{p, 0.f}); }
    __device__ __forceinline__ double reduce(Red r) {
        const KSum k = block_ksum(s, r.k);
        return (double)k.s + (double)k.c;
    }
};

// three FP64 sums with one pair of barriers
__device__ __forceinline__ void block_sum3(double &x, double &y, double &z, double *red) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    x = wave_sum(x);
    y = wave_sum(y);
    z = wave_sum(z);
    __syncthreads();
    if (lane == 0) {
        red[wid] = x;
        red[nw + wid] = y;
        red[2 * nw + wid] = z;
    }
    __syncthreads();
    double a = 0, b = 0, c = 0;
    for (int w = 0; w < nw; ++w) {
        a += red[w];
        b += red[nw + w];
        c += red[2 * nw + w];
    }
    x = a;
    y = b;
    z = c;
}

// K FP64 sums with one pair of barriers; each equals block_sum of its value
template <int K>
__device__ __forceinline__ void block_sumk(double *v, double *red) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
#pragma unroll
    for (int j = 0; j < K; ++j) v[j] = wave_sum(v[j]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < K; ++j) red[j * nw + wid] = v[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < K; ++j) {
        double a = 0;
        for (int w = 0; w < nw; ++w) a += red[j * nw + w];
        v[j] = a;
    }
}

template <int U = kU, typename T>
__device__ __forceinline__ double wg_dot(const T *a, const T *b, int n, Red r) {
    Acc<T> acc;
    stream<U>(n, [&](int i) { return P2<T>{a[i], b[i]}; },
              [&](int, P2<T> v) { acc.add(v.a * v.b); });
    return acc.reduce(r);
}

template <int U = kU, typename T>
__device__ __forceinline__ double wg_maxabs(const T *a, int n, Red r) {
    double m = 0;
    stream<U>(n, [&](int i) { return a[i]; }, [&](int, T v) {
        const double f = fabs((double)v);
        m = f > m ? f : m;
    });
    return block_max(m, r.d);
}

// ---- objective and gradient of one line ----
// table gathers in flight per thread: 16 coefficients each, held as FP64 for
// FP64 state
template <typename T>
constexpr int kUG = sizeof(T) == 8 ? 2 : 4;

template <typename T, typename C, bool MID = false>
__device__ double wg_objective(const LineGeom &G, const T *x, const T *d, double alpha, Red r) {
    const int n = G.n;
    double acc = 0;
    if constexpr (sizeof(T) == 8) {
        // src/qew_model.h:53-99 (TrialAt: x0(i) + alpha * d(i))
        stream<kUG<T>>(n, [&](int i) {
                   const int ip = (i + 1 == n) ? 0 : i + 1;
                   return P2<double>{d ? x[i] + alpha * d[i] : x[i],
                                     d ? x[ip] + alpha * d[ip] : x[ip]};
               },
               [&](int i, P2<double> v) {
                   const int ip = (i + 1 == n) ? 0 : i + 1;
                   const double hi = v.a, hip = v.b;
                   const double dh = div1(hip - hi, G.dx);
                   const double le =
                       (G.model == 0) ? G.lt * dh * dh * 0.5 : G.lt * (sqrt(1.0 + dh * dh) - 1.0);
                   const double xc = (double)(i + ip) * 0.5 * G.dx;
                   const double hc = (hip + hi) * 0.5;
                   double vv;
                   if constexpr (MID) vv = mid_v64<double>(G.nd, mid_column(G.nd, i, ip), hc);
                   else vv = sample_v64<C>(G.nd, xc, hc);
                   acc += G.dx * (le + vv - G.f * hi);
               });
    } else {
        // The sufficient-decrease test compares energies whose difference
        // shrinks like |g|^2 while the energy itself is O(n): FP32 terms
        // (even compensated) leave a rounding floor far above that, so the
        // energy of the FP32 trial point (the same FP32 values the commit
        // writes) is evaluated in FP64 on the FP32 table.
        const float al = (float)alpha;
        stream<kUG<T>>(n, [&](int i) {
                   const int ip = (i + 1 == n) ? 0 : i + 1;
                   return P2<double>{d ? (double)__builtin_fmaf(al, d[i], x[i]) : (double)x[i],
                                     d ? (double)__builtin_fmaf(al, d[ip], x[ip]) : (double)x[ip]};
               },
               [&](int i, P2<double> v) {
                   const int ip = (i + 1 == n) ? 0 : i + 1;
                   const double oi = v.a, oip = v.b;
                   const double hi = G.anchor + oi, hip = G.anchor + oip;
                   const double dh = div1(oip - oi, G.dx);
                   const double s2 = dh * dh;
                   const double le =
                       (G.model == 0) ? G.lt * s2 * 0.5 : G.lt * (s2 / (sqrt(1.0 + s2) + 1.0));
                   const double hc = (hip + hi) * 0.5;
                   double vv;
                   if constexpr (MID) vv = mid_v64(G.nd, mid_column(G.nd, i, ip), hc);
                   else vv = sample_v64<float>(G.nd, (double)(i + ip) * 0.5 * G.dx, hc);
                   acc += G.dx * (le + vv - G.f * hi);
               });
    }
    return block_sum(acc, r.d);
}

// Table potentials: the kLS trial heights of one edge nearly always lie in
// one y cell (trial steps are small against a cell), so the edge's 4 x 4
// block is gathered once, for trial 0, and every trial in that cell takes V
// from it (same locate, weights and summation order as sample_v64, so the
// same value); a trial in another cell is sampled on its own. One block per
// edge instead of kLS: FP32 state (FP32 blocks) keeps two edges in flight,
// FP64 state one (measured on the 512-line batch: +14 % and +5 % minima/s;
// more edges spill). Per-thread accumulation runs over the thread's edges in
// order, as in the one-edge-at-a-time loop.
template <typename T>
constexpr int kUT = sizeof(T) == 8 ? 1 : 2;

// FP32 state on a midpoint-contracted table (dtype 3, its own kernel
// instantiation): an edge's V needs the 4 values of one column, so the held
// block is 4 floats; trials in the held y cell take V from it, others read
// their own column values. Energies in FP64 as on FP32 tables. Two edges in
// flight per thread measured fastest on the 512-line batch (5650 minima/s
// against 5480 with four, whose spills grew from 224 to 400 B per lane).
constexpr int kUTM = 2;

template <typename T>
__device__ __forceinline__ void objective_trials_mid(const LineGeom &G, const T *x, const T *d,
                                                     const double *alpha, double (&acc)[kLS]) {
    const int n = G.n;
    constexpr int K = kLS;
    // FP32 state: FP32 offsets from the anchor on an FP32 midpoint table;
    // FP64 state: heights (anchor 0) on an FP64 midpoint table
    struct Item {
        P4<T> v;
        int iy;
        T c[4];
    };
    auto offsets = [&](const P4<T> &v, int j, double &oi, double &oip) {
        if constexpr (sizeof(T) == 8) {
            oi = v.a + alpha[j] * v.b;
            oip = v.c + alpha[j] * v.d;
        } else {
            const float al = (float)alpha[j];
            oi = (double)__builtin_fmaf(al, v.b, v.a);
            oip = (double)__builtin_fmaf(al, v.d, v.c);
        }
    };
    stream<kUTM>(
        n,
        [&](int i) {
            const int ip = (i + 1 == n) ? 0 : i + 1;
            Item it;
            it.v = P4<T>{x[i], d[i], x[ip], d[ip]};
            double oi, oip, fy;
            offsets(it.v, 0, oi, oip);
            int iyn[4];
            const double a0 = (sizeof(T) == 8) ? 0.0 : G.anchor;
            mid_locate_y64(G.nd, ((a0 + oip) + (a0 + oi)) * 0.5, iyn, fy);
            it.iy = iyn[1];
            mid_column4(G.nd, mid_column(G.nd, i, ip), iyn, it.c);
            return it;
        },
        [&](int i, const Item &it) {
            const int ip = (i + 1 == n) ? 0 : i + 1;
            const int ix = mid_column(G.nd, i, ip);
#pragma unroll
            for (int j = 0; j < K; ++j) {
                double oi, oip;
                offsets(it.v, j, oi, oip);
                const double a0 = (sizeof(T) == 8) ? 0.0 : G.anchor;
                const double hi = a0 + oi, hip = a0 + oip;
                int iyn[4];
                double fy;
                mid_locate_y64(G.nd, (hip + hi) * 0.5, iyn, fy);
                double vv;
                if (iyn[1] == it.iy) {
                    vv = mid_v_from(it.c, fy);
                } else {
                    T c[4];
                    mid_column4(G.nd, ix, iyn, c);
                    vv = mid_v_from(c, fy);
                }
                const double dh = div1(oip - oi, G.dx);
                const double s2 = dh * dh;
                double le;
                if constexpr (sizeof(T) == 8)  // src/qew_model.h:53-99
                    le = (G.model == 0) ? G.lt * dh * dh * 0.5 : G.lt * (sqrt(1.0 + dh * dh) - 1.0);
                else
                    le = (G.model == 0) ? G.lt * s2 * 0.5 : G.lt * (s2 / (sqrt(1.0 + s2) + 1.0));
                acc[j] += G.dx * (le + vv - G.f * hi);
            }
        });
}

template <typename T, typename C, bool MID = false>
__device__ __forceinline__ void objective_trials_table(const LineGeom &G, const T *x, const T *d,
                                                       const double *alpha, double (&acc)[kLS]) {
    const int n = G.n;
    constexpr int K = kLS;
    if constexpr (MID) {
        objective_trials_mid<T>(G, x, d, alpha, acc);
        return;
    }
    using B = typename std::conditional<sizeof(T) == 8, C, float>::type;  // block storage
    // trial heights of edge (i, ip) for step alpha[j]: FP64 state adds
    // alpha d; FP32 state forms the offset in FP32 and adds the anchor
    auto heights = [&](const P4<T> &v, int j, double &hi, double &hip) {
        if constexpr (sizeof(T) == 8) {
            hi = v.a + alpha[j] * v.b;
            hip = v.c + alpha[j] * v.d;
        } else {
            const float al = (float)alpha[j];
            hi = G.anchor + (double)__builtin_fmaf(al, v.b, v.a);
            hip = G.anchor + (double)__builtin_fmaf(al, v.d, v.c);
        }
    };
    struct Item {
        P4<T> v;
        int iy;
        B c[4][4];
    };
    stream<kUT<T>>(
        n,
        [&](int i) {
            const int ip = (i + 1 == n) ? 0 : i + 1;
            Item it;
            it.v = P4<T>{x[i], d[i], x[ip], d[ip]};
            double hi, hip, fx, fy;
            heights(it.v, 0, hi, hip);
            int ixn[4], iyn[4];
            locate64(G.nd, (double)(i + ip) * 0.5 * G.dx, (hip + hi) * 0.5, ixn, iyn, fx, fy);
            it.iy = iyn[1];
            gather_block64<B>(G.nd, ixn[1], iyn, it.c);
            return it;
        },
        [&](int i, const Item &it) {
            const int ip = (i + 1 == n) ? 0 : i + 1;
            const double xc = (double)(i + ip) * 0.5 * G.dx;
            // trials in the held cell first; the others are sampled after the
            // held block is dead (each trial has its own accumulator)
            double vt[K];
            bool held[K];
#pragma unroll
            for (int j = 0; j < K; ++j) {
                double hi, hip;
                heights(it.v, j, hi, hip);
                int ixn[4], iyn[4];
                double fx, fy;
                locate64(G.nd, xc, (hip + hi) * 0.5, ixn, iyn, fx, fy);
                held[j] = iyn[1] == it.iy;
                vt[j] = 0;
                if (held[j]) {
                    double wx[4], wy[4];
                    bspline_w(fx, wx);
                    bspline_w(fy, wy);
                    vt[j] = v_from_block(it.c, wx, wy);
                }
            }
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if (!held[j]) {
                    double hi, hip;
                    heights(it.v, j, hi, hip);
                    vt[j] = sample_v64<B>(G.nd, xc, (hip + hi) * 0.5);
                }
            }
#pragma unroll
            for (int j = 0; j < K; ++j) {
                double hi, hip;
                heights(it.v, j, hi, hip);
                const double vv = vt[j];
                double le;
                if constexpr (sizeof(T) == 8) {
                    // src/qew_model.h:53-99 (TrialAt: x0(i) + alpha * d(i))
                    const double dh = div1(hip - hi, G.dx);
                    le = (G.model == 0) ? G.lt * dh * dh * 0.5 : G.lt * (sqrt(1.0 + dh * dh) - 1.0);
                } else {
                    const float al = (float)alpha[j];
                    const double oi = (double)__builtin_fmaf(al, it.v.b, it.v.a);
                    const double oip = (double)__builtin_fmaf(al, it.v.d, it.v.c);
                    const double dh = div1(oip - oi, G.dx);
                    const double s2 = dh * dh;
                    le = (G.model == 0) ? G.lt * s2 * 0.5 : G.lt * (s2 / (sqrt(1.0 + s2) + 1.0));
                }
                acc[j] += G.dx * (le + vv - G.f * hi);
            }
        });
}

// Energies at the kLS trial points x + alpha[j] d in one pass over the line.
// Each site term, each thread's summation order and the workgroup sum are
// those of wg_objective, so phi[j] equals wg_objective(G, x, d, alpha[j]).
template <typename T, typename C, bool MID = false>
__device__ __forceinline__ void wg_objective_trials(const LineGeom &G, const T *x, const T *d,
                                    const double *alpha, double *phi, Red r) {
    const int n = G.n;
    constexpr int K = kLS;
    double acc[K];
#pragma unroll
    for (int j = 0; j < K; ++j) acc[j] = 0;
    if (G.nd.kind == TABLE) {
        objective_trials_table<T, C, MID>(G, x, d, alpha, acc);
        block_sumk<K>(acc, r.d);
#pragma unroll
        for (int j = 0; j < K; ++j) phi[j] = acc[j];
        return;
    }
    stream<1>(
        n, [&](int i) {
            const int ip = (i + 1 == n) ? 0 : i + 1;
            return P4<T>{x[i], d[i], x[ip], d[ip]};
        },
        [&](int i, P4<T> v) {
            const int ip = (i + 1 == n) ? 0 : i + 1;
            const double xc = (double)(i + ip) * 0.5 * G.dx;
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if constexpr (sizeof(T) == 8) {
                    // src/qew_model.h:53-99 (TrialAt: x0(i) + alpha * d(i))
                    const double hi = v.a + alpha[j] * v.b, hip = v.c + alpha[j] * v.d;
                    const double dh = div1(hip - hi, G.dx);
                    const double le = (G.model == 0) ? G.lt * dh * dh * 0.5
                                                     : G.lt * (sqrt(1.0 + dh * dh) - 1.0);
                    const double hc = (hip + hi) * 0.5;
                    acc[j] += G.dx * (le + sample_v64<C>(G.nd, xc, hc) - G.f * hi);
                } else {
                    const float al = (float)alpha[j];
                    const double oi = (double)__builtin_fmaf(al, v.b, v.a);
                    const double oip = (double)__builtin_fmaf(al, v.d, v.c);
                    const double hi = G.anchor + oi, hip = G.anchor + oip;
                    const double dh = div1(oip - oi, G.dx);
                    const double s2 = dh * dh;
                    const double le = (G.model == 0) ? G.lt * s2 * 0.5
                                                     : G.lt * (s2 / (sqrt(1.0 + s2) + 1.0));
                    acc[j] += G.dx * (le + sample_v64<float>(G.nd, xc, (hip + hi) * 0.5) - G.f * hi);
                }
            }
        });
    block_sumk<K>(acc, r.d);
#pragma unroll
    for (int j = 0; j < K; ++j) phi[j] = acc[j];
}

template <typename T, typename C, bool MID = false>
__device__ __forceinline__ void wg_gradient(const LineGeom &G, const T *x, T *g, T *edy, T *esl) {
    const int n = G.n;
    const bool arc = G.model == 1;
    stream<kUG<T>>(n, [&](int i) {
               const int ip = (i + 1 == n) ? 0 : i + 1;
               return P2<T>{x[i], x[ip]};
           },
           [&](int i, P2<T> v) {
               const int ip = (i + 1 == n) ? 0 : i + 1;
               if constexpr (sizeof(T) == 8) {
                   const double dh = div1(v.b - v.a, G.dx);
                   if constexpr (MID)
                       edy[i] = mid_dy64<double>(G.nd, mid_column(G.nd, i, ip), (v.b + v.a) * 0.5);
                   else
                       edy[i] = sample_dy64<C>(G.nd, (double)(i + ip) * 0.5 * G.dx, (v.b + v.a) * 0.5);
                   if (arc) esl[i] = dh / sqrt(1.0 + dh * dh);
               } else {
                   float vv, dy;
                   if constexpr (MID) edge32_mid<false, true>(G, i, ip, v.a, v.b, vv, dy);
                   else edge32_grid<false, true>(G, i, ip, v.a, v.b, vv, dy);
                   edy[i] = dy;
                   const float dh = (v.b - v.a) * G.inv_dx;
                   if (arc) esl[i] = dh * rsqrtf(1.0f + dh * dh);
               }
           });
    __syncthreads();
    struct Nb {
        T xm, xi, xp, em, ei, sm, si;
    };
    stream<4>(n, [&](int i) {
               const int ip = (i + 1 == n) ? 0 : i + 1;
               const int im = (i == 0) ? n - 1 : i - 1;
               Nb b;
               b.em = edy[im];
               b.ei = edy[i];
               if (!arc) {
                   b.xm = x[im];
                   b.xi = x[i];
                   b.xp = x[ip];
                   b.sm = b.si = 0;
               } else {
                   b.sm = esl[im];
                   b.si = esl[i];
                   b.xm = b.xi = b.xp = 0;
               }
               return b;
           },
           [&](int i, Nb b) {
               if constexpr (sizeof(T) == 8) {
                   double lf;
                   if (!arc) lf = div1(-G.lt * (b.xm + b.xp - 2.0 * b.xi), G.dx * G.dx);
                   else lf = div1(G.lt * (b.sm - b.si), G.dx);
                   g[i] = G.dx * (lf + 0.5 * (b.em + b.ei) - G.f);
               } else {
                   float lf;
                   const float lt = (float)G.lt;
                   if (!arc) lf = -lt * ((b.xm + b.xp) - 2.0f * b.xi) * (G.inv_dx * G.inv_dx);
                   else lf = lt * (b.sm - b.si) * G.inv_dx;
                   g[i] = G.dxf * (lf + 0.5f * (b.em + b.ei) - (float)G.f);
               }
           });
    __syncthreads();
}

// ---- in-place LDS Stockham FFT of N = 2^a 3^b 5^c 7^d complex values ----
// tw[j] = exp(2 pi i j / M) with M = 4 N; the stage twiddle
// exp(sign 2 pi i q k / (Ns R)) is tw[q k M / (Ns R)] and the radix-R root
// exp(sign 2 pi i / R) is tw[M / R], both conjugated for sign < 0.
// One radix-R stage. N <= kMaxFftN and kLB threads: at most
// ceil(kMaxFftN / (R kLB)) butterflies per thread; the loops have
// compile-time bounds so the values stay in registers.
template <typename T, int R>
__device__ __forceinline__ void fft_stage(C2<T> *buf, int N, int Ns, int sign, const C2<T> *tw,
                                          int M) {
    constexpr int CB = (kMaxFftN / R + kLB - 1) / kLB;
    const int m = N / R;
    const int step = M / (Ns * R);
    C2<T> v[CB][R];
#pragma unroll
    for (int c = 0; c < CB; ++c) {
        const int j = threadIdx.x + c * kLB;
        if (j < m)
#pragma unroll
            for (int q = 0; q < R; ++q) v[c][q] = buf[j + q * m];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < CB; ++c) {
        const int j = threadIdx.x + c * kLB;
        if (j >= m) continue;
        const int k = j % Ns;
        C2<T> u[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            u[q] = v[c][q];
            if (q > 0 && k > 0) {
                C2<T> w = tw[q * k * step];
                if (sign < 0) w.y = -w.y;
                u[q] = cm(u[q], w);
            }
        }
        const int idxD = (j / Ns) * Ns * R + k;
        if constexpr (R == 3 || R == 5 || R == 7) {
            // direct R-point DFT: out_s = sum_q u_q W_R^(q s)
            const int rs = M / R;
#pragma unroll
            for (int o = 0; o < R; ++o) {
                C2<T> acc = u[0];
#pragma unroll
                for (int q = 1; q < R; ++q) {
                    C2<T> w = tw[((q * o) % R) * rs];
                    if (sign < 0) w.y = -w.y;
                    const C2<T> t = cm(u[q], w);
                    acc = {acc.x + t.x, acc.y + t.y};
                }
                buf[idxD + o * Ns] = acc;
            }
        } else if constexpr (R == 2) {
            buf[idxD] = {u[0].x + u[1].x, u[0].y + u[1].y};
            buf[idxD + Ns] = {u[0].x - u[1].x, u[0].y - u[1].y};
        } else {
            const C2<T> a = {u[0].x + u[2].x, u[0].y + u[2].y};
            const C2<T> b = {u[0].x - u[2].x, u[0].y - u[2].y};
            const C2<T> cc = {u[1].x + u[3].x, u[1].y + u[3].y};
            const C2<T> e = {u[1].x - u[3].x, u[1].y - u[3].y};
            const C2<T> je = (sign < 0) ? C2<T>{e.y, -e.x} : C2<T>{-e.y, e.x};
            buf[idxD] = {a.x + cc.x, a.y + cc.y};
            buf[idxD + Ns] = {b.x + je.x, b.y + je.y};
            buf[idxD + 2 * Ns] = {a.x - cc.x, a.y - cc.y};
            buf[idxD + 3 * Ns] = {b.x - je.x, b.y - je.y};
        }
    }
    __syncthreads();
}

// Radices in the order 4, 2 and, in the MIXED instantiation (lengths with
// factors 3, 5 or 7; its larger butterflies would raise the register pressure
// of the power-of-two kernels), 3, 5, 7. The host accepts only lengths with
// no other prime factor; any other remainder stops the loop rather than spin.
template <typename T, bool MIXED>
__device__ void lds_fft(C2<T> *buf, int N, int sign, const C2<T> *tw, int M) {
    int Ns = 1;
    if constexpr (!MIXED) {
        while (Ns < N) {
            if ((N / Ns) % 4 == 0) {
                fft_stage<T, 4>(buf, N, Ns, sign, tw, M);
                Ns *= 4;
            } else {
                fft_stage<T, 2>(buf, N, Ns, sign, tw, M);
                Ns *= 2;
            }
        }
        return;
    }
    while (Ns < N) {
        const int rest = N / Ns;
        if (rest % 4 == 0) {
            fft_stage<T, 4>(buf, N, Ns, sign, tw, M);
            Ns *= 4;
        } else if (rest % 2 == 0) {
            fft_stage<T, 2>(buf, N, Ns, sign, tw, M);
            Ns *= 2;
        } else if constexpr (MIXED) {
            if (rest % 3 == 0) {
                fft_stage<T, 3>(buf, N, Ns, sign, tw, M);
                Ns *= 3;
            } else if (rest % 5 == 0) {
                fft_stage<T, 5>(buf, N, Ns, sign, tw, M);
                Ns *= 5;
            } else if (rest % 7 == 0) {
                fft_stage<T, 7>(buf, N, Ns, sign, tw, M);
                Ns *= 7;
            } else {
                break;
            }
        } else {
            break;
        }
    }
}

// d = irfft(rfft(q) / eig), eig[k] = (kappa/dx) 4 sin^2(pi k / n) + shift
// (src/lbfgs.h:116-146, 173-179); n = 2N. `buf` holds the packed q on entry
// and d, as n reals, on exit.
template <typename T, bool MIXED>
__device__ void wg_precondition(int n, double kdx, double shift, C2<T> *buf, const C2<T> *tw,
                                int M) {
    const int N = n / 2;
    lds_fft<T, MIXED>(buf, N, -1, tw, M);
    const int e_w = M / n;        // W_n^k = conj(tw[k e_w])
    const int e_s = M / (2 * n);  // sin(pi k / n) = tw[k e_s].y
    auto eig = [&](int k) {
        // sin(pi k / n) for k in [0, N]; k e_s <= M / 4
        const double s = (double)tw[k * e_s].y;
        return kdx * 4 * s * s + shift;
    };
    for (int k = threadIdx.x; k <= N / 2; k += blockDim.x) {
        const int nk = (N - k) % N;
        const C2<T> Zk = buf[k], Zn = buf[nk];
        if (k == 0) {
            // X_0 = Re Z0 + Im Z0, X_N = Re Z0 - Im Z0 (both real)
            const T y0 = (Zk.x + Zk.y) / (T)eig(0);
            const T yN = (Zk.x - Zk.y) / (T)eig(N);
            buf[0] = {(y0 + yN) * (T)0.5, (y0 - yN) * (T)0.5};
            continue;
        }
        // E_k = (Z_k + conj Z_{N-k}) / 2, O_k = (Z_k - conj Z_{N-k}) / (2i)
        const C2<T> E = {(Zk.x + Zn.x) * (T)0.5, (Zk.y - Zn.y) * (T)0.5};
        const C2<T> O = {(Zk.y + Zn.y) * (T)0.5, -(Zk.x - Zn.x) * (T)0.5};
        const C2<T> t = tw[k * e_w];
        const C2<T> w = {t.x, -t.y};  // W_n^k
        const C2<T> wO = cm(w, O);
        const T ek = (T)eig(k), enk = (T)eig(N - k);
        const C2<T> Yk = {(E.x + wO.x) / ek, (E.y + wO.y) / ek};
        // X_{N-k} = conj(E_k) - conj(W_n^k O_k)
        const C2<T> Yn = {(E.x - wO.x) / enk, (-E.y + wO.y) / enk};
        // repack: Z'_k = E'_k + i O'_k, E' = (Y_k + conj Y_{N-k})/2,
        // O' = (Y_k - conj Y_{N-k}) W_n^{-k} / 2
        const C2<T> winv = {w.x, -w.y};
        const C2<T> Ep = {(Yk.x + Yn.x) * (T)0.5, (Yk.y - Yn.y) * (T)0.5};
        const C2<T> Op = cm(C2<T>{(Yk.x - Yn.x) * (T)0.5, (Yk.y + Yn.y) * (T)0.5}, winv);
        buf[k] = {Ep.x - Op.y, Ep.y + Op.x};
        if (nk != k) {
            // Z'_{N-k} = conj(E'_k) + i conj(O'_k)
            buf[nk] = {Ep.x + Op.y, -Ep.y + Op.x};
        }
    }
    __syncthreads();
    lds_fft<T, MIXED>(buf, N, +1, tw, M);
    // complex j is reals 2j, 2j + 1: each thread scales its own entries in place
    const T inv = (T)(1.0 / (double)N);
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const C2<T> z = buf[j];
        buf[j] = {z.x * inv, z.y * inv};
    }
    __syncthreads();
}

template <typename T, typename C, bool MID = false, bool MIXED = false>
__device__ __forceinline__ void solve_line(const BatchArgs &a, int line, Red r, double *rho, double *la,
                           int *slot, int *sbuf, C2<T> *fftbuf) {
    // elements per thread with loads in flight in the vector passes: 16 for
    // the FP32 midpoint-table kernel (its line search frees the registers), 8
    // elsewhere (wider passes spill); the per-thread order is the same
    constexpr int UV = (MID && sizeof(T) == 4) ? 16 : kU;
    const qew_lbfgs_params &lp = a.lp;
    const LineDesc L = a.lines[line];
    const NoiseDesc nd = a.noises[L.noise];
    const int n = L.n;
    const int m = lp.m;
    const int tid = threadIdx.x;
    const C2<T> *tw = static_cast<const C2<T> *>(a.tw);
    if constexpr (MIXED)
        if (lp.precondition) tw += a.tw_off[line];
    double *h = a.h + L.off;
    const double anchor = (sizeof(T) == 8) ? 0.0 : h[0];
    const LineGeom G = make_geom(L, nd, anchor);
    // per-line workspace: x g gn d q edy esl, then (m + 1) S and (m + 1) Y buffers
    T *ws = static_cast<T *>(a.work) + (size_t)L.off * (size_t)(7 + 2 * (m + 1));
    T *x = ws, *g = ws + n, *gn = ws + 2 * (size_t)n, *d = ws + 3 * (size_t)n,
      *q = ws + 4 * (size_t)n, *edy = ws + 5 * (size_t)n, *esl = ws + 6 * (size_t)n;
    auto Sb = [&](int k) { return ws + (size_t)(7 + k) * n; };
    auto Yb = [&](int k) { return ws + (size_t)(7 + (m + 1) + k) * n; };
    for (int i = tid; i < n; i += blockDim.x) x[i] = (T)(h[i] - anchor);
    if (tid == 0)
        for (int k = 0; k <= m; ++k) sbuf[k] = k;  // sbuf[m] is the spare (s_tmp / y_tmp)
    __syncthreads();

    double kdx = 0, shift = 0;
    if (lp.precondition) {
        const double pi = 3.14159265358979323846;
        kdx = L.lt / G.dx;
        shift = L.shift > 0 ? L.shift : lp.precond_shift;
        if (shift <= 0) {
            const double s1 = sin(pi / (double)n);
            shift = kdx * 4 * s1 * s1;
        }
    }

    wg_gradient<T, C, MID>(G, x, g, edy, esl);
    qew_lbfgs_result res{};
    int nslots = 0;  // slot[0..nslots) chronological rows, slot[nslots-1] newest
    double gamma = 1.0;
    int ls_fail_streak = 0;
    double phi = 0;
    bool finished = false;
    for (int it = 0; it < lp.max_iter; ++it) {
        const int hs = nslots;
        // With the preconditioner, q and then d live in the LDS FFT buffer
        // (n reals) from the first pass to the second loop's last pass, which
        // writes d to global memory: the two-loop passes read only their
        // S and Y rows from memory. Without it, q and d are global vectors.
        // Values and each thread's operation order are the same either way.
        const bool pre = lp.precondition != 0;
        T *fb = reinterpret_cast<T *>(fftbuf);
        // one pass: max|g|, q = g and the first two-loop dot S_newest . q
        double gmax = 0, al_dot = 0;
        auto first_pass = [&](T *qv) {
            const T *S0 = hs > 0 ? Sb(sbuf[slot[hs - 1]]) : nullptr;
            Acc<T> acc;
            stream<UV>(n, [&](int i) { return P2<T>{g[i], S0 ? S0[i] : (T)0}; },
                   [&](int i, P2<T> v) {
                       const double f = fabs((double)v.a);
                       gmax = f > gmax ? f : gmax;
                       qv[i] = v.a;
                       if (S0) acc.add(v.b * v.a);
                   });
            gmax = block_max(gmax, r.d);
            if (S0) al_dot = acc.reduce(r);
        };
        if (pre) first_pass(fb);
        else first_pass(q);
        if (gmax < lp.ftol) {
            res.iterations = it;
            res.max_force = gmax;
            res.converged = 1;
            finished = true;
            break;
        }
        // two-loop recursion, first loop: q -= al Y for the newest pair first;
        // each pass also forms the next (older) pair's S . q
        auto first_loop = [&](T *qv) {
            for (int idx = hs - 1; idx >= 0; --idx) {
                const int row = slot[idx];
                const double al = rho[row] * al_dot;
                if (tid == 0) la[idx] = al;
                const T c = (T)(-al);
                const T *Y = Yb(sbuf[row]);
                const T *Sn = idx > 0 ? Sb(sbuf[slot[idx - 1]]) : nullptr;
                Acc<T> acc;
                stream<UV>(n, [&](int i) { return P3<T>{qv[i], Y[i], Sn ? Sn[i] : (T)0}; },
                       [&](int i, P3<T> v) {
                           const T qn = v.a + c * v.b;
                           qv[i] = qn;
                           if (Sn) acc.add(v.c * qn);
                       });
                if (Sn) al_dot = acc.reduce(r);
                else __syncthreads();
            }
        };
        if (pre) {
            first_loop(fb);
            wg_precondition<T, MIXED>(n, kdx, shift, fftbuf, tw, MIXED ? 2 * n : a.tw_m);
        } else {
            first_loop(q);
            const T g0 = (T)(nslots == 0 ? 1.0 : gamma);
            stream<UV>(n, [&](int i) { return q[i]; }, [&](int i, T v) { d[i] = g0 * v; });
            __syncthreads();
        }
        // second loop, oldest pair first: d += (la - b) S with b = rho Y . d;
        // the last pass also negates d, forms g . d and stores d to memory
        double dphi0 = 0;
        auto second_loop = [&](T *dv) {
            double b_dot = 0;
            if (hs > 0) b_dot = wg_dot<UV>(Yb(sbuf[slot[0]]), dv, n, r);
            for (int idx = 0; idx < hs; ++idx) {
                const int row = slot[idx];
                const double bb = rho[row] * b_dot;
                const T c = (T)(la[idx] - bb);
                const T *S = Sb(sbuf[row]);
                const bool last = idx == hs - 1;
                const T *Yn = last ? g : Yb(sbuf[slot[idx + 1]]);
                Acc<T> acc;
                // separate call sites for the two destinations keep each
                // store in its own address space
                auto pass = [&](T *dout, bool neg) {
                    stream<UV>(n, [&](int i) { return P3<T>{dv[i], S[i], Yn[i]}; },
                           [&](int i, P3<T> v) {
                               T dn = v.a + c * v.b;
                               if (neg) dn = -dn;
                               dout[i] = dn;
                               acc.add(v.c * dn);
                           });
                };
                if (last) pass(d, true);
                else pass(dv, false);
                const double s = acc.reduce(r);
                if (last) dphi0 = s;
                else b_dot = s;
            }
            if (hs == 0) {
                Acc<T> acc;
                stream<UV>(n, [&](int i) { return P2<T>{dv[i], g[i]}; },
                       [&](int i, P2<T> v) {
                           const T dn = -v.a;
                           d[i] = dn;
                           acc.add(v.b * dn);
                       });
                dphi0 = acc.reduce(r);
            }
        };
        if (pre) second_loop(fb);
        else second_loop(d);
        if (dphi0 >= 0) {
            nslots = 0;
            Acc<T> acc;
            stream<UV>(n, [&](int i) { return g[i]; }, [&](int i, T v) {
                d[i] = -v;
                acc.add(v * v);
            });
            dphi0 = -acc.reduce(r);
        }
        if (it == 0) {
            phi = wg_objective<T, C, MID>(G, x, nullptr, 0.0, r);
            ++res.n_func_evals;
        }
        const double phi0 = phi;
        double alpha = 1.0;
        if (!lp.precondition && nslots == 0) alpha = 1.0 / fmax(1.0, sqrt(wg_dot<UV>(g, g, n, r)));
        // Armijo backtracking alpha, alpha s, alpha s^2, ... (the sequence and
        // the first accepted step of the reference loop); the energies of K
        // consecutive trials come from one pass, the evaluation count is the
        // sequential one
        constexpr int K = kLS;
        bool ok = false;
        for (int ls = 0; ls < lp.max_ls && !ok;) {
            double al[K], ph[K];
            al[0] = alpha;
#pragma unroll
            for (int j = 1; j < K; ++j) al[j] = al[j - 1] * lp.ls_shrink;
            wg_objective_trials<T, C, MID>(G, x, d, al, ph, r);
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if (ok || ls >= lp.max_ls) continue;
                ++ls;
                ++res.n_func_evals;
                if (ph[j] <= phi0 + lp.c1 * al[j] * dphi0) {
                    phi = ph[j];
                    alpha = al[j];
                    ok = true;
                }
            }
            if (!ok) alpha = al[K - 1] * lp.ls_shrink;
        }
        if (!ok) {
            if (nslots == 0 || ++ls_fail_streak > 2) {
                res.iterations = it;
                res.max_force = gmax;
                res.converged = 0;
                finished = true;
                break;
            }
            nslots = 0;
            ++res.n_ls_restarts;
            continue;
        }
        ls_fail_streak = 0;
        const T alT = (T)alpha;
        // curvature pair into the spare buffers. FP64: s = alpha d as in the
        // reference. FP32: s is the step the state actually took,
        // x_new - x_old (exact: both are FP32 values within a factor 2 of
        // each other, or the difference is exactly representable), because
        // near convergence alpha d falls below the offsets' resolution and
        // the intended step no longer matches the realised one.
        T *st = Sb(sbuf[m]), *yt = Yb(sbuf[m]);
        stream<UV>(n, [&](int i) { return P2<T>{x[i], d[i]}; }, [&](int i, P2<T> v) {
            const T xn = v.a + alT * v.b;
            x[i] = xn;
            st[i] = (sizeof(T) == 8) ? alT * v.b : xn - v.a;
        });
        __syncthreads();
        wg_gradient<T, C, MID>(G, x, gn, edy, esl);
        double sy = 0, ss = 0, yy = 0;
        stream<UV>(n, [&](int i) { return P3<T>{st[i], gn[i], g[i]}; }, [&](int i, P3<T> v) {
            const T y = v.b - v.c;
            yt[i] = y;
            sy += (double)v.a * (double)y;
            ss += (double)v.a * (double)v.a;
            yy += (double)y * (double)y;
        });
        block_sum3(sy, ss, yy, r.d);
        if (sizeof(T) == 4 && ss == 0.0) {
            // FP32 state: the accepted step is below the offsets' resolution
            // everywhere, so the line can no longer move; stop (not
            // converged) instead of iterating in place until max_iter
            res.iterations = it + 1;
            res.max_force = wg_maxabs<UV>(gn, n, r);
            res.converged = res.max_force < lp.ftol ? 1 : 0;
            g = gn;
            finished = true;
            break;
        }
        if (sy > lp.curvature_eps * sqrt(ss * yy)) {
            int row;
            __syncthreads();
            if (nslots < m) {
                row = nslots;
                if (tid == 0) slot[nslots] = row;
                ++nslots;
            } else {
                row = slot[0];
                if (tid == 0) {
                    for (int k = 0; k < m - 1; ++k) slot[k] = slot[k + 1];
                    slot[m - 1] = row;
                }
            }
            if (tid == 0) {
                const int tmp = sbuf[row];
                sbuf[row] = sbuf[m];
                sbuf[m] = tmp;
                rho[row] = 1.0 / sy;
            }
            gamma = sy / yy;
            __syncthreads();
        } else {
            ++res.n_curvature_skips;
        }
        T *t = g;
        g = gn;
        gn = t;
    }
    if (!finished) {
        res.iterations = lp.max_iter;
        res.max_force = wg_maxabs<UV>(g, n, r);
        res.converged = 0;
    }
    for (int i = tid; i < n; i += blockDim.x) h[i] = anchor + (double)x[i];
    if (tid == 0) a.results[line] = res;
    __syncthreads();
}

template <typename T, typename C, bool MID = false, bool MIXED = false>
__global__ __launch_bounds__(kLB) void lbfgs_batch_kernel(BatchArgs a) {
    extern __shared__ unsigned char lds[];
    __shared__ double s_red[kLS * (kLB / kWave)];
    __shared__ KSum s_kred[kLB / kWave];
    __shared__ double rho[kMaxM], la[kMaxM];
    __shared__ int slot[kMaxM], sbuf[kMaxM + 1];
    __shared__ int s_line;
    Red r{s_red, s_kred};
    for (;;) {
        if (threadIdx.x == 0) s_line = (int)atomicAdd(a.next, 1u);
        __syncthreads();
        const int qi = s_line;
        __syncthreads();
        if (qi >= a.nlines) return;
        solve_line<T, C, MID, MIXED>(a, a.order[qi], r, rho, la, slot, sbuf,
                                     reinterpret_cast<C2<T> *>(lds));
    }
}

// tw[j] = exp(2 pi i j / M), FP64 sincospi rounded to T
template <typename T>
__global__ void twiddle_kernel(C2<T> *tw, int M) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    double s, c;
    sincospi(2.0 * (double)j / (double)M, &s, &c);
    tw[j] = {(T)c, (T)s};
}

__global__ void zero_u32(unsigned *p) { *p = 0u; }

// ---- FIRE (src/fire.h:41-131), one workgroup per line ----
// The reference's passes per iteration are max|F|; v += dt F; P = F.v;
// (v = 0 on P <= 0); |v|^2, |F|^2; mix; h += dt v; gradient. Here |F|^2 is
// summed in the max|F| pass (F does not change in between), |v|^2 in the P
// pass (v changes after it only when it is zeroed, and then |v|^2 = 0), and
// the zeroing, the mix and the position step are one element-wise pass.
// Every element update is the reference's expression; the sums are
// workgroup tree sums.
struct FireArgs {
    const LineDesc *lines;
    const NoiseDesc *noises;
    const int *order;
    int nlines;
    unsigned *next;
    double *h;
    double *work;
    qew_fire_result *results;
    qew_fire_params fp;
};

template <typename C>
__device__ __forceinline__ void fire_line(const FireArgs &a, int line, Red r) {
    const qew_fire_params &fp = a.fp;
    const LineDesc L = a.lines[line];
    const NoiseDesc nd = a.noises[L.noise];
    const int n = L.n;
    const LineGeom G = make_geom(L, nd, 0.0);
    double *x = a.h + L.off;
    double *ws = a.work + (size_t)L.off * 4;
    double *g = ws, *v = ws + n, *edy = ws + 2 * (size_t)n, *esl = ws + 3 * (size_t)n;
    // every pass maps element i to thread i mod kLB, so element-wise passes
    // need no barrier between them; the gradient reads neighbours
    stream(n, [&](int) { return 0; }, [&](int i, int) { v[i] = 0.0; });
    double dt = fp.dt_init > 0 ? fp.dt_init : 0.1 * fp.dt_max;
    double alpha = fp.alpha_start;
    int n_pos = 0;
    __syncthreads();
    wg_gradient<double, C>(G, x, g, edy, esl);
    qew_fire_result res{};
    res.iterations = fp.max_iter;
    res.converged = 0;
    bool done = false;
    for (int it = 0; it < fp.max_iter; ++it) {
        double max_f = 0, nf2 = 0;
        stream(n, [&](int i) { return g[i]; }, [&](int, double gi) {
            const double f = fabs(gi);
            max_f = f > max_f ? f : max_f;
            nf2 += gi * gi;
        });
        max_f = block_max(max_f, r.d);
        if (max_f < fp.ftol) {
            res.iterations = it;
            res.max_force = max_f;
            res.converged = 1;
            done = true;
            break;
        }
        nf2 = block_sum(nf2, r.d);
        double P = 0, nv2 = 0, unused = 0;
        stream(n, [&](int i) { return P2<double>{g[i], v[i]}; }, [&](int i, P2<double> e) {
            const double vi = e.b - dt * e.a;
            v[i] = vi;
            P += -e.a * vi;
            nv2 += vi * vi;
        });
        block_sum3(P, nv2, unused, r.d);
        const bool zero = !(P > 0);
        if (!zero) {
            if (n_pos > fp.n_min) {
                dt = fmin(dt * fp.f_inc, fp.dt_max);
                alpha *= fp.f_alpha;
            }
            ++n_pos;
        } else {
            dt *= fp.f_dec;
            alpha = fp.alpha_start;
            n_pos = 0;
            nv2 = 0.0;
        }
        const double nf = sqrt(nf2);
        const bool mix = nf > 0;
        const double am = alpha;
        const double scale = mix ? alpha * sqrt(nv2) / nf : 0.0;
        stream(n, [&](int i) { return P3<double>{g[i], v[i], x[i]}; },
               [&](int i, P3<double> e) {
                   double vi = zero ? 0.0 : e.b;
                   if (mix) vi = (1.0 - am) * vi - scale * e.a;
                   v[i] = vi;
                   x[i] = e.c + dt * vi;
               });
        __syncthreads();
        wg_gradient<double, C>(G, x, g, edy, esl);
    }
    if (!done) res.max_force = wg_maxabs(g, n, r);
    if (threadIdx.x == 0) a.results[line] = res;
    __syncthreads();
}

template <typename C>
__global__ __launch_bounds__(kLB) void fire_batch_kernel(FireArgs a) {
    __shared__ double s_red[kLS * (kLB / kWave)];
    __shared__ KSum s_kred[kLB / kWave];
    __shared__ int s_line;
    Red r{s_red, s_kred};
    for (;;) {
        if (threadIdx.x == 0) s_line = (int)atomicAdd(a.next, 1u);
        __syncthreads();
        const int qi = s_line;
        __syncthreads();
        if (qi >= a.nlines) return;
        fire_line<C>(a, a.order[qi], r);
    }
}

// Launches one persistent minimisation kernel: as many workgroups as fit at
// once (each takes lines from the queue), at most one per line.
template <class K>
void launch_persistent(K kernel, const BatchArgs &a, int nlines, size_t lds, qew_ctx ctx) {
    if (lds > 64 * 1024)
        QEW_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    QEW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kLB, lds));
    const int grid = std::max(1, std::min(nlines, std::max(1, per_cu) * ctx->num_cus));
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(kLB), lds, ctx->stream, a);
    QEW_HIP(hipGetLastError());
}

}  // namespace

namespace qew_host {
// The MIXED kernels (radix-3/5/7 preconditioner) are instantiated in their own
// translation unit, qew_lbfgs_mixed.hip, which compiles this file with
// QEW_LBFGS_MIXED_UNIT defined: instantiated beside the power-of-two kernels,
// they change how the compiler inlines the device code all of them share, and
// the power-of-two kernels then spill more (FP64: 200 -> 296 bytes per lane).
// `args` is a BatchArgs of this file; variant 0..4: FP64 state on FP64
// midpoint tables, FP32 on FP32 midpoint tables, FP32 on FP32 tables, FP64 on
// FP32 tables, FP64 on FP64 tables.
void launch_lbfgs_mixed(int variant, const void *args, int nlines, size_t lds, qew_ctx ctx);
}  // namespace qew_host

#ifdef QEW_LBFGS_MIXED_UNIT

void qew_host::launch_lbfgs_mixed(int variant, const void *args, int nlines, size_t lds,
                                  qew_ctx ctx) {
    const BatchArgs &a = *static_cast<const BatchArgs *>(args);
    switch (variant) {
        case 0: launch_persistent(lbfgs_batch_kernel<double, double, true, true>, a, nlines, lds, ctx); break;
        case 1: launch_persistent(lbfgs_batch_kernel<float, float, true, true>, a, nlines, lds, ctx); break;
        case 2: launch_persistent(lbfgs_batch_kernel<float, float, false, true>, a, nlines, lds, ctx); break;
        case 3: launch_persistent(lbfgs_batch_kernel<double, float, false, true>, a, nlines, lds, ctx); break;
        default: launch_persistent(lbfgs_batch_kernel<double, double, false, true>, a, nlines, lds, ctx); break;
    }
}

#else  // the power-of-two kernels and the entry points

using namespace qew_host;

extern "C" qew_status qew_batch_lbfgs(qew_batch b, int dtype, const qew_lbfgs_params *lp,
                                      double *h, qew_lbfgs_result *results) {
    return guard([&] {
        QEW_REQUIRE(b && lp && h && results, "null pointer");
        QEW_REQUIRE(dtype == QEW_F64 || dtype == QEW_F32, "dtype must be QEW_F64 or QEW_F32");
        QEW_REQUIRE(lp->m >= 1 && lp->m <= kMaxM, "history length m must be in [1, 32]");
        if (dtype == QEW_F32) {
            // FP32 tables, or midpoint-contracted ones on lines of their geometry
            for (const auto &nd : b->noises)
                QEW_REQUIRE(nd.kind == TABLE && (nd.dtype == QEW_F32 || nd.dtype == QEW_F32_MID),
                            "QEW_F32 minimisation needs FP32 or QEW_F32_MID coefficient tables");
            require_f32_tables(b->lines, b->noises, "QEW_F32 minimisation");
        } else if (b->noises[0].kind == TABLE && b->noises[0].dtype == QEW_F64_MID) {
            // FP64 midpoint-contracted tables on lines of their geometry
            for (const auto &l : b->lines) {
                const auto &nd = b->noises[l.noise];
                QEW_REQUIRE(nd.kind == TABLE && nd.dtype == QEW_F64_MID,
                            "all tables of a batch must share one storage type");
                QEW_REQUIRE(l.n % 2 == 0 && l.L / (double)l.n == nd.dx,
                            "a QEW_F64_MID table needs lines of even length whose site spacing "
                            "L / n equals the table's x spacing");
            }
        } else {
            require_wide_tables(b->noises);
        }
        const int tdt = b->noises[0].kind == TABLE ? b->noises[0].dtype : 0;
        for (const auto &nd : b->noises)
            QEW_REQUIRE(nd.kind != TABLE || nd.dtype == tdt,
                        "all tables of a batch must share one storage type");
        if (lp->precondition)
            for (const auto &l : b->lines) {
                int r = l.n / 2;
                for (const int p : {2, 3, 5, 7})
                    while (r > 1 && r % p == 0) r /= p;
                QEW_REQUIRE(l.n >= 4 && l.n % 2 == 0 && l.n / 2 <= kMaxFftN && r == 1,
                            "the batched preconditioner needs even line lengths in [4, 16384] "
                            "whose half has no prime factor above 7");
            }
        qew_ctx ctx = b->ctx;
        use_device(ctx);
        const size_t es = dtype == QEW_F32 ? sizeof(float) : sizeof(double);
        const size_t nvec = 7 + 2 * ((size_t)lp->m + 1);
        int maxN = 1;
        if (lp->precondition)
            for (const auto &l : b->lines) maxN = std::max(maxN, l.n / 2);
        // one twiddle table of 4 N entries per distinct transform length N, so
        // that the stage twiddles, the radix roots, W_n^k (n = 2N) and
        // sin(pi k / n) are all entries of the line's table
        std::vector<int64_t> tw_off(b->nlines, 0);
        std::vector<std::pair<int, int64_t>> tables;  // (N, offset)
        int64_t tw_total = 0;
        // lengths whose half is not a power of two need the odd-radix stages
        bool mixed = false;
        if (lp->precondition)
            for (const auto &l : b->lines) mixed = mixed || ((l.n / 2) & (l.n / 2 - 1)) != 0;
        if (lp->precondition && !mixed) {
            tables.emplace_back(maxN, 0);  // one table of 4 * max N entries serves them all
            tw_total = 4 * (int64_t)maxN;
        } else if (lp->precondition)
            for (int l = 0; l < b->nlines; ++l) {
                const int N = b->lines[l].n / 2;
                auto it = std::find_if(tables.begin(), tables.end(),
                                       [&](const std::pair<int, int64_t> &t) { return t.first == N; });
                if (it == tables.end()) {
                    tables.emplace_back(N, tw_total);
                    tw_off[l] = tw_total;
                    tw_total += 4 * (int64_t)N;
                } else {
                    tw_off[l] = it->second;
                }
            }
        // layout: [order nlines ints][counter][results][table offsets][twiddles][work]
        auto up = [](size_t v) { return (v + 255) / 256 * 256; };
        const size_t off_next = up(sizeof(int) * b->nlines);
        const size_t off_res = off_next + 256;
        const size_t off_twoff = off_res + up(sizeof(qew_lbfgs_result) * b->nlines);
        const size_t off_tw = off_twoff + up(sizeof(int64_t) * b->nlines);
        const size_t off_work = off_tw + up(2 * es * (size_t)std::max<int64_t>(tw_total, 1));
        b->d_scratch.ensure(off_work + es * nvec * (size_t)b->total);
        char *w = b->d_scratch.as<char>();
        std::vector<int> order(b->nlines);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return b->lines[x].n > b->lines[y].n; });
        QEW_HIP(hipMemcpyAsync(w, order.data(), sizeof(int) * b->nlines, hipMemcpyHostToDevice,
                               ctx->stream));
        hipLaunchKernelGGL(zero_u32, dim3(1), dim3(1), 0, ctx->stream,
                           reinterpret_cast<unsigned *>(w + off_next));
        QEW_HIP(hipMemcpyAsync(w + off_twoff, tw_off.data(), sizeof(int64_t) * b->nlines,
                               hipMemcpyHostToDevice, ctx->stream));
        for (const auto &t : tables) {
            const int M = 4 * t.first;
            if (dtype == QEW_F32)
                hipLaunchKernelGGL(twiddle_kernel<float>, dim3((M + 255) / 256), dim3(256), 0,
                                   ctx->stream, reinterpret_cast<C2<float> *>(w + off_tw) + t.second,
                                   M);
            else
                hipLaunchKernelGGL(twiddle_kernel<double>, dim3((M + 255) / 256), dim3(256), 0,
                                   ctx->stream,
                                   reinterpret_cast<C2<double> *>(w + off_tw) + t.second, M);
        }
        QEW_HIP(hipGetLastError());
        BatchArgs a{};
        a.lines = b->d_lines.as<LineDesc>();
        a.noises = b->d_noises.as<NoiseDesc>();
        a.order = reinterpret_cast<int *>(w);
        a.nlines = b->nlines;
        a.next = reinterpret_cast<unsigned *>(w + off_next);
        a.h = h;
        a.work = w + off_work;
        a.results = reinterpret_cast<qew_lbfgs_result *>(w + off_res);
        a.tw = w + off_tw;
        a.tw_m = (int)(4 * (int64_t)maxN);
        a.tw_off = reinterpret_cast<const int64_t *>(w + off_twoff);
        a.lp = *lp;
        // the FFT buffer doubles as the packed-q buffer: n reals = N complex
        const size_t lds = (size_t)maxN * 2 * es;
        auto launch = [&](auto kernel) { launch_persistent(kernel, a, b->nlines, lds, ctx); };
        if (mixed) {
            const int variant = (dtype == QEW_F64 && tdt == QEW_F64_MID)   ? 0
                                : (dtype == QEW_F32 && tdt == QEW_F32_MID) ? 1
                                : (dtype == QEW_F32)                       ? 2
                                : (tdt == 1)                               ? 3
                                                                           : 4;
            launch_lbfgs_mixed(variant, &a, b->nlines, lds, ctx);
        } else {
            if (dtype == QEW_F64 && tdt == QEW_F64_MID) launch(lbfgs_batch_kernel<double, double, true>);
            else if (dtype == QEW_F32 && tdt == QEW_F32_MID) launch(lbfgs_batch_kernel<float, float, true>);
            else if (dtype == QEW_F32) launch(lbfgs_batch_kernel<float, float>);
            else if (tdt == 1) launch(lbfgs_batch_kernel<double, float>);
            else launch(lbfgs_batch_kernel<double, double>);
        }
        QEW_HIP(hipMemcpyAsync(results, a.results, sizeof(qew_lbfgs_result) * b->nlines,
                               hipMemcpyDeviceToHost, ctx->stream));
        QEW_HIP(hipStreamSynchronize(ctx->stream));
    });
}

qew_fire_params qew_fire_default_params(void) {
    qew_fire_params p;
    p.dt_max = 0;
    p.dt_init = 0;
    p.ftol = 1e-6;
    p.max_iter = 100000;
    p.n_min = 5;
    p.f_inc = 1.1;
    p.f_dec = 0.5;
    p.alpha_start = 0.1;
    p.f_alpha = 0.99;
    return p;
}

extern "C" qew_status qew_batch_fire(qew_batch b, const qew_fire_params *fp, double *h,
                                     qew_fire_result *results) {
    return guard([&] {
        QEW_REQUIRE(b && fp && h && results, "null pointer");
        QEW_REQUIRE(fp->dt_max > 0, "FireParams.dt_max must be > 0");
        QEW_REQUIRE(fp->max_iter >= 0, "FireParams.max_iter must be >= 0");
        require_wide_tables(b->noises);
        const int tdt = b->noises[0].kind == TABLE ? b->noises[0].dtype : 0;
        for (const auto &nd : b->noises)
            QEW_REQUIRE(nd.kind != TABLE || nd.dtype == tdt,
                        "all tables of a batch must share one storage type");
        qew_ctx ctx = b->ctx;
        use_device(ctx);
        // layout: [order nlines ints][counter][results][work: g v edy esl]
        auto up = [](size_t v) { return (v + 255) / 256 * 256; };
        const size_t off_next = up(sizeof(int) * b->nlines);
        const size_t off_res = off_next + 256;
        const size_t off_work = off_res + up(sizeof(qew_fire_result) * b->nlines);
        b->d_scratch.ensure(off_work + sizeof(double) * 4 * (size_t)b->total);
        char *w = b->d_scratch.as<char>();
        std::vector<int> order(b->nlines);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return b->lines[x].n > b->lines[y].n; });
        QEW_HIP(hipMemcpyAsync(w, order.data(), sizeof(int) * b->nlines, hipMemcpyHostToDevice,
                               ctx->stream));
        hipLaunchKernelGGL(zero_u32, dim3(1), dim3(1), 0, ctx->stream,
                           reinterpret_cast<unsigned *>(w + off_next));
        QEW_HIP(hipGetLastError());
        FireArgs a{};
        a.lines = b->d_lines.as<LineDesc>();
        a.noises = b->d_noises.as<NoiseDesc>();
        a.order = reinterpret_cast<int *>(w);
        a.nlines = b->nlines;
        a.next = reinterpret_cast<unsigned *>(w + off_next);
        a.h = h;
        a.work = reinterpret_cast<double *>(w + off_work);
        a.results = reinterpret_cast<qew_fire_result *>(w + off_res);
        a.fp = *fp;
        auto launch = [&](auto kernel) {
            int per_cu = 0;
            QEW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kLB, 0));
            const int grid = std::max(1, std::min(b->nlines, std::max(1, per_cu) * ctx->num_cus));
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(kLB), 0, ctx->stream, a);
            QEW_HIP(hipGetLastError());
        };
        if (tdt == 1) launch(fire_batch_kernel<float>);
        else launch(fire_batch_kernel<double>);
        QEW_HIP(hipMemcpyAsync(results, a.results, sizeof(qew_fire_result) * b->nlines,
                               hipMemcpyDeviceToHost, ctx->stream));
        QEW_HIP(hipStreamSynchronize(ctx->stream));
    });
}

extern "C" qew_status qew_fire_minimize(qew_ctx ctx, qew_noise noise, const qew_params *p,
                                        const qew_fire_params *fp, int n, double *h,
                                        qew_fire_result *result) {
    return guard([&] {
        ctx_check(ctx);
        QEW_REQUIRE(noise && p && fp && h && result, "null pointer");
        QEW_REQUIRE(n >= 1, "line length must be >= 1");
        QEW_REQUIRE(fp->dt_max > 0, "FireParams.dt_max must be > 0");
        auto check = [](qew_status s) {
            if (s != QEW_OK) throw Error(s, qew_last_error());
        };
        qew_batch b = nullptr;
        check(qew_batch_create(ctx, 1, &n, &noise, p, &b));
        struct Release {
            qew_batch b;
            ~Release() { (void)qew_batch_destroy(b); }
        } release{b};
        check(qew_batch_fire(b, fp, h, result));
    });
}

#endif  // QEW_LBFGS_MIXED_UNIT
