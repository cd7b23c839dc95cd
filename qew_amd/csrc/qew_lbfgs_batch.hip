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
// Stockham FFT (radix 8 for FP32 state on FP32 tables, else radix 4, then at
// most one radix-4 or radix-2 stage), unpacked to the
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
#include <chrono>
#include <functional>
#include <cstdlib>
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

// Workspace vectors per line for history length m: x g gn d q edy esl, m + 1
// S and m + 1 Y rows, and two for a team's vector exchange.
__host__ __device__ constexpr int lbfgs_nvec(int m) { return 9 + 2 * (m + 1); }

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
    // slots: slot s is member slot_member[s] of the team of team_size[l]
    // workgroups that solves line l = slot_line[s]; the members of a team
    // are consecutive slots, teams ordered longest line first
    const int *slot_line;
    const int *slot_member;
    const int *team_size;
    int nslots;
    unsigned long long *xchg;  // per line [2][kMaxTeam][kXchgVals][2] tagged words
    int *abort;
};

template <typename T>
__device__ __forceinline__ C2<T> cm(C2<T> a, C2<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// for (i = tid; i < n; i += kLB) use(i, ld(i)), with the loads of kU
// consecutive iterations issued before any of their uses. Full chunks run
// without per-element guards, so the uses (and any loads inside them, such
// as table gathers) share one basic block and can overlap.
// The same over the index range [i0, i1) (i = i0 + tid, i0 + tid + kLB, ...).
template <int U = kU, typename Ld, typename Use>
__device__ __forceinline__ void stream(int i0, int i1, Ld ld, Use use) {
    for (int b = i0 + (int)threadIdx.x; b < i1; b += U * kLB) {
        decltype(ld(0)) v[U];
        if (b + (U - 1) * kLB < i1) {
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = ld(b + u * kLB);
#pragma unroll
            for (int u = 0; u < U; ++u) use(b + u * kLB, v[u]);
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = b + u * kLB;
                if (i < i1) v[u] = ld(i);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = b + u * kLB;
                if (i < i1) use(i, v[u]);
            }
        }
    }
}

template <int U = kU, typename Ld, typename Use>
__device__ __forceinline__ void stream(int n, Ld ld, Use use) {
    stream<U>(0, n, ld, use);
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

// Team exchange area and device-scope (cross-XCD coherent) scalar access,
// used by the team machinery below.
constexpr int kMaxTeam = 8;
constexpr int kXchgVals = 8;  // values one member publishes per exchange

template <typename T>
__device__ __forceinline__ void st_dev(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T ld_dev(const T *p) {
    return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- workgroup reductions (uniform results) ----
struct Red {
    double *d;
    KSum *k;
};

// dot-product accumulator: FP64 sum, or a Kahan pair for FP32 vectors
template <typename T>
struct Acc;
// With a team (below), reduce(r, tm) combines the members' workgroup
// partials; part() gives this workgroup's partial as two doubles and
// combine() the team value from every member's pair at offset j of an
// exchange's values (member k's at k * kXchgVals).
struct Team;
__device__ __forceinline__ double team_sum1(Team &tm, double v);
__device__ __forceinline__ KSum team_ksum1(Team &tm, KSum k);
template <>
struct Acc<double> {
    double s = 0;
    __device__ void add(double p) { s += p; }
    __device__ __forceinline__ double reduce(Red r) { return block_sum(s, r.d); }
    __device__ __forceinline__ double reduce(Red r, Team &tm) {
        return team_sum1(tm, block_sum(s, r.d));
    }
    __device__ __forceinline__ void part(Red r, double *o) {
        o[0] = block_sum(s, r.d);
        o[1] = 0.0;
    }
    __device__ static double combine(const double *x, int size, int j) {
        double acc = x[j];
        for (int k = 1; k < size; ++k) acc += x[k * kXchgVals + j];
        return acc;
    }
};
template <>
struct Acc<float> {
    KSum s{0.f, 0.f};
    __device__ void add(float p) { s = ksum_add(s, KSum{p, 0.f}); }
    __device__ __forceinline__ double reduce(Red r) {
        const KSum k = block_ksum(s, r.k);
        return (double)k.s + (double)k.c;
    }
    __device__ __forceinline__ double reduce(Red r, Team &tm) {
        const KSum k = team_ksum1(tm, block_ksum(s, r.k));
        return (double)k.s + (double)k.c;
    }
    __device__ __forceinline__ void part(Red r, double *o) {
        const KSum k = block_ksum(s, r.k);
        o[0] = (double)k.s;
        o[1] = (double)k.c;
    }
    __device__ static double combine(const double *x, int size, int j) {
        KSum t{(float)x[j], (float)x[j + 1]};
        for (int k = 1; k < size; ++k)
            t = ksum_add(t, KSum{(float)x[k * kXchgVals + j], (float)x[k * kXchgVals + j + 1]});
        return (double)t.s + (double)t.c;
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
__device__ __forceinline__ double wg_maxabs_seg(const T *a, int i0, int i1, Red r) {
    double m = 0;
    stream<U>(i0, i1, [&](int i) { return a[i]; }, [&](int, T v) {
        const double f = fabs((double)v);
        m = f > m ? f : m;
    });
    return block_max(m, r.d);
}

template <int U = kU, typename T>
__device__ __forceinline__ double wg_maxabs(const T *a, int n, Red r) {
    return wg_maxabs_seg<U>(a, 0, n, r);
}

// ---- teams: several workgroups on one line ----
// A line may be solved by a team of `size` workgroups, member k owning the
// sites [k n / size, (k + 1) n / size). Decisions stay uniform across the
// team: every workgroup reduces its own sites, publishes the partial, and
// every member combines all partials in member order, so all members see the
// same value.
//
// Data crosses workgroups (and XCDs, whose L2 caches are not coherent with
// each other) only in 64-bit words written and read with device-scope
// atomic stores and loads, each word holding a 32-bit tag (the team's
// exchange count, the same on every member) beside 32 payload bits. A reader
// polls each word until its tag is the one it waits for. Device-scope atomics
// are coherent per location, so a word with the right tag carries the
// payload written with it; no ordering between different words (which across
// XCDs would need L2 write-backs and invalidations) is relied on. Every wait
// is bounded: a member that gives up raises the launch's abort flag, after
// which no member waits any more and the host reports the launch as failed.
typedef unsigned long long u64;

struct Team {
    int size, member, s0, s1;
    u64 *xchg;     // this line's small exchanges: [2][kMaxTeam][kXchgVals][2 words]
    u64 *bulk;     // this line's vector exchange (preconditioner), 2n words
    int *abort;    // launch-wide flag: a wait gave up
    int epoch;     // exchanges on this line so far (the next tag is epoch + 1)
    int *s_ok;     // LDS: 1 while no wait has failed
    double *s_x;   // LDS [2][kMaxTeam * kXchgVals]: the last two small exchanges
};

// Polls of one wait before it gives up: each poll is a device-scope load
// (~1 us) and a short sleep, so this is seconds, longer than any line's
// whole minimisation (a member's first wait can last until its team is
// complete, i.e. until another line's solve ends).
constexpr long kTeamMaxPolls = 1L << 22;

__device__ __forceinline__ u64 tag_word(unsigned tag, unsigned payload) {
    return ((u64)tag << 32) | (u64)payload;
}

// The payload of word w once its tag is `tag`; `polls` counts this thread's
// polls in the current wait; gives up (ok = false, abort raised) past the
// limit or once any other wait of the launch gave up.
__device__ __forceinline__ unsigned wait_word(const u64 *w, unsigned tag, long &polls, bool &ok,
                                              int *abort) {
    for (;;) {
        const u64 u = ld_dev(w);
        if ((unsigned)(u >> 32) == tag) return (unsigned)u;
        if (!ok) return 0u;  // after a failed wait: zeros, never stale payloads
        if (++polls > kTeamMaxPolls || ((polls & 63) == 0 && ld_dev(abort) != 0)) {
            st_dev(abort, 1);
            ok = false;
            return 0u;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

__device__ __forceinline__ void put_f64(u64 *w, unsigned tag, double v) {
    const u64 b = (u64)__double_as_longlong(v);
    st_dev(w, tag_word(tag, (unsigned)b));
    st_dev(w + 1, tag_word(tag, (unsigned)(b >> 32)));
}

__device__ __forceinline__ double get_f64(const u64 *w, unsigned tag, long &polls, bool &ok,
                                          int *abort) {
    const unsigned lo = wait_word(w, tag, polls, ok, abort);
    const unsigned hi = wait_word(w + 1, tag, polls, ok, abort);
    return __longlong_as_double((long long)(((u64)hi << 32) | lo));
}

// One line element of type T through the vector exchange (1 word for FP32,
// 2 for FP64, element i at i or 2i).
template <typename T>
__device__ __forceinline__ void put_elem(u64 *xw, int i, unsigned tag, T v) {
    if constexpr (sizeof(T) == 4) st_dev(xw + i, tag_word(tag, __float_as_uint(v)));
    else put_f64(xw + 2 * (size_t)i, tag, v);
}
template <typename T>
__device__ __forceinline__ T get_elem(const u64 *xw, int i, unsigned tag, long &polls, bool &ok,
                                      int *abort) {
    if constexpr (sizeof(T) == 4) return __uint_as_float(wait_word(xw + i, tag, polls, ok, abort));
    else return get_f64(xw + 2 * (size_t)i, tag, polls, ok, abort);
}

// Every member's v[0..NV) of this exchange, for all threads of the
// workgroup (in LDS, member k's value j at k * kXchgVals + j). Two slots
// alternate, in memory and in LDS: a member publishes into a memory slot
// only after it has read every member's values of the exchange before, which
// every member publishes only after reading the slot's previous contents;
// a wave writes an LDS slot only after every wave has passed the barrier of
// the exchange before, i.e. finished reading the slot's previous contents.
template <int NV>
__device__ __forceinline__ const double *team_exchange(Team &tm, const double *v) {
    const unsigned tag = (unsigned)++tm.epoch;
    u64 *sl = tm.xchg + (size_t)(tag & 1u) * kMaxTeam * kXchgVals * 2;
    double *out = tm.s_x + (tag & 1u) * kMaxTeam * kXchgVals;
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < NV; ++j) put_f64(sl + 2 * (tm.member * kXchgVals + j), tag, v[j]);
    if (threadIdx.x < kWave) {
        bool ok = *tm.s_ok != 0;
        long polls = 0;
        for (int idx = threadIdx.x; idx < tm.size * NV; idx += kWave) {
            const int k = idx / NV, j = idx % NV;
            out[k * kXchgVals + j] = get_f64(sl + 2 * (k * kXchgVals + j), tag, polls, ok, tm.abort);
        }
        if (!ok) *tm.s_ok = 0;
    }
    __syncthreads();
    return out;
}

// v[j] <- sum over members (in member order) of their v[j], j < NV.
template <int NV>
__device__ __forceinline__ void team_sum(Team &tm, double *v) {
    if (tm.size == 1) return;
    const double *x = team_exchange<NV>(tm, v);
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        double acc = x[j];
        for (int k = 1; k < tm.size; ++k) acc += x[k * kXchgVals + j];
        v[j] = acc;
    }
}

__device__ __forceinline__ double team_max(Team &tm, double a) {
    if (tm.size == 1) return a;
    const double *x = team_exchange<1>(tm, &a);
    double m = x[0];
    for (int k = 1; k < tm.size; ++k) m = x[k * kXchgVals] > m ? x[k * kXchgVals] : m;
    return m;
}

__device__ __forceinline__ double team_sum1(Team &tm, double v) {
    team_sum<1>(tm, &v);
    return v;
}

// Kahan pairs combined in member order
__device__ __forceinline__ KSum team_ksum(Team &tm, KSum k) {
    if (tm.size == 1) return k;
    const double v[2] = {(double)k.s, (double)k.c};
    const double *x = team_exchange<2>(tm, v);
    KSum t{(float)x[0], (float)x[1]};
    for (int m = 1; m < tm.size; ++m)
        t = ksum_add(t, KSum{(float)x[m * kXchgVals], (float)x[m * kXchgVals + 1]});
    return t;
}

__device__ __forceinline__ KSum team_ksum1(Team &tm, KSum k) { return team_ksum(tm, k); }

// ---- objective and gradient of one line ----
// table gathers in flight per thread: 16 coefficients each, held as FP64 for
// FP64 state
template <typename T>
constexpr int kUG = sizeof(T) == 8 ? 2 : 4;

// A team member's x and d are private copies of its segment with one halo
// site on each side: the pointers it passes index sites s0 - 1 .. s1 (its
// left neighbour's last site, its own, its right neighbour's first), so the
// right site of its edge i is at i + 1 even for the line's last edge. One
// workgroup on the whole line indexes the wrapped neighbour ip.
#define QEW_XR(i, ip) (x[TEAM ? (i) + 1 : (ip)])
#define QEW_DR(i, ip) (d[TEAM ? (i) + 1 : (ip)])

// This workgroup's share (sites [s0, s1)) of the energy at x + alpha d.
template <typename T, typename C, bool MID = false, bool TEAM = false>
__device__ double wg_objective(const LineGeom &G, const T *x, const T *d, double alpha, Red r,
                               const Team &tm) {
    const int n = G.n;
    double acc = 0;
    if constexpr (sizeof(T) == 8) {
        // src/qew_model.h:53-99 (TrialAt: x0(i) + alpha * d(i))
        stream<kUG<T>>((TEAM ? tm.s0 : 0), (TEAM ? tm.s1 : n), [&](int i) {
                   const int ip = (i + 1 == n) ? 0 : i + 1;
                   return P2<double>{d ? x[i] + alpha * d[i] : x[i],
                                     d ? QEW_XR(i, ip) + alpha * QEW_DR(i, ip) : QEW_XR(i, ip)};
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
        stream<kUG<T>>((TEAM ? tm.s0 : 0), (TEAM ? tm.s1 : n), [&](int i) {
                   const int ip = (i + 1 == n) ? 0 : i + 1;
                   return P2<double>{d ? (double)__builtin_fmaf(al, d[i], x[i]) : (double)x[i],
                                     d ? (double)__builtin_fmaf(al, QEW_DR(i, ip), QEW_XR(i, ip))
                                       : (double)QEW_XR(i, ip)};
               },
               [&](int i, P2<double> v) {
                   const int ip = (i + 1 == n) ? 0 : i + 1;
                   const double oi = v.a, oip = v.b;
                   const double hi = G.anchor + oi, hip = G.anchor + oip;
                   const double dh = divr(oip - oi, G.dx, G.rdx);
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

template <typename T, bool TEAM>
__device__ __forceinline__ void objective_trials_mid(const LineGeom &G, const T *x, const T *d,
                                                     const double *alpha, double (&acc)[kLS],
                                                     const Team &tm) {
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
        (TEAM ? tm.s0 : 0), (TEAM ? tm.s1 : n),
        [&](int i) {
            const int ip = (i + 1 == n) ? 0 : i + 1;
            Item it;
            it.v = P4<T>{x[i], d[i], QEW_XR(i, ip), QEW_DR(i, ip)};
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
                const double dh = divr(oip - oi, G.dx, G.rdx);
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

template <typename T, typename C, bool MID = false, bool TEAM = false>
__device__ __forceinline__ void objective_trials_table(const LineGeom &G, const T *x, const T *d,
                                                       const double *alpha, double (&acc)[kLS],
                                                       const Team &tm) {
    const int n = G.n;
    constexpr int K = kLS;
    if constexpr (MID) {
        objective_trials_mid<T, TEAM>(G, x, d, alpha, acc, tm);
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
        (TEAM ? tm.s0 : 0), (TEAM ? tm.s1 : n),
        [&](int i) {
            const int ip = (i + 1 == n) ? 0 : i + 1;
            Item it;
            it.v = P4<T>{x[i], d[i], QEW_XR(i, ip), QEW_DR(i, ip)};
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
                    const double dh = divr(hip - hi, G.dx, G.rdx);
                    le = (G.model == 0) ? G.lt * dh * dh * 0.5 : G.lt * (sqrt(1.0 + dh * dh) - 1.0);
                } else {
                    const float al = (float)alpha[j];
                    const double oi = (double)__builtin_fmaf(al, it.v.b, it.v.a);
                    const double oip = (double)__builtin_fmaf(al, it.v.d, it.v.c);
                    const double dh = divr(oip - oi, G.dx, G.rdx);
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
template <typename T, typename C, bool MID = false, bool TEAM = false>
__device__ __forceinline__ void wg_objective_trials(const LineGeom &G, const T *x, const T *d,
                                    const double *alpha, double *phi, Red r, Team &tm) {
    const int n = G.n;
    constexpr int K = kLS;
    double acc[K];
#pragma unroll
    for (int j = 0; j < K; ++j) acc[j] = 0;
    if (G.nd.kind == TABLE) {
        objective_trials_table<T, C, MID, TEAM>(G, x, d, alpha, acc, tm);
        block_sumk<K>(acc, r.d);
        if constexpr (TEAM) team_sum<K>(tm, acc);
#pragma unroll
        for (int j = 0; j < K; ++j) phi[j] = acc[j];
        return;
    }
    stream<1>(
        (TEAM ? tm.s0 : 0), (TEAM ? tm.s1 : n), [&](int i) {
            const int ip = (i + 1 == n) ? 0 : i + 1;
            return P4<T>{x[i], d[i], QEW_XR(i, ip), QEW_DR(i, ip)};
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
                    const double dh = divr(oip - oi, G.dx, G.rdx);
                    const double s2 = dh * dh;
                    const double le = (G.model == 0) ? G.lt * s2 * 0.5
                                                     : G.lt * (s2 / (sqrt(1.0 + s2) + 1.0));
                    acc[j] += G.dx * (le + sample_v64<float>(G.nd, xc, (hip + hi) * 0.5) - G.f * hi);
                }
            }
        });
    block_sumk<K>(acc, r.d);
    if constexpr (TEAM) team_sum<K>(tm, acc);
#pragma unroll
    for (int j = 0; j < K; ++j) phi[j] = acc[j];
}

// Gradient at the sites [s0, s1) of the line (the whole line for one
// workgroup). Edge pass: dV/dy (and the arclength slope) of the edges (i,
// i + 1); a team member also forms its left halo edge (s0 - 1, s0) from its
// halo site, into its private edy / esl copies (indexed like x, from s0 - 1).
// Site pass: g(i) from the edges i - 1 and i.
template <typename T, typename C, bool MID = false, bool TEAM = false>
__device__ __forceinline__ void wg_gradient(const LineGeom &G, const T *x, T *g, T *edy, T *esl,
                                            const Team &tm) {
    const int n = G.n;
    const bool arc = G.model == 1;
    const int s0 = TEAM ? tm.s0 : 0, s1 = TEAM ? tm.s1 : n;
    stream<kUG<T>>(TEAM ? s0 - 1 : s0, s1, [&](int e) {
               const int ip = (e + 1 == n) ? 0 : e + 1;
               return P2<T>{x[e], QEW_XR(e, ip)};
           },
           [&](int e, P2<T> v) {
               // the edge's place on the line (e = -1 is the wrap edge n - 1)
               const int i = e < 0 ? n - 1 : e;
               const int ip = (i + 1 == n) ? 0 : i + 1;
               if constexpr (sizeof(T) == 8) {
                   const double dh = divr(v.b - v.a, G.dx, G.rdx);
                   if constexpr (MID)
                       edy[e] = mid_dy64<double>(G.nd, mid_column(G.nd, i, ip), (v.b + v.a) * 0.5);
                   else
                       edy[e] = sample_dy64<C>(G.nd, (double)(i + ip) * 0.5 * G.dx, (v.b + v.a) * 0.5);
                   if (arc) esl[e] = dh / sqrt(1.0 + dh * dh);
               } else {
                   float vv, dy;
                   if constexpr (MID) edge32_mid<false, true>(G, i, ip, v.a, v.b, vv, dy);
                   else edge32_grid<false, true>(G, i, ip, v.a, v.b, vv, dy);
                   edy[e] = dy;
                   const float dh = (v.b - v.a) * G.inv_dx;
                   if (arc) esl[e] = dh * rsqrtf(1.0f + dh * dh);
               }
           });
    __syncthreads();
    struct Nb {
        T xm, xi, xp, em, ei, sm, si;
    };
    stream<4>(s0, s1, [&](int i) {
               const int ip = (i + 1 == n) ? 0 : i + 1;
               const int im = TEAM ? i - 1 : (i == 0) ? n - 1 : i - 1;
               Nb b;
               b.em = edy[im];
               b.ei = edy[i];
               if (!arc) {
                   b.xm = x[im];
                   b.xi = x[i];
                   b.xp = QEW_XR(i, ip);
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
                   if (!arc) lf = divr(-G.lt * (b.xm + b.xp - 2.0 * b.xi), G.dx * G.dx, G.rdx2);
                   else lf = divr(G.lt * (b.sm - b.si), G.dx, G.rdx);
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
        } else if constexpr (R == 8) {
            // radix 2 x 4: a = u_q + u_{q+4} and b = (u_q - u_{q+4}) W_8^q
            // (q = 0..3) give the even and the odd outputs as 4-point DFTs
            const T h = (T)0.70710678118654752440;  // cos(pi / 4)
            auto rot = [&](const C2<T> &e) {  // e times W_4 = e^(sign i pi / 2)
                return (sign < 0) ? C2<T>{e.y, -e.x} : C2<T>{-e.y, e.x};
            };
            C2<T> a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a[q] = {u[q].x + u[q + 4].x, u[q].y + u[q + 4].y};
                b[q] = {u[q].x - u[q + 4].x, u[q].y - u[q + 4].y};
            }
            // W_8 = h (1 + sign i), W_8^2 = W_4, W_8^3 = h (-1 + sign i)
            {
                const C2<T> r1 = rot(b[1]);
                b[1] = {h * (b[1].x + r1.x), h * (b[1].y + r1.y)};
                b[2] = rot(b[2]);
                const C2<T> r3 = rot(b[3]);
                b[3] = {h * (r3.x - b[3].x), h * (r3.y - b[3].y)};
            }
            auto dft4 = [&](const C2<T> (&x)[4], int o0) {
                const C2<T> p = {x[0].x + x[2].x, x[0].y + x[2].y};
                const C2<T> m = {x[0].x - x[2].x, x[0].y - x[2].y};
                const C2<T> c = {x[1].x + x[3].x, x[1].y + x[3].y};
                const C2<T> e = rot(C2<T>{x[1].x - x[3].x, x[1].y - x[3].y});
                buf[idxD + o0 * Ns] = {p.x + c.x, p.y + c.y};
                buf[idxD + (o0 + 2) * Ns] = {m.x + e.x, m.y + e.y};
                buf[idxD + (o0 + 4) * Ns] = {p.x - c.x, p.y - c.y};
                buf[idxD + (o0 + 6) * Ns] = {m.x - e.x, m.y - e.y};
            };
            dft4(a, 0);
            dft4(b, 1);
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

// Radix-8 stages for FP32 state on FP32 tables: there the transform's five
// rounding stages instead of seven bring the directions close enough to
// FP64's that the line search needs fewer trials. FP64 state measured no
// faster with them and keeps its iterates; FP32 state on midpoint tables
// keeps radix 4, with which every minimum of the batched workload lands in
// its FP64 basin (with radix 8 one line settles in a nearby one).
template <typename T, bool MID>
constexpr bool kR8 = sizeof(T) == 4 && !MID;

// Radices in the order 8 (R8), 4, 2 and, in the MIXED instantiation
// (lengths with factors 3, 5 or 7; its larger butterflies would raise the
// register pressure of the power-of-two kernels), then 3, 5, 7. The host accepts only lengths with
// no other prime factor; any other remainder stops the loop rather than spin.
template <typename T, bool MIXED, bool R8>
__device__ void lds_fft(C2<T> *buf, int N, int sign, const C2<T> *tw, int M) {
    int Ns = 1;
    if constexpr (!MIXED) {
        while (Ns < N) {
            if (R8 && (N / Ns) % 8 == 0) {
                fft_stage<T, 8>(buf, N, Ns, sign, tw, M);
                Ns *= 8;
            } else if ((N / Ns) % 4 == 0) {
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
        if (R8 && rest % 8 == 0) {
            fft_stage<T, 8>(buf, N, Ns, sign, tw, M);
            Ns *= 8;
        } else if (rest % 4 == 0) {
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
template <typename T, bool MIXED, bool R8>
__device__ void wg_precondition(int n, double kdx, double shift, C2<T> *buf, const C2<T> *tw,
                                int M) {
    const int N = n / 2;
    lds_fft<T, MIXED, R8>(buf, N, -1, tw, M);
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
    lds_fft<T, MIXED, R8>(buf, N, +1, tw, M);
    // complex j is reals 2j, 2j + 1: each thread scales its own entries in place
    const T inv = (T)(1.0 / (double)N);
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const C2<T> z = buf[j];
        buf[j] = {z.x * inv, z.y * inv};
    }
    __syncthreads();
}

// Team version of wg_precondition: the members send their q sites through
// the line's vector exchange, the leader (member 0) receives the rest of the
// line into its LDS buffer, transforms it and sends back the others' sites of
// d, which they receive into their own buffers.
template <typename T, bool MIXED, bool R8>
__device__ __forceinline__ void team_precondition(Team &tm, int n, double kdx, double shift,
                                                  C2<T> *buf, const C2<T> *tw, int M) {
    T *fb = reinterpret_cast<T *>(buf);
    const unsigned tag_q = (unsigned)++tm.epoch, tag_d = (unsigned)++tm.epoch;
    bool ok = *tm.s_ok != 0;
    long polls = 0;
    if (tm.member == 0) {
        for (int i = tm.s1 + (int)threadIdx.x; i < n; i += kLB)
            fb[i] = get_elem<T>(tm.bulk, i, tag_q, polls, ok, tm.abort);
        if (!ok) *tm.s_ok = 0;
        __syncthreads();
        wg_precondition<T, MIXED, R8>(n, kdx, shift, buf, tw, M);
        for (int i = tm.s1 + (int)threadIdx.x; i < n; i += kLB) put_elem<T>(tm.bulk, i, tag_d, fb[i]);
    } else {
        for (int i = tm.s0 + (int)threadIdx.x; i < tm.s1; i += kLB)
            put_elem<T>(tm.bulk, i, tag_q, fb[i]);
        for (int i = tm.s0 + (int)threadIdx.x; i < tm.s1; i += kLB)
            fb[i] = get_elem<T>(tm.bulk, i, tag_d, polls, ok, tm.abort);
        if (!ok) *tm.s_ok = 0;
        __syncthreads();
    }
}

// The first and last site values of d and g of every member, exchanged so
// that each member holds its neighbours' d (the halo of its boundary edges)
// and g (for a restart with d = -g).
template <typename T>
__device__ __forceinline__ void team_halo_dg(Team &tm, T *d, const T *g, T *hal_g) {
    const double v[4] = {(double)d[tm.s0], (double)d[tm.s1 - 1], (double)g[tm.s0],
                         (double)g[tm.s1 - 1]};
    const double *x = team_exchange<4>(tm, v);
    if (threadIdx.x == 0) {
        const double *lft = x + ((tm.member + tm.size - 1) % tm.size) * kXchgVals;
        const double *rgt = x + ((tm.member + 1) % tm.size) * kXchgVals;
        d[tm.s0 - 1] = (T)lft[1];
        d[tm.s1] = (T)rgt[0];
        hal_g[0] = (T)lft[3];
        hal_g[1] = (T)rgt[2];
    }
    __syncthreads();
}

// Sites [s0, s1) of one line, solved by one workgroup (TEAM false: the
// whole line) or as member tm.member of a team. A member's x, d, edy and esl
// are private copies of its segment with a halo site on each side (each
// line's workspace vectors have 2 kMaxTeam spare elements for them); it forms
// its neighbours' x updates itself (x + alpha d, the same arithmetic as
// theirs, so the same values) and gets their d (and g, in hal_g = {g[s0 - 1],
// g[s1]}, for a restart with d = -g) at one exchange per iteration.
template <typename T, typename C, bool MID = false, bool MIXED = false, bool TEAM = false>
__device__ __forceinline__ void solve_line(const BatchArgs &a, int line, Red r, double *rho, double *la,
                           int *slot, int *sbuf, C2<T> *fftbuf, Team &tm, T *hal_g) {
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
    const int s0 = TEAM ? tm.s0 : 0, s1 = TEAM ? tm.s1 : n;
    constexpr bool team = TEAM;
    // this workgroup's share of a sum, combined over the team
    auto red = [&](Acc<T> &acc) {
        if constexpr (TEAM) return acc.reduce(r, tm);
        else return acc.reduce(r);
    };
    const C2<T> *tw = static_cast<const C2<T> *>(a.tw);
    if constexpr (MIXED)
        if (lp.precondition) tw += a.tw_off[line];
    double *h = a.h + L.off;
    const double anchor = (sizeof(T) == 8) ? 0.0 : h[0];
    const LineGeom G = make_geom(L, nd, anchor);
    // per-line workspace: x g gn d q edy esl, (m + 1) S and (m + 1) Y
    // buffers and two for a team's vector exchange, each of n + 2 kMaxTeam
    // elements (lines back to back, in batch order); a team member's x, d,
    // edy, esl start at its private block
    const size_t np = (size_t)n + 2 * kMaxTeam;
    T *ws = static_cast<T *>(a.work) +
            ((size_t)L.off + 2 * (size_t)kMaxTeam * (size_t)line) * (size_t)lbfgs_nvec(m);
    T *x = ws, *g = ws + np, *gn = ws + 2 * np, *d = ws + 3 * np, *q = ws + 4 * np,
      *edy = ws + 5 * np, *esl = ws + 6 * np;
    if constexpr (TEAM) {
        const ptrdiff_t pb = (ptrdiff_t)tm.member * (tm.s1 - tm.s0 + 2) + 1 - tm.s0;
        x += pb;
        d += pb;
        edy += pb;
        esl += pb;
        // 8-byte aligned words in the last two vectors: n words for FP32
        // elements, 2n for FP64 (the two vectors hold 2 np elements)
        const uintptr_t bw = reinterpret_cast<uintptr_t>(ws + (size_t)(lbfgs_nvec(m) - 2) * np);
        tm.bulk = reinterpret_cast<u64 *>((bw + 7) & ~(uintptr_t)7);
    }
    auto Sb = [&](int k) { return ws + (size_t)(7 + k) * np; };
    auto Yb = [&](int k) { return ws + (size_t)(7 + (m + 1) + k) * np; };
    for (int i = s0 + tid; i < s1; i += blockDim.x) x[i] = (T)(h[i] - anchor);
    if (tid == 0) {
        for (int k = 0; k <= m; ++k) sbuf[k] = k;  // sbuf[m] is the spare (s_tmp / y_tmp)
        if (team) {
            x[s0 - 1] = (T)(h[(s0 + n - 1) % n] - anchor);
            x[s1] = (T)(h[s1 % n] - anchor);
        }
    }
    __syncthreads();

    double kdx = 0, shift = 0;
    if (lp.precondition) {
        const double pi = 3.14159265358979323846;
        kdx = L.lt / G.dx;
        shift = L.shift > 0 ? L.shift : lp.precond_shift;
        if (shift <= 0) {
            const double s1n = sin(pi / (double)n);
            shift = kdx * 4 * s1n * s1n;
        }
    }

    wg_gradient<T, C, MID, TEAM>(G, x, g, edy, esl, tm);
    qew_lbfgs_result res{};
    int nslots = 0;  // slot[0..nslots) chronological rows, slot[nslots-1] newest
    double gamma = 1.0;
    int ls_fail_streak = 0;
    double phi = 0;
    bool finished = false;
    for (int it = 0; it < lp.max_iter; ++it) {
        if (TEAM && !*tm.s_ok) {  // a team wait timed out: the launch reports failure
            res.iterations = it;
            res.converged = 0;
            finished = true;
            break;
        }
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
            stream<UV>(s0, s1, [&](int i) { return P2<T>{g[i], S0 ? S0[i] : (T)0}; },
                   [&](int i, P2<T> v) {
                       const double f = fabs((double)v.a);
                       gmax = f > gmax ? f : gmax;
                       qv[i] = v.a;
                       if (S0) acc.add(v.b * v.a);
                   });
            gmax = block_max(gmax, r.d);
            if constexpr (!TEAM) {
                if (S0) al_dot = acc.reduce(r);
            } else {
                double o[3] = {gmax, 0.0, 0.0};
                acc.part(r, o + 1);
                const double *xv = team_exchange<3>(tm, o);
                gmax = xv[0];
                for (int k = 1; k < tm.size; ++k) {
                    const double gk = xv[k * kXchgVals];
                    gmax = gk > gmax ? gk : gmax;
                }
                if (S0) al_dot = Acc<T>::combine(xv, tm.size, 1);
            }
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
                stream<UV>(s0, s1, [&](int i) { return P3<T>{qv[i], Y[i], Sn ? Sn[i] : (T)0}; },
                       [&](int i, P3<T> v) {
                           const T qn = v.a + c * v.b;
                           qv[i] = qn;
                           if (Sn) acc.add(v.c * qn);
                       });
                if (Sn) al_dot = red(acc);
                else __syncthreads();
            }
        };
        if (pre) {
            first_loop(fb);
            if constexpr (TEAM)
                team_precondition<T, MIXED, kR8<T, MID>>(tm, n, kdx, shift, fftbuf, tw, MIXED ? 2 * n : a.tw_m);
            else
                wg_precondition<T, MIXED, kR8<T, MID>>(n, kdx, shift, fftbuf, tw, MIXED ? 2 * n : a.tw_m);
        } else {
            first_loop(q);
            const T g0 = (T)(nslots == 0 ? 1.0 : gamma);
            stream<UV>(s0, s1, [&](int i) { return q[i]; }, [&](int i, T v) { d[i] = g0 * v; });
            __syncthreads();
        }
        // second loop, oldest pair first: d += (la - b) S with b = rho Y . d;
        // the last pass also negates d, forms g . d and stores d to memory
        double dphi0 = 0;
        auto second_loop = [&](T *dv) {
            double b_dot = 0;
            if (hs > 0) {
                const T *Y0 = Yb(sbuf[slot[0]]);
                Acc<T> acc;
                stream<UV>(s0, s1, [&](int i) { return P2<T>{Y0[i], dv[i]}; },
                           [&](int, P2<T> v) { acc.add(v.a * v.b); });
                b_dot = red(acc);
            }
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
                    stream<UV>(s0, s1, [&](int i) { return P3<T>{dv[i], S[i], Yn[i]}; },
                           [&](int i, P3<T> v) {
                               T dn = v.a + c * v.b;
                               if (neg) dn = -dn;
                               dout[i] = dn;
                               acc.add(v.c * dn);
                           });
                };
                if (last) pass(d, true);
                else pass(dv, false);
                const double s = red(acc);
                if (last) dphi0 = s;
                else b_dot = s;
            }
            if (hs == 0) {
                Acc<T> acc;
                stream<UV>(s0, s1, [&](int i) { return P2<T>{dv[i], g[i]}; },
                       [&](int i, P2<T> v) {
                           const T dn = -v.a;
                           d[i] = dn;
                           acc.add(v.b * dn);
                       });
                dphi0 = red(acc);
            }
        };
        if (pre) second_loop(fb);
        else second_loop(d);
        if constexpr (TEAM) team_halo_dg(tm, d, g, hal_g);
        if (dphi0 >= 0) {
            nslots = 0;
            Acc<T> acc;
            stream<UV>(s0, s1, [&](int i) { return g[i]; }, [&](int i, T v) {
                d[i] = -v;
                acc.add(v * v);
            });
            dphi0 = -red(acc);
            if (team && tid == 0) {
                d[s0 - 1] = -hal_g[0];
                d[s1] = -hal_g[1];
            }
            __syncthreads();
        }
        if (it == 0) {
            phi = wg_objective<T, C, MID, TEAM>(G, x, nullptr, 0.0, r, tm);
            if constexpr (TEAM) phi = team_sum1(tm, phi);
            ++res.n_func_evals;
        }
        const double phi0 = phi;
        double alpha = 1.0;
        if (!lp.precondition && nslots == 0) {
            Acc<T> acc;
            stream<UV>(s0, s1, [&](int i) { return g[i]; }, [&](int, T v) { acc.add(v * v); });
            alpha = 1.0 / fmax(1.0, sqrt(red(acc)));
        }
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
            wg_objective_trials<T, C, MID, TEAM>(G, x, d, al, ph, r, tm);
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
        stream<UV>(s0, s1, [&](int i) { return P2<T>{x[i], d[i]}; }, [&](int i, P2<T> v) {
            const T xn = v.a + alT * v.b;
            x[i] = xn;
            st[i] = (sizeof(T) == 8) ? alT * v.b : xn - v.a;
        });
        if (team && tid == 0) {
            // the neighbours' boundary sites take the same step
            x[s0 - 1] = x[s0 - 1] + alT * d[s0 - 1];
            x[s1] = x[s1] + alT * d[s1];
        }
        __syncthreads();
        wg_gradient<T, C, MID, TEAM>(G, x, gn, edy, esl, tm);
        double sy3[3] = {0, 0, 0};
        stream<UV>(s0, s1, [&](int i) { return P3<T>{st[i], gn[i], g[i]}; }, [&](int i, P3<T> v) {
            const T y = v.b - v.c;
            yt[i] = y;
            sy3[0] += (double)v.a * (double)y;
            sy3[1] += (double)v.a * (double)v.a;
            sy3[2] += (double)y * (double)y;
        });
        block_sum3(sy3[0], sy3[1], sy3[2], r.d);
        if constexpr (TEAM) team_sum<3>(tm, sy3);
        const double sy = sy3[0], ss = sy3[1], yy = sy3[2];
        if (sizeof(T) == 4 && ss == 0.0) {
            // FP32 state: the accepted step is below the offsets' resolution
            // everywhere, so the line can no longer move; stop (not
            // converged) instead of iterating in place until max_iter
            res.iterations = it + 1;
            res.max_force = wg_maxabs_seg<UV>(gn, s0, s1, r);
            if constexpr (TEAM) res.max_force = team_max(tm, res.max_force);
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
        res.max_force = wg_maxabs_seg<UV>(g, s0, s1, r);
        if constexpr (TEAM) res.max_force = team_max(tm, res.max_force);
        res.converged = 0;
    }
    for (int i = s0 + tid; i < s1; i += blockDim.x) h[i] = anchor + (double)x[i];
    if (tid == 0 && (!TEAM || tm.member == 0)) a.results[line] = res;
    __syncthreads();
}

template <typename T, typename C, bool MID = false, bool MIXED = false>
__global__ __launch_bounds__(kLB) void lbfgs_batch_kernel(BatchArgs a) {
    extern __shared__ unsigned char lds[];
    __shared__ double s_red[kLS * (kLB / kWave)];
    __shared__ KSum s_kred[kLB / kWave];
    __shared__ double rho[kMaxM], la[kMaxM];
    __shared__ int slot[kMaxM], sbuf[kMaxM + 1];
    __shared__ T s_hal_g[2];
    __shared__ double s_x[2 * kMaxTeam * kXchgVals];
    __shared__ int s_slot, s_ok;
    Red r{s_red, s_kred};
    for (;;) {
        if (threadIdx.x == 0) s_slot = (int)atomicAdd(a.next, 1u);
        __syncthreads();
        // wave-uniform values in scalar registers: the slot, its line and
        // team, and everything indexed by them
        const int si = __builtin_amdgcn_readfirstlane(s_slot);
        __syncthreads();
        if (si >= a.nslots) return;
        // slot si: member slot_member[si] of the team of team_size[line]
        // workgroups that solves line slot_line[si]
        const int line = __builtin_amdgcn_readfirstlane(a.slot_line[si]);
        const int member = __builtin_amdgcn_readfirstlane(a.slot_member[si]);
        const int size = __builtin_amdgcn_readfirstlane(a.team_size[line]);
        const int seg = __builtin_amdgcn_readfirstlane(a.lines[line].n) / size;
        Team tm{size,
                member,
                member * seg,
                (member + 1) * seg,
                a.xchg + (size_t)line * 2 * kMaxTeam * kXchgVals * 2,
                nullptr,  // the vector exchange, in the line's workspace
                a.abort,
                0,
                &s_ok,
                s_x};
        if (threadIdx.x == 0) s_ok = 1;
        __syncthreads();
        if (size == 1)
            solve_line<T, C, MID, MIXED, false>(a, line, r, rho, la, slot, sbuf,
                                                reinterpret_cast<C2<T> *>(lds), tm, s_hal_g);
        else
            solve_line<T, C, MID, MIXED, true>(a, line, r, rho, la, slot, sbuf,
                                               reinterpret_cast<C2<T> *>(lds), tm, s_hal_g);
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
__global__ void zero_u32_n(unsigned *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Workgroups per line in the batched L-BFGS, from a policy of (minimum
// length, team size) rows, longest first: a line takes the team size of the
// first row whose minimum it reaches (1 if none), if that size divides it
// into segments of at least 64 sites. QEW_LBFGS_TEAM=k (read per call) sets
// k for every line; QEW_LBFGS_TEAMS="n1:k1,n2:k2,..." sets the rows.
struct TeamPolicy {
    std::vector<std::pair<int, int>> rows;  // (min length, team size), decreasing length
    int forced = 0;
};

TeamPolicy lbfgs_team_policy() {
    TeamPolicy p;
    // measured best on the 512-line batch of lengths 1024 ... 16384
    p.rows = {{16384, 2}};
    if (const char *e = std::getenv("QEW_LBFGS_TEAM")) p.forced = std::atoi(e);
    if (const char *e = std::getenv("QEW_LBFGS_TEAMS")) {
        p.rows.clear();
        const char *c = e;
        while (*c) {
            char *end = nullptr;
            const long len = std::strtol(c, &end, 10);
            if (end == c || *end != ':') break;
            const long k = std::strtol(end + 1, &end, 10);
            p.rows.emplace_back((int)len, (int)k);
            c = (*end == ',') ? end + 1 : end;
            if (!*end) break;
        }
        std::sort(p.rows.begin(), p.rows.end(),
                  [](const std::pair<int, int> &x, const std::pair<int, int> &y) {
                      return x.first > y.first;
                  });
    }
    return p;
}

// Policies the tuner times for a batch: one workgroup per line first, then
// teams of 2 or 4 on the longest length present and of 2 on the next one
// (lengths of at least 4096 sites; shorter lines gain nothing from a team).
std::vector<TeamPolicy> lbfgs_team_candidates(const std::vector<LineDesc> &lines) {
    std::vector<int> lens;
    for (const auto &l : lines) lens.push_back(l.n);
    std::sort(lens.begin(), lens.end(), std::greater<int>());
    lens.erase(std::unique(lens.begin(), lens.end()), lens.end());
    std::vector<TeamPolicy> out(1);  // all single
    if (lens.empty() || lens[0] < 4096) return out;
    const int l1 = lens[0], l2 = lens.size() > 1 && lens[1] >= 4096 ? lens[1] : 0;
    for (int k1 : {2, 4}) {
        TeamPolicy p;
        p.rows = {{l1, k1}};
        out.push_back(p);
        if (l2) {
            p.rows = {{l1, k1}, {l2, 2}};
            out.push_back(p);
        }
    }
    return out;
}

// Per-line team sizes from the site-iterations n (iterations + 1) each line
// took on one workgroup: 4 workgroups on the lines within a factor f4 of the
// largest, 2 within f2, 1 otherwise (where the size divides the line into
// segments of at least 64 sites), for (f4, f2) in a few splits.
std::vector<std::vector<int>> lbfgs_work_team_candidates(const std::vector<LineDesc> &lines,
                                                         const std::vector<qew_lbfgs_result> &r) {
    std::vector<double> work(lines.size());
    double wmax = 0;
    for (size_t l = 0; l < lines.size(); ++l) {
        work[l] = (double)lines[l].n * (double)(r[l].iterations + 1);
        wmax = std::max(wmax, work[l]);
    }
    auto fits = [](int n, int k) { return n % k == 0 && n / k >= 64; };
    std::vector<std::vector<int>> out;
    for (const auto &f : {std::make_pair(0.7, 0.35), std::make_pair(0.5, 0.3),
                          std::make_pair(2.0, 0.45)}) {
        std::vector<int> t(lines.size(), 1);
        for (size_t l = 0; l < lines.size(); ++l) {
            const int n = lines[l].n;
            if (work[l] >= f.first * wmax && fits(n, 4)) t[l] = 4;
            else if (work[l] >= f.second * wmax && fits(n, 2)) t[l] = 2;
        }
        out.push_back(t);
    }
    return out;
}

int lbfgs_team_size(int n, const TeamPolicy &p) {
    auto fits = [&](int k) { return k >= 1 && k <= kMaxTeam && n % k == 0 && n / k >= 64; };
    if (p.forced > 0) return fits(p.forced) ? p.forced : 1;
    for (const auto &row : p.rows)
        if (n >= row.first) return fits(row.second) ? row.second : 1;
    return 1;
}

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
    // one workgroup per line: a team of one, no halo
    const Team solo{1, 0, 0, n, nullptr, nullptr, nullptr, 0, nullptr};
    // every pass maps element i to thread i mod kLB, so element-wise passes
    // need no barrier between them; the gradient reads neighbours
    stream(n, [&](int) { return 0; }, [&](int i, int) { v[i] = 0.0; });
    double dt = fp.dt_init > 0 ? fp.dt_init : 0.1 * fp.dt_max;
    double alpha = fp.alpha_start;
    int n_pos = 0;
    __syncthreads();
    wg_gradient<double, C>(G, x, g, edy, esl, solo);
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
        wg_gradient<double, C>(G, x, g, edy, esl, solo);
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
// Launches a persistent kernel with as many workgroups as fit on the device at
// once (at most one per line); returns that number.
template <class K>
int launch_persistent(K kernel, const BatchArgs &a, int nlines, size_t lds, qew_ctx ctx) {
    if (lds > 64 * 1024)
        QEW_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    QEW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kLB, lds));
    const int grid = std::max(1, std::min(nlines, std::max(1, per_cu) * ctx->num_cus));
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(kLB), lds, ctx->stream, a);
    QEW_HIP(hipGetLastError());
    return grid;
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
int launch_lbfgs_mixed(int variant, const void *args, int nlines, size_t lds, qew_ctx ctx);
}  // namespace qew_host

#ifdef QEW_LBFGS_MIXED_UNIT

int qew_host::launch_lbfgs_mixed(int variant, const void *args, int nlines, size_t lds,
                                 qew_ctx ctx) {
    const BatchArgs &a = *static_cast<const BatchArgs *>(args);
    switch (variant) {
        case 0: return launch_persistent(lbfgs_batch_kernel<double, double, true, true>, a, nlines, lds, ctx);
        case 1: return launch_persistent(lbfgs_batch_kernel<float, float, true, true>, a, nlines, lds, ctx);
        case 2: return launch_persistent(lbfgs_batch_kernel<float, float, false, true>, a, nlines, lds, ctx);
        case 3: return launch_persistent(lbfgs_batch_kernel<double, float, false, true>, a, nlines, lds, ctx);
        default: return launch_persistent(lbfgs_batch_kernel<double, double, false, true>, a, nlines, lds, ctx);
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
        const size_t nvec = (size_t)lbfgs_nvec(lp->m);
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
        // Teams: the longest lines may be split over several workgroups, each
        // owning a contiguous segment, so that they stop setting the launch's
        // length (one line on one CU otherwise); slots are queued longest
        // line first, a team's members consecutively.
        std::vector<int> order(b->nlines);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return b->lines[x].n > b->lines[y].n; });
        const int max_slots = kMaxTeam * b->nlines;
        // layout: [slot lines][slot members][team sizes][counter][abort]
        // [results][table offsets][team exchange][twiddles][work]
        auto up = [](size_t v) { return (v + 255) / 256 * 256; };
        const size_t off_smem = up(sizeof(int) * max_slots);
        const size_t off_team = off_smem + up(sizeof(int) * max_slots);
        const size_t off_next = off_team + up(sizeof(int) * b->nlines);
        const size_t off_abort = off_next + 256;
        const size_t off_res = off_abort + 256;
        const size_t off_twoff = off_res + up(sizeof(qew_lbfgs_result) * b->nlines);
        const size_t off_xchg = off_twoff + up(sizeof(int64_t) * b->nlines);
        // tagged words of the small exchanges, zeroed (no tag) every launch
        const int xchg_u32 = 2 * 2 * kMaxTeam * kXchgVals * 2 * b->nlines;
        const size_t off_tw = off_xchg + up(sizeof(unsigned) * (size_t)xchg_u32);
        const size_t off_work = off_tw + up(2 * es * (size_t)std::max<int64_t>(tw_total, 1));
        // each line's workspace vectors hold n + 2 kMaxTeam elements
        b->d_scratch.ensure(off_work + es * nvec *
                                           ((size_t)b->total + 2 * (size_t)kMaxTeam * b->nlines));
        char *w = b->d_scratch.as<char>();
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
        a.order = nullptr;
        a.slot_line = reinterpret_cast<const int *>(w);
        a.slot_member = reinterpret_cast<const int *>(w + off_smem);
        a.team_size = reinterpret_cast<const int *>(w + off_team);
        a.xchg = reinterpret_cast<unsigned long long *>(w + off_xchg);
        a.abort = reinterpret_cast<int *>(w + off_abort);
        a.nlines = b->nlines;
        a.next = reinterpret_cast<unsigned *>(w + off_next);
        a.work = w + off_work;
        a.results = reinterpret_cast<qew_lbfgs_result *>(w + off_res);
        a.tw = w + off_tw;
        a.tw_m = (int)(4 * (int64_t)maxN);
        a.tw_off = reinterpret_cast<const int64_t *>(w + off_twoff);
        a.lp = *lp;
        // the FFT buffer doubles as the packed-q buffer: n reals = N complex
        const size_t lds = (size_t)maxN * 2 * es;
        // One launch on line state hh with the team sizes of `policy`;
        // returns its workgroups.
        auto teams_of = [&](const TeamPolicy &policy) {
            std::vector<int> team(b->nlines);
            for (int l = 0; l < b->nlines; ++l) team[l] = lbfgs_team_size(b->lines[l].n, policy);
            return team;
        };
        auto run = [&](const std::vector<int> &team, double *hh) {
            std::vector<int> slot_line, slot_member;
            for (const int l : order)
                for (int k = 0; k < team[l]; ++k) {
                    slot_line.push_back(l);
                    slot_member.push_back(k);
                }
            const int nslots = (int)slot_line.size();
            QEW_HIP(hipMemcpyAsync(w, slot_line.data(), sizeof(int) * nslots,
                                   hipMemcpyHostToDevice, ctx->stream));
            QEW_HIP(hipMemcpyAsync(w + off_smem, slot_member.data(), sizeof(int) * nslots,
                                   hipMemcpyHostToDevice, ctx->stream));
            QEW_HIP(hipMemcpyAsync(w + off_team, team.data(), sizeof(int) * b->nlines,
                                   hipMemcpyHostToDevice, ctx->stream));
            hipLaunchKernelGGL(zero_u32, dim3(1), dim3(1), 0, ctx->stream,
                               reinterpret_cast<unsigned *>(w + off_next));
            hipLaunchKernelGGL(zero_u32, dim3(1), dim3(1), 0, ctx->stream,
                               reinterpret_cast<unsigned *>(w + off_abort));
            hipLaunchKernelGGL(zero_u32_n, dim3((xchg_u32 + 255) / 256), dim3(256), 0,
                               ctx->stream, reinterpret_cast<unsigned *>(w + off_xchg), xchg_u32);
            BatchArgs ar = a;
            ar.nslots = nslots;
            ar.h = hh;
            int grid = 0;
            auto launch = [&](auto kernel) {
                grid = launch_persistent(kernel, ar, nslots, lds, ctx);
            };
            if (mixed) {
                const int variant = (dtype == QEW_F64 && tdt == QEW_F64_MID)   ? 0
                                    : (dtype == QEW_F32 && tdt == QEW_F32_MID) ? 1
                                    : (dtype == QEW_F32)                       ? 2
                                    : (tdt == 1)                               ? 3
                                                                               : 4;
                grid = launch_lbfgs_mixed(variant, &ar, nslots, lds, ctx);
            } else {
                if (dtype == QEW_F64 && tdt == QEW_F64_MID)
                    launch(lbfgs_batch_kernel<double, double, true>);
                else if (dtype == QEW_F32 && tdt == QEW_F32_MID)
                    launch(lbfgs_batch_kernel<float, float, true>);
                else if (dtype == QEW_F32) launch(lbfgs_batch_kernel<float, float>);
                else if (tdt == 1) launch(lbfgs_batch_kernel<double, float>);
                else launch(lbfgs_batch_kernel<double, double>);
            }
            int aborted = 0;
            QEW_HIP(hipMemcpyAsync(&aborted, a.abort, sizeof(int), hipMemcpyDeviceToHost,
                                   ctx->stream));
            QEW_HIP(hipStreamSynchronize(ctx->stream));
            if (aborted)
                throw Error(QEW_ERR_HIP,
                            "batched L-BFGS: a team workgroup gave up waiting for its team "
                            "(not all of the launch's workgroups were resident)");
            return grid;
        };
        // Team sizes: forced by QEW_LBFGS_TEAM(S), else the batch's tuned
        // choice for this precision and preconditioner setting, picked on
        // the first call by timing every candidate on a copy of the line
        // state (each candidate reaches the same minima within the
        // reductions' rounding; only time differs). Candidates: the
        // length-based policies of lbfgs_team_candidates, then, from the
        // iteration counts the one-workgroup run measured, teams on the lines
        // with the most site-iterations (the launch's critical path).
        const TeamPolicy env_policy = lbfgs_team_policy();
        const bool forced = env_policy.forced > 0 || std::getenv("QEW_LBFGS_TEAMS") != nullptr;
        const int key = (dtype == QEW_F32 ? 1 : 0) + (lp->precondition ? 2 : 0);
        std::vector<int> team;
        if (forced) {
            team = teams_of(env_policy);
        } else {
            auto &tuned = b->lbfgs_teams[key];
            if (!tuned.tuned) {
                const std::vector<TeamPolicy> cands = lbfgs_team_candidates(b->lines);
                tuned.tuned = true;
                tuned.team = teams_of(cands.front());
                if (cands.size() > 1) {
                    const size_t bytes = sizeof(double) * (size_t)b->total;
                    b->d_tune.ensure(bytes);
                    double *copy = b->d_tune.as<double>();
                    double best = 1e300;
                    auto time_one = [&](const std::vector<int> &t) {
                        QEW_HIP(hipMemcpyAsync(copy, h, bytes, hipMemcpyDeviceToDevice,
                                               ctx->stream));
                        QEW_HIP(hipStreamSynchronize(ctx->stream));
                        const auto t0 = std::chrono::steady_clock::now();
                        run(t, copy);
                        const double dt = std::chrono::duration<double>(
                                              std::chrono::steady_clock::now() - t0)
                                              .count();
                        if (dt < best) {
                            best = dt;
                            tuned.team = t;
                        }
                    };
                    std::vector<qew_lbfgs_result> solo(b->nlines);
                    for (size_t c = 0; c < cands.size(); ++c) {
                        time_one(teams_of(cands[c]));
                        if (c == 0) {  // one workgroup per line: the lines' iterations
                            QEW_HIP(hipMemcpyAsync(solo.data(), a.results,
                                                   sizeof(qew_lbfgs_result) * b->nlines,
                                                   hipMemcpyDeviceToHost, ctx->stream));
                            QEW_HIP(hipStreamSynchronize(ctx->stream));
                        }
                    }
                    for (const auto &t : lbfgs_work_team_candidates(b->lines, solo))
                        time_one(t);
                }
            }
            team = tuned.team;
        }
        const int grid = run(team, h);
        QEW_HIP(hipMemcpyAsync(results, a.results, sizeof(qew_lbfgs_result) * b->nlines,
                               hipMemcpyDeviceToHost, ctx->stream));
        QEW_HIP(hipStreamSynchronize(ctx->stream));
        b->lbfgs_workgroups = grid;
        b->lbfgs_last_team = team;
    });
}

extern "C" qew_status qew_batch_lbfgs_plan(qew_batch b, int *workgroups, int *team_per_line) {
    return guard([&] {
        QEW_REQUIRE(b && workgroups, "null pointer");
        QEW_REQUIRE((int)b->lbfgs_last_team.size() == b->nlines,
                    "no qew_batch_lbfgs call on this batch yet");
        *workgroups = b->lbfgs_workgroups;
        if (team_per_line)
            for (int l = 0; l < b->nlines; ++l) team_per_line[l] = b->lbfgs_last_team[l];
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
