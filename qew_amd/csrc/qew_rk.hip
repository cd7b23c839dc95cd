// Rosso-Krauth no-passing construction (src/rosso_krauth.h:52-194) for
// batches of lines.
//
// A sweep updates colour 0 then colour 1 in place; within a colour the sites
// are independent (their neighbours are the other colour). Two execution
// paths give bit-identical results:
//  * per-colour launches over every still-active line (any length), with a
//    per-line status kernel that applies the runaway / blocked / timeout
//    decision on the device; the host reads back only an "any line still
//    running" flag every few sweeps;
//  * one workgroup per line with the line state in LDS, running every sweep
//    of the call in one launch (lines up to kRkLdsMaxN sites; used when the
//    batch has enough lines to fill the device).
// Both skip sites whose own height and both neighbours are unchanged since
// the site was last found blocked (its velocity would be the same v <= 0),
// and resolve the moving sites with lane groups that march and bisect
// cooperatively (forward_zero_group). The line state, the advance sums, the
// v <= 0 tests and the bisection are FP64; the tables may be FP64 or FP32.
#include <algorithm>
#include <type_traits>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "qew_internal.hpp"

using namespace qew_dev;

namespace {

constexpr int kRkBlock = 256;

struct RkState {
    int active;
    int sweeps;
    int blocked;
    int runaway;
    double cum_adv;
    double total_adv;
    double mean_h0;
    double mean_h;
    // site-velocity evaluations the reference algorithm makes in this call
    // (skipped sites included), added to by every wave
    unsigned long long vel_evals;
    // site-velocity evaluations this call performed (skipped sites excluded,
    // every lane of a group counted), added to by every wave
    unsigned long long vel_done;
};

struct RkArgs {
    const LineDesc *lines;
    const NoiseDesc *noises;
    double *h;
    unsigned char *moved;  // per site: moved in its last colour pass
    double *partial;       // [line][tile][colour] advance sums of the current sweep
    int *pmoved;           // [line][tile][colour] sites moved in the current sweep
    int tiles;             // workgroups per colour per line
    RkState *state;
    qew_rk_params rp;
};

// Coefficient-type tag of midpoint-contracted FP32 tables (dtype 3): one
// value per cell, the bicubic's x-sum at x-fraction 1/2, which is every edge
// midpoint's fraction on the lines such tables are accepted for.
struct MidF32 {};

// Column of the x cell of an edge midpoint (locate64's x part).
__device__ __forceinline__ int mid_cell_x(const NoiseDesc &nd, double x) {
    double gx = divr(x, nd.dx, nd.rdx);
    gx = wrap_cells(gx, nd.nx);
    int ix = (int)floor(gx);
    if (ix >= nd.nx) ix -= nd.nx;
    return ix;
}

// dV/dy at an edge midpoint: sample_dy64 on FP64/FP32 tables and analytic
// potentials; on a midpoint-contracted table the y weights of locate64 /
// bspline_dw applied to the 4 values of the column, summed b = 0..3 in FP64.
template <typename C>
__device__ __forceinline__ double edge_dy(const NoiseDesc &nd, double x, double y) {
    if constexpr (std::is_same<C, MidF32>::value) {
        int iyn[4];
        double fy, dwy[4];
        mid_locate_y64(nd, y, iyn, fy);
        bspline_dw(fy, dwy);
        float c[4];
        mid_column4(nd, mid_cell_x(nd, x), iyn, c);
        double d = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) d += (double)c[b] * dwy[b];
        return divr(d, nd.dy, nd.rdy);
    } else {
        return sample_dy64<C>(nd, x, y);
    }
}

// rk_site_velocity (src/rosso_krauth.h:52-82)
template <typename C>
__device__ __forceinline__ double site_velocity(int i, int n, double dx, double lt, double f,
                                                int model, double hi, double hl, double hr,
                                                const NoiseDesc &nd) {
    const int im = (i == 0) ? n - 1 : i - 1;
    const int ip = (i + 1 == n) ? 0 : i + 1;
    double line_force;
    if (model == 0) {
        line_force = div1(-lt * (hl + hr - 2.0 * hi), dx * dx);
    } else {
        const double dhr = div1(hr - hi, dx);
        const double dhl = div1(hi - hl, dx);
        line_force = div1(lt * (dhl / sqrt(1.0 + dhl * dhl) - dhr / sqrt(1.0 + dhr * dhr)), dx);
    }
    const double xcl = (double)(im + i) * 0.5 * dx;
    const double hcl = (hl + hi) * 0.5;
    const double xcr = (double)(i + ip) * 0.5 * dx;
    const double hcr = (hi + hr) * 0.5;
    const double dvl = edge_dy<C>(nd, xcl, hcl);
    const double dvr = edge_dy<C>(nd, xcr, hcr);
    const double grad = dx * (line_force + 0.5 * (dvl + dvr) - f);
    return -grad;
}

// dV/dy on one edge of one site (fixed x) at successive trial heights, with
// the 4x4 coefficient block of the last y cell kept in registers: a march or
// bisection mostly stays in one cell, so most evaluations gather nothing.
// Same locate, weights and summation order as sample_dy64, so the values
// are bit-identical.
template <typename C>
struct EdgeDy {
    int iy = -1;  // cell of the held block (-1: none)
    C c[4][4];    // c[a][b]: x-row a, y-level b

    __device__ __forceinline__ double eval(const NoiseDesc &nd, double x, double y) {
        if (nd.kind != TABLE) return sample_dy64<C>(nd, x, y);
        int ixn[4], iyn[4];
        double fx, fy;
        locate64(nd, x, y, ixn, iyn, fx, fy);
        double wx[4], dwy[4];
        bspline_w(fx, wx);
        bspline_dw(fy, dwy);
        if (iyn[1] != iy) {
            iy = iyn[1];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                double r[4];
                row4_64<C>(nd, ixn[1], iyn[b], r);
#pragma unroll
                for (int a = 0; a < 4; ++a) c[a][b] = (C)r[a];
            }
        }
        double d = 0;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            double row_dy = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b) row_dy += (double)c[a][b] * dwy[b];
            d += wx[a] * row_dy;
        }
        return divr(d, nd.dy, nd.rdy);
    }
};

// The same for a midpoint-contracted table: the held block is the column's
// 4 values at the last y cell (same locate, weights and order as edge_dy).
template <>
struct EdgeDy<MidF32> {
    int iy = -1;
    float c[4];

    __device__ __forceinline__ double eval(const NoiseDesc &nd, double x, double y) {
        int iyn[4];
        double fy, dwy[4];
        mid_locate_y64(nd, y, iyn, fy);
        bspline_dw(fy, dwy);
        if (iyn[1] != iy) {
            iy = iyn[1];
            mid_column4(nd, mid_cell_x(nd, x), iyn, c);
        }
        double d = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) d += (double)c[b] * dwy[b];
        return divr(d, nd.dy, nd.rdy);
    }
};

// site_velocity with the two edges' blocks held across calls for one site.
template <typename C>
__device__ __forceinline__ double site_velocity_held(int i, int n, double dx, double lt, double f,
                                                     int model, double hi, double hl, double hr,
                                                     const NoiseDesc &nd, EdgeDy<C> &el,
                                                     EdgeDy<C> &er) {
    const int im = (i == 0) ? n - 1 : i - 1;
    const int ip = (i + 1 == n) ? 0 : i + 1;
    double line_force;
    if (model == 0) {
        line_force = div1(-lt * (hl + hr - 2.0 * hi), dx * dx);
    } else {
        const double dhr = div1(hr - hi, dx);
        const double dhl = div1(hi - hl, dx);
        line_force = div1(lt * (dhl / sqrt(1.0 + dhl * dhl) - dhr / sqrt(1.0 + dhr * dhr)), dx);
    }
    const double dvl = el.eval(nd, (double)(im + i) * 0.5 * dx, (hl + hi) * 0.5);
    const double dvr = er.eval(nd, (double)(i + ip) * 0.5 * dx, (hi + hr) * 0.5);
    const double grad = dx * (line_force + 0.5 * (dvl + dvr) - f);
    return -grad;
}

// The lanes [base, base + g) of a wave, as the low g bits of a ballot.
__device__ __forceinline__ unsigned long long group_bits(unsigned long long m, int base, int g) {
    return (m >> base) & (g == kWave ? ~0ull : ((1ull << g) - 1ull));
}

// rk_forward_zero (src/rosso_krauth.h:88-124) for a site whose velocity at hi
// is already known to be positive, computed by a group of g lanes (g a power
// of two, 1..64, lanes base..base+g-1 of the wave; `li` = lane in the group).
// The march evaluates g consecutive steps at once: lane j repeats the
// sequential additions a += dstep, advanced += dstep j times, so its trial
// point is bit-identical to step j of the loop, and the first step with
// v <= 0 is the sequential one. The bisection evaluates the next d =
// log2(g) levels at once: lane li owns node li + 1 of the heap-ordered
// subtree below the current bracket (children 2x: v <= 0 -> hi = mid,
// 2x + 1: v > 0 -> lo = mid), recomputes its bracket by the same halvings
// and checks the same loop condition, so the path walked afterwards is the
// sequential path. Every lane of the group must call this with the same
// site; all return the same height. `evals` (every lane) grows by the number
// of evaluations the sequential loop makes, march and bisection; `done` (per
// lane) by the evaluations this lane performs.
template <typename C>
__device__ double forward_zero_group(int i, int n, double dx, double lt, double f, int model,
                                     double hi, double hl, double hr, const NoiseDesc &nd,
                                     const qew_rk_params &rp, int g, int li, int base,
                                     unsigned &evals, unsigned &done) {
    // one evaluation site for both phases keeps a single inlined copy of the
    // velocity (two gathers) in the kernel
    double a = hi, advanced = 0;  // march state
    double lo = 0, hic = 0;       // bisection bracket
    int k = 0;                    // bisections done
    bool marching = true;
    const int d = (g == 1) ? 1 : 31 - __clz(g);
    const int node = li + 1;
    const int depth = 31 - __clz(node);
    EdgeDy<C> el, er;
    for (;;) {
        double trial, aa = 0, ad = 0;
        bool want;
        if (marching) {
            if (!(advanced < rp.max_advance)) return a;
            aa = a;
            ad = advanced;
            for (int j = 0; j < li; ++j) {
                aa = aa + rp.dstep;
                ad += rp.dstep;
            }
            // dstep > 0 (checked on the host), so the steps reached form a prefix
            want = ad < rp.max_advance;
            trial = aa + rp.dstep;
        } else {
            if (!(k < 60 && (hic - lo) > rp.root_tol)) return 0.5 * (lo + hic);
            want = node < (1 << d);
            double l = lo, h = hic;
            for (int t = depth - 1; t >= 0 && want; --t) {
                if (!(k + (depth - 1 - t) < 60 && (h - l) > rp.root_tol)) {
                    want = false;
                    break;
                }
                const double mid = 0.5 * (l + h);
                if ((node >> t) & 1) l = mid;
                else h = mid;
            }
            want = want && (k + depth < 60) && ((h - l) > rp.root_tol);
            trial = 0.5 * (l + h);
        }
        double v = 0;
        if (want) {
            v = site_velocity_held<C>(i, n, dx, lt, f, model, trial, hl, hr, nd, el, er);
            ++done;
        }
        if (marching) {
            const unsigned long long hm = group_bits(__ballot(want && v <= 0), base, g);
            const unsigned long long rm = group_bits(__ballot(want), base, g);
            if (hm) {
                const int j = __builtin_ctzll(hm);
                evals += j + 1;
                lo = __shfl(aa, base + j);
                hic = __shfl(trial, base + j);
                marching = false;
                continue;
            }
            const int steps = __popcll(rm);
            evals += steps;
            const int last = base + (steps > 0 ? steps - 1 : 0);
            const double b_last = __shfl(trial, last);
            const double ad_last = __shfl(ad + rp.dstep, last);
            if (steps == 0) return a;
            a = b_last;
            advanced = ad_last;
            if (steps < g) return a;
        } else {
            const unsigned long long pm = group_bits(__ballot(want && v > 0), base, g);
            int x = 1;
            for (int t = 0; t < d; ++t) {
                if (!(k < 60 && (hic - lo) > rp.root_tol)) break;
                const double mid = 0.5 * (lo + hic);
                const bool p = (pm >> (x - 1)) & 1ull;
                if (p) lo = mid;
                else hic = mid;
                ++k;
                ++evals;
                x = 2 * x + (p ? 1 : 0);
            }
        }
    }
}

// Lanes per moving site for m moving sites shared by `threads` lanes.
__device__ __forceinline__ int group_size(int m, int threads) {
    int g = kWave;
    while (g > 1 && g * m > threads) g >>= 1;
    return g;
}

// The runaway / blocked / timeout decision after a sweep
// (src/rosso_krauth.h:168-188). Returns whether the line keeps sweeping.
__device__ bool sweep_status(RkState &st, double total, int n, const qew_rk_params &rp) {
    st.cum_adv += total;
    const double mean_adv = st.cum_adv / (double)n;
    st.sweeps += 1;
    st.total_adv = total;
    st.mean_h = st.mean_h0 + mean_adv;
    if (mean_adv > rp.runaway) {
        st.runaway = 1;
        st.blocked = 0;
        return false;
    }
    if (total < rp.sweep_tol) {
        st.blocked = 1;
        st.runaway = 0;
        return false;
    }
    return st.sweeps < rp.max_sweeps;  // else timeout
}

// One colour pass over 256 sites of one line per workgroup. A workgroup
// whose sites and neighbours all kept their heights since their last pass
// only records a zero advance. Otherwise every changed site's velocity is
// evaluated, and the sites that move are listed in LDS and resolved by lane
// groups, so a workgroup with few moving sites finishes in a few dependent
// evaluation rounds instead of the ~40 of a sequential march and bisection.
// The advance sum is formed per site owner in a fixed order. Registers are
// capped for 3 waves per SIMD (2 uncapped): the extra latency hiding in
// sweeps where most sites move outweighs the few spills (measured on the
// 64 x 16384 depinning batch: 1.5-6 % less time, same probes and sweeps; a
// 4-wave cap spills more and loses).
template <typename C>
__global__ __launch_bounds__(kRkBlock, 3) void rk_color_kernel(RkArgs a, int color) {
    __shared__ int s_list[kRkBlock];
    __shared__ double s_new[kRkBlock];
    __shared__ double red[kRkBlock / kWave];
    __shared__ int s_count;
    const int line = blockIdx.y;
    if (!a.state[line].active) return;
    const LineDesc L = a.lines[line];
    const int n = L.n;
    const int tid = threadIdx.x;
    const int half = (n - color + 1) / 2;  // sites of this colour: i = 2 j + color < n
    const int j = blockIdx.x * kRkBlock + tid;
    const int i = 2 * j + color;
    const bool valid = j < half;
    const unsigned char *mv = a.moved + L.off;
    bool dirty = false;
    int im = 0, ip = 0;
    if (valid) {
        im = (i == 0) ? n - 1 : i - 1;
        ip = (i + 1 == n) ? 0 : i + 1;
        dirty = (mv[i] | mv[im] | mv[ip]) != 0;
    }
    const size_t pidx = ((size_t)line * a.tiles + blockIdx.x) * 2 + color;
    double *part = a.partial + pidx;
    if (!__syncthreads_or(dirty)) {
        if (tid == 0) {
            *part = 0.0;
            a.pmoved[pidx] = 0;
            const int cnt = std::min(kRkBlock, std::max(0, half - (int)blockIdx.x * kRkBlock));
            if (cnt) atomicAdd(&a.state[line].vel_evals, (unsigned long long)cnt);
        }
        return;
    }
    const NoiseDesc nd = a.noises[L.noise];
    const double dx = L.L / (double)n;
    double *h = a.h + L.off;
    if (tid == 0) s_count = 0;
    __syncthreads();
    unsigned evals = valid ? 1u : 0u;
    unsigned done = dirty ? 1u : 0u;
    double hi = 0;
    if (dirty) {
        hi = h[i];
        const double v0 = site_velocity<C>(i, n, dx, L.lt, L.f, L.model, hi, h[im], h[ip], nd);
        if (v0 > 0) s_list[atomicAdd(&s_count, 1)] = tid;
        else s_new[tid] = hi;
    }
    __syncthreads();
    const int m = s_count;
    const int g = group_size(m, kRkBlock);
    const int li = tid & (g - 1);
    const int base = (tid & (kWave - 1)) & ~(g - 1);
    for (int idx = tid / g; idx < m; idx += kRkBlock / g) {
        const int t = s_list[idx];
        const int s = 2 * (blockIdx.x * kRkBlock + t) + color;
        const int sm = (s == 0) ? n - 1 : s - 1;
        const int sp = (s + 1 == n) ? 0 : s + 1;
        unsigned e = 0;
        const double hnew = forward_zero_group<C>(s, n, dx, L.lt, L.f, L.model, h[s], h[sm], h[sp],
                                                  nd, a.rp, g, li, base, e, done);
        if (li == 0) {
            evals += e;
            s_new[t] = hnew;
        }
    }
    __syncthreads();
    double adv = 0;
    bool moved = false;
    if (dirty) {
        const double hn = s_new[tid];
        adv = hn - hi;
        moved = hn != hi;
        if (moved) h[i] = hn;
        a.moved[L.off + i] = moved ? 1 : 0;
    }
    const double total = block_sum(adv, red);
    const int nmoved = __syncthreads_count(moved);
    if (tid == 0) {
        *part = total;
        a.pmoved[pidx] = nmoved;
    }
    // every lane of the wave reaches the reduction (no early return above)
    const unsigned ev = wave_sum(evals);
    if ((tid & (kWave - 1)) == 0 && ev) atomicAdd(&a.state[line].vel_evals, (unsigned long long)ev);
    const unsigned dn = wave_sum(done);
    if ((tid & (kWave - 1)) == 0 && dn) atomicAdd(&a.state[line].vel_done, (unsigned long long)dn);
}

// One workgroup per line: the sweep's advance (per-workgroup sums of both
// colours, added in a fixed order) and the status decision.
// running[0] counts the lines that keep sweeping, running[1] the sites they
// moved in this sweep.
__global__ __launch_bounds__(kRkBlock) void rk_status_kernel(RkArgs a, unsigned *running) {
    __shared__ double red[kRkBlock / kWave];
    const int line = blockIdx.x;
    RkState st = a.state[line];
    if (!st.active) return;
    const LineDesc L = a.lines[line];
    const double *part = a.partial + (size_t)line * a.tiles * 2;
    const int *pm = a.pmoved + (size_t)line * a.tiles * 2;
    double acc = 0, mc = 0;
    for (int t = threadIdx.x; t < 2 * a.tiles; t += blockDim.x) {
        acc += part[t];
        mc += (double)pm[t];
    }
    const double total = block_sum(acc, red);
    const double nmoved = block_sum(mc, red);
    if (threadIdx.x != 0) return;
    const bool go = sweep_status(st, total, L.n, a.rp);
    // vel_evals is added to by the colour kernels: leave it out of the store
    a.state[line].active = go ? 1 : 0;
    a.state[line].sweeps = st.sweeps;
    a.state[line].blocked = st.blocked;
    a.state[line].runaway = st.runaway;
    a.state[line].cum_adv = st.cum_adv;
    a.state[line].total_adv = st.total_adv;
    a.state[line].mean_h = st.mean_h;
    if (go) {
        atomicAdd(running, 1u);
        atomicAdd(running + 1, (unsigned)nmoved);
    }
}

__global__ __launch_bounds__(kRkBlock) void rk_init_kernel(RkArgs a) {
    __shared__ double red[kRkBlock / kWave];
    const int line = blockIdx.x;
    const LineDesc L = a.lines[line];
    double acc = 0;
    for (int i = threadIdx.x; i < L.n; i += blockDim.x) {
        acc += a.h[L.off + i];
        a.moved[L.off + i] = 1;  // every site is evaluated in the first sweep
    }
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) {
        RkState st{};
        st.active = a.rp.max_sweeps > 0 ? 1 : 0;
        st.mean_h0 = s / (double)L.n;
        st.mean_h = 0;
        a.state[line] = st;
    }
}

// Whole rk_block calls for lines held in LDS: one workgroup per line runs
// every sweep (both colours, the advance sum and the status decision) with
// the line state and the moved flags in LDS, so a call is one launch however
// many sweeps it takes. Sites of a colour are taken in chunks of
// kRkLineBlock; a chunk with no changed site costs one barrier.
// 512 threads: two FP64 gathers in flight take ~170 VGPRs, more than a
// 1024-thread workgroup can give each wave.
constexpr int kRkLineBlock = 512;
constexpr int kRkLdsMaxN = 16384;

// With `cont`, the kernel continues rk_block calls that the per-colour
// launches or an earlier launch of this kernel began: the line state, moved
// flags and RkState (sweeps, advance, counters) are taken from global memory,
// lines that already stopped are left alone, and the sweeps go on exactly as
// one uninterrupted run would take them. A launch does at most `sweep_cap`
// sweeps per line; a line still sweeping then saves its moved flags and stays
// active for the next launch.
template <typename C>
__global__ __launch_bounds__(kRkLineBlock) void rk_line_kernel(RkArgs a, int lds_n, int cont,
                                                               int sweep_cap) {
    extern __shared__ double s_h[];
    unsigned char *s_mv = reinterpret_cast<unsigned char *>(s_h + lds_n);
    __shared__ int s_list[kRkLineBlock];
    __shared__ double s_new[kRkLineBlock];
    __shared__ double red[kRkLineBlock / kWave];
    __shared__ int s_count;
    __shared__ int s_go;
    const int line = blockIdx.x;
    const LineDesc L = a.lines[line];
    const NoiseDesc nd = a.noises[L.noise];
    const int n = L.n;
    const double dx = L.L / (double)n;
    const int tid = threadIdx.x;
    double *hg = a.h + L.off;
    RkState st{};
    if (cont) {
        st = a.state[line];
        if (!st.active) return;  // uniform: every thread read the same flag
        for (int i = tid; i < n; i += kRkLineBlock) {
            s_h[i] = hg[i];
            s_mv[i] = a.moved[L.off + i];
        }
        __syncthreads();
    } else {
        double acc = 0;
        for (int i = tid; i < n; i += kRkLineBlock) {
            const double v = hg[i];
            s_h[i] = v;
            s_mv[i] = 1;
            acc += v;
        }
        st.mean_h0 = block_sum(acc, red) / (double)n;
    }
    unsigned long long evals = 0, done = 0;
    bool go = cont ? true : a.rp.max_sweeps > 0;
    int chunk = 0;
    bool capped = false;
    while (go) {
        double my_adv = 0;
        for (int color = 0; color < 2; ++color) {
            const int half = (n - color + 1) / 2;
            for (int c0 = 0; c0 < half; c0 += kRkLineBlock) {
                const int i = 2 * (c0 + tid) + color;
                const bool valid = c0 + tid < half;
                evals += valid ? 1u : 0u;
                bool dirty = false;
                int im = 0, ip = 0;
                if (valid) {
                    im = (i == 0) ? n - 1 : i - 1;
                    ip = (i + 1 == n) ? 0 : i + 1;
                    dirty = (s_mv[i] | s_mv[im] | s_mv[ip]) != 0;
                }
                // own flags are written by their owner only; the neighbours'
                // belong to the other colour, fixed during this pass
                if (!__syncthreads_or(dirty)) continue;
                if (tid == 0) s_count = 0;
                __syncthreads();
                double hi = 0;
                if (dirty) {
                    ++done;
                    hi = s_h[i];
                    const double v0 =
                        site_velocity<C>(i, n, dx, L.lt, L.f, L.model, hi, s_h[im], s_h[ip], nd);
                    if (v0 > 0) s_list[atomicAdd(&s_count, 1)] = tid;
                    else s_new[tid] = hi;
                }
                __syncthreads();
                const int m = s_count;
                const int g = group_size(m, kRkLineBlock);
                const int li = tid & (g - 1);
                const int base = (tid & (kWave - 1)) & ~(g - 1);
                for (int idx = tid / g; idx < m; idx += kRkLineBlock / g) {
                    const int t = s_list[idx];
                    const int s = 2 * (c0 + t) + color;
                    const int sm = (s == 0) ? n - 1 : s - 1;
                    const int sp = (s + 1 == n) ? 0 : s + 1;
                    unsigned e = 0, dn = 0;
                    const double hnew = forward_zero_group<C>(s, n, dx, L.lt, L.f, L.model, s_h[s],
                                                              s_h[sm], s_h[sp], nd, a.rp, g, li,
                                                              base, e, dn);
                    done += dn;
                    if (li == 0) {
                        evals += e;
                        s_new[t] = hnew;
                    }
                }
                __syncthreads();
                if (dirty) {
                    const double hn = s_new[tid];
                    my_adv += hn - hi;
                    s_h[i] = hn;
                    s_mv[i] = hn != hi ? 1 : 0;
                }
            }
            __syncthreads();
        }
        // each thread's sites are fixed, so the sum is reproducible
        const double total = block_sum(my_adv, red);
        if (tid == 0) s_go = sweep_status(st, total, n, a.rp) ? 1 : 0;
        __syncthreads();
        go = s_go != 0;
        if (go && ++chunk >= sweep_cap) {  // uniform: s_go and chunk are
            capped = true;
            go = false;
        }
    }
    for (int i = tid; i < n; i += kRkLineBlock) hg[i] = s_h[i];
    if (capped)
        for (int i = tid; i < n; i += kRkLineBlock) a.moved[L.off + i] = s_mv[i];
    const unsigned long long ev = wave_sum(evals);
    if ((tid & (kWave - 1)) == 0 && ev) atomicAdd(&a.state[line].vel_evals, ev);
    const unsigned long long dn = wave_sum(done);
    if ((tid & (kWave - 1)) == 0 && dn) atomicAdd(&a.state[line].vel_done, dn);
    if (tid == 0) {
        // the counter is added to by every wave: leave it out of the store
        a.state[line].active = capped ? 1 : 0;
        a.state[line].sweeps = st.sweeps;
        a.state[line].blocked = st.blocked;
        a.state[line].runaway = st.runaway;
        a.state[line].cum_adv = st.cum_adv;
        a.state[line].total_adv = st.total_adv;
        a.state[line].mean_h0 = st.mean_h0;
        a.state[line].mean_h = st.mean_h;
    }
}

__global__ void rk_velocity_kernel(NoiseDesc nd, int n, double dx, double lt, double f, int model,
                                   int m, const int *idx, const double *hi, const double *hl,
                                   const double *hr, double *v) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    if (nd.dtype == 1)
        v[k] = site_velocity<float>(idx[k], n, dx, lt, f, model, hi[k], hl[k], hr[k], nd);
    else
        v[k] = site_velocity<double>(idx[k], n, dx, lt, f, model, hi[k], hl[k], hr[k], nd);
}

// rk_forward_zero (src/rosso_krauth.h:88-124) at m independent sites, one
// thread each, in the reference's sequential order: the velocity at hi; if
// positive, march b = a + dstep (a = b, advanced += dstep) until v(b) <= 0 or
// advanced reaches max_advance, then at most 60 bisections while the bracket
// is wider than root_tol. The march length is bounded on the host.
template <typename C>
__global__ void rk_forward_zero_kernel(NoiseDesc nd, int n, double dx, double lt, double f,
                                       int model, qew_rk_params rp, int m, const int *idx,
                                       const double *hi, const double *hl, const double *hr,
                                       double *out, int *found, int64_t *evals) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int i = idx[k];
    const double h0 = hi[k], l = hl[k], r = hr[k];
    int64_t ev = 1;
    int fd = 1;
    double res = h0;
    if (site_velocity<C>(i, n, dx, lt, f, model, h0, l, r, nd) > 0) {
        double a = h0, advanced = 0;
        fd = 0;
        res = a;
        while (advanced < rp.max_advance) {
            const double b = a + rp.dstep;
            ++ev;
            if (site_velocity<C>(i, n, dx, lt, f, model, b, l, r, nd) <= 0) {
                double lo = a, hc = b;
                for (int t = 0; t < 60 && (hc - lo) > rp.root_tol; ++t) {
                    const double mid = 0.5 * (lo + hc);
                    ++ev;
                    if (site_velocity<C>(i, n, dx, lt, f, model, mid, l, r, nd) > 0) lo = mid;
                    else hc = mid;
                }
                fd = 1;
                res = 0.5 * (lo + hc);
                break;
            }
            a = b;
            advanced += rp.dstep;
            res = a;
        }
    }
    out[k] = res;
    if (found) found[k] = fd;
    if (evals) evals[k] = ev;
}

// A march of max_advance / dstep steps must end: dstep positive, and few
// enough steps that `advanced += dstep` keeps growing (far fewer than 2^52).
void check_march(const qew_rk_params &rp) {
    QEW_REQUIRE(rp.dstep > 0 || !(rp.max_advance > 0),
                "rk_block: dstep must be positive when max_advance > 0");
    QEW_REQUIRE(!(rp.max_advance > 0) || rp.max_advance / rp.dstep <= (double)(1 << 26),
                "rk_block: max_advance / dstep exceeds 2^26 march steps per site");
}

// Device scratch of run_rk for nlines lines of total length `total`.
struct RkScratch {
    RkState *state;
    unsigned *running;  // [0] lines still sweeping, [1] sites they moved
    unsigned char *moved;
    double *partial;
    int *pmoved;
};

size_t rk_scratch_bytes(int nlines, int64_t total, int tiles) {
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    return up(sizeof(RkState) * nlines) + 256 + up((size_t)total) +
           up(sizeof(double) * 2 * (size_t)tiles * nlines) +
           up(sizeof(int) * 2 * (size_t)tiles * nlines);
}

RkScratch rk_scratch(char *w, int nlines, int64_t total, int tiles) {
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    RkScratch s;
    s.state = reinterpret_cast<RkState *>(w);
    w += up(sizeof(RkState) * nlines);
    s.running = reinterpret_cast<unsigned *>(w);
    w += 256;
    s.moved = reinterpret_cast<unsigned char *>(w);
    w += up((size_t)total);
    s.partial = reinterpret_cast<double *>(w);
    w += up(sizeof(double) * 2 * (size_t)tiles * nlines);
    s.pmoved = reinterpret_cast<int *>(w);
    return s;
}

int rk_tiles(int max_n) { return std::max(1, ((max_n + 1) / 2 + kRkBlock - 1) / kRkBlock); }

// Whether lines of up to max_n sites fit the LDS path on this device.
bool line_kernel_fits(qew_ctx ctx, int max_n) {
    int lds_max = 0, lds_optin = 0;
    QEW_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
    QEW_HIP(hipDeviceGetAttribute(&lds_optin, hipDeviceAttributeSharedMemPerBlockOptin, ctx->device));
    lds_max = std::max(lds_max, lds_optin);
    const size_t need = (sizeof(double) + 1) * (size_t)max_n + 8192;
    return max_n <= kRkLdsMaxN && need <= (size_t)lds_max;
}

// QEW_RK_KERNEL=line|launch|hybrid forces a path (tests, tuning): 1 line,
// 2 launch, 3 hybrid switching at its first poll, 0 not set.
int rk_env_path() {
    const char *env = std::getenv("QEW_RK_KERNEL");
    if (env && std::strcmp(env, "line") == 0) return 1;
    if (env && std::strcmp(env, "launch") == 0) return 2;
    if (env && std::strcmp(env, "hybrid") == 0) return 3;
    return 0;
}

// How run_rk sweeps: all in the LDS line kernel, all with per-colour
// launches, or per-colour launches until few sites move and then the line
// kernel for the rest of the call (switch_per_line: the switch happens at a
// poll where the lines still sweeping moved at most this many sites per line
// in their last sweep; a negative value switches at the first poll).
enum class RkMode { Line, Launch, Hybrid };

// Sites moved per still-sweeping line below which the line kernel takes over:
// a sweep of the line kernel scans its line in LDS (32 chunk barriers for
// 16384 sites) and resolves the moving sites with lane groups, while every
// sweep of the launches costs three launches over all lines' tiles.
constexpr double kRkSwitchMovedPerLine = 128.0;

// Sweeps per line kernel launch: between launches the host reads the lines'
// states (and may report progress), so a call of many sweeps never runs as
// one kernel of unbounded length. QEW_RK_LINE_SWEEPS sets another positive
// value (the tests use a small one to cross many launch boundaries).
int rk_line_sweeps_per_launch() {
    const char *env = std::getenv("QEW_RK_LINE_SWEEPS");
    const int v = env ? std::atoi(env) : 0;
    return v > 0 ? v : 2048;
}

// QEW_RK_PROGRESS=<seconds>: rk_block writes a line to stderr at most this
// often while it sweeps (lines still sweeping, their largest sweep count).
double rk_progress_interval() {
    const char *env = std::getenv("QEW_RK_PROGRESS");
    return env ? std::atof(env) : -1.0;
}

struct RkProgress {
    double every = rk_progress_interval();
    std::chrono::steady_clock::time_point last = std::chrono::steady_clock::now();
    void report(int active, int max_sweeps) {
        if (every < 0) return;
        const auto now = std::chrono::steady_clock::now();
        if (std::chrono::duration<double>(now - last).count() < every) return;
        last = now;
        std::fprintf(stderr, "rk_block: %d lines sweeping, %d sweeps\n", active, max_sweeps);
        std::fflush(stderr);
    }
};

// Runs the sweeps of the lines whose descriptors are on the device in the LDS
// line kernel, rk_line_sweeps_per_launch() sweeps per launch until every line
// has stopped; `cont` = 0 starts the call, 1 continues the states in `a`.
// Leaves the final states in `out`.
void run_line_kernel(hipStream_t s, const RkArgs &a, int nlines, int max_n, int table_dtype,
                     int cont, std::vector<RkState> &out) {
    const size_t lds = (sizeof(double) + 1) * (size_t)max_n;
    const int cap = rk_line_sweeps_per_launch();
    auto go = [&](auto kernel, int c) {
        QEW_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(kernel, dim3(nlines), dim3(kRkLineBlock), lds, s, a, max_n, c, cap);
    };
    RkProgress progress;
    out.resize(nlines);
    for (;;) {
        if (table_dtype == QEW_F32_MID) go(rk_line_kernel<MidF32>, cont);
        else if (table_dtype == 1) go(rk_line_kernel<float>, cont);
        else go(rk_line_kernel<double>, cont);
        QEW_HIP(hipGetLastError());
        QEW_HIP(hipMemcpyAsync(out.data(), a.state, sizeof(RkState) * nlines,
                               hipMemcpyDeviceToHost, s));
        QEW_HIP(hipStreamSynchronize(s));
        int active = 0, most = 0;
        for (const RkState &st : out) {
            active += st.active ? 1 : 0;
            most = std::max(most, st.sweeps);
        }
        if (!active) return;
        progress.report(active, most);
        cont = 1;
    }
}

void run_rk(qew_ctx ctx, const LineDesc *d_lines, const NoiseDesc *d_noises, int nlines, int max_n,
            int table_dtype, const qew_rk_params &rp, double *h, const RkScratch &w,
            unsigned *running_host, std::vector<RkState> &out, RkMode mode,
            double switch_per_line = kRkSwitchMovedPerLine) {
    // a march step of zero (or NaN) with a positive advance cap never ends;
    // the reference would spin forever, the device must not
    check_march(rp);
    hipStream_t s = ctx->stream;
    const int tiles = rk_tiles(max_n);
    RkArgs a{d_lines, d_noises, h, w.moved, w.partial, w.pmoved, tiles, w.state, rp};
    if (mode == RkMode::Line) {
        QEW_HIP(hipMemsetAsync(w.state, 0, sizeof(RkState) * (size_t)nlines, s));
        run_line_kernel(s, a, nlines, max_n, table_dtype, 0, out);
        return;
    } else {
        hipLaunchKernelGGL(rk_init_kernel, dim3(nlines), dim3(kRkBlock), 0, s, a);
        QEW_HIP(hipGetLastError());
        const dim3 grid(tiles, nlines);
        // check the running flag every `poll` sweeps; lines that finished in
        // between are frozen, so the extra launches are no-ops for them
        int poll = 4;
        RkProgress progress;
        for (int sweep = 0; sweep < rp.max_sweeps;) {
            const int todo = std::min(poll, rp.max_sweeps - sweep);
            for (int k = 0; k < todo; ++k) {
                for (int color = 0; color < 2; ++color) {
                    if (table_dtype == QEW_F32_MID)
                        hipLaunchKernelGGL(rk_color_kernel<MidF32>, grid, dim3(kRkBlock), 0, s, a,
                                           color);
                    else if (table_dtype == 1)
                        hipLaunchKernelGGL(rk_color_kernel<float>, grid, dim3(kRkBlock), 0, s, a,
                                           color);
                    else
                        hipLaunchKernelGGL(rk_color_kernel<double>, grid, dim3(kRkBlock), 0, s, a,
                                           color);
                }
                if (k == todo - 1) QEW_HIP(hipMemsetAsync(w.running, 0, 2 * sizeof(unsigned), s));
                hipLaunchKernelGGL(rk_status_kernel, dim3(nlines), dim3(kRkBlock), 0, s, a,
                                   w.running);
            }
            QEW_HIP(hipGetLastError());
            sweep += todo;
            QEW_HIP(hipMemcpyAsync(running_host, w.running, 2 * sizeof(unsigned),
                                   hipMemcpyDeviceToHost, s));
            QEW_HIP(hipStreamSynchronize(s));
            if (running_host[0] == 0) break;
            if (mode == RkMode::Hybrid &&
                (switch_per_line < 0 ||
                 (double)running_host[1] <= switch_per_line * (double)running_host[0])) {
                // the lines still sweeping finish in the LDS kernel, from the
                // state and moved flags the launches left
                run_line_kernel(s, a, nlines, max_n, table_dtype, 1, out);
                return;
            }
            progress.report((int)running_host[0], sweep);
            poll = std::min(poll * 2, 256);
        }
    }
    out.resize(nlines);
    QEW_HIP(hipMemcpyAsync(out.data(), w.state, sizeof(RkState) * nlines, hipMemcpyDeviceToHost, s));
    QEW_HIP(hipStreamSynchronize(s));
}

qew_rk_result to_result(const RkState &st) {
    qew_rk_result r{};
    r.blocked = st.blocked;
    r.runaway = st.runaway;
    r.sweeps = st.sweeps;
    r.total_advance = st.total_adv;
    r.mean_h = st.sweeps > 0 ? st.mean_h : 0.0;
    r.n_velocity_evals = (long)st.vel_evals;
    r.n_velocity_evals_done = (long)st.vel_done;
    return r;
}

}  // namespace

using namespace qew_host;

extern "C" {

qew_status qew_rk_block(qew_ctx ctx, qew_noise noise, const qew_params *p,
                        const qew_rk_params *rp, int n, double *h, qew_rk_result *result) {
    return guard([&] {
        ctx_check(ctx);
        QEW_REQUIRE(noise && p && rp && h && result, "null pointer");
        QEW_REQUIRE(n >= 1, "line length must be >= 1");
        if (n % 2 != 0)
            throw Error(QEW_ERR_INVALID_ARGUMENT,
                        "rk_block: line length must be even for the red-black sweep");
        require_wide_table(noise->desc);
        use_device(ctx);
        LineDesc l = single_line(n, p);
        // workspace: [LineDesc][NoiseDesc][run_rk scratch]
        const size_t off_noise = 256, off_rk = 512;
        ctx->work.ensure(off_rk + rk_scratch_bytes(1, n, rk_tiles(n)));
        char *w = ctx->work.as<char>();
        ctx->pinned.ensure(sizeof(double) * 64);
        QEW_HIP(hipMemcpyAsync(w, &l, sizeof(LineDesc), hipMemcpyHostToDevice, ctx->stream));
        QEW_HIP(hipMemcpyAsync(w + off_noise, &noise->desc, sizeof(NoiseDesc), hipMemcpyHostToDevice,
                               ctx->stream));
        std::vector<RkState> out;
        const int dt = noise->desc.kind == TABLE ? noise->desc.dtype : 0;
        // one line: a single workgroup leaves the device idle unless forced
        const int env = rk_env_path();
        const bool fits = line_kernel_fits(ctx, n);
        const RkMode mode = (env == 1 && fits) ? RkMode::Line
                            : (env == 3 && fits) ? RkMode::Hybrid : RkMode::Launch;
        run_rk(ctx, reinterpret_cast<LineDesc *>(w), reinterpret_cast<NoiseDesc *>(w + off_noise), 1, n,
               dt, *rp, h, rk_scratch(w + off_rk, 1, n, rk_tiles(n)), ctx->pinned.as<unsigned>(), out,
               mode, env == 3 ? -1.0 : kRkSwitchMovedPerLine);
        *result = to_result(out[0]);
    });
}

qew_status qew_batch_rk_block(qew_batch b, const qew_rk_params *rp, double *h,
                              qew_rk_result *results) {
    return guard([&] {
        QEW_REQUIRE(b && rp && h && results, "null pointer");
        for (const auto &l : b->lines)
            if (l.n % 2 != 0)
                throw Error(QEW_ERR_INVALID_ARGUMENT,
                            "rk_block: line length must be even for the red-black sweep");
        for (const auto &nd : b->noises)
            QEW_REQUIRE(nd.kind != TABLE || nd.dtype == b->noises[0].dtype,
                        "all tables of a batch must share one storage type");
        // FP64 or FP32 tables (or analytic potentials), or midpoint-contracted
        // FP32 tables on lines of their geometry
        if (b->noises[0].kind == TABLE && b->noises[0].dtype == QEW_F32_MID)
            require_f32_tables(b->lines, b->noises, "rk_block");
        else
            require_wide_tables(b->noises);
        qew_ctx ctx = b->ctx;
        use_device(ctx);
        b->d_scratch.ensure(rk_scratch_bytes(b->nlines, b->total, rk_tiles(b->max_n)));
        ctx->pinned.ensure(sizeof(double) * 64);
        std::vector<RkState> out;
        const int dt = b->noises[0].kind == TABLE ? b->noises[0].dtype : 0;
        const RkScratch sc =
            rk_scratch(b->d_scratch.as<char>(), b->nlines, b->total, rk_tiles(b->max_n));
        auto run = [&](double *hh, const qew_rk_params &p, RkMode mode, double sw) {
            run_rk(ctx, b->d_lines.as<LineDesc>(), b->d_noises.as<NoiseDesc>(), b->nlines, b->max_n,
                   dt, p, hh, sc, ctx->pinned.as<unsigned>(), out, mode, sw);
        };
        // Path: forced by QEW_RK_KERNEL, else per-colour launches when the
        // lines do not fit in LDS, else the faster of the two timed on the
        // batch's first call (16 sweeps each on a copy of the line state; all
        // paths give identical results) and kept for later calls; when the
        // launches win, the call still hands its lines to the LDS kernel once
        // few of their sites move (hybrid).
        RkMode mode = RkMode::Launch;
        double sw = kRkSwitchMovedPerLine;
        const int env = rk_env_path();
        const bool fits = line_kernel_fits(ctx, b->max_n);
        if (env) {
            mode = (env == 1 && fits) ? RkMode::Line : (env == 3 && fits) ? RkMode::Hybrid
                                                                         : RkMode::Launch;
            if (env == 3) sw = -1.0;
        } else if (fits) {
            if (b->rk_path == 0) {
                qew_rk_params trial = *rp;
                trial.max_sweeps = std::min(rp->max_sweeps, 16);
                const size_t bytes = sizeof(double) * (size_t)b->total;
                b->d_tune.ensure(bytes);
                double *copy = b->d_tune.as<double>();
                double t_path[2];
                for (int k = 0; k < 2; ++k) {
                    QEW_HIP(hipMemcpyAsync(copy, h, bytes, hipMemcpyDeviceToDevice, ctx->stream));
                    QEW_HIP(hipStreamSynchronize(ctx->stream));
                    const auto t0 = std::chrono::steady_clock::now();
                    run(copy, trial, k == 0 ? RkMode::Line : RkMode::Launch, sw);
                    t_path[k] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
                                    .count();
                }
                b->rk_path = t_path[0] < t_path[1] ? 1 : 2;
            }
            mode = b->rk_path == 1 ? RkMode::Line : RkMode::Hybrid;
        }
        run(h, *rp, mode, sw);
        for (int l = 0; l < b->nlines; ++l) results[l] = to_result(out[l]);
    });
}

qew_status qew_rk_forward_zero(qew_ctx ctx, qew_noise noise, const qew_params *p,
                               const qew_rk_params *rp, int n, int m, const int *i,
                               const double *hi, const double *hl, const double *hr, double *out,
                               int *found, int64_t *evals) {
    return guard([&] {
        ctx_check(ctx);
        QEW_REQUIRE(noise && p && rp, "null pointer");
        QEW_REQUIRE(n >= 1 && m >= 0, "bad sizes");
        check_march(*rp);
        if (m == 0) return;
        QEW_REQUIRE(i && hi && hl && hr && out, "null pointer");
        require_wide_table(noise->desc);
        use_device(ctx);
        const double dx = p->physical_size / (double)n;
        auto launch = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3((m + 127) / 128), dim3(128), 0, ctx->stream,
                               noise->desc, n, dx, p->line_tension, p->driving_force, p->model,
                               *rp, m, i, hi, hl, hr, out, found, evals);
        };
        if (noise->desc.kind == TABLE && noise->desc.dtype == 1) launch(rk_forward_zero_kernel<float>);
        else launch(rk_forward_zero_kernel<double>);
        QEW_HIP(hipGetLastError());
    });
}

qew_status qew_rk_site_velocity(qew_ctx ctx, qew_noise noise, const qew_params *p, int n, int m,
                                const int *i, const double *hi, const double *hl,
                                const double *hr, double *v) {
    return guard([&] {
        ctx_check(ctx);
        QEW_REQUIRE(noise && p, "null pointer");
        QEW_REQUIRE(n >= 1 && m >= 0, "bad sizes");
        if (m == 0) return;
        QEW_REQUIRE(i && hi && hl && hr && v, "null pointer");
        require_wide_table(noise->desc);
        use_device(ctx);
        hipLaunchKernelGGL(rk_velocity_kernel, dim3((m + 255) / 256), dim3(256), 0, ctx->stream,
                           noise->desc, n, p->physical_size / (double)n, p->line_tension,
                           p->driving_force, p->model, m, i, hi, hl, hr, v);
        QEW_HIP(hipGetLastError());
    });
}

}  // extern "C"
