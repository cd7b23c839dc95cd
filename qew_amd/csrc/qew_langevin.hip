// Overdamped Langevin step fused with the gradient (src/dynamics.h:34-64):
//   h_i <- h_i + (-drift * dE/dh_i + diff * N(0,1))
// One launch per step over every line of a batch; the step reads the old line
// state and writes a second buffer (the stencil needs the old neighbours), and
// the buffers swap between steps. The normal for (seed, step, site) comes from
// Philox4x32-10 computed in registers: no RNG state in memory, no pool locks,
// and the same stream on any grid shape. diff == 0 skips the draw entirely.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "qew_internal.hpp"

using namespace qew_dev;

namespace {

struct StepArgs {
    const LineDesc *lines;
    const NoiseDesc *noises;
    const int *tile_line;
    const void *h_in;
    void *h_out;
    const double *anchor;
    double drift, diff;
    uint64_t seed;  // line l uses seed + l
    uint64_t step;
};

template <typename C>
__global__ __launch_bounds__(kTile) void langevin64_kernel(StepArgs a, LineDesc one_line,
                                                           NoiseDesc one_noise, int single) {
    __shared__ double s_dvdy[kTile + 1];
    __shared__ double s_slope[kTile + 1];
    __shared__ double s_h[kTile + 1];
    const int tid = threadIdx.x;
    const int tile = blockIdx.x;
    int line;
    LineDesc L;
    NoiseDesc N;
    if (single) {
        line = 0;
        L = one_line;
        N = one_noise;
        L.tile0 = 0;
    } else {
        line = a.tile_line[tile];
        L = a.lines[line];
        N = a.noises[L.noise];
    }
    const int n = L.n;
    // edge (i, i + 1 mod n), i = first - 1 + tid: site i for tid >= 1, the
    // tile's left halo for tid 0
    const int gl = (tile - L.tile0) * kTileSites - 1 + tid;
    const bool owns = tid >= 1 && gl < n;
    const int i = gl < 0 ? n - 1 : gl;
    const double *h = static_cast<const double *>(a.h_in) + L.off;
    const double dx = L.L / (double)n;
    const double rdx = uniform_f64(1.0 / dx), rdx2 = uniform_f64(1.0 / (dx * dx));
    const double lt = L.lt, f = L.f;
    const bool arc = (L.model == 1);
    double hi = 0, hip = 0;
    if (owns || tid == 0) {
        const int ip = (i + 1 == n) ? 0 : i + 1;
        hi = h[i];
        hip = h[ip];
        const double xc = (double)(i + ip) * 0.5 * dx;
        const double hc = (hip + hi) * 0.5;
        s_dvdy[tid] = sample_dy64<C>(N, xc, hc);
        s_h[tid] = hi;
        if (arc) {
            const double dh = divr(hip - hi, dx, rdx);
            s_slope[tid] = dh / sqrt(1.0 + dh * dh);
        }
    }
    __syncthreads();
    if (!owns) return;
    double line_force;
    if (!arc) {
        const double d2h = s_h[tid - 1] + hip - 2.0 * hi;
        line_force = divr(-lt * d2h, dx * dx, rdx2);
    } else {
        line_force = divr(lt * (s_slope[tid - 1] - s_slope[tid]), dx, rdx);
    }
    const double g = dx * (line_force + 0.5 * (s_dvdy[tid - 1] + s_dvdy[tid]) - f);
    double hn;
    if (a.diff > 0)
        hn = hi + (-a.drift * g + a.diff * philox_normal(a.seed + (uint64_t)line, a.step, (uint64_t)i));
    else
        hn = hi + -a.drift * g;
    static_cast<double *>(a.h_out)[L.off + i] = hn;
}

// FP32 line state (offsets from the per-line anchor) on FP32 tables.
__global__ __launch_bounds__(kTile) void langevin32_kernel(StepArgs a) {
    __shared__ float s_dvdy[kTile + 1];
    __shared__ float s_slope[kTile + 1];
    __shared__ float s_h[kTile + 1];
    const int tid = threadIdx.x;
    const int tile = blockIdx.x;
    const int line = a.tile_line[tile];
    const LineDesc L = a.lines[line];
    const NoiseDesc N = a.noises[L.noise];
    const int n = L.n;
    const int gl = (tile - L.tile0) * kTileSites - 1 + tid;
    const bool owns = tid >= 1 && gl < n;
    const int i = gl < 0 ? n - 1 : gl;
    const float *h = static_cast<const float *>(a.h_in) + L.off;
    const double anchor = a.anchor ? a.anchor[line] : 0.0;
    const LineGeom G = make_geom(L, N, anchor);
    const float dxf = G.dxf, inv_dx = G.inv_dx;
    const float lt = (float)L.lt, f = (float)L.f;
    const bool arc = (L.model == 1);
    auto edge_dy = [&](int e, int ep, float he, float hep) -> float {
        float v, dy;
        edge32<false, true>(G, e, ep, he, hep, v, dy);
        return dy;
    };

    float hi = 0.f, hip = 0.f;
    if (owns || tid == 0) {
        const int ip = (i + 1 == n) ? 0 : i + 1;
        hi = h[i];
        hip = h[ip];
        s_dvdy[tid] = edge_dy(i, ip, hi, hip);
        s_h[tid] = hi;
        const float dh = (hip - hi) * inv_dx;
        s_slope[tid] = dh * rsqrtf(1.0f + dh * dh);
    }
    __syncthreads();
    if (!owns) return;
    float lf;
    if (!arc) lf = -lt * ((s_h[tid - 1] + hip) - 2.0f * hi) * (inv_dx * inv_dx);
    else lf = lt * (s_slope[tid - 1] - s_slope[tid]) * inv_dx;
    const float g = dxf * (lf + 0.5f * (s_dvdy[tid - 1] + s_dvdy[tid]) - f);
    float hn = hi - (float)a.drift * g;
    if (a.diff > 0)
        hn += (float)(a.diff * state_normal<float>(a.seed + (uint64_t)line, a.step, (uint64_t)i));
    static_cast<float *>(a.h_out)[L.off + i] = hn;
}

// Several steps per launch. A workgroup holds an extended tile of kLvExt
// consecutive sites (mod n) in LDS and advances all of them; after step s
// the outer s + 1 sites at each end depend on sites outside the tile, so
// after K steps the central kLvExt - 2K sites are exact and are the tile's
// output. Each site's update is the single-step kernel's (same operands,
// operation order and Philox key (seed + line, step, site)), so a launch
// equals K single steps bit for bit, whatever K; the halo work is redundant
// (kLvExt / (kLvExt - 2K)) and saves K - 1 launches and round trips through
// memory. K and the tile kind (these workgroup tiles or the wave tiles
// below) are picked per batch by timing (lv_plan).
// Short lines (n < kLvExt) alias sites within the tile: the copies compute
// identical values and each site is written once.
constexpr int kLvExt = 256;
constexpr int kLvMaxSteps = 32;
template <int K>
constexpr int lv_out() { return kLvExt - 2 * K; }

struct MultiArgs {
    const LineDesc *lines;
    const NoiseDesc *noises;
    const int2 *tiles;  // (line, first output site) per workgroup
    const void *h_in;
    void *h_out;
    const double *anchor;
    double drift[kLvMaxSteps], diff[kLvMaxSteps];
    int nsteps;     // steps of this launch, 1..K
    int ntiles;     // entries of tiles
    uint64_t seed;  // line l uses seed + l
    uint64_t step0;
};

// extended position t of a tile (halo K) whose first output site is `out0`
template <int K>
__device__ __forceinline__ int ext_site(int out0, int t, int n) {
    int gi = (out0 - K + t) % n;
    return gi < 0 ? gi + n : gi;
}

// Kicks diff[s] * N(0,1) of steps [s0, s1) of this launch for the tile's
// sites, passed to store(s, t, kick) for every step with diff[s] > 0. Sites
// 2p and 2p + 1 share Philox block (key, step, p): when threads t and t ^ 1
// hold the two sites of one pair they split the steps (even thread: even
// steps, odd thread: odd steps) and each block feeds both sites (cos to the
// even site, sin to the odd one), so a thread draws half as many blocks as
// it has steps, and the draws of different steps are independent. A thread
// whose partner holds another pair (the line wraps inside the tile) draws
// its own site's blocks. s0 is even.
template <typename S, class Store>
__device__ __forceinline__ void draw_kicks(const MultiArgs &a, uint64_t key, int t, int gi, int gj,
                                           int s0, int s1, Store store) {
    if (gi != gj && (gi >> 1) == (gj >> 1)) {
        const int te = (gi & 1) ? (t ^ 1) : t;  // thread holding the even site
        for (int s2 = s0; s2 < s1; s2 += 2) {
            // s2 is even and below s1 <= kLvMaxSteps, so s2 + 1 < kLvMaxSteps
            const int s = s2 + (t & 1);
            const double df = (t & 1) ? a.diff[s2 + 1] : a.diff[s2];
            if (s < s1 && df > 0) {
                double zc, zs;
                state_normal_pair<S>(key, a.step0 + (uint64_t)s, (uint64_t)(gi >> 1), zc, zs);
                store(s, te, df * zc);
                store(s, te ^ 1, df * zs);
            }
        }
    } else {
        for (int s = s0; s < s1; ++s) {
            const double df = a.diff[s];
            if (df > 0) store(s, t, df * state_normal<S>(key, a.step0 + (uint64_t)s, (uint64_t)gi));
        }
    }
}

// Steps whose kicks one draw pass keeps in LDS: all K for FP32 state (4 B
// per kick), at most 16 for FP64 state (8 B), so the kick table stays within
// 32 KB per workgroup.
template <int K, int BYTES>
constexpr int lv_kick_chunk() { return (BYTES == 4 || K < 16) ? K : 16; }

// Step loop shared by both precisions (two barriers per step): kicks are
// drawn at the start of each chunk of KC steps, heights ping-pong between
// two LDS rows so a step's writes never meet the reads of the same step.
template <typename C, int K>
__global__ __launch_bounds__(kLvExt) void langevin64_multi(MultiArgs a) {
    constexpr int KC = lv_kick_chunk<K, 8>();
    __shared__ double s_h[2][kLvExt];
    __shared__ double s_dvdy[kLvExt];   // [t + 1]: edge between positions t, t + 1
    __shared__ double s_slope[kLvExt];
    __shared__ double s_kick[KC][kLvExt];
    const int t = threadIdx.x;
    const int2 tl = a.tiles[blockIdx.x];
    const int line = tl.x;
    const LineDesc L = a.lines[line];
    const NoiseDesc N = a.noises[L.noise];
    const int n = L.n;
    const double dx = L.L / (double)n;
    const double rdx = uniform_f64(1.0 / dx), rdx2 = uniform_f64(1.0 / (dx * dx));
    const double lt = L.lt, f = L.f;
    const bool arc = (L.model == 1);
    const int gi = ext_site<K>(tl.y, t, n);
    const int gj = ext_site<K>(tl.y, t ^ 1, n);
    const int gp = (gi + 1 == n) ? 0 : gi + 1;
    const double xc = (double)(gi + gp) * 0.5 * dx;
    const uint64_t key = a.seed + (uint64_t)line;
    auto store = [&](int s, int u, double k) { s_kick[s % KC][u] = k; };
    s_h[0][t] = static_cast<const double *>(a.h_in)[L.off + gi];
    for (int s = 0; s < a.nsteps; ++s) {
        const int cur = s & 1;
        if (s % KC == 0) {
            if (s > 0) __syncthreads();  // the previous chunk's kicks have been read
            draw_kicks<double>(a, key, t, gi, gj, s, min(s + KC, a.nsteps), store);
        }
        __syncthreads();
        const double hi = s_h[cur][t];
        if (t + 1 < kLvExt) {
            const double hp = s_h[cur][t + 1];
            s_dvdy[t + 1] = sample_dy64<C>(N, xc, (hp + hi) * 0.5);
            if (arc) {
                const double dh = divr(hp - hi, dx, rdx);
                s_slope[t + 1] = dh / sqrt(1.0 + dh * dh);
            }
        }
        __syncthreads();
        double hn = hi;
        if (t >= 1 && t + 1 < kLvExt) {
            double line_force;
            if (!arc) line_force = divr(-lt * (s_h[cur][t - 1] + s_h[cur][t + 1] - 2.0 * hi), dx * dx, rdx2);
            else line_force = divr(lt * (s_slope[t] - s_slope[t + 1]), dx, rdx);
            const double g = dx * (line_force + 0.5 * (s_dvdy[t] + s_dvdy[t + 1]) - f);
            const double dr = a.drift[s];
            if (a.diff[s] > 0) hn = hi + (-dr * g + s_kick[s % KC][t]);
            else hn = hi + -dr * g;
        }
        s_h[cur ^ 1][t] = hn;
    }
    // each thread reads back the row it wrote last
    const int o = t - K;
    if (o >= 0 && o < lv_out<K>() && tl.y + o < n)
        static_cast<double *>(a.h_out)[L.off + tl.y + o] = s_h[a.nsteps & 1][t];
}

template <int K>
__global__ __launch_bounds__(kLvExt) void langevin32_multi(MultiArgs a) {
    constexpr int KC = lv_kick_chunk<K, 4>();
    __shared__ float s_h[2][kLvExt];
    __shared__ float s_dvdy[kLvExt];
    __shared__ float s_slope[kLvExt];
    __shared__ float s_kick[KC][kLvExt];
    const int t = threadIdx.x;
    const int2 tl = a.tiles[blockIdx.x];
    const int line = tl.x;
    const LineDesc L = a.lines[line];
    const NoiseDesc N = a.noises[L.noise];
    const int n = L.n;
    const double anchor = a.anchor ? a.anchor[line] : 0.0;
    const LineGeom G = make_geom(L, N, anchor);
    const float dxf = G.dxf, inv_dx = G.inv_dx;
    const float lt = (float)L.lt, f = (float)L.f;
    const bool arc = (L.model == 1);
    const int gi = ext_site<K>(tl.y, t, n);
    const int gj = ext_site<K>(tl.y, t ^ 1, n);
    const int gp = (gi + 1 == n) ? 0 : gi + 1;
    const uint64_t key = a.seed + (uint64_t)line;
    auto store = [&](int s, int u, double k) { s_kick[s % KC][u] = (float)k; };
    s_h[0][t] = static_cast<const float *>(a.h_in)[L.off + gi];
    for (int s = 0; s < a.nsteps; ++s) {
        const int cur = s & 1;
        if (s % KC == 0) {
            if (s > 0) __syncthreads();  // the previous chunk's kicks have been read
            draw_kicks<float>(a, key, t, gi, gj, s, min(s + KC, a.nsteps), store);
        }
        __syncthreads();
        const float hi = s_h[cur][t];
        if (t + 1 < kLvExt) {
            const float hp = s_h[cur][t + 1];
            float v, dy;
            edge32<false, true>(G, gi, gp, hi, hp, v, dy);
            s_dvdy[t + 1] = dy;
            const float dh = (hp - hi) * inv_dx;
            s_slope[t + 1] = dh * rsqrtf(1.0f + dh * dh);
        }
        __syncthreads();
        float hn = hi;
        if (t >= 1 && t + 1 < kLvExt) {
            float lf;
            if (!arc) lf = -lt * ((s_h[cur][t - 1] + s_h[cur][t + 1]) - 2.0f * hi) * (inv_dx * inv_dx);
            else lf = lt * (s_slope[t] - s_slope[t + 1]) * inv_dx;
            const float g = dxf * (lf + 0.5f * (s_dvdy[t] + s_dvdy[t + 1]) - f);
            hn = hi - (float)a.drift[s] * g;
            if (a.diff[s] > 0) hn += s_kick[s % KC][t];
        }
        s_h[cur ^ 1][t] = hn;
    }
    const int o = t - K;
    if (o >= 0 && o < lv_out<K>() && tl.y + o < n)
        static_cast<float *>(a.h_out)[L.off + tl.y + o] = s_h[a.nsteps & 1][t];
}

// Wave tiles: one wavefront holds kWave consecutive sites (mod n) in
// registers -- kWave - 2K output sites and a K-site halo each side -- and
// advances them K steps. Neighbour heights, edge derivatives and slopes come
// from cross-lane shuffles, so a step has no barrier and no LDS round trip; a
// workgroup is four independent wave tiles. Each site's update uses the same
// operations, operand order and Philox keys as the workgroup tiles above
// (bit-identical results). Kicks are held in registers, so the step loops
// are unrolled.
template <int K>
constexpr int lv_wave_out() { return kWave - 2 * K; }
constexpr int kLvWavesPerBlock = kLvExt / kWave;

// kick[s] = diff[s] * N(0,1) for steps 0 .. K-1 of this launch (entries of
// steps at or beyond nsteps are not used). Pairs of lanes holding sites 2p
// and 2p + 1 split the steps as in draw_kicks and swap the second normal of
// each block through a lane-xor shuffle; other lanes draw their own.
template <int K, typename T>
__device__ __forceinline__ void wave_kicks(const MultiArgs &a, uint64_t key, int lane, int gi,
                                           int gj, T (&kick)[K]) {
    const bool paired = gi != gj && (gi >> 1) == (gj >> 1);
    const int odd = lane & 1;
#pragma unroll
    for (int s2 = 0; s2 < K; s2 += 2) {
        const int s = paired ? s2 + odd : s2;
        double zc, zs;
        state_normal_pair<T>(key, a.step0 + (uint64_t)s, (uint64_t)(gi >> 1), zc, zs);
        const double mine = (gi & 1) ? zs : zc;   // this lane's site, step s
        const double other = (gi & 1) ? zc : zs;  // the pair's other site, step s
        const double recv = __shfl_xor(other, 1); // paired: this site, step s ^ 1
        const double d0 = a.diff[s2], d1 = a.diff[s2 + 1];
        if (paired) {
            kick[s2] = (T)(d0 * (odd ? recv : mine));
            kick[s2 + 1] = (T)(d1 * (odd ? mine : recv));
        } else {
            kick[s2] = (T)(d0 * mine);
            kick[s2 + 1] = (T)(d1 * state_normal<T>(key, a.step0 + (uint64_t)(s2 + 1), (uint64_t)gi));
        }
    }
}

template <int K>
__global__ __launch_bounds__(kLvExt) void langevin32_wave(MultiArgs a) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wt = blockIdx.x * kLvWavesPerBlock + (int)(threadIdx.x / kWave);
    if (wt >= a.ntiles) return;  // the whole wavefront
    const int2 tl = a.tiles[wt];
    const int line = tl.x;
    const LineDesc L = a.lines[line];
    const NoiseDesc N = a.noises[L.noise];
    const int n = L.n;
    const double anchor = a.anchor ? a.anchor[line] : 0.0;
    const LineGeom G = make_geom(L, N, anchor);
    const float dxf = G.dxf, inv_dx = G.inv_dx;
    const float lt = (float)L.lt, f = (float)L.f;
    const bool arc = (L.model == 1);
    const int gi = ext_site<K>(tl.y, lane, n);
    const int gj = ext_site<K>(tl.y, lane ^ 1, n);
    const int gp = (gi + 1 == n) ? 0 : gi + 1;
    const bool upd = lane >= 1 && lane + 1 < kWave;
    float kick[K];
    wave_kicks<K>(a, a.seed + (uint64_t)line, lane, gi, gj, kick);
    float h = static_cast<const float *>(a.h_in)[L.off + gi];
#pragma unroll
    for (int s = 0; s < K; ++s) {
        if (s >= a.nsteps) break;
        // lane 63 has no right neighbour: its edge is evaluated and discarded
        const float hp = __shfl_down(h, 1);
        const float hm = __shfl_up(h, 1);
        float v, dy;
        edge32<false, true>(G, gi, gp, h, hp, v, dy);
        const float dh = (hp - h) * inv_dx;
        const float dyl = __shfl_up(dy, 1);
        float lf;
        if (!arc) {
            lf = -lt * ((hm + hp) - 2.0f * h) * (inv_dx * inv_dx);
        } else {
            const float sl = dh * rsqrtf(1.0f + dh * dh);
            lf = lt * (__shfl_up(sl, 1) - sl) * inv_dx;
        }
        const float g = dxf * (lf + 0.5f * (dyl + dy) - f);
        float hn = h - (float)a.drift[s] * g;
        if (a.diff[s] > 0) hn += kick[s];
        if (upd) h = hn;
    }
    const int o = lane - K;
    if (o >= 0 && o < lv_wave_out<K>() && tl.y + o < n)
        static_cast<float *>(a.h_out)[L.off + tl.y + o] = h;
}

template <typename C, int K>
__global__ __launch_bounds__(kLvExt) void langevin64_wave(MultiArgs a) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wt = blockIdx.x * kLvWavesPerBlock + (int)(threadIdx.x / kWave);
    if (wt >= a.ntiles) return;  // the whole wavefront
    const int2 tl = a.tiles[wt];
    const int line = tl.x;
    const LineDesc L = a.lines[line];
    const NoiseDesc N = a.noises[L.noise];
    const int n = L.n;
    const double dx = L.L / (double)n;
    const double rdx = uniform_f64(1.0 / dx), rdx2 = uniform_f64(1.0 / (dx * dx));
    const double lt = L.lt, f = L.f;
    const bool arc = (L.model == 1);
    const int gi = ext_site<K>(tl.y, lane, n);
    const int gj = ext_site<K>(tl.y, lane ^ 1, n);
    const int gp = (gi + 1 == n) ? 0 : gi + 1;
    const double xc = (double)(gi + gp) * 0.5 * dx;
    const bool upd = lane >= 1 && lane + 1 < kWave;
    double kick[K];
    wave_kicks<K>(a, a.seed + (uint64_t)line, lane, gi, gj, kick);
    double h = static_cast<const double *>(a.h_in)[L.off + gi];
#pragma unroll
    for (int s = 0; s < K; ++s) {
        if (s >= a.nsteps) break;
        const double hp = __shfl_down(h, 1);
        const double hm = __shfl_up(h, 1);
        const double dy = sample_dy64<C>(N, xc, (hp + h) * 0.5);
        const double dyl = __shfl_up(dy, 1);
        double line_force;
        if (!arc) {
            line_force = divr(-lt * (hm + hp - 2.0 * h), dx * dx, rdx2);
        } else {
            const double dh = divr(hp - h, dx, rdx);
            const double sl = dh / sqrt(1.0 + dh * dh);
            line_force = divr(lt * (__shfl_up(sl, 1) - sl), dx, rdx);
        }
        const double g = dx * (line_force + 0.5 * (dyl + dy) - f);
        const double dr = a.drift[s];
        double hn;
        if (a.diff[s] > 0) hn = h + (-dr * g + kick[s]);
        else hn = h + -dr * g;
        if (upd) h = hn;
    }
    const int o = lane - K;
    if (o >= 0 && o < lv_wave_out<K>() && tl.y + o < n)
        static_cast<double *>(a.h_out)[L.off + tl.y + o] = h;
}

// (line, first output site) per tile of `out` output sites
void ensure_tiles(qew_batch b, int out) {
    if (b->ltile_out == out) return;
    std::vector<int2> tiles;
    for (int l = 0; l < b->nlines; ++l)
        for (int o = 0; o < b->lines[l].n; o += out) tiles.push_back(make_int2(l, o));
    b->d_ltile.ensure(sizeof(int2) * tiles.size());
    QEW_HIP(hipMemcpyAsync(b->d_ltile.ptr, tiles.data(), sizeof(int2) * tiles.size(),
                           hipMemcpyHostToDevice, b->ctx->stream));
    QEW_HIP(hipStreamSynchronize(b->ctx->stream));
    b->ltiles = (int)tiles.size();
    b->ltile_out = out;
}

// Launch plan: plan > 0 runs `plan` steps per launch on workgroup tiles,
// plan < 0 runs -plan steps per launch on wave tiles.
int plan_steps(int plan) { return plan > 0 ? plan : -plan; }
int plan_out(int plan) { return plan > 0 ? kLvExt - 2 * plan : kWave + 2 * plan; }

template <int K>
void launch_block(qew_batch b, int dtype, const MultiArgs &m) {
    const dim3 grid(b->ltiles), block(kLvExt);
    hipStream_t s = b->ctx->stream;
    if (dtype == QEW_F32) hipLaunchKernelGGL(langevin32_multi<K>, grid, block, 0, s, m);
    else if (b->noises[0].dtype == 1) hipLaunchKernelGGL((langevin64_multi<float, K>), grid, block, 0, s, m);
    else hipLaunchKernelGGL((langevin64_multi<double, K>), grid, block, 0, s, m);
    QEW_HIP(hipGetLastError());
}

template <int K>
void launch_wave(qew_batch b, int dtype, const MultiArgs &m) {
    const dim3 grid((b->ltiles + kLvWavesPerBlock - 1) / kLvWavesPerBlock), block(kLvExt);
    hipStream_t s = b->ctx->stream;
    if (dtype == QEW_F32) hipLaunchKernelGGL(langevin32_wave<K>, grid, block, 0, s, m);
    else if (b->noises[0].dtype == 1) hipLaunchKernelGGL((langevin64_wave<float, K>), grid, block, 0, s, m);
    else hipLaunchKernelGGL((langevin64_wave<double, K>), grid, block, 0, s, m);
    QEW_HIP(hipGetLastError());
}

// nsteps steps with launch plan `plan`, ping-ponging between bufs[0] (input)
// and bufs[1]; returns the buffer that holds the result
void *multi_steps(qew_batch b, int plan, int dtype, void *bufs[2], const double *anchor, int nsteps,
                  const double *drift, const double *diff, uint64_t seed, uint64_t step0) {
    const int K = plan_steps(plan);
    ensure_tiles(b, plan_out(plan));
    MultiArgs m{};
    m.lines = b->d_lines.as<LineDesc>();
    m.noises = b->d_noises.as<NoiseDesc>();
    m.tiles = b->d_ltile.as<int2>();
    m.ntiles = b->ltiles;
    m.anchor = anchor;
    m.seed = seed;
    int launches = 0;
    for (int k0 = 0; k0 < nsteps; k0 += K, ++launches) {
        m.nsteps = std::min(K, nsteps - k0);
        for (int s = 0; s < m.nsteps; ++s) {
            m.drift[s] = drift[k0 + s];
            m.diff[s] = diff[k0 + s];
        }
        m.step0 = step0 + (uint64_t)k0;
        m.h_in = bufs[launches & 1];
        m.h_out = bufs[(launches + 1) & 1];
        switch (plan) {
            case 8: launch_block<8>(b, dtype, m); break;
            case 16: launch_block<16>(b, dtype, m); break;
            case kLvMaxSteps: launch_block<kLvMaxSteps>(b, dtype, m); break;
            case -8: launch_wave<8>(b, dtype, m); break;
            case -16: launch_wave<16>(b, dtype, m); break;
            default: throw qew_host::Error(QEW_ERR_INVALID_ARGUMENT, "bad Langevin launch plan");
        }
    }
    return bufs[launches & 1];
}

// Launch plan for this batch. QEW_LANGEVIN_TILE ("block" or "wave") and
// QEW_LANGEVIN_STEPS_PER_LAUNCH (8, 16 or 32; 32 for block tiles only) pin it
// (16 steps when only the tile kind is given, block tiles when only the step
// count is); otherwise the plan timed on the batch's first call of at least
// 96 steps (32 steps of each candidate on a copy of the line state, so the
// caller's state is untouched; every plan gives identical results); 16
// steps on block tiles before that.
int lv_plan(qew_batch b, int dtype, const void *h, const double *anchor, int nsteps,
            const double *drift, const double *diff, uint64_t seed, uint64_t step0) {
    const char *env_k = std::getenv("QEW_LANGEVIN_STEPS_PER_LAUNCH");
    const char *env_t = std::getenv("QEW_LANGEVIN_TILE");
    if (env_k || env_t) {
        const int k = env_k ? std::atoi(env_k) : 16;
        QEW_REQUIRE(k == 8 || k == 16 || k == 32, "QEW_LANGEVIN_STEPS_PER_LAUNCH must be 8, 16 or 32");
        const std::string tile = env_t ? env_t : "block";
        QEW_REQUIRE(tile == "block" || tile == "wave", "QEW_LANGEVIN_TILE must be block or wave");
        if (tile == "block") return k;
        QEW_REQUIRE(k != 32, "wave tiles take 8 or 16 steps per launch");
        return -k;
    }
    if (b->lv_k) return b->lv_k;
    if (nsteps < 3 * kLvMaxSteps) return 16;
    hipStream_t s = b->ctx->stream;
    const size_t es = dtype == QEW_F32 ? sizeof(float) : sizeof(double);
    const size_t bytes = es * (size_t)b->total;
    b->d_tune.ensure(2 * bytes);
    char *t0 = b->d_tune.as<char>();
    void *tb[2] = {t0, t0 + bytes};
    const qew_host::EventPair ev;
    int best = 16;
    float best_ms = 0;
    for (int plan : {8, 16, kLvMaxSteps, -8, -16}) {
        QEW_HIP(hipMemcpyAsync(tb[0], h, bytes, hipMemcpyDeviceToDevice, s));
        ensure_tiles(b, plan_out(plan));
        QEW_HIP(hipEventRecord(ev.start, s));
        multi_steps(b, plan, dtype, tb, anchor, kLvMaxSteps, drift, diff, seed, step0);
        QEW_HIP(hipEventRecord(ev.stop, s));
        QEW_HIP(hipEventSynchronize(ev.stop));
        const float ms = ev.elapsed_ms();
        if (best_ms == 0 || ms < best_ms) {
            best_ms = ms;
            best = plan;
        }
    }
    b->lv_k = best;
    return best;
}

}  // namespace

using namespace qew_host;

extern "C" {

qew_status qew_langevin_step(qew_ctx ctx, qew_noise noise, const qew_params *p, int n,
                             double *h, double drift, double diff, uint64_t seed,
                             uint64_t step) {
    return guard([&] {
        ctx_check(ctx);
        QEW_REQUIRE(noise && p && h, "null pointer");
        QEW_REQUIRE(n >= 1, "line length must be >= 1");
        require_wide_table(noise->desc);
        use_device(ctx);
        const LineDesc l = single_line(n, p);
        ctx->work.ensure(sizeof(double) * (size_t)n);
        StepArgs a{};
        a.h_in = h;
        a.h_out = ctx->work.ptr;
        a.drift = drift;
        a.diff = diff;
        a.seed = seed;
        a.step = step;
        if (noise->desc.dtype == 1)
            hipLaunchKernelGGL(langevin64_kernel<float>, dim3(l.ntiles), dim3(kTile), 0, ctx->stream, a,
                               l, noise->desc, 1);
        else
            hipLaunchKernelGGL(langevin64_kernel<double>, dim3(l.ntiles), dim3(kTile), 0, ctx->stream,
                               a, l, noise->desc, 1);
        QEW_HIP(hipGetLastError());
        QEW_HIP(hipMemcpyAsync(h, ctx->work.ptr, sizeof(double) * (size_t)n,
                               hipMemcpyDeviceToDevice, ctx->stream));
    });
}

qew_status qew_batch_langevin(qew_batch b, int dtype, void *h, const double *anchor, int nsteps,
                              const double *drift, const double *diff, uint64_t seed,
                              uint64_t step0) {
    return guard([&] {
        QEW_REQUIRE(b != nullptr, "null batch");
        QEW_REQUIRE(h && drift && diff, "null pointer");
        QEW_REQUIRE(nsteps >= 0, "negative step count");
        QEW_REQUIRE(dtype == QEW_F64 || dtype == QEW_F32, "dtype must be QEW_F64 or QEW_F32");
        if (nsteps == 0) return;
        qew_ctx ctx = b->ctx;
        use_device(ctx);
        const size_t es = dtype == QEW_F32 ? sizeof(float) : sizeof(double);
        if (dtype == QEW_F32) {
            require_f32_tables(b->lines, b->noises, "QEW_F32 dynamics");
        } else {
            QEW_REQUIRE(!anchor, "anchors apply to QEW_F32 line state only");
            require_wide_tables(b->noises);
            for (const auto &nd : b->noises)
                QEW_REQUIRE(nd.kind != TABLE || nd.dtype == b->noises[0].dtype,
                            "all tables of a batch must share one storage type");
        }
        b->d_scratch.ensure(es * (size_t)b->total);
        void *bufs[2] = {h, b->d_scratch.ptr};
        if (nsteps >= 2) {
            const int plan = lv_plan(b, dtype, h, anchor, nsteps, drift, diff, seed, step0);
            void *res = multi_steps(b, plan, dtype, bufs, anchor, nsteps, drift, diff, seed, step0);
            if (res != h)
                QEW_HIP(hipMemcpyAsync(h, res, es * (size_t)b->total, hipMemcpyDeviceToDevice,
                                       ctx->stream));
            return;
        }
        StepArgs a{};
        a.lines = b->d_lines.as<LineDesc>();
        a.noises = b->d_noises.as<NoiseDesc>();
        a.tile_line = b->d_tile_line.as<int>();
        a.anchor = anchor;
        a.seed = seed;
        const int table_dtype = b->noises[0].dtype;
        for (int k = 0; k < nsteps; ++k) {
            a.h_in = bufs[k & 1];
            a.h_out = bufs[(k + 1) & 1];
            a.drift = drift[k];
            a.diff = diff[k];
            a.step = step0 + (uint64_t)k;
            if (dtype == QEW_F32)
                hipLaunchKernelGGL(langevin32_kernel, dim3(b->ntiles), dim3(kTile), 0, ctx->stream, a);
            else if (table_dtype == 1)
                hipLaunchKernelGGL(langevin64_kernel<float>, dim3(b->ntiles), dim3(kTile), 0,
                                   ctx->stream, a, LineDesc{}, NoiseDesc{}, 0);
            else
                hipLaunchKernelGGL(langevin64_kernel<double>, dim3(b->ntiles), dim3(kTile), 0,
                                   ctx->stream, a, LineDesc{}, NoiseDesc{}, 0);
            QEW_HIP(hipGetLastError());
        }
        if (nsteps & 1)
            QEW_HIP(hipMemcpyAsync(h, b->d_scratch.ptr, es * (size_t)b->total,
                                   hipMemcpyDeviceToDevice, ctx->stream));
    });
}

}  // extern "C"
