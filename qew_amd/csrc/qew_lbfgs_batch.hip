// Batched L-BFGS in one persistent launch (src/lbfgs.h:94-287, restated per
// line). One workgroup owns one line at a time and runs the whole
// minimisation -- two-loop recursion, FFT preconditioner, Armijo line search,
// curvature-pair update -- with every scalar decision taken on the device:
// reductions are workgroup-wide and return the same value to every thread, so
// control flow stays uniform and no host round trip happens per iteration.
// Workgroups pull lines from a queue ordered longest first; a finished line
// retires and its workgroup takes the next, which balances ragged batches.
//
// The preconditioner d = irfft(rfft(q) / eig) runs in LDS: the real line of
// length n is packed as n/2 complex values, transformed by an in-place
// Stockham FFT (radix 4, one radix-2 stage for odd log2), unpacked to the
// half spectrum, divided by eig, repacked and transformed back.
//
// Precision modes: FP64 (the reference's arithmetic; sums are workgroup tree
// sums) and FP32 (line offsets from a per-line anchor, FP32 tables, FP32
// vectors, Kahan-compensated FP32 dot products and energies) in which the
// Armijo test, the energies it compares and the curvature pair s.y, s.s, y.y
// stay FP64.
#include <algorithm>
#include <numeric>
#include <vector>

#include "qew_internal.hpp"

using namespace qew_dev;

namespace {

constexpr int kLB = 512;      // threads per workgroup
constexpr int kMaxM = 32;     // history length limit
constexpr int kMaxFftN = 8192;  // complex LDS transform length (line length 16384)

struct BatchArgs {
    const LineDesc *lines;
    const NoiseDesc *noises;
    const int *order;
    int nlines;
    unsigned *next;
    double *h;
    void *work;
    qew_lbfgs_result *results;
    qew_lbfgs_params lp;
};

template <typename T>
struct C2 {
    T x, y;
};

template <typename T>
__device__ __forceinline__ C2<T> cm(C2<T> a, C2<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

template <typename T>
__device__ __forceinline__ C2<T> cis(double num, double den) {
    // exp(i pi * 2 num / den)
    double s, c;
    sincospi(2.0 * num / den, &s, &c);
    return {(T)c, (T)s};
}

// ---- workgroup reductions (uniform results) ----
struct Red {
    double *d;
    KSum *k;
};

__device__ __forceinline__ double wg_sum(double v, Red r) { return block_sum(v, r.d); }

template <typename T>
__device__ double wg_dot(const T *a, const T *b, int n, Red r) {
    if constexpr (sizeof(T) == 8) {
        double acc = 0;
        for (int i = threadIdx.x; i < n; i += blockDim.x) acc += a[i] * b[i];
        return block_sum(acc, r.d);
    } else {
        KSum acc{0.f, 0.f};
        for (int i = threadIdx.x; i < n; i += blockDim.x) acc = ksum_add(acc, KSum{a[i] * b[i], 0.f});
        acc = block_ksum(acc, r.k);
        return (double)acc.s + (double)acc.c;
    }
}

template <typename T>
__device__ double wg_maxabs(const T *a, int n, Red r) {
    double m = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double v = fabs((double)a[i]);
        m = v > m ? v : m;
    }
    return block_max(m, r.d);
}

// ---- objective and gradient of one line ----
template <typename T, typename C>
__device__ double wg_objective(const LineGeom &G, const T *x, const T *d, double alpha, Red r) {
    const int n = G.n;
    if constexpr (sizeof(T) == 8) {
        // src/qew_model.h:53-99 (TrialAt: x0(i) + alpha * d(i))
        double acc = 0;
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const int ip = (i + 1 == n) ? 0 : i + 1;
            const double hi = d ? x[i] + alpha * d[i] : x[i];
            const double hip = d ? x[ip] + alpha * d[ip] : x[ip];
            const double dh = (hip - hi) / G.dx;
            const double le = (G.model == 0) ? G.lt * dh * dh * 0.5 : G.lt * (sqrt(1.0 + dh * dh) - 1.0);
            const double xc = (double)(i + ip) * 0.5 * G.dx;
            const double hc = (hip + hi) * 0.5;
            acc += G.dx * (le + sample_v64<C>(G.nd, xc, hc) - G.f * hi);
        }
        return block_sum(acc, r.d);
    } else {
        // The sufficient-decrease test compares energies whose difference
        // shrinks like |g|^2 while the energy itself is O(n): FP32 terms
        // (even compensated) leave a rounding floor far above that, so the
        // energy of the FP32 trial point (the same FP32 values the commit
        // writes) is evaluated in FP64 on the FP32 table.
        const float al = (float)alpha;
        double acc = 0;
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const int ip = (i + 1 == n) ? 0 : i + 1;
            const double oi = d ? (double)__builtin_fmaf(al, d[i], x[i]) : (double)x[i];
            const double oip = d ? (double)__builtin_fmaf(al, d[ip], x[ip]) : (double)x[ip];
            const double hi = G.anchor + oi, hip = G.anchor + oip;
            const double dh = (oip - oi) / G.dx;
            const double s2 = dh * dh;
            const double le = (G.model == 0) ? G.lt * s2 * 0.5 : G.lt * (s2 / (sqrt(1.0 + s2) + 1.0));
            const double xc = (double)(i + ip) * 0.5 * G.dx;
            acc += G.dx * (le + sample_v64<float>(G.nd, xc, (hip + hi) * 0.5) - G.f * hi);
        }
        return block_sum(acc, r.d);
    }
}

template <typename T, typename C>
__device__ void wg_gradient(const LineGeom &G, const T *x, T *g, T *edy, T *esl) {
    const int n = G.n;
    const bool arc = G.model == 1;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int ip = (i + 1 == n) ? 0 : i + 1;
        if constexpr (sizeof(T) == 8) {
            const double dh = (x[ip] - x[i]) / G.dx;
            edy[i] = sample_dy64<C>(G.nd, (double)(i + ip) * 0.5 * G.dx, (x[ip] + x[i]) * 0.5);
            if (arc) esl[i] = dh / sqrt(1.0 + dh * dh);
        } else {
            float v, dy;
            edge32<false, true>(G, i, ip, x[i], x[ip], v, dy);
            edy[i] = dy;
            const float dh = (x[ip] - x[i]) * G.inv_dx;
            if (arc) esl[i] = dh * rsqrtf(1.0f + dh * dh);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int ip = (i + 1 == n) ? 0 : i + 1;
        const int im = (i == 0) ? n - 1 : i - 1;
        if constexpr (sizeof(T) == 8) {
            double lf;
            if (!arc) lf = -G.lt * (x[im] + x[ip] - 2.0 * x[i]) / (G.dx * G.dx);
            else lf = G.lt * (esl[im] - esl[i]) / G.dx;
            g[i] = G.dx * (lf + 0.5 * (edy[im] + edy[i]) - G.f);
        } else {
            float lf;
            const float lt = (float)G.lt;
            if (!arc) lf = -lt * ((x[im] + x[ip]) - 2.0f * x[i]) * (G.inv_dx * G.inv_dx);
            else lf = lt * (esl[im] - esl[i]) * G.inv_dx;
            g[i] = G.dxf * (lf + 0.5f * (edy[im] + edy[i]) - (float)G.f);
        }
    }
    __syncthreads();
}

// ---- in-place LDS Stockham FFT of N = 2^k complex values ----
template <typename T>
__device__ void lds_fft(C2<T> *buf, int N, int sign) {
    int Ns = 1;
    while (Ns < N) {
        const int R = ((N / Ns) % 4 == 0) ? 4 : 2;
        const int m = N / R;
        // N <= 8192 and 512 threads: at most 4 radix-4 or 8 radix-2
        // butterflies per thread, 16 values either way
        C2<T> v[16];
        int c = 0;
        for (int j = threadIdx.x; j < m; j += blockDim.x, ++c)
            for (int q = 0; q < R; ++q) v[c * R + q] = buf[j + q * m];
        __syncthreads();
        c = 0;
        for (int j = threadIdx.x; j < m; j += blockDim.x, ++c) {
            const int k = j % Ns;
            C2<T> u[4];
            for (int q = 0; q < R; ++q) {
                u[q] = v[c * R + q];
                if (q > 0 && k > 0) u[q] = cm(u[q], cis<T>((double)sign * q * k, (double)Ns * R));
            }
            const int idxD = (j / Ns) * Ns * R + k;
            if (R == 2) {
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
        Ns *= R;
    }
}

// d = irfft(rfft(q) / eig), eig[k] = (kappa/dx) 4 sin^2(pi k / n) + shift
// (src/lbfgs.h:116-146, 173-179); n = 2N.
template <typename T>
__device__ void wg_precondition(const T *q, T *d, int n, double kdx, double shift, C2<T> *buf) {
    const int N = n / 2;
    for (int j = threadIdx.x; j < N; j += blockDim.x) buf[j] = {q[2 * j], q[2 * j + 1]};
    __syncthreads();
    lds_fft(buf, N, -1);
    const double pi = 3.14159265358979323846;
    auto eig = [&](int k) {
        const double s = sin(pi * (double)k / (double)n);
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
        const C2<T> w = cis<T>(-(double)k, (double)n);  // W_n^k
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
    lds_fft(buf, N, +1);
    const T inv = (T)(1.0 / (double)N);
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        d[2 * j] = buf[j].x * inv;
        d[2 * j + 1] = buf[j].y * inv;
    }
    __syncthreads();
}

template <typename T, typename C>
__device__ void solve_line(const BatchArgs &a, int line, Red r, double *rho, double *la,
                           int *slot, int *sbuf, C2<T> *fftbuf) {
    const qew_lbfgs_params &lp = a.lp;
    const LineDesc L = a.lines[line];
    const NoiseDesc nd = a.noises[L.noise];
    const int n = L.n;
    const int m = lp.m;
    const int tid = threadIdx.x;
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

    wg_gradient<T, C>(G, x, g, edy, esl);
    qew_lbfgs_result res{};
    int nslots = 0;  // slot[0..nslots) chronological rows, slot[nslots-1] newest
    double gamma = 1.0;
    int ls_fail_streak = 0;
    double phi = 0;
    bool finished = false;
    for (int it = 0; it < lp.max_iter; ++it) {
        const double gmax = wg_maxabs(g, n, r);
        if (gmax < lp.ftol) {
            res.iterations = it;
            res.max_force = gmax;
            res.converged = 1;
            finished = true;
            break;
        }
        // two-loop recursion; q starts as g
        for (int i = tid; i < n; i += blockDim.x) q[i] = g[i];
        __syncthreads();
        const int hs = nslots;
        for (int idx = hs - 1; idx >= 0; --idx) {
            const int row = slot[idx];
            const T *S = Sb(sbuf[row]), *Y = Yb(sbuf[row]);
            const double al = rho[row] * wg_dot(S, q, n, r);
            if (tid == 0) la[idx] = al;
            const T c = (T)(-al);
            for (int i = tid; i < n; i += blockDim.x) q[i] += c * Y[i];
            __syncthreads();
        }
        if (lp.precondition) {
            wg_precondition<T>(q, d, n, kdx, shift, fftbuf);
        } else {
            const T g0 = (T)(nslots == 0 ? 1.0 : gamma);
            for (int i = tid; i < n; i += blockDim.x) d[i] = g0 * q[i];
            __syncthreads();
        }
        for (int idx = 0; idx < hs; ++idx) {
            const int row = slot[idx];
            const T *S = Sb(sbuf[row]), *Y = Yb(sbuf[row]);
            const double b = rho[row] * wg_dot(Y, d, n, r);
            const T c = (T)(la[idx] - b);
            for (int i = tid; i < n; i += blockDim.x) d[i] += c * S[i];
            __syncthreads();
        }
        for (int i = tid; i < n; i += blockDim.x) d[i] = -d[i];
        __syncthreads();
        double dphi0 = wg_dot(g, d, n, r);
        if (dphi0 >= 0) {
            nslots = 0;
            for (int i = tid; i < n; i += blockDim.x) d[i] = -g[i];
            __syncthreads();
            dphi0 = -wg_dot(g, g, n, r);
        }
        if (it == 0) {
            phi = wg_objective<T, C>(G, x, nullptr, 0.0, r);
            ++res.n_func_evals;
        }
        const double phi0 = phi;
        double alpha = 1.0;
        if (!lp.precondition && nslots == 0) alpha = 1.0 / fmax(1.0, sqrt(wg_dot(g, g, n, r)));
        bool ok = false;
        for (int ls = 0; ls < lp.max_ls; ++ls) {
            const double pt = wg_objective<T, C>(G, x, d, alpha, r);
            ++res.n_func_evals;
            if (pt <= phi0 + lp.c1 * alpha * dphi0) {
                phi = pt;
                ok = true;
                break;
            }
            alpha *= lp.ls_shrink;
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
        for (int i = tid; i < n; i += blockDim.x) {
            if constexpr (sizeof(T) == 8) {
                x[i] += alT * d[i];
            } else {
                const T xo = x[i];
                const T xn = xo + alT * d[i];
                x[i] = xn;
                st[i] = xn - xo;
            }
        }
        __syncthreads();
        wg_gradient<T, C>(G, x, gn, edy, esl);
        double sy = 0, ss = 0, yy = 0;
        for (int i = tid; i < n; i += blockDim.x) {
            const T s = (sizeof(T) == 8) ? alT * d[i] : st[i];
            const T y = gn[i] - g[i];
            st[i] = s;
            yt[i] = y;
            sy += (double)s * (double)y;
            ss += (double)s * (double)s;
            yy += (double)y * (double)y;
        }
        sy = wg_sum(sy, r);
        ss = wg_sum(ss, r);
        yy = wg_sum(yy, r);
        if (sizeof(T) == 4 && ss == 0.0) {
            // FP32 state: the accepted step is below the offsets' resolution
            // everywhere, so the line can no longer move; stop (not
            // converged) instead of iterating in place until max_iter
            res.iterations = it + 1;
            res.max_force = wg_maxabs(gn, n, r);
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
        res.max_force = wg_maxabs(g, n, r);
        res.converged = 0;
    }
    for (int i = tid; i < n; i += blockDim.x) h[i] = anchor + (double)x[i];
    if (tid == 0) a.results[line] = res;
    __syncthreads();
}

template <typename T, typename C>
__global__ __launch_bounds__(kLB) void lbfgs_batch_kernel(BatchArgs a) {
    extern __shared__ unsigned char lds[];
    __shared__ double s_red[16];
    __shared__ KSum s_kred[16];
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
        solve_line<T, C>(a, a.order[qi], r, rho, la, slot, sbuf, reinterpret_cast<C2<T> *>(lds));
    }
}

__global__ void zero_u32(unsigned *p) { *p = 0u; }

}  // namespace

using namespace qew_host;

extern "C" qew_status qew_batch_lbfgs(qew_batch b, int dtype, const qew_lbfgs_params *lp,
                                      double *h, qew_lbfgs_result *results) {
    return guard([&] {
        QEW_REQUIRE(b && lp && h && results, "null pointer");
        QEW_REQUIRE(dtype == QEW_F64 || dtype == QEW_F32, "dtype must be QEW_F64 or QEW_F32");
        QEW_REQUIRE(lp->m >= 1 && lp->m <= kMaxM, "history length m must be in [1, 32]");
        const int tdt = b->noises[0].kind == TABLE ? b->noises[0].dtype : 0;
        for (const auto &nd : b->noises) {
            QEW_REQUIRE(nd.kind != TABLE || nd.dtype == tdt,
                        "all tables of a batch must share one storage type");
            if (dtype == QEW_F32)
                QEW_REQUIRE(nd.kind == TABLE && nd.dtype == 1,
                            "QEW_F32 minimisation needs FP32 coefficient tables");
        }
        if (lp->precondition)
            for (const auto &l : b->lines)
                QEW_REQUIRE(l.n >= 4 && (l.n & (l.n - 1)) == 0 && l.n / 2 <= kMaxFftN,
                            "the batched preconditioner needs power-of-two line lengths in "
                            "[4, 16384]");
        qew_ctx ctx = b->ctx;
        use_device(ctx);
        const size_t es = dtype == QEW_F32 ? sizeof(float) : sizeof(double);
        const size_t nvec = 7 + 2 * ((size_t)lp->m + 1);
        // layout: [order nlines ints][counter][results][work]
        const size_t off_next = ((sizeof(int) * b->nlines + 255) / 256) * 256;
        const size_t off_res = off_next + 256;
        const size_t off_work = off_res + ((sizeof(qew_lbfgs_result) * b->nlines + 255) / 256) * 256;
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
        BatchArgs a{};
        a.lines = b->d_lines.as<LineDesc>();
        a.noises = b->d_noises.as<NoiseDesc>();
        a.order = reinterpret_cast<int *>(w);
        a.nlines = b->nlines;
        a.next = reinterpret_cast<unsigned *>(w + off_next);
        a.h = h;
        a.work = w + off_work;
        a.results = reinterpret_cast<qew_lbfgs_result *>(w + off_res);
        a.lp = *lp;
        int maxN = 1;
        if (lp->precondition)
            for (const auto &l : b->lines) maxN = std::max(maxN, l.n / 2);
        const size_t lds = (size_t)maxN * 2 * es;
        auto launch = [&](auto kernel) {
            if (lds > 64 * 1024)
                QEW_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            int per_cu = 0;
            QEW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kLB, lds));
            const int grid = std::max(1, std::min(b->nlines, std::max(1, per_cu) * ctx->num_cus));
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(kLB), lds, ctx->stream, a);
            QEW_HIP(hipGetLastError());
        };
        if (dtype == QEW_F32) launch(lbfgs_batch_kernel<float, float>);
        else if (tdt == 1) launch(lbfgs_batch_kernel<double, float>);
        else launch(lbfgs_batch_kernel<double, double>);
        QEW_HIP(hipMemcpyAsync(results, a.results, sizeof(qew_lbfgs_result) * b->nlines,
                               hipMemcpyDeviceToHost, ctx->stream));
        QEW_HIP(hipStreamSynchronize(ctx->stream));
    });
}
