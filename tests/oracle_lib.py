"""ctypes binding of the CPU oracle (oracle/qew_oracle.cpp) for the tests.

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it as the checker.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")

TABLE, SINCOS, CONST, ZERO, HARMONIC, WASHBOARD, MIDTABLE = range(7)
LINEAR, ARCLENGTH = 0, 1


class Noise(C.Structure):
    _fields_ = [("kind", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double),
                ("coef", C.POINTER(C.c_double)), ("p", C.c_double * 4)]


class Params(C.Structure):
    _fields_ = [("physical_size", C.c_double), ("line_tension", C.c_double),
                ("driving_force", C.c_double), ("model", C.c_int)]


class LbfgsParams(C.Structure):
    _fields_ = [("m", C.c_int), ("ftol", C.c_double), ("max_iter", C.c_int),
                ("c1", C.c_double), ("ls_shrink", C.c_double), ("max_ls", C.c_int),
                ("curvature_eps", C.c_double), ("precondition", C.c_int),
                ("precond_shift", C.c_double)]


class LbfgsResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("max_force", C.c_double),
                ("converged", C.c_int), ("n_func_evals", C.c_long),
                ("n_curvature_skips", C.c_long), ("n_ls_restarts", C.c_long)]


class FireParams(C.Structure):
    """qew::FireParams (src/fire.h:23-33) with its defaults; dt_max has none."""
    _fields_ = [("dt_max", C.c_double), ("dt_init", C.c_double), ("ftol", C.c_double),
                ("max_iter", C.c_int), ("n_min", C.c_int), ("f_inc", C.c_double),
                ("f_dec", C.c_double), ("alpha_start", C.c_double), ("f_alpha", C.c_double)]

    def __init__(self, dt_max, **kw):
        d = dict(dt_max=dt_max, dt_init=0.0, ftol=1e-6, max_iter=100000, n_min=5, f_inc=1.1,
                 f_dec=0.5, alpha_start=0.1, f_alpha=0.99)
        d.update(kw)
        super().__init__(**d)


class FireResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("max_force", C.c_double), ("converged", C.c_int)]


class RkParams(C.Structure):
    _fields_ = [("dstep", C.c_double), ("max_advance", C.c_double),
                ("root_tol", C.c_double), ("sweep_tol", C.c_double),
                ("runaway", C.c_double), ("max_sweeps", C.c_int)]


class RkResult(C.Structure):
    _fields_ = [("blocked", C.c_int), ("runaway", C.c_int), ("sweeps", C.c_int),
                ("total_advance", C.c_double), ("mean_h", C.c_double),
                ("n_velocity_evals", C.c_long)]


_DP = C.POINTER(C.c_double)
_libs = {}


def load(omp=False):
    name = "liboracle_omp.so" if omp else "liboracle.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(ORACLE_DIR, name)
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR, name], check=True)
    lib = C.CDLL(path)
    P = C.POINTER
    lib.orc_bspline_weights.argtypes = [C.c_double, _DP, _DP]
    lib.orc_sample_points.argtypes = [P(Noise), C.c_int, _DP, _DP, _DP, _DP, _DP]
    lib.orc_sample_value.argtypes = [P(Noise), C.c_double, C.c_double]
    lib.orc_sample_value.restype = C.c_double
    lib.orc_sample_dy.argtypes = [P(Noise), C.c_double, C.c_double]
    lib.orc_sample_dy.restype = C.c_double
    lib.orc_objective.argtypes = [C.c_int, _DP, P(Params), P(Noise)]
    lib.orc_objective.restype = C.c_double
    lib.orc_objective_trial.argtypes = [C.c_int, _DP, _DP, C.c_double, P(Params), P(Noise)]
    lib.orc_objective_trial.restype = C.c_double
    lib.orc_gradient.argtypes = [C.c_int, _DP, P(Params), P(Noise), _DP]
    lib.orc_gradient_edges.argtypes = [C.c_int, _DP, P(Params), P(Noise), _DP, _DP, _DP]
    lib.orc_fft.argtypes = [C.c_int, _DP, C.c_int]
    lib.orc_fft2.argtypes = [C.c_int, C.c_int, _DP, C.c_int]
    lib.orc_prefilter.argtypes = [C.c_int, C.c_int, _DP, _DP]
    lib.orc_filtered_noise.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_uint64, _DP, _DP]
    lib.orc_precond_eig.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _DP]
    lib.orc_precondition.argtypes = [C.c_int, _DP, _DP, _DP]
    lib.orc_lbfgs.argtypes = [C.c_int, _DP, P(Params), P(Noise), P(LbfgsParams), P(LbfgsResult)]
    lib.orc_fire.argtypes = [C.c_int, _DP, P(Params), P(Noise), P(FireParams), P(FireResult)]
    lib.orc_philox4x32_10.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
    lib.orc_normal.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
    lib.orc_normal.restype = C.c_double
    lib.orc_langevin_step.argtypes = [C.c_int, _DP, P(Params), P(Noise), C.c_double, C.c_double,
                                      C.c_uint64, C.c_uint64]
    lib.orc_rk_site_velocity.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                         C.c_int, C.c_double, C.c_double, C.c_double, P(Noise)]
    lib.orc_rk_site_velocity.restype = C.c_double
    lib.orc_rk_forward_zero.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_int, C.c_double, C.c_double, C.c_double, P(Noise),
                                        P(RkParams), P(C.c_int), P(C.c_long)]
    lib.orc_rk_forward_zero.restype = C.c_double
    lib.orc_rk_block.argtypes = [C.c_int, _DP, P(Params), P(Noise), P(RkParams), P(RkResult)]
    lib.orc_rk_block.restype = C.c_int
    lib.orc_langevin_run.argtypes = [C.c_int, _DP, P(Params), P(Noise), C.c_int, _DP, _DP,
                                     C.c_uint64, C.c_uint64]
    lib.orc_lbfgs_batch.argtypes = [C.c_int, P(C.c_int64), _DP, P(Params), P(Noise),
                                    P(LbfgsParams), _DP, P(C.c_int), P(LbfgsResult)]
    lib.orc_num_threads.restype = C.c_int
    lib.orc_set_num_threads.argtypes = [C.c_int]
    _libs[name] = lib
    return lib


def dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


class OracleNoise:
    """Owns the arrays an orc_noise points at."""

    def __init__(self, kind, coef=None, nx=0, ny=0, dx=1.0, dy=1.0, p=(0, 0, 0, 0)):
        self.coef = None if coef is None else np.ascontiguousarray(coef, dtype=np.float64)
        self.s = Noise(kind, nx, ny, dx, dy,
                       dp(self.coef) if self.coef is not None else _DP(),
                       (C.c_double * 4)(*p))

    @classmethod
    def table(cls, coef, Lx, Ly):
        nx, ny = coef.shape
        return cls(TABLE, coef, nx, ny, Lx / nx, Ly / ny)

    @classmethod
    def mid_table(cls, values, Lx, Ly):
        """Midpoint-contracted values T[ix, iy] (row-major) of a table of
        spacing Lx / nx, Ly / ny: sampled only at x-fraction 1/2."""
        nx, ny = values.shape
        return cls(MIDTABLE, values, nx, ny, Lx / nx, Ly / ny)

    @property
    def ref(self):
        return C.byref(self.s)


def params(L, lt, f, model):
    return Params(L, lt, f, model)


def lbfgs_params(m=8, ftol=1e-6, max_iter=10000, c1=1e-4, ls_shrink=0.5, max_ls=30,
                 curvature_eps=1e-12, precondition=False, precond_shift=0.0):
    return LbfgsParams(m, ftol, max_iter, c1, ls_shrink, max_ls, curvature_eps,
                       int(precondition), precond_shift)


def rk_params(dstep=0.0, max_advance=0.0, root_tol=1e-10, sweep_tol=1e-9, runaway=0.0,
              max_sweeps=200000):
    return RkParams(dstep, max_advance, root_tol, sweep_tol, runaway, max_sweeps)


# ---- convenience wrappers ------------------------------------------------

def sample_points(noise, x, y):
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    v, dvx, dvy = (np.empty_like(x) for _ in range(3))
    lib.orc_sample_points(noise.ref, len(x), dp(x), dp(y), dp(v), dp(dvx), dp(dvy))
    return v, dvx, dvy


def objective(h, p, noise, omp=False):
    h = np.ascontiguousarray(h, dtype=np.float64)
    return load(omp).orc_objective(len(h), dp(h), C.byref(p), noise.ref)


def objective_trial(x0, d, alpha, p, noise):
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    return load().orc_objective_trial(len(x0), dp(x0), dp(d), alpha, C.byref(p), noise.ref)


def gradient(h, p, noise, omp=False):
    h = np.ascontiguousarray(h, dtype=np.float64)
    g = np.empty_like(h)
    load(omp).orc_gradient(len(h), dp(h), C.byref(p), noise.ref, dp(g))
    return g


def gradient_edges(h, p, noise):
    """(grad, edge dV/dy, edge slope or None) of the reference's two passes."""
    h = np.ascontiguousarray(h, dtype=np.float64)
    g, e = np.empty_like(h), np.empty_like(h)
    s = np.empty_like(h) if p.model == 1 else None
    load().orc_gradient_edges(len(h), dp(h), C.byref(p), noise.ref, dp(g), dp(e),
                              dp(s) if s is not None else _DP())
    return g, e, s


def fft(x, sign=-1):
    x = np.asarray(x, dtype=np.complex128)
    buf = np.ascontiguousarray(np.stack([x.real, x.imag], axis=-1).reshape(-1))
    load().orc_fft(len(x), dp(buf), sign)
    return buf[0::2] + 1j * buf[1::2]


def fft2(a, sign=-1):
    a = np.asarray(a, dtype=np.complex128)
    nx, ny = a.shape
    buf = np.ascontiguousarray(np.stack([a.real, a.imag], axis=-1).reshape(-1))
    load().orc_fft2(nx, ny, dp(buf), sign)
    return (buf[0::2] + 1j * buf[1::2]).reshape(nx, ny)


def prefilter(field):
    field = np.ascontiguousarray(field, dtype=np.float64)
    coef = np.empty_like(field)
    load().orc_prefilter(field.shape[0], field.shape[1], dp(field), dp(coef))
    return coef


def filtered_noise(nx, ny, Lx, Ly, amp, xi_x, xi_y, seed):
    field = np.empty((nx, ny))
    coef = np.empty((nx, ny))
    load().orc_filtered_noise(nx, ny, Lx, Ly, amp, xi_x, xi_y, seed, dp(field), dp(coef))
    return field, coef


def precond_eig(n, kappa, L, shift):
    eig = np.empty(n // 2 + 1)
    load().orc_precond_eig(n, kappa, L, shift, dp(eig))
    return eig


def precondition(q, eig):
    q = np.ascontiguousarray(q, dtype=np.float64)
    d = np.empty_like(q)
    load().orc_precondition(len(q), dp(q), dp(np.ascontiguousarray(eig)), dp(d))
    return d


def lbfgs(h, p, noise, lp, omp=False):
    h = np.array(h, dtype=np.float64)
    r = LbfgsResult()
    load(omp).orc_lbfgs(len(h), dp(h), C.byref(p), noise.ref, C.byref(lp), C.byref(r))
    return h, r


def fire(h, p, noise, fp):
    h = np.array(h, dtype=np.float64)
    r = FireResult()
    load().orc_fire(len(h), dp(h), C.byref(p), noise.ref, C.byref(fp), C.byref(r))
    return h, r


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    load().orc_philox4x32_10(c, k, o)
    return list(o)


def normal(seed, step, site):
    return load().orc_normal(seed, step, site)


def langevin_step(h, p, noise, a, b, seed, step):
    load().orc_langevin_step(len(h), dp(h), C.byref(p), noise.ref, a, b, seed, step)


def rk_site_velocity(i, n, dx, lt, f, model, hi, hl, hr, noise):
    return load().orc_rk_site_velocity(i, n, dx, lt, f, model, hi, hl, hr, noise.ref)


def rk_forward_zero(i, n, dx, lt, f, model, hi, hl, hr, noise, rp):
    """(height, found, evaluations) of the oracle's rk_forward_zero."""
    found = C.c_int()
    evals = C.c_long()
    h = load().orc_rk_forward_zero(i, n, dx, lt, f, model, hi, hl, hr, noise.ref, C.byref(rp),
                                   C.byref(found), C.byref(evals))
    return h, found.value, evals.value


def rk_block(h, p, noise, rp, omp=False):
    r = RkResult()
    rc = load(omp).orc_rk_block(len(h), dp(h), C.byref(p), noise.ref, C.byref(rp), C.byref(r))
    if rc != 0:
        raise ValueError("rk_block: line length must be even for the red-black sweep")
    return r


def langevin_run(h, p, noise, a, b, seed, step0=0, omp=False):
    """len(a) Langevin steps of h in place with per-step coefficients a[k], b[k]."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    assert len(a) == len(b)
    load(omp).orc_langevin_run(len(h), dp(h), C.byref(p), noise.ref, len(a), dp(a), dp(b),
                               seed, step0)


def lbfgs_batch(lines, params, noises, lp, shifts=None, longest_first=True, omp=True):
    """orc_lbfgs on every line (a list of arrays), lines spread over OpenMP
    threads; returns (minimised lines, results)."""
    n = len(lines)
    off = np.concatenate([[0], np.cumsum([len(x) for x in lines])]).astype(np.int64)
    h = np.ascontiguousarray(np.concatenate(lines), dtype=np.float64)
    P = (Params * n)(*params)
    N = (Noise * n)(*[nz.s for nz in noises])
    R = (LbfgsResult * n)()
    sh = None if shifts is None else np.ascontiguousarray(shifts, dtype=np.float64)
    order = None
    if longest_first:
        order = np.ascontiguousarray(np.argsort([-len(x) for x in lines], kind="stable"),
                                     dtype=np.int32)
    load(omp).orc_lbfgs_batch(n, off.ctypes.data_as(C.POINTER(C.c_int64)), dp(h), P, N,
                              C.byref(lp), _DP() if sh is None else dp(sh),
                              C.POINTER(C.c_int)() if order is None else
                              order.ctypes.data_as(C.POINTER(C.c_int)), R)
    return [h[off[k]:off[k + 1]].copy() for k in range(n)], list(R)


def fc_search(tables, n, y0, rp, L=None, lt=1.0, model=LINEAR, f_start=0.125, k_way=4,
              rel_tol=1e-10, workers=16):
    """The K-way depinning-threshold search (relaxation at f = 0, ramp
    f_start * 2^k until not blocked, then k_way forces per round between the
    bracket ends, each started from the highest blocked configuration; a
    timeout counts as not blocked, src/qew_runs.h:224-241 and
    tests/test_rosso_krauth.cpp:76-88) on the sequential oracle's rk_block.
    The lines, and the probes of a round, run on a thread pool (each probe is
    an independent sequential call, so the results do not depend on the
    scheduling). Returns [(lo, hi, h_lo, rounds)] per line."""
    from concurrent.futures import ThreadPoolExecutor
    L = float(n if L is None else L)
    pool = ThreadPoolExecutor(max_workers=workers)

    def probe(h0, f, nz):
        h = h0.copy()
        return rk_block(h, params(L, lt, f, model), nz, rp).blocked, h

    def one(nz):
        h_lo = np.full(n, float(y0))
        assert rk_block(h_lo, params(L, lt, 0.0, model), nz, rp).blocked, \
            "a line does not block at f = 0"
        lo, hi, f = 0.0, np.inf, f_start
        while np.isinf(hi):
            f_probe = max(2 * lo, f)
            blocked, h = probe(h_lo, f_probe, nz)
            if blocked:
                lo, h_lo = f_probe, h
            else:
                hi = f_probe
            f *= 2
        rounds = 0
        while hi - lo > rel_tol * max(abs(lo), abs(hi)):
            fs = [lo + j * (hi - lo) / (k_way + 1) for j in range(1, k_way + 1)]
            got = list(pool.map(lambda fj: probe(h_lo, fj, nz), fs))
            new_lo, new_h, new_hi = lo, h_lo, hi
            for fj, (blocked, h) in zip(fs, got):
                if not blocked:
                    new_hi = fj
                    break
                new_lo, new_h = fj, h
            lo, h_lo, hi = new_lo, new_h, new_hi
            rounds += 1
        return lo, hi, h_lo, rounds

    with ThreadPoolExecutor(max_workers=len(tables)) as lines:
        out = list(lines.map(one, tables))
    pool.shutdown()
    return out
