"""qew_amd -- the qEW hot path on AMD MI355X (gfx950), Python host layer.

Every call goes through the C ABI of libqew_amd.so (include/qew_amd.h); torch
tensors on the GPU are only the device buffers handed to it. The names follow
the reference's C++ API:

  Params, Model                     src/qew_model.h:18-27
  objective, objective_trial,
  gradient                          src/qew_model.h:101-201
  FilteredNoise / analytic noises   src/filtered_noise.h:183-191 and the
                                    reference tests' Noise functors
  LbfgsParams, LbfgsResult,
  lbfgs_minimize                    src/lbfgs.h:32-287
  FireParams, FireResult,
  fire_minimize                     src/fire.h:21-131
  LangevinDynamics                  src/dynamics.h:27-70
  RkParams, RkResult, rk_block,
  rk_site_velocity                  src/rosso_krauth.h:34-194
  Batch                             (no reference counterpart: many lines per launch)

There is no fallback: importing works without a GPU, but every call needs the
built library and a HIP device.
"""
import ctypes as C
from dataclasses import dataclass, field as dc_field
from enum import IntEnum

import numpy as np

from . import _lib
from ._lib import (ARCLENGTH, BF16, F32, F32_MID, F64, F64_MID, LINEAR, QewError, QewInvalidArgument, check,
                   load)

__all__ = [
    "Model", "Params", "LbfgsParams", "LbfgsResult", "FireParams", "FireResult",
    "fire_minimize", "RkParams", "RkResult",
    "Context", "context", "FilteredNoise", "AnalyticNoise", "ZeroNoise", "ConstNoise",
    "HarmonicWell", "SinusoidNoise", "objective", "objective_trial", "gradient",
    "evaluate", "lbfgs_minimize", "LangevinDynamics", "rk_block", "rk_site_velocity",
    "rk_forward_zero", "gradient_edges",
    "Batch", "F32", "F32_MID", "F64", "F64_MID", "BF16", "QewError", "QewInvalidArgument",
]


class Model(IntEnum):
    Linear = LINEAR
    Arclength = ARCLENGTH


@dataclass
class Params:
    """qew::Params (src/qew_model.h:20-27)."""
    physical_size: float = 1.0
    line_tension: float = 1.0
    driving_force: float = 0.0
    model: Model = Model.Linear

    def c(self):
        return _lib.Params(float(self.physical_size), float(self.line_tension),
                           float(self.driving_force), int(self.model))


@dataclass
class LbfgsParams:
    """qew::LbfgsParams (src/lbfgs.h:32-52), same defaults."""
    m: int = 8
    ftol: float = 1e-6
    max_iter: int = 10000
    c1: float = 1e-4
    ls_shrink: float = 0.5
    max_ls: int = 30
    curvature_eps: float = 1e-12
    precondition: bool = False
    precond_shift: float = 0.0

    def c(self):
        return _lib.LbfgsParams(self.m, self.ftol, self.max_iter, self.c1, self.ls_shrink,
                                self.max_ls, self.curvature_eps, int(self.precondition),
                                self.precond_shift)


@dataclass
class LbfgsResult:
    """qew::LbfgsResult (src/lbfgs.h:54-61)."""
    iterations: int = 0
    max_force: float = 0.0
    converged: bool = False
    n_func_evals: int = 0
    n_curvature_skips: int = 0
    n_ls_restarts: int = 0

    @classmethod
    def from_c(cls, r):
        return cls(r.iterations, r.max_force, bool(r.converged), r.n_func_evals,
                   r.n_curvature_skips, r.n_ls_restarts)


@dataclass
class FireParams:
    """qew::FireParams (src/fire.h:23-33), same defaults; dt_max has no
    default there either and must be set > 0."""
    dt_max: float = 0.0
    dt_init: float = 0.0
    ftol: float = 1e-6
    max_iter: int = 100000
    n_min: int = 5
    f_inc: float = 1.1
    f_dec: float = 0.5
    alpha_start: float = 0.1
    f_alpha: float = 0.99

    def c(self):
        return _lib.FireParams(self.dt_max, self.dt_init, self.ftol, self.max_iter, self.n_min,
                               self.f_inc, self.f_dec, self.alpha_start, self.f_alpha)


@dataclass
class FireResult:
    """qew::FireResult (src/fire.h:35-39)."""
    iterations: int = 0
    max_force: float = 0.0
    converged: bool = False

    @classmethod
    def from_c(cls, r):
        return cls(r.iterations, r.max_force, bool(r.converged))


@dataclass
class RkParams:
    """qew::RkParams (src/rosso_krauth.h:34-41), same defaults."""
    dstep: float = 0.0
    max_advance: float = 0.0
    root_tol: float = 1e-10
    sweep_tol: float = 1e-9
    runaway: float = 0.0
    max_sweeps: int = 200000

    def c(self):
        return _lib.RkParams(self.dstep, self.max_advance, self.root_tol, self.sweep_tol,
                             self.runaway, self.max_sweeps)


@dataclass
class RkResult:
    """qew::RkResult (src/rosso_krauth.h:43-49)."""
    blocked: bool = False
    runaway: bool = False
    sweeps: int = 0
    total_advance: float = 0.0
    mean_h: float = 0.0
    # site-velocity evaluations the reference algorithm makes (no reference field)
    n_velocity_evals: int = 0
    # evaluations performed on the device (unchanged sites skipped)
    n_velocity_evals_done: int = 0

    @classmethod
    def from_c(cls, r):
        return cls(bool(r.blocked), bool(r.runaway), r.sweeps, r.total_advance, r.mean_h,
                   r.n_velocity_evals, r.n_velocity_evals_done)


def _torch():
    import torch
    return torch


class Context:
    """A qew context on one device, ordered on that device's default stream
    (the stream torch uses unless told otherwise)."""

    def __init__(self, device=0):
        load()
        self.device = device
        h = C.c_void_p()
        check(_lib.load().qew_ctx_create(device, None, C.byref(h)))
        self.handle = h

    def synchronize(self):
        check(_lib.load().qew_ctx_synchronize(self.handle))

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                _lib.load().qew_ctx_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


_contexts = {}


def context(device=0):
    if device not in _contexts:
        _contexts[device] = Context(device)
    return _contexts[device]


def _ptr(t, dtype=None):
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("device buffers must be CUDA (HIP) torch tensors")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"expected {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError("device buffers must be contiguous")
    return C.c_void_p(t.data_ptr())


def _dev(t):
    return t.device.index if t.device.index is not None else 0


# ---------------------------------------------------------------------------
# potentials
# ---------------------------------------------------------------------------
class _Noise:
    handle = None
    ctx = None

    def __del__(self):
        try:
            if self.handle:
                _lib.load().qew_noise_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def info(self):
        k, dt, nx, ny = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        dx, dy = C.c_double(), C.c_double()
        check(_lib.load().qew_noise_info(self.handle, C.byref(k), C.byref(dt), C.byref(nx),
                                         C.byref(ny), C.byref(dx), C.byref(dy)))
        return dict(kind=k.value, dtype=dt.value, nx=nx.value, ny=ny.value, dx=dx.value,
                    dy=dy.value)

    def sample(self, x, y):
        """DeviceNoise::sample at device points: returns (v, dv/dx, dv/dy)."""
        torch = _torch()
        x = x.contiguous().double()
        y = y.contiguous().double()
        v, dvx, dvy = (torch.empty_like(x) for _ in range(3))
        check(_lib.load().qew_noise_sample(self.ctx.handle, self.handle, x.numel(), _ptr(x),
                                           _ptr(y), _ptr(v), _ptr(dvx), _ptr(dvy)))
        return v, dvx, dvy

    def _part(self, name, x, y):
        torch = _torch()
        x = x.contiguous().double()
        y = y.contiguous().double()
        out = torch.empty_like(x)
        check(getattr(_lib.load(), name)(self.ctx.handle, self.handle, x.numel(), _ptr(x),
                                         _ptr(y), _ptr(out)))
        return out

    def sample_value(self, x, y):
        """DeviceNoise::sample_value at device points (the objective's sampler)."""
        return self._part("qew_noise_sample_value", x, y)

    def sample_dy(self, x, y):
        """DeviceNoise::sample_dy at device points (the gradient's sampler)."""
        return self._part("qew_noise_sample_dy", x, y)


class FilteredNoise(_Noise):
    """FilteredNoise (src/filtered_noise.h:183-191). `dtype` is the device
    table's storage type (F64 or F32)."""

    def __init__(self, nx, ny, Lx, Ly, amplitude, xi_x, xi_y, seed, dtype=F64, device=0):
        self.ctx = context(device)
        h = C.c_void_p()
        check(_lib.load().qew_noise_generate(self.ctx.handle, nx, ny, Lx, Ly, amplitude, xi_x,
                                             xi_y, seed, dtype, C.byref(h)))
        self.handle = h

    @classmethod
    def from_field(cls, field, Lx, Ly, dtype=F64, device=0):
        self = cls.__new__(cls)
        self.ctx = context(device)
        f = np.ascontiguousarray(field, dtype=np.float64)
        h = C.c_void_p()
        check(_lib.load().qew_noise_from_field(self.ctx.handle, f.ctypes.data_as(_lib.DP),
                                               f.shape[0], f.shape[1], Lx, Ly, dtype,
                                               C.byref(h)))
        self.handle = h
        return self

    @classmethod
    def from_coefficients(cls, coef, dx, dy, dtype=F64, device=0):
        self = cls.__new__(cls)
        self.ctx = context(device)
        c = np.ascontiguousarray(coef, dtype=np.float64)
        h = C.c_void_p()
        check(_lib.load().qew_noise_from_coefficients(self.ctx.handle, c.ctypes.data_as(_lib.DP),
                                                      c.shape[0], c.shape[1], dx, dy, dtype,
                                                      C.byref(h)))
        self.handle = h
        return self

    @classmethod
    def generate_batch(cls, specs, dtype=F64, device=0):
        """One table per spec (nx, ny, Lx, Ly, amplitude, xi_x, xi_y, seed); the
        independent phase streams are drawn on parallel host threads."""
        ctx = context(device)
        n = len(specs)
        col = lambda k, ct: (ct * n)(*[s[k] for s in specs])
        out = (C.c_void_p * n)()
        check(_lib.load().qew_noise_generate_batch(
            ctx.handle, n, col(0, C.c_int), col(1, C.c_int), col(2, C.c_double),
            col(3, C.c_double), col(4, C.c_double), col(5, C.c_double), col(6, C.c_double),
            col(7, C.c_uint64), dtype, out))
        res = []
        for k in range(n):
            self = cls.__new__(cls)
            self.ctx = ctx
            self.handle = C.c_void_p(out[k])
            res.append(self)
        return res

    def convert(self, dtype):
        """A copy of this potential with its table stored as `dtype`."""
        other = type(self).__new__(type(self))
        other.ctx = self.ctx
        h = C.c_void_p()
        check(_lib.load().qew_noise_convert(self.handle, dtype, C.byref(h)))
        other.handle = h
        return other

    def field(self):
        i = self.info()
        out = np.empty((i["nx"], i["ny"]))
        check(_lib.load().qew_noise_field(self.handle, out.ctypes.data_as(_lib.DP)))
        return out

    def coefficients(self):
        i = self.info()
        out = np.empty((i["nx"], i["ny"]))
        check(_lib.load().qew_noise_coefficients(self.handle, out.ctypes.data_as(_lib.DP)))
        return out


class AnalyticNoise(_Noise):
    """V = amp sin(kx x) cos(ky y) (reference tests/test_qew_model.cpp:22-35)."""
    KIND = _lib.SINCOS

    def __init__(self, *p, device=0):
        self.ctx = context(device)
        arr = (C.c_double * 4)(*(list(map(float, p)) + [0.0] * (4 - len(p))))
        h = C.c_void_p()
        check(_lib.load().qew_noise_analytic(self.ctx.handle, self.KIND, arr, C.byref(h)))
        self.handle = h
        self.params = p


class ConstNoise(AnalyticNoise):
    KIND = _lib.CONST


class ZeroNoise(AnalyticNoise):
    KIND = _lib.ZERO


class HarmonicWell(AnalyticNoise):
    """V = k (y - y0)^2 / 2 (reference tests/test_lbfgs.cpp:29-38)."""
    KIND = _lib.HARMONIC


class SinusoidNoise(AnalyticNoise):
    """V = A cos(k y) (reference tests/test_rosso_krauth.cpp:24-31)."""
    KIND = _lib.WASHBOARD


# ---------------------------------------------------------------------------
# one line (FP64 device tensors)
# ---------------------------------------------------------------------------
def objective(h, p, noise, d=None, alpha=None):
    """objective(h, p, noise); with d and alpha: objective(x0=h, d, alpha)."""
    torch = _torch()
    e = C.c_double()
    pc = p.c()
    if d is None:
        check(_lib.load().qew_objective(noise.ctx.handle, noise.handle, C.byref(pc), h.numel(),
                                        _ptr(h, torch.float64), C.byref(e)))
    else:
        check(_lib.load().qew_objective_trial(noise.ctx.handle, noise.handle, C.byref(pc),
                                              h.numel(), _ptr(h, torch.float64),
                                              _ptr(d, torch.float64), float(alpha), C.byref(e)))
    return e.value


def objective_trial(x0, d, alpha, p, noise):
    return objective(x0, p, noise, d=d, alpha=alpha)


def gradient(h, p, noise, grad=None):
    torch = _torch()
    if grad is None:
        grad = torch.empty_like(h)
    pc = p.c()
    check(_lib.load().qew_gradient(noise.ctx.handle, noise.handle, C.byref(pc), h.numel(),
                                   _ptr(h, torch.float64), _ptr(grad, torch.float64)))
    return grad


def roughness(h, L, acf=True, psd=True):
    """Roughness of a stack of periodic profiles on the device (h: CUDA FP64
    tensor (n,) or (m, n); box length L): returns (ell, A, q, C, w2) with the
    height-difference ACF A(ell) and the PSD C(q) averaged over the profiles
    (None when not asked for) and each profile's squared width w2, all torch
    tensors on h's device; conventions of qew_amd.analysis."""
    torch = _torch()
    h2 = h.reshape(1, -1) if h.dim() == 1 else h
    m, n = h2.shape
    h2 = h2.contiguous()
    kw = dict(dtype=torch.float64, device=h.device)
    a = torch.empty(n // 2, **kw) if acf else None
    c = torch.empty(n // 2, **kw) if psd else None
    w2 = torch.empty(m, **kw)
    check(_lib.load().qew_roughness(context(_dev(h)).handle, m, n, float(L),
                                    _ptr(h2, torch.float64),
                                    None if a is None else _ptr(a, torch.float64),
                                    None if c is None else _ptr(c, torch.float64),
                                    _ptr(w2, torch.float64)))
    k = torch.arange(1, n // 2 + 1, **kw)
    ell = k * (float(L) / n)
    q = 2.0 * np.pi * k / float(L)
    return ell, a, q, c, w2


def gradient_edges(h, p, noise):
    """gradient(h, p, noise, grad, ws) with the GradientWorkspace's per-edge
    quantities (src/qew_model.h:126-201): returns (grad, edge dV/dy, edge slope;
    the slope is None for the linear model)."""
    torch = _torch()
    grad, edvdy = torch.empty_like(h), torch.empty_like(h)
    eslope = torch.empty_like(h) if int(p.model) == ARCLENGTH else None
    pc = p.c()
    check(_lib.load().qew_gradient_edges(noise.ctx.handle, noise.handle, C.byref(pc), h.numel(),
                                         _ptr(h, torch.float64), _ptr(grad, torch.float64),
                                         _ptr(edvdy, torch.float64),
                                         None if eslope is None else _ptr(eslope, torch.float64)))
    return grad, edvdy, eslope


def evaluate(h, p, noise, grad=None):
    """Fused gradient and objective in one pass: returns (E, grad)."""
    torch = _torch()
    if grad is None:
        grad = torch.empty_like(h)
    e = C.c_double()
    pc = p.c()
    check(_lib.load().qew_eval(noise.ctx.handle, noise.handle, C.byref(pc), h.numel(),
                               _ptr(h, torch.float64), _ptr(grad, torch.float64), C.byref(e)))
    return e.value, grad


def lbfgs_minimize(h, p, noise, lp=None):
    """lbfgs_minimize (src/lbfgs.h:94): relaxes h in place (plain host-driven
    FP64 path)."""
    torch = _torch()
    lp = lp or LbfgsParams()
    r = _lib.LbfgsResult()
    pc, lc = p.c(), lp.c()
    check(_lib.load().qew_lbfgs_minimize(noise.ctx.handle, noise.handle, C.byref(pc), C.byref(lc),
                                         h.numel(), _ptr(h, torch.float64), C.byref(r)))
    return LbfgsResult.from_c(r)


def fire_minimize(h, p, noise, fp):
    """fire_minimize (src/fire.h:41): relaxes h (FP64, on the GPU) in place;
    the whole relaxation is one launch."""
    torch = _torch()
    r = _lib.FireResult()
    pc, fc = p.c(), fp.c()
    check(_lib.load().qew_fire_minimize(noise.ctx.handle, noise.handle, C.byref(pc), C.byref(fc),
                                        h.numel(), _ptr(h, torch.float64), C.byref(r)))
    return FireResult.from_c(r)


class LangevinDynamics:
    """LangevinDynamics(n, seed) (src/dynamics.h:27-70); normals from
    Philox4x32-10 keyed by (seed, step, site)."""

    def __init__(self, n, seed):
        self.n = n
        self.seed = seed
        self.steps = 0

    def step(self, h, p, noise, drift_coeff, diff_coeff):
        torch = _torch()
        if h.numel() != self.n:
            raise ValueError("line length differs from the integrator's")
        pc = p.c()
        check(_lib.load().qew_langevin_step(noise.ctx.handle, noise.handle, C.byref(pc), self.n,
                                            _ptr(h, torch.float64), float(drift_coeff),
                                            float(diff_coeff), self.seed, self.steps))
        self.steps += 1


def rk_block(h, p, noise, rp):
    """rk_block (src/rosso_krauth.h:128-194). Raises ValueError for odd n."""
    torch = _torch()
    r = _lib.RkResult()
    pc, rc = p.c(), rp.c()
    check(_lib.load().qew_rk_block(noise.ctx.handle, noise.handle, C.byref(pc), C.byref(rc),
                                   h.numel(), _ptr(h, torch.float64), C.byref(r)))
    return RkResult.from_c(r)


def rk_site_velocity(i, n, p, noise, hi, hl, hr):
    """rk_site_velocity (src/rosso_krauth.h:52-82) at device arrays of sites."""
    torch = _torch()
    v = torch.empty_like(hi)
    pc = p.c()
    i = i.to(torch.int32).contiguous()
    check(_lib.load().qew_rk_site_velocity(noise.ctx.handle, noise.handle, C.byref(pc), n,
                                           i.numel(), _ptr(i), _ptr(hi, torch.float64),
                                           _ptr(hl, torch.float64), _ptr(hr, torch.float64),
                                           _ptr(v, torch.float64)))
    return v


def rk_forward_zero(i, n, p, noise, rp, hi, hl, hr):
    """rk_forward_zero (src/rosso_krauth.h:88-124) at device arrays of sites:
    returns (heights, found, evaluation counts)."""
    torch = _torch()
    out = torch.empty_like(hi)
    found = torch.empty(hi.numel(), dtype=torch.int32, device=hi.device)
    evals = torch.empty(hi.numel(), dtype=torch.int64, device=hi.device)
    pc, rc = p.c(), rp.c()
    i = i.to(torch.int32).contiguous()
    check(_lib.load().qew_rk_forward_zero(noise.ctx.handle, noise.handle, C.byref(pc), C.byref(rc),
                                          n, i.numel(), _ptr(i), _ptr(hi, torch.float64),
                                          _ptr(hl, torch.float64), _ptr(hr, torch.float64),
                                          _ptr(out, torch.float64), _ptr(found), _ptr(evals)))
    return out, found, evals


# ---------------------------------------------------------------------------
# batches
# ---------------------------------------------------------------------------
class Batch:
    """Independent lines evaluated/minimised/driven together. Line l has
    lengths[l] sites, potential noises[l] and parameters params[l]; the lines
    are stored back to back (see `offsets`)."""

    def __init__(self, lengths, noises, params, device=0):
        self.ctx = context(device)
        n = len(lengths)
        self.nlines = n
        self.noises = list(noises)  # keep the potentials alive
        L = (C.c_int * n)(*lengths)
        N = (C.c_void_p * n)(*[nz.handle.value for nz in noises])
        P = (_lib.Params * n)(*[p.c() for p in params])
        h = C.c_void_p()
        check(_lib.load().qew_batch_create(self.ctx.handle, n, L, N, P, C.byref(h)))
        self.handle = h
        off = (C.c_int64 * (n + 1))()
        check(_lib.load().qew_batch_offsets(h, off))
        self.offsets = np.array(off[:], dtype=np.int64)
        self.total = int(self.offsets[-1])

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                _lib.load().qew_batch_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def set_precond_shifts(self, shifts):
        s = (C.c_double * self.nlines)(*map(float, shifts))
        check(_lib.load().qew_batch_set_precond_shifts(self.handle, s))

    def set_forces(self, forces):
        f = (C.c_double * self.nlines)(*map(float, forces))
        check(_lib.load().qew_batch_set_forces(self.handle, f))

    def eval(self, h, grad=None, energy=None, dtype=F64, anchor=None):
        torch = _torch()
        tdt = torch.float32 if dtype == F32 else torch.float64
        if grad is None:
            grad = torch.empty_like(h)
        if energy is None:
            energy = torch.empty(self.nlines, dtype=torch.float64, device=h.device)
        check(_lib.load().qew_batch_eval(self.handle, dtype, _ptr(h, tdt),
                                         None if anchor is None else _ptr(anchor, torch.float64),
                                         _ptr(grad, tdt), _ptr(energy, torch.float64)))
        return energy, grad

    def lbfgs(self, h, lp=None, dtype=F64):
        torch = _torch()
        lp = lp or LbfgsParams()
        res = (_lib.LbfgsResult * self.nlines)()
        lc = lp.c()
        check(_lib.load().qew_batch_lbfgs(self.handle, dtype, C.byref(lc), _ptr(h, torch.float64),
                                          res))
        return [LbfgsResult.from_c(r) for r in res]

    def lbfgs_plan(self):
        """(workgroups, team sizes) of the last lbfgs() launch: the persistent
        workgroups and how many of them solved each line (numpy int array)."""
        wg = C.c_int(0)
        t = np.zeros(self.nlines, dtype=np.int32)
        check(_lib.load().qew_batch_lbfgs_plan(self.handle, C.byref(wg),
                                               t.ctypes.data_as(C.POINTER(C.c_int))))
        return wg.value, t

    def lbfgs_workgroups(self):
        """Persistent workgroups of the last lbfgs() launch."""
        return self.lbfgs_plan()[0]

    def fire(self, h, fp):
        """FIRE on every line (FP64 line state h, one persistent launch)."""
        torch = _torch()
        res = (_lib.FireResult * self.nlines)()
        fc = fp.c()
        check(_lib.load().qew_batch_fire(self.handle, C.byref(fc), _ptr(h, torch.float64), res))
        return [FireResult.from_c(r) for r in res]

    def langevin(self, h, drift, diff, seed, step0=0, dtype=F64, anchor=None):
        torch = _torch()
        tdt = torch.float32 if dtype == F32 else torch.float64
        drift = np.ascontiguousarray(drift, dtype=np.float64)
        diff = np.ascontiguousarray(diff, dtype=np.float64)
        check(_lib.load().qew_batch_langevin(self.handle, dtype, _ptr(h, tdt),
                                             None if anchor is None else _ptr(anchor, torch.float64),
                                             len(drift), drift.ctypes.data_as(_lib.DP),
                                             diff.ctypes.data_as(_lib.DP), seed, step0))

    def rk_block(self, h, rp):
        torch = _torch()
        res = (_lib.RkResult * self.nlines)()
        rc = rp.c()
        check(_lib.load().qew_batch_rk_block(self.handle, C.byref(rc), _ptr(h, torch.float64), res))
        return [RkResult.from_c(r) for r in res]
