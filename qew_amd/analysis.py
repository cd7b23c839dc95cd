"""Roughness analysis of line profiles (host-side numpy post-processing).

The reference computes these quantities outside its C++ code: its static
post-processing (tools/postprocess_static.py:20-47) calls
`height_difference_acf(h, L)` and `psd(h, L)` from a module that is not part
of the reference checkout, and writes their outputs as `ell`, `acf`, `q` and
`psd` beside each relaxed profile `h`. The functions here take the same
arguments and return the same pairs; their conventions are stated per
function (the reference module's own conventions cannot be read, so they
follow the spectral-characterisation definitions of Jacobs, Junge and
Pastewka, Surf. Topogr.: Metrol. Prop. 5, 013001 (2017)). A profile is a
periodic line of n heights on a grid of spacing dx = L / n.

`roughness_exponent` fits the Hurst (roughness) exponent zeta of a
self-affine profile, A(ell) ~ ell^(2 zeta) or C(q) ~ q^-(1 + 2 zeta);
`width2` is the squared interface width W^2 = <(h - <h>)^2>.
"""
import numpy as np

__all__ = ["height_difference_acf", "psd", "width2", "roughness_exponent"]


def _profiles(h):
    a = np.asarray(h, dtype=np.float64)
    if a.ndim not in (1, 2) or a.shape[-1] < 2:
        raise ValueError("h must be one profile (n,) or a stack of profiles (m, n), n >= 2")
    return a


def height_difference_acf(h, L):
    """Height-difference autocorrelation A(ell) = 1/2 <(h(x + ell) - h(x))^2>_x
    of a periodic profile, for ell = r dx, r = 1 .. n // 2.

    h: (n,) or (m, n) heights (a stack is averaged over its profiles).
    Returns (ell, acf). Computed through the circular autocorrelation:
    A(r dx) = R(0) - R(r), R(r) = (1/n) sum_x h(x) h(x + r)."""
    a = _profiles(h)
    n = a.shape[-1]
    a = a - a.mean(axis=-1, keepdims=True)
    f = np.fft.rfft(a, axis=-1)
    R = np.fft.irfft(np.abs(f) ** 2, n=n, axis=-1) / n
    if R.ndim == 2:
        R = R.mean(axis=0)
    r = np.arange(1, n // 2 + 1)
    return r * (float(L) / n), R[0] - R[r]


def psd(h, L):
    """One-dimensional power spectral density of a periodic profile,
    C(q_k) = (L / n^2) |sum_x h(x) exp(-i q_k x)|^2 at q_k = 2 pi k / L,
    k = 1 .. n // 2 (the mean, k = 0, is left out). With this normalisation
    (1 / L) sum over all k != 0 (both signs) of C(q_k) is the variance of h.

    h: (n,) or (m, n) heights (a stack is averaged over its profiles).
    Returns (q, C)."""
    a = _profiles(h)
    n = a.shape[-1]
    f = np.fft.rfft(a, axis=-1)
    C = np.abs(f[..., 1:n // 2 + 1]) ** 2 * (float(L) / n ** 2)
    if C.ndim == 2:
        C = C.mean(axis=0)
    k = np.arange(1, n // 2 + 1)
    return 2.0 * np.pi * k / float(L), C


def width2(h):
    """Squared interface width W^2 = <(h - <h>)^2> per profile."""
    a = _profiles(h)
    return a.var(axis=-1)


def roughness_exponent(h, L, lmin, lmax, method="acf"):
    """Least-squares fit of the roughness exponent zeta on a log-log scale.

    method "acf": A(ell) ~ ell^(2 zeta) over lmin <= ell <= lmax;
    method "psd": C(q) ~ q^-(1 + 2 zeta) over 2 pi / lmax <= q <= 2 pi / lmin.
    h may be a stack of profiles, whose A or C is averaged before the fit.
    Raises ValueError when the window holds fewer than 3 points."""
    if method == "acf":
        x, y = height_difference_acf(h, L)
        sel = (x >= lmin) & (x <= lmax) & (y > 0)
    elif method == "psd":
        x, y = psd(h, L)
        sel = (x >= 2 * np.pi / lmax) & (x <= 2 * np.pi / lmin) & (y > 0)
    else:
        raise ValueError("method must be 'acf' or 'psd'")
    if sel.sum() < 3:
        raise ValueError("the fit window holds fewer than 3 points")
    slope = np.polyfit(np.log(x[sel]), np.log(y[sel]), 1)[0]
    return slope / 2.0 if method == "acf" else (-slope - 1.0) / 2.0
