"""Host set-up for a quantum impurity: the X-He kernel uimp_k.

The reference fills uimp_k(:,:,:,k) (Dynamics_quantum_impurities/
potenimpini.f90:19-29) with the 3-D Fourier transform of a radial X-He
potential,

    V^(p) = (4 pi / q) int_0^inf r V(r) sin(q r) dr,   q = 2 pi |p|,

evaluated by FT_V_spline (FT_V_spline.f90:14-275) on the half spectrum's |p|
(the same p-grid as the other spectral kernels, BCN4HeDFT.f90:337-371).  The
potential is capped below r_cutoff: V = umax there, or, with the reference's
parabolic-fit option (FT_V_spline.f90:132-146), the inverted parabola
a r^2 + c through (r_cutoff, umax) and (r_cutoff + 0.05, V(r_cutoff + 0.05)).

This module computes the same transform with its own quadrature: the core in
closed form, the tail r >= r_cutoff by Filon's method for sine integrals
(Abramowitz & Stegun 25.4.47) on two uniform segments, fine near the core and
coarse in the far tail.  Only distinct |p| values of a grid are transformed.
The table is set-up data for hedft_set_quantum_impurity; nothing here runs on
the time-step path.
"""
from __future__ import annotations

import numpy as np

PARABOLA_DR = 0.05  # dr_FT_V_Spline of FT_V_spline.f90:4


def lennard_jones(eps: float, sigma: float):
    """V(r) = 4 eps ((sigma/r)^12 - (sigma/r)^6), vectorised."""
    def v(r):
        s6 = (sigma / np.asarray(r, dtype=np.float64)) ** 6
        return 4.0 * eps * (s6 * s6 - s6)
    return v


def _filon_weights(theta: np.ndarray):
    """alpha, beta, gamma of Filon's sine rule (A&S 25.4.47), series for small theta."""
    t = np.asarray(theta, dtype=np.float64)
    small = t < 0.2
    ts = np.where(small, 1.0, t)
    s, c = np.sin(ts), np.cos(ts)
    a = 1.0 / ts + s * c / ts**2 - 2.0 * s * s / ts**3
    b = 2.0 * ((1.0 + c * c) / ts**2 - 2.0 * s * c / ts**3)
    g = 4.0 * (s / ts**3 - c / ts**2)
    t2 = t * t
    a_s = t * t2 * (2.0 / 45.0 + t2 * (-2.0 / 315.0 + t2 * (2.0 / 4725.0)))
    b_s = 2.0 / 3.0 + t2 * (2.0 / 15.0 + t2 * (-4.0 / 105.0 + t2 * (2.0 / 567.0)))
    g_s = 4.0 / 3.0 + t2 * (-2.0 / 15.0 + t2 * (1.0 / 210.0 + t2 * (-1.0 / 11340.0)))
    return np.where(small, a_s, a), np.where(small, b_s, b), np.where(small, g_s, g)


def _filon_sin(f: np.ndarray, x0: float, h: float, q: np.ndarray, chunk: int = 64) -> np.ndarray:
    """int_{x0}^{x0+2n h} f(x) sin(q x) dx for each q; f sampled at 2n+1 uniform nodes."""
    n2 = f.size - 1
    assert n2 % 2 == 0 and n2 > 0
    x = x0 + h * np.arange(f.size)
    alpha, beta, gamma = _filon_weights(q * h)
    out = np.empty(q.size)
    fe, fo = f[0::2], f[1::2]
    xe, xo = x[0::2], x[1::2]
    for i0 in range(0, q.size, chunk):
        qq = q[i0:i0 + chunk, None]
        se = np.sin(qq * xe) @ fe
        se -= 0.5 * (f[-1] * np.sin(qq[:, 0] * x[-1]) + f[0] * np.sin(qq[:, 0] * x[0]))
        so = np.sin(qq * xo) @ fo
        ends = f[0] * np.cos(qq[:, 0] * x[0]) - f[-1] * np.cos(qq[:, 0] * x[-1])
        sl = slice(i0, i0 + qq.shape[0])
        out[sl] = h * (alpha[sl] * ends + beta[sl] * se + gamma[sl] * so)
    return out


def _simpson(f: np.ndarray, h: float) -> float:
    return h / 3.0 * (f[0] + f[-1] + 4.0 * f[1:-1:2].sum() + 2.0 * f[2:-1:2].sum())


def _nodes(a: float, b: float, h: float):
    n = max(2, int(np.ceil((b - a) / h)))
    n += n % 2
    return a, (b - a) / n, n + 1


def radial_ft(v, p, r_cutoff: float, umax: float, parabolic: bool = False, r_max: float = 5000.0,
              r_split: float = 50.0, h_near: float = 0.005, h_far: float = 0.5) -> np.ndarray:
    """V^(p) of the capped radial potential v (a vectorised callable of r) at each |p|."""
    p = np.atleast_1d(np.asarray(p, dtype=np.float64))
    q = 2.0 * np.pi * np.abs(p)
    r0 = float(r_cutoff)
    # core r < r0: a r^2 + c  (a = 0, c = umax without the parabolic fit)
    if parabolic:
        r1 = r0 + PARABOLA_DR
        f0, f1 = float(umax), float(v(np.array([r1]))[0])
        a = (f0 - f1) / (r0 * r0 - r1 * r1)
        c = (r0 * r0 * f1 - r1 * r1 * f0) / (r0 * r0 - r1 * r1)
        if a > 0.0:
            raise ValueError("parabolic core fit is not an inverted parabola")
    else:
        a, c = 0.0, float(umax)
    out = np.empty_like(q)
    pos = q > 0.0
    qp = q[pos]
    x = qp * r0
    s, cs = np.sin(x), np.cos(x)
    # int_0^r0 r sin(q r) dr and int_0^r0 r^3 sin(q r) dr
    i1 = (s - x * cs) / qp**2
    i3 = ((3.0 * x * x - 6.0) * s - (x**3 - 6.0 * x) * cs) / qp**4
    core = c * i1 + a * i3
    core0 = c * r0**3 / 3.0 + a * r0**5 / 5.0  # q -> 0: int_0^r0 r^2 (a r^2 + c) dr
    tail = np.zeros_like(qp)
    tail0 = 0.0
    for lo, hi, h in ((r0, min(r_split, r_max), h_near), (min(r_split, r_max), r_max, h_far)):
        if hi <= lo:
            continue
        x0, hh, nn = _nodes(lo, hi, h)
        r = x0 + hh * np.arange(nn)
        vr = np.asarray(v(r), dtype=np.float64)
        tail += _filon_sin(r * vr, x0, hh, qp)
        tail0 += _simpson(r * r * vr, hh)
    out[pos] = 4.0 * np.pi * (core + tail) / qp
    out[~pos] = 4.0 * np.pi * (core0 + tail0)
    return out


def pmod_grid(nx: int, ny: int, nz: int, xmax: float, ymax: float, zmax: float) -> np.ndarray:
    """|p| on the half spectrum (nz, ny, nx//2+1) in FFT order (BCN4HeDFT.f90:337-371)."""
    def axis(n, xm):
        h = 2.0 * xm / n
        return np.fft.fftfreq(n, d=h)
    px = np.abs(axis(nx, xmax)[: nx // 2 + 1])
    py, pz = axis(ny, ymax), axis(nz, zmax)
    return np.sqrt(px[None, None, :] ** 2 + py[None, :, None] ** 2 + pz[:, None, None] ** 2)


def uimp_k_table(nx: int, ny: int, nz: int, xmax: float, ymax: float, zmax: float, v, r_cutoff: float,
                 umax: float, parabolic: bool = False, **kw) -> np.ndarray:
    """uimp_k on the half spectrum (nz, ny, nx//2+1): V^(|p|) of the capped potential v."""
    pm = pmod_grid(nx, ny, nz, xmax, ymax, zmax)
    uniq, inv = np.unique(pm.ravel(), return_inverse=True)
    vals = radial_ft(v, uniq, r_cutoff, umax, parabolic=parabolic, **kw)
    return vals[inv].reshape(pm.shape)
