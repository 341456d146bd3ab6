"""Seeded synthetic droplets and the impurity pair-potential table.

The reference reads its initial density from a file (readen.f90:74-84) and
tabulates the He-impurity potential with Select_Pot (potenimpini.f90:31-59).
The workloads here use:

* a Woods-Saxon droplet rho(r) = 0.0218/(1+exp((r-R)/0.7)), R = 2.22 N^(1/3),
  times (1 + 1e-3 U) with U ~ Uniform[0,1) drawn by numpy's default_rng(seed)
  in C order over (iz, iy, ix), normalised to N atoms;
* a Lennard-Jones He-X pair potential V = 4 eps ((s/r)^12 - (s/r)^6) with the
  Select_Pot core convention (r < r_cutoff -> umax, He_statics/V_impur.f90:
  189-194), an entry the reference's library does not have.
"""
from __future__ import annotations

import numpy as np

PSTOK = 7.63822291        # ps <-> K^-1 conversion of the reference (DFT4He3d.f90:242)
MP_U = 0.02061484         # proton mass in K^-1 A^-2 (Dynamics_TDDFT/modules.f90: mp_u)


def grid_axes(n: int, xmax: float) -> np.ndarray:
    """x_i = -xmax + h (i-1), h = 2 xmax / n (BCN4HeDFT.f90:293-327)."""
    h = 2.0 * xmax / n
    return -xmax + h * np.arange(n)


def woods_saxon(nx, ny, nz, xmax, ymax, zmax, n4, seed=0, rho0=0.0218, a=0.7, noise=1e-3,
                center=(0.0, 0.0, 0.0)) -> np.ndarray:
    x, y, z = grid_axes(nx, xmax), grid_axes(ny, ymax), grid_axes(nz, zmax)
    r = np.sqrt((x[None, None, :] - center[0]) ** 2 + (y[None, :, None] - center[1]) ** 2
                + (z[:, None, None] - center[2]) ** 2)
    R = 2.22 * n4 ** (1.0 / 3.0)
    rho = rho0 / (1.0 + np.exp((r - R) / a))
    rng = np.random.default_rng(seed)
    rho = rho * (1.0 + noise * rng.random(rho.shape))
    dxyz = (2 * xmax / nx) * (2 * ymax / ny) * (2 * zmax / nz)
    return np.ascontiguousarray(rho * (n4 / (rho.sum() * dxyz)))


def select_pot_lj(r, eps=20.0, sigma=4.0, r_cutoff=3.0, umax=None):
    """LJ entry with Select_Pot semantics: umax inside r_cutoff.  By default
    umax = V(r_cutoff), so the capped core is continuous."""
    r = np.asarray(r, dtype=np.float64)
    if umax is None:
        sc6 = (sigma / r_cutoff) ** 6
        umax = 4.0 * eps * (sc6 * sc6 - sc6)
    with np.errstate(divide="ignore", invalid="ignore"):
        sr6 = (sigma / np.where(r > 0, r, 1.0)) ** 6
        v = 4.0 * eps * (sr6 * sr6 - sr6)
    return np.where(r < r_cutoff, umax, v)


def select_pot_lj_dr(r, eps=20.0, sigma=4.0, r_cutoff=3.0, drumax=0.0):
    """dV/dr of the LJ entry, the kind of function the reference's drselec_gs_k_k
    names for the pair force (potenimp.f90:44): 24 eps/r ((s/r)^6 - 2 (s/r)^12)
    outside r_cutoff, drumax inside."""
    r = np.asarray(r, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        rr = np.where(r > 0, r, 1.0)
        sr6 = (sigma / rr) ** 6
        dv = 24.0 * eps / rr * (sr6 - 2.0 * sr6 * sr6)
    return np.where(r < r_cutoff, drumax, dv)


def pair_tables(nimp: int, rmax: float, npair: int = 200001, pairs=None):
    """Impurity-impurity tables on r_i = i*delpair over [0, rmax], one row per pair
    k < m in the order (1,2),(1,3),..,(2,3),..: (vpair, dvpair, delpair).
    pairs: list of LJ keyword dicts (eps, sigma, r_cutoff) per pair, default
    one LJ (eps 20 K, sigma 4 A, r_cutoff 3 A) for every pair."""
    npairs = nimp * (nimp - 1) // 2
    pairs = pairs or [{}] * npairs
    delpair = rmax / (npair - 1)
    r = delpair * np.arange(npair)
    v = np.stack([select_pot_lj(r, **kw) for kw in pairs])
    dv = np.stack([select_pot_lj_dr(r, **{k: x for k, x in kw.items() if k != "umax"}) for kw in pairs])
    return np.ascontiguousarray(v), np.ascontiguousarray(dv), delpair


def impurity_table(n_max: int, box_max: float, **lj):
    """potenimpini (potenimpini.f90:31-59): npot = 100 max(n) + 1 points on
    [0, 4 max(xmax)], DelInter = rmax/(npot-1).  Returns (potion, delinter, rmax)."""
    npot = 100 * n_max + 1
    rmax = 4.0 * box_max
    delinter = rmax / (npot - 1)
    r = delinter * np.arange(npot)
    return np.ascontiguousarray(select_pot_lj(r, **lj)), delinter, rmax


def tddft_setup(drop, n, xmax, n4, seed=0, mass_amu=108.0, r0=(0.0, 0.0, -20.0), v_aps=(0.0, 0.0, 3.0),
                lj=None):
    """Load a Woods-Saxon droplet and one LJ impurity into a context (a Droplet or
    an oracle.Oracle: both expose the same methods).  Velocities in A/ps are
    converted with PSTOK (DFT4He3d.f90:299)."""
    lj = lj or {}
    den = woods_saxon(n, n, n, xmax, xmax, xmax, n4, seed=seed)
    drop.set("DEN", den)
    potion, delinter, rmax = impurity_table(n, xmax, **lj)
    drop.set_impurities(potion[None, :], delinter, rmax, [mass_amu * MP_U])
    drop.set_impurity_state(np.array(r0)[None, :], np.array(v_aps)[None, :] * PSTOK, np.zeros((1, 3)))
    return den


def gaussian_packet(nx, ny, nz, xmax, ymax, zmax, r0=(0.0, 0.0, 0.0), width=1.5, k=(0.0, 0.0, 0.0)):
    """Normalised wave packet exp(-|r-r0|^2/(2 width^2) + i k.r): sum |psix|^2 dxyz = 1
    (the reference reads psix from denX files, readenc.f90)."""
    x, y, z = grid_axes(nx, xmax), grid_axes(ny, ymax), grid_axes(nz, zmax)
    X, Y, Z = x[None, None, :], y[None, :, None], z[:, None, None]
    r2 = (X - r0[0]) ** 2 + (Y - r0[1]) ** 2 + (Z - r0[2]) ** 2
    psi = np.exp(-r2 / (2.0 * width * width) + 1j * (k[0] * X + k[1] * Y + k[2] * Z))
    dxyz = (2 * xmax / nx) * (2 * ymax / ny) * (2 * zmax / nz)
    psi /= np.sqrt((np.abs(psi) ** 2).sum() * dxyz)
    return np.ascontiguousarray(psi)


def quantum_setup(drop, n, xmax, n4, uimp_k, seed=0, m_imp_u=4.002603, mu4=0.0, r0=(0.0, 0.0, 0.0),
                  width=1.5, k=(0.0, 0.0, 0.0)):
    """Woods-Saxon droplet plus one quantum impurity (a wave packet) in a context
    (a Droplet or an oracle.Oracle): den, uimp_k, mass, psix."""
    den = woods_saxon(n, n, n, xmax, xmax, xmax, n4, seed=seed)
    drop.set("DEN", den)
    drop.set_quantum_impurity(uimp_k, m_imp_u, mu4)
    drop.set("PSIX", gaussian_packet(n, n, n, xmax, xmax, xmax, r0, width, k))
    return den


def ensemble(count=48, seed=3, n_lo=200.0, n_hi=2000.0, v_lo=1.0, v_hi=8.0, grid=128, h=0.4, z0=-20.0,
             grids=None):
    """Seeded droplet ensemble: N log-uniform in [n_lo, n_hi], impurity speed
    uniform in [v_lo, v_hi] A/ps toward the centre along +z from z0, impact
    parameter b uniform in [0, R] along x, R = 2.22 N^(1/3).  The grid edge is
    clamped to `grid` (2R+16 A exceeds 128 h for N > ~500; those droplets are
    truncated by the box), or, with `grids` (e.g. (128, 192)), the smallest
    edge that holds 2R+16 A (the largest one when none does: clamped).
    Returns one dict per droplet."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n4 = float(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
        v = float(rng.uniform(v_lo, v_hi))
        R = 2.22 * n4 ** (1.0 / 3.0)
        b = float(rng.uniform(0.0, R))
        need = int(np.ceil((2 * R + 16.0) / h))
        g = grid if grids is None else next((e for e in sorted(grids) if e >= need), max(grids))
        out.append({"n4": n4, "v": v, "b": b, "grid": g, "xmax": 0.5 * g * h,
                    "clamped": need > g, "seed": int(rng.integers(0, 2**31 - 1))})
    return out

