"""Pins of the oracle's quantum-impurity restatement (Dynamics_quantum_impurities)
and of the package's uimp_k transform, on the CPU.

* radial transform: closed form of a Gaussian; the reference's own FT_V_spline
  (compiled into oracle/_ref) for two of its potentials;
* potenimp: a constant kernel reduces the convolutions to products; a brute-force
  periodic convolution on a small grid fixes normalisation and orientation;
  the He-X interaction energy computed through potx4 equals the one through upotx;
* steppers: a free impurity plane wave rotates by exp(-i h2o2mimp kappa t), with
  kappa the periodic (icon 8) 13-point stencil's symbol, to round-off (a scheme below fourth order,
  a wrong sign of i or a wrong mass factor misses by orders of magnitude);
  coupled real-time evolution conserves both norms;
* energy: ekinx of a plane wave, eHeX of a constant kernel; centroid of a
  symmetric packet;
* several impurities: constant X-He and X-X kernels give potx4 and each upotx
  in closed form (and fix which impurity of a pair feels uimp_k_k(i,j)); a
  brute-force convolution fixes the pair term; two free plane waves of
  different masses rotate independently; ekinx, eHeX and the pair energy of
  plane waves in closed form; swapping two identical impurities swaps the
  trajectory; each impurity keeps its norm.
"""
import numpy as np
import pytest

from oracle import oracle as O
from he_dft_2024_amd import inputs as I
from he_dft_2024_amd import quantum as Q
from tests.conftest import relerr

H2O2M4 = 6.05969638298


def test_radial_ft_gaussian_closed_form():
    A, b = 50.0, 1.3
    v = lambda r: A * np.exp(-(np.asarray(r) / b) ** 2)
    p = np.linspace(0.0, 3.0, 31)
    rc = 1e-3
    got = Q.radial_ft(v, p, rc, float(v(rc)), r_max=30.0, r_split=30.0)
    exact = A * np.pi ** 1.5 * b ** 3 * np.exp(-(np.pi * b * p) ** 2)
    assert relerr(got, exact) < 1e-9


@pytest.mark.skipif(O.ref_q() is None, reason="oracle/_ref/libhedft_refq.so not built")
@pytest.mark.parametrize("selec,rc,umax,parabolic", [("He2S_fci", 3.17502, 580.0065, True),
                                                     ("LJ_OT", 2.190323254008688, 157.4659520268255, False)])
def test_radial_ft_vs_reference_ft_v_spline(selec, rc, umax, parabolic):
    """Values at the reference's own p-table nodes (where its p-spline is exact).
    The transform here is converged (halving its steps changes it by < 1e-8); what
    is left against the reference is the reference's own quadrature (a cubic
    spline of V on 0.05 A steps with finite-difference end slopes), up to 5e-5
    of max |V^| for the steep r^-12 core of LJ_OT."""
    hq, pmax = 0.05, 50.0
    p = np.arange(0, 60) * hq
    ref = O.ref_ft_v_spline(selec, p, pmax, hq, rc, umax, parabolic=parabolic)
    v = lambda r: O.ref_select_pot(selec, r, rc, umax)
    mine = Q.radial_ft(v, p, rc, umax, parabolic=parabolic)
    finer = Q.radial_ft(v, p, rc, umax, parabolic=parabolic, h_near=0.0025, h_far=0.25, r_split=100.0)
    assert relerr(mine, finer) < 1e-8
    assert relerr(mine, ref) < 1e-4


def test_uimp_k_table_is_radial():
    n, xmax = 16, 4.0
    v = Q.lennard_jones(10.22, 2.556)
    t = Q.uimp_k_table(n, n, n, xmax, xmax, xmax, v, 2.0, float(v(2.0)))
    pm = Q.pmod_grid(n, n, n, xmax, xmax, xmax)
    assert t.shape == (n, n, n // 2 + 1)
    # equal |p| -> equal value; p = 0 entry is 4 pi int r^2 V
    sel = np.isclose(pm, pm[0, 0, 1])
    assert np.ptp(t[sel]) == 0.0
    assert np.isclose(t[0, 0, 0], Q.radial_ft(v, [0.0], 2.0, float(v(2.0)))[0])


def _quantum_oracle(n=16, xmax=3.2, uimp_k=0.0, m=4.002603, otc=0, psix=None, den=None, **kw):
    o = O.Oracle(n, n, n, xmax, xmax, xmax, O.default_params(core_otc=otc, **kw))
    o.set("DEN", den if den is not None else I.woods_saxon(n, n, n, xmax, xmax, xmax, 20.0, seed=1))
    o.set_quantum_impurity(uimp_k, m)
    o.set("PSIX", psix if psix is not None else I.gaussian_packet(n, n, n, xmax, xmax, xmax, width=1.0))
    return o


def test_potenimp_constant_kernel():
    c = 37.5
    o = _quantum_oracle(uimp_k=c)
    o.potenimp()
    assert relerr(o.field("POTX4"), c * np.abs(o.field("PSIX")) ** 2) < 1e-13
    assert relerr(o.field("UPOTX"), c * o.field("DEN")) < 1e-13


def test_potenimp_brute_force_periodic_convolution():
    n, xmax = 8, 2.0
    rng = np.random.default_rng(4)
    # a real kernel symmetric under r -> -r (periodically) has a real spectrum
    k = rng.standard_normal((n, n, n))
    k = 0.5 * (k + np.roll(k[::-1, ::-1, ::-1], 1, axis=(0, 1, 2)))
    uk = np.fft.rfftn(k)
    assert np.abs(uk.imag).max() < 1e-12
    psix = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
    den = rng.random((n, n, n))
    o = _quantum_oracle(n, xmax, uimp_k=uk.real, psix=psix, den=den)
    o.potenimp()
    rx = np.abs(psix) ** 2
    brute = np.zeros_like(rx)
    bden = np.zeros_like(rx)
    for s in np.ndindex(n, n, n):  # sum_y f(y) K(x - y)
        brute += rx[s] * np.roll(k, s, axis=(0, 1, 2))
        bden += den[s] * np.roll(k, s, axis=(0, 1, 2))
    assert relerr(o.field("POTX4"), brute) < 1e-12
    assert relerr(o.field("UPOTX"), bden) < 1e-12


def test_interaction_energy_both_ways():
    n, xmax = 16, 3.2
    v = Q.lennard_jones(10.22, 2.556)
    uk = Q.uimp_k_table(n, n, n, xmax, xmax, xmax, v, 2.0, float(v(2.0)))
    o = _quantum_oracle(n, xmax, uimp_k=uk)
    o.potenimp()
    a = (o.field("POTX4") * o.field("DEN")).sum()
    b = (o.field("UPOTX") * np.abs(o.field("PSIX")) ** 2).sum()
    assert abs(a - b) <= 1e-12 * abs(a)


def _plane_wave(n, xmax, m):
    x = I.grid_axes(n, xmax)
    kx = 2.0 * np.pi * m / (2.0 * xmax)
    return np.ascontiguousarray(np.broadcast_to(np.exp(1j * kx * x)[None, None, :], (n, n, n))).copy()


def test_free_impurity_plane_wave_phase():
    n, xmax, m_imp = 16, 3.2, 2 * 4.002603
    psi0 = _plane_wave(n, xmax, 4)
    h = 2 * xmax / n
    kappa = -(O.derivnd(2, psi0, h, 1, 8) / psi0).real.mean()  # periodic stencil's symbol
    den = np.full((n, n, n), 0.02)
    mu4 = 3.7
    o = O.Oracle(n, n, n, xmax, xmax, xmax, O.default_params(icon=8))
    o.set("DEN", den)
    o.set_quantum_impurity(0.0, m_imp, mu4)
    o.set("PSIX", psi0)
    o.tddft_init()
    pot4 = o.field("POT4").mean()  # uniform helium: a constant mean field
    assert np.ptp(o.field("POT4")) < 1e-12 * abs(pot4)
    dt = 5e-4 / I.PSTOK
    for it in range(1, 13):
        o.tddft_iteration(dt, it, 3)
    w = H2O2M4 * 4.002603 / m_imp * kappa
    exact = psi0 * np.exp(-1j * w * 12 * dt)
    assert w * dt > 5e-4  # a second-order scheme would miss by ~(w dt)^3 per step
    assert relerr(o.field("PSIX"), exact) < 1e-11
    # the helium rotates with (pot4 - mu4) (Dynamics_quantum_impurities/steprk.f90:68)
    he = np.sqrt(den) * np.exp(-1j * (pot4 - mu4) * 12 * dt)
    assert relerr(o.field("PSI"), he) < 1e-11


def test_coupled_evolution_conserves_norms():
    n, xmax = 16, 3.2
    v = Q.lennard_jones(10.22, 2.556)
    uk = Q.uimp_k_table(n, n, n, xmax, xmax, xmax, v, 2.0, float(v(2.0)))
    o = _quantum_oracle(n, xmax, uimp_k=uk, psix=I.gaussian_packet(n, n, n, xmax, xmax, xmax, width=1.0,
                                                                   k=(0.5, 0.0, 0.0)))
    o.tddft_init()
    dv = (2 * xmax / n) ** 3
    nx0, nh0 = (np.abs(o.field("PSIX")) ** 2).sum() * dv, o.field("DEN").sum() * dv
    dt = 2e-4 / I.PSTOK
    for it in range(1, 13):
        o.tddft_iteration(dt, it, 3)
    nx1, nh1 = (np.abs(o.field("PSIX")) ** 2).sum() * dv, o.field("DEN").sum() * dv
    assert abs(nx1 - nx0) < 1e-7 * nx0
    assert abs(nh1 - nh0) < 1e-7 * nh0
    assert np.abs(o.field("PSIX") - I.gaussian_packet(n, n, n, xmax, xmax, xmax, width=1.0,
                                                      k=(0.5, 0.0, 0.0))).max() > 1e-4  # it moved


def test_energy_terms_closed_forms():
    n, xmax, m_imp = 16, 3.2, 3.0
    psi0 = _plane_wave(n, xmax, 3) * 0.7
    h = 2 * xmax / n
    k1 = (O.derivnd(1, psi0, h, 1, 8) / psi0).mean()  # i * first-derivative symbol
    c = 11.0
    o = _quantum_oracle(n, xmax, uimp_k=c, m=m_imp, psix=psi0, icon=8)
    o.tddft_init()
    e = o.energy()
    dv = h ** 3
    nrm = (np.abs(psi0) ** 2).sum() * dv
    assert abs(e["ekinx"] - H2O2M4 * 4.002603 / m_imp * abs(k1) ** 2 * nrm) < 1e-11 * e["ekinx"]
    assert abs(e["eHeX"] - c * (np.abs(psi0) ** 2 * o.field("DEN")).sum() * dv) < 1e-11 * abs(e["eHeX"])
    assert abs(e["eimpu"] - (e["ekinx"] + e["eHeX"])) < 1e-12 * abs(e["eimpu"])
    assert abs(e["etot"] - (e["etot4"] + e["eimpu"])) < 1e-12 * abs(e["etot"])


def test_centroid_of_symmetric_packet():
    n, xmax = 32, 6.4
    h = 2 * xmax / n
    r0 = (-6.4 + 14 * h, -6.4 + 17 * h, -6.4 + 15 * h)  # grid points
    o = _quantum_oracle(n, xmax, uimp_k=0.0, psix=I.gaussian_packet(n, n, n, xmax, xmax, xmax, r0=r0, width=1.2))
    nrm, rc = o.quantum_centroid()
    assert abs(nrm - 1.0) < 1e-12
    assert np.abs(rc - np.array(r0)).max() < 1e-10


# ---- several quantum impurities (potenimp.f90:13-39, steprk.f90:84-106,
# steppc.f90:64-139, energy.f90:102-132) ------------------------------------------
def _two_oracle(n, xmax, uks, masses, psixs, uimp_kk=None, den=None, **kw):
    o = O.Oracle(n, n, n, xmax, xmax, xmax, O.default_params(core_otc=0, **kw))
    o.set("DEN", den if den is not None else I.woods_saxon(n, n, n, xmax, xmax, xmax, 20.0, seed=1))
    o.set_quantum_impurities(uks, masses, uimp_kk=uimp_kk)
    for k, p in enumerate(psixs):
        o.select_quantum(k)
        o.set("PSIX", p)
    return o


def _fields(o, k, *names):
    o.select_quantum(k)
    return [o.field(n) for n in names]


def test_two_impurities_potenimp_constant_kernels():
    """Constant kernels turn every convolution into a product: potx4 = c0|psix0|^2
    + c1|psix1|^2, upotx_i = c_i den + a_ij |psix_j|^2 with row i the impurity that
    feels uimp_k_k(i,j) (a swapped pair index gives the other constant)."""
    n, xmax = 16, 3.2
    p0 = I.gaussian_packet(n, n, n, xmax, xmax, xmax, r0=(0.5, 0.0, 0.0), width=1.0)
    p1 = I.gaussian_packet(n, n, n, xmax, xmax, xmax, r0=(-0.8, 0.3, 0.0), width=1.3, k=(0.0, 0.4, 0.0))
    c0, c1, a01, a10 = 37.5, -12.0, 5.25, 80.0
    o = _two_oracle(n, xmax, [c0, c1], [4.002603, 20.0], [p0, p1], uimp_kk={(0, 1): a01, (1, 0): a10})
    o.potenimp()
    d0, d1, den = np.abs(p0) ** 2, np.abs(p1) ** 2, o.field("DEN")
    assert relerr(o.field("POTX4"), c0 * d0 + c1 * d1) < 1e-13
    assert relerr(_fields(o, 0, "UPOTX")[0], c0 * den + a01 * d1) < 1e-13
    assert relerr(_fields(o, 1, "UPOTX")[0], c1 * den + a10 * d0) < 1e-13
    # a 'null' pair (no table) couples nothing
    o2 = _two_oracle(n, xmax, [c0, c1], [4.002603, 20.0], [p0, p1], uimp_kk={(0, 1): a01})
    o2.potenimp()
    assert relerr(_fields(o2, 1, "UPOTX")[0], c1 * den) < 1e-13


def test_pair_term_brute_force_periodic_convolution():
    n, xmax = 8, 2.0
    rng = np.random.default_rng(7)
    k = rng.standard_normal((n, n, n))
    k = 0.5 * (k + np.roll(k[::-1, ::-1, ::-1], 1, axis=(0, 1, 2)))
    uk = np.fft.rfftn(k).real
    p0 = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
    p1 = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
    o = _two_oracle(n, xmax, [0.0, 0.0], [4.0, 4.0], [p0, p1], uimp_kk={(0, 1): uk}, den=rng.random((n, n, n)))
    o.potenimp()
    d1 = np.abs(p1) ** 2
    brute = np.zeros_like(d1)
    for s in np.ndindex(n, n, n):
        brute += d1[s] * np.roll(k, s, axis=(0, 1, 2))
    assert relerr(_fields(o, 0, "UPOTX")[0], brute) < 1e-12


def test_two_free_plane_waves_rotate_independently():
    """No coupling: each impurity rotates by exp(-i h2o2mimp_k kappa_k t) with its
    own mass and wave vector (one slot rotation shared by both histories)."""
    n, xmax = 16, 3.2
    h = 2 * xmax / n
    ms, waves = (2 * 4.002603, 7.0), (_plane_wave(n, xmax, 4), _plane_wave(n, xmax, 2) * 0.5)
    kap = [-(O.derivnd(2, w, h, 1, 8) / w).real.mean() for w in waves]
    o = _two_oracle(n, xmax, [0.0, 0.0], ms, waves, den=np.full((n, n, n), 0.02), icon=8)
    o.tddft_init()
    dt = 5e-4 / I.PSTOK
    for it in range(1, 13):
        o.tddft_iteration(dt, it, 3)
    for k in range(2):
        w = H2O2M4 * 4.002603 / ms[k] * kap[k]
        assert relerr(_fields(o, k, "PSIX")[0], waves[k] * np.exp(-1j * w * 12 * dt)) < 1e-11


def test_two_impurity_energies_closed_forms():
    n, xmax = 16, 3.2
    h = 2 * xmax / n
    w0, w1 = _plane_wave(n, xmax, 3) * 0.7, _plane_wave(n, xmax, 1) * 1.1
    ms, c0, c1, a01, a10 = (3.0, 9.0), 11.0, -4.0, 2.5, 6.0
    o = _two_oracle(n, xmax, [c0, c1], ms, [w0, w1], uimp_kk={(0, 1): a01, (1, 0): a10}, icon=8)
    o.tddft_init()
    e = o.energy()
    dv = h ** 3
    den = o.field("DEN")
    ekx = 0.0
    for w, m in ((w0, ms[0]), (w1, ms[1])):
        k1 = (O.derivnd(1, w, h, 1, 8) / w).mean()
        ekx += H2O2M4 * 4.002603 / m * abs(k1) ** 2 * (np.abs(w) ** 2).sum()
    assert abs(e["ekinx"] - ekx * dv) < 1e-11 * e["ekinx"]
    ehx = ((c0 * np.abs(w0) ** 2 + c1 * np.abs(w1) ** 2) * den).sum() * dv
    assert abs(e["eHeX"] - ehx) < 1e-11 * abs(ehx)
    # pairs k < m with uimp_k_k(k,m) only (energy.f90:122-129)
    eii = a01 * (np.abs(w0) ** 2 * np.abs(w1) ** 2).sum() * dv
    assert abs(e["eimpu_impu"] - eii) < 1e-12 * abs(eii)
    assert abs(e["eimpu"] - (e["ekinx"] + e["eHeX"] + e["eimpu_impu"])) < 1e-12 * abs(e["eimpu"])


def test_swapping_identical_impurities_swaps_the_trajectory():
    """Two impurities of one species (same mass, X-He kernel and a symmetric X-X
    kernel): starting from (A, B) and from (B, A) gives the same helium and the
    impurity wave functions swapped."""
    n, xmax = 16, 3.2
    v = Q.lennard_jones(10.22, 2.556)
    uk = Q.uimp_k_table(n, n, n, xmax, xmax, xmax, v, 2.0, float(v(2.0)))
    vxx = Q.lennard_jones(20.0, 3.0)
    ukk = Q.uimp_k_table(n, n, n, xmax, xmax, xmax, vxx, 2.5, float(vxx(2.5)))
    A = I.gaussian_packet(n, n, n, xmax, xmax, xmax, r0=(1.0, 0.0, 0.5), width=1.0, k=(0.5, 0.0, 0.0))
    B = I.gaussian_packet(n, n, n, xmax, xmax, xmax, r0=(-1.2, 0.4, 0.0), width=1.1, k=(0.0, -0.3, 0.2))
    runs = []
    for pair in ((A, B), (B, A)):
        o = _two_oracle(n, xmax, [uk, uk], [4.0, 4.0], pair, uimp_kk={(0, 1): ukk, (1, 0): ukk})
        o.tddft_init()
        dt = 2e-4 / I.PSTOK
        for it in range(1, 9):
            o.tddft_iteration(dt, it, 3)
        runs.append(o)
    a, b = runs
    assert relerr(a.field("PSI"), b.field("PSI")) < 1e-12
    assert relerr(_fields(a, 0, "PSIX")[0], _fields(b, 1, "PSIX")[0]) < 1e-12
    assert relerr(_fields(a, 1, "PSIX")[0], _fields(b, 0, "PSIX")[0]) < 1e-12
    assert np.abs(_fields(a, 0, "PSIX")[0] - A).max() > 1e-4  # it moved
    dv = (2 * xmax / n) ** 3
    for k, p in enumerate((A, B)):  # each impurity keeps its norm
        assert abs((np.abs(_fields(a, k, "PSIX")[0]) ** 2).sum() * dv - (np.abs(p) ** 2).sum() * dv) < 1e-7
