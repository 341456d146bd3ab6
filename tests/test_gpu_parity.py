"""GPU parity: the HIP library (through the C-ABI) against the FP64 oracle on the
same seeded inputs.  Tolerances are relative to the largest magnitude of the
compared field; FP64 throughout (the drop-in path)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import relerr
from he_dft_2024_amd import inputs as I

pytestmark = pytest.mark.gpu

RTOL_OP = 1e-12      # single operator (FFT, stencil, poten) vs oracle
RTOL_TRAJ = 1e-9     # a few integrator steps


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("shape", [(32, 32, 32), (48, 64, 32), (96, 48, 64), (128, 128, 32), (192, 192, 32),
                                   (64, 32, 256), (64, 64, 128), (128, 128, 128)])
def test_fft_roundtrip_vs_numpy(gpu, shape):
    nx, ny, nz = shape
    _say("fft", shape)
    d = gpu.Droplet(nx, ny, nz, 0.2 * nx, 0.2 * ny, 0.2 * nz, mode=gpu.MODE_STATICS)
    a = np.random.default_rng(1).standard_normal((nz, ny, nx))
    d.set("WORK1", a)
    d.fftfw("WORK1", "WORKK1")
    s = d.field("WORKK1")
    assert relerr(s, np.fft.rfftn(a)) < 1e-14
    d.fftbk("WORKK1", "WORK2")
    assert relerr(d.field("WORK2"), a) < 1e-14
    d.close()


@pytest.mark.parametrize("icon", [13, 8, 18, 0, 24])
def test_derivnd_vs_oracle(gpu, icon):
    nx, ny, nz = 32, 48, 64
    _say("derivnD icon", icon)
    p = gpu.default_params(icon=icon)
    d = gpu.Droplet(nx, ny, nz, 6.4, 9.6, 12.8, p, mode=gpu.MODE_TDDFT)
    rng = np.random.default_rng(icon)
    f = rng.standard_normal((nz, ny, nx))
    fc = f + 1j * rng.standard_normal((nz, ny, nx))
    h = (0.4, 0.4, 0.4)
    for id_ in (1, 2):
        for iv in (1, 2, 3):
            d.set("WORK1", f)
            d.derivnd(id_, iv, "WORK1", "WORK2")
            ref = O.derivnd(id_, f, h[iv - 1], iv, icon)
            assert relerr(d.field("WORK2"), ref) < RTOL_OP, (id_, iv)
            d.set("PSI", fc)
            d.derivnd(id_, iv, "PSI", "WORKC1")
            refc = O.derivnd(id_, fc, h[iv - 1], iv, icon)
            assert relerr(d.field("WORKC1"), refc) < RTOL_OP, (id_, iv, "complex")
    d.close()


def test_kernel_tables_vs_oracle(gpu):
    _say("kernels 48x32x64")
    p = gpu.default_params(core_otc=1)
    d = gpu.Droplet(48, 32, 64, 9.6, 6.4, 12.8, p)
    o = O.Oracle(48, 32, 64, 9.6, 6.4, 12.8, O.default_params(core_otc=1))
    # fvlj4 is a difference of nearly equal terms at large |p| (fforma.f90:96-100);
    # both sides evaluate it in long double and agree far below the reference's
    # own kernel error (its s13adf Si is off by up to 6e-14 for x > 16)
    assert relerr(d.field("FVLJ4"), o.field("FVLJ4")) < 1e-10
    for name in ("WCGK", "KALFS"):
        assert relerr(d.field(name), o.field(name)) < 1e-14, name
    d.close()


def _pair(n, xmax, mode, **kw):
    p = gpu_params = None
    import he_dft_2024_amd as H
    gp = H.default_params(**kw)
    op = O.default_params(**kw)
    d = H.Droplet(n, n, n, xmax, xmax, xmax, gp, mode=mode)
    o = O.Oracle(n, n, n, xmax, xmax, xmax, op, mode=mode)
    return d, o


MF_FIELDS = ["DENCG", "DELJ4", "DXDEN", "DYDEN", "DZDEN", "POT4"]
ALPHA_FIELDS = ["DENALF", "FALFS", "INTXALF", "INTYALF", "INTZALF", "UALPHAS"]


@pytest.mark.parametrize("otc,lsolid", [(1, 0), (0, 0), (0, 1)])
def test_poten_tddft_vs_oracle(gpu, otc, lsolid):
    n, xmax, n4 = 32, 6.4, 60.0
    _say("poten tddft otc", otc, "lsolid", lsolid)
    d, o = _pair(n, xmax, gpu.MODE_TDDFT, core_otc=otc, lsolid=lsolid)
    den = I.woods_saxon(n, n, n, xmax, xmax, xmax, n4, seed=5)
    for x in (d, o):
        x.set("DEN", den)
        x.set_impurities(np.zeros((1, 10)), 1.0, 100.0, [1.0])
        x.set_impurity_state(np.zeros((1, 3)), np.zeros((1, 3)), np.zeros((1, 3)))
        x.poten()
    names = MF_FIELDS + (ALPHA_FIELDS if otc else [])
    for name in names:
        assert relerr(d.field(name), o.field(name)) < RTOL_OP, name
    d.close()


def test_poten_statics_stale_state_vs_oracle(gpu):
    """Statics poten uses the caller's fden/dencg/denalf/grad(den) and a fresh den."""
    n, xmax, n4 = 32, 6.4, 60.0
    _say("poten statics stale")
    d, o = _pair(n, xmax, gpu.MODE_STATICS, core_otc=1)
    den0 = I.woods_saxon(n, n, n, xmax, xmax, xmax, n4, seed=6)
    den1 = I.woods_saxon(n, n, n, xmax, xmax, xmax, n4, seed=7) * 1.01
    for x in (d, o):
        x.set("DEN", den0)
        x.statics_init(n4)       # fden, dencg, denalf, grad from den0
        x.set("DEN", den1)       # fresh den only
        x.poten()
    for name in MF_FIELDS + ALPHA_FIELDS:
        assert relerr(d.field(name), o.field(name)) < RTOL_OP, name
    d.close()


@pytest.mark.parametrize("log_psi", [False, True])
def test_statics_steppers_read_the_drivers_fden(gpu, log_psi):
    """Direct STEPRKR/STEPPCR calls with the Lennard-Jones term alone: their inner
    poten rebuilds delj4 from the driver's fden (He_statics/poten.f90:27-30,
    refreshed only by the driver, BCN4HeDFT.f90:1128-1131), so pot4 = delj4 + uext
    stays one field and the run equals the run with eps4 = 0 and uext + delj4 as
    external potential, bit for bit, and the oracle's run."""
    n, xm, n4 = 32, 6.4, 60.0
    _say("statics stale fden: LJ alone vs functional off + uext, log psi", log_psi)
    h = 2 * xm / n
    dt = 0.02 * h * h / (4 * 6.05969638298)
    x = I.grid_axes(n, xm)
    uext = 0.4 * (x[None, None, :] ** 2 + x[None, :, None] ** 2 + x[:, None, None] ** 2) - 3.0
    den = I.woods_saxon(n, n, n, xm, xm, xm, n4, seed=4, noise=0.0 if log_psi else 1e-3)
    names = ("PSI", "DEN", "POT4", "PSIOLD1", "PSIOLD3", "HPSIOLD2") + (("PSIL",) if log_psi else ("TIMEC",))

    def run(mod, eps4_off, ue):
        p = mod.default_params(cp4=0.0, cpp4=0.0, core_otc=0, lsolid=0, lexternal=1,
                               **({"eps4": 0.0} if eps4_off else {}))
        s = (gpu.Droplet if mod is gpu else O.Oracle)(n, n, n, xm, xm, xm, p, mode=mod.MODE_STATICS)
        s.set("UEXT", ue)
        s.set("DEN", den)
        mu = s.statics_init(n4)
        dlj = np.array(s.field("DELJ4"))
        rk, pc = (s.steprkrl, s.steppcrl) if log_psi else (s.steprkr, s.steppcr)
        if log_psi:
            s.statics_log_start()
        log = []
        for it in range(1, 11):
            r = rk(dt, mu, n4) if it <= 4 else pc(dt, mu, n4)
            mu = r[0]
            log.append(r)
        out = [np.array(s.field(k)) for k in names]
        s.close()
        return dlj, log, out

    dlj, log_a, a = run(gpu, False, uext)
    _, log_b, b = run(gpu, True, uext + dlj)
    _, log_o, o = run(O, False, uext)
    assert np.abs(dlj).max() > 1.0
    assert log_a == log_b
    for k, u, v, w in zip(names, a, b, o):
        assert np.array_equal(u, v), k
        if k in ("PSI", "DEN", "POT4", "PSIL", "TIMEC"):
            assert relerr(u, w) < RTOL_TRAJ, k
    assert relerr(np.array([r[:2] for r in log_a]), np.array([r[:2] for r in log_o])) < RTOL_TRAJ


def test_statics_alpha_s_field_and_affine_poten(gpu):
    """term_alfa reads only the driver's denalf and grad(den) (He_statics/term_alfa.f90):
    direct stepper calls with the OTC core equal the 'OT ' run with uext + ualphas as
    external potential; and the statics poten, with the driver's arrays held, is affine
    in den (its only read of the current den is the correlation convolution,
    He_statics/poten.f90:42-52), which a refresh from the current den breaks."""
    n, xm, n4 = 32, 6.4, 60.0
    _say("statics stale denalf/grad and affine poten")
    h = 2 * xm / n
    dt = 0.02 * h * h / (4 * 6.05969638298)
    x = I.grid_axes(n, xm)
    uext = 0.4 * (x[None, None, :] ** 2 + x[None, :, None] ** 2 + x[:, None, None] ** 2) - 3.0
    den = I.woods_saxon(n, n, n, xm, xm, xm, n4, seed=4)
    res = []
    for otc in (1, 0):
        d = gpu.Droplet(n, n, n, xm, xm, xm, gpu.default_params(core_otc=otc, lexternal=1), mode=gpu.MODE_STATICS)
        d.set("UEXT", uext if otc else uext + ua)
        d.set("DEN", den)
        mu = d.statics_init(n4)
        if otc:
            ua = np.array(d.field("UALPHAS"))
        for it in range(1, 11):
            mu = (d.steprkr(dt, mu, n4) if it <= 4 else d.steppcr(dt, mu, n4))[0]
        res.append((mu, [np.array(d.field(k)) for k in ("PSI", "POT4", "PSIOLD2", "HPSIOLD1")]))
        d.close()
    assert np.abs(ua).max() > 0.1
    assert abs(res[0][0] - res[1][0]) < 1e-13 * abs(res[0][0])
    for u, v in zip(res[0][1], res[1][1]):
        assert np.abs(u - v).max() < 1e-12 * np.abs(v).max()
    d = gpu.Droplet(n, n, n, xm, xm, xm, gpu.default_params(core_otc=1, lexternal=1), mode=gpu.MODE_STATICS)
    d.set("UEXT", uext)
    d.set("DEN", I.woods_saxon(n, n, n, xm, xm, xm, n4, seed=6))
    d.statics_init(n4)
    d1 = I.woods_saxon(n, n, n, xm, xm, xm, n4, seed=7)
    d2 = 0.6 * I.woods_saxon(n, n, n, xm, xm, xm, 0.7 * n4, seed=8)
    pots = []
    for dd in (d1, d2, 0.5 * (d1 + d2)):
        d.set("DEN", dd)
        d.poten()
        pots.append(np.array(d.field("POT4")))
    d.close()
    assert np.abs(pots[2] - 0.5 * (pots[0] + pots[1])).max() < 1e-13 * np.abs(pots[2]).max()


def test_statics_trajectory_vs_oracle(gpu):
    n, xmax, n4 = 32, 6.4, 50.0
    _say("statics 4 RK + 8 PC")
    d, o = _pair(n, xmax, gpu.MODE_STATICS, core_otc=1)
    den = I.woods_saxon(n, n, n, xmax, xmax, xmax, n4, seed=0)
    dt = 0.01 * 0.4 ** 2 / (4 * 6.05969638298)
    mus = []
    for x in (d, o):
        x.set("DEN", den)
        mu = x.statics_init(n4)
        hist = [mu]
        for it in range(1, 13):
            r = x.statics_iteration(dt, it, n4, mu)
            mu = r[0]
            hist.append(mu)
        mus.append(hist)
    assert relerr(mus[0], mus[1]) < RTOL_TRAJ
    for name in ("PSI", "DEN", "POT4", "PC", "PSIOLD1", "PSIOLD2", "PSIOLD3", "HPSIOLD1", "HPSIOLD2",
                 "TIMEC"):
        assert relerr(d.field(name), o.field(name)) < RTOL_TRAJ, name
    eg, eo = d.energy(), o.energy()
    for k in ("ekin4", "elj4", "ecor4", "ealphas", "etot", "norm"):
        assert abs(eg[k] - eo[k]) <= 1e-10 * abs(eo[k]) + 1e-12, k
    d.close()


def test_c0_statics_relaxation_vs_oracle(gpu):
    """C0 at its size: the N=300 droplet on 64^3 (h = 0.4 A) from the seeded
    Woods-Saxon start with 1e-3 noise, 200 imaginary-time iterations (4 RK-Gill
    + 196 Hamming PC, dtau = 1e-4 with the local time step), GPU through
    hedft_statics_run (graph replay) against the oracle's per-iteration loop:
    E/N, mu, the norm and the density."""
    import he_dft_2024_amd as H
    n, xm, n4, niter = 64, 12.8, 300.0, 200
    _say("C0 statics 64^3 N=300, 200 iterations vs oracle")
    den = I.woods_saxon(n, n, n, xm, xm, xm, n4, seed=0)
    d = H.Droplet(n, n, n, xm, xm, xm, H.default_params(core_otc=1), mode=H.MODE_STATICS)
    o = O.Oracle(n, n, n, xm, xm, xm, O.default_params(core_otc=1), mode=O.MODE_STATICS)
    d.set("DEN", den)
    o.set("DEN", den)
    d.statics_init(n4)
    mu_o = o.statics_init(n4)
    mu_g = d.statics_run(1e-4, 1, niter, n4)[0]
    for it in range(1, niter + 1):
        mu_o = o.statics_iteration(1e-4, it, n4, mu_o)[0]
    assert abs(mu_g - mu_o) <= 1e-9 * abs(mu_o), (mu_g, mu_o)
    eg, eo = d.energy(), o.energy()
    for k in ("ekin4", "elj4", "ecor4", "ealphas", "etot", "norm"):
        assert abs(eg[k] - eo[k]) <= 1e-9 * abs(eo[k]) + 1e-12, (k, eg[k], eo[k])
    dg, do = d.field("DEN"), o.field("DEN")
    assert np.linalg.norm(dg - do) / np.linalg.norm(do) < 1e-9
    d.close()


def test_statics_run_equals_iterations(gpu):
    """hedft_statics_run (iterations queued without host syncs, PC iterations
    replayed from a captured graph) leaves every field, mu4 and its error equal
    bit for bit to hedft_statics_iteration called once per iteration, across
    runs that start in the RK start and at other slot phases."""
    import he_dft_2024_amd as H
    n, xmax, n4 = 32, 6.4, 50.0
    _say("statics run vs iterations")
    den = I.woods_saxon(n, n, n, xmax, xmax, xmax, n4, seed=0)
    dt = 0.01 * 0.4 ** 2 / (4 * 6.05969638298)
    names = ("PSI", "DEN", "POT4", "PC", "PSIOLD1", "PSIOLD2", "PSIOLD3", "HPSIOLD1", "HPSIOLD2", "DENCG",
             "DENALF", "TIMEC")
    out = []
    for mode in ("iterations", "graph", "eager"):
        d = H.Droplet(n, n, n, xmax, xmax, xmax, H.default_params(core_otc=1), mode=H.MODE_STATICS)
        d.set("DEN", den)
        mu = d.statics_init(n4)
        snaps = []
        it = 1
        for k in (3, 17, 1, 13, 8):
            if mode == "iterations":
                for j in range(it, it + k):
                    prev = mu
                    mu, err = d.statics_iteration(dt, j, n4, mu)
                snaps.append([mu, err])
            else:
                m, e, pc, ms = d.statics_run(dt, it, k, n4, use_graph=(mode == "graph"))
                assert ms > 0
                snaps.append([m, e])
            snaps[-1] += [d.field(x) for x in names]
            it += k
        out.append(snaps)
        d.close()
    for v in (1, 2):
        for i, (a, b) in enumerate(zip(out[v], out[0])):
            assert a[0] == b[0] and a[1] == b[1], (v, i)
            for name, x, y in zip(names, a[2:], b[2:]):
                assert np.array_equal(x, y), (v, i, name)


# the absorbing layer of the reference's TDDFT example (Lambdah = 2, t*mean =
# box half-width - 2 A, t*surf = 2 A; Dynamics_TDDFT/input_files/
# xe-he1000-hoc-200mps.input:2-4), here on all three axes
def _absorber(xmax, ymax=None, zmax=None):
    ymax = xmax if ymax is None else ymax
    zmax = xmax if zmax is None else zmax
    return {"lambdah": 2.0, "tmean": (xmax - 2.0, ymax - 2.0, zmax - 2.0), "tsurf": (2.0, 2.0, 2.0)}


@pytest.mark.parametrize("otc,absorb", [(1, False), (0, False), (1, True), (0, True)])
def test_tddft_trajectory_vs_oracle(gpu, otc, absorb):
    n, xmax, n4 = 32, 6.4, 60.0
    _say("tddft 3 RK + 9 PC otc", otc, "absorber", absorb)
    d, o = _pair(n, xmax, gpu.MODE_TDDFT, core_otc=otc, **(_absorber(xmax) if absorb else {}))
    dt = 5e-4 / I.PSTOK
    errs = []
    for x in (d, o):
        I.tddft_setup(x, n, xmax, n4, seed=2, r0=(0.0, 0.5, -4.0), v_aps=(0.0, 0.0, 3.0))
        x.tddft_init()
        e = []
        for it in range(1, 13):
            e.append(x.tddft_iteration(dt, it, 3))
        errs.append(np.array(e))
    assert relerr(errs[0][3:, 0], errs[1][3:, 0]) < RTOL_TRAJ
    for name in ("PSI", "DEN", "POT4", "PC", "PSIOLD1", "PSIOLD2", "PSIOLD3", "HPSIOLD1", "HPSIOLD2",
                 "UIMP"):
        assert relerr(d.field(name), o.field(name)) < RTOL_TRAJ, name
    tc = d.field("TIMEC")
    assert np.array_equal(tc, o.field("TIMEC"))
    if absorb:   # the layer is active and removes helium at the walls
        assert tc.real.max() > 1.0
        assert d.energy()["norm"] < n4 * (1 - 1e-9)
    else:
        assert tc.real.max() == 0.0
    # RK-Gill's q returns to zero after every full 4-stage step in exact arithmetic
    # (stage 4: a=1/6, b=2, c=1/2), so what remains is round-off on both sides
    kscale = np.abs(o.field("HPSIOLD1")).max()
    assert np.abs(d.field("Q")).max() < 1e-12 * kscale
    assert np.abs(o.field("Q")).max() < 1e-12 * kscale
    sg, so = d.impurity_state(), o.impurity_state()
    for k in ("rimp", "vimp", "aimp", "force"):
        assert relerr(sg[k], so[k]) < RTOL_TRAJ, k
    eg, eo = d.energy(), o.energy()
    for k in ("ekin4", "elj4", "ecor4", "ealphas", "eHeX", "ekinx", "etot", "norm"):
        assert abs(eg[k] - eo[k]) <= 1e-10 * abs(eo[k]) + 1e-12, k
    d.close()


# 128 x 128 planes (FP64) take their own plane-FFT kernels; a thin slab in z
# keeps the oracle's share of these tests to seconds
SLAB = (128, 128, 32, 25.6, 25.6, 6.4)


def _slab_pair(mode, absorb=False, **kw):
    import he_dft_2024_amd as H
    nx, ny, nz, xm, ym, zm = SLAB
    if absorb:
        kw.update(_absorber(xm, ym, zm))
    d = H.Droplet(nx, ny, nz, xm, ym, zm, H.default_params(**kw), mode=mode)
    o = O.Oracle(nx, ny, nz, xm, ym, zm, O.default_params(**kw), mode=mode)
    return d, o


def _slab_setup(x, seed):
    nx, ny, nz, xm, ym, zm = SLAB
    x.set("DEN", I.woods_saxon(nx, ny, nz, xm, ym, zm, 40.0, seed=seed, center=(1.3, -0.7, 0.0)))
    potion, delinter, rmax = I.impurity_table(nx, xm)
    x.set_impurities(potion[None, :], delinter, rmax, [108.0 * I.MP_U])
    x.set_impurity_state(np.array([[-2.0, 1.0, 0.2]]), np.array([[3.0, 0.0, 0.0]]) * I.PSTOK, np.zeros((1, 3)))


@pytest.mark.parametrize("otc", [1, 0])
def test_poten_tddft_128_planes_vs_oracle(gpu, otc):
    _say("poten tddft 128x128x32 otc", otc)
    d, o = _slab_pair(gpu.MODE_TDDFT, core_otc=otc)
    for x in (d, o):
        _slab_setup(x, seed=11)
        x.poten()
    for name in MF_FIELDS + (ALPHA_FIELDS if otc else []):
        assert relerr(d.field(name), o.field(name)) < RTOL_OP, name
    d.close()


@pytest.mark.parametrize("absorb", [False, True])
def test_tddft_trajectory_128_planes_vs_oracle(gpu, absorb):
    _say("tddft 128x128x32 3 RK + 3 PC otc absorber", absorb)
    d, o = _slab_pair(gpu.MODE_TDDFT, absorb=absorb, core_otc=1)
    dt = 5e-4 / I.PSTOK
    errs = []
    for x in (d, o):
        _slab_setup(x, seed=12)
        x.tddft_init()
        errs.append(np.array([x.tddft_iteration(dt, it, 3) for it in range(1, 7)]))
    assert relerr(errs[0][3:, 0], errs[1][3:, 0]) < RTOL_TRAJ
    for name in ("PSI", "DEN", "POT4", "PSIOLD1", "HPSIOLD1", "UIMP"):
        assert relerr(d.field(name), o.field(name)) < RTOL_TRAJ, name
    sg, so = d.impurity_state(), o.impurity_state()
    for k in ("rimp", "vimp", "force"):
        assert relerr(sg[k], so[k]) < RTOL_TRAJ, k
    eg, eo = d.energy(), o.energy()
    for k in ("ekin4", "elj4", "ecor4", "ealphas", "eHeX", "etot"):
        assert abs(eg[k] - eo[k]) <= 1e-10 * abs(eo[k]) + 1e-12, k
    d.close()


def test_tddft_trajectory_long_z_vs_oracle(gpu):
    """z pencils of length 256 (the 256^3 grids' column tiles) in the fused steppers."""
    _say("tddft 32x32x256 3 RK + 3 PC otc")
    import he_dft_2024_amd as H
    nx, ny, nz, xm, ym, zm = 32, 32, 256, 6.4, 6.4, 51.2
    d = H.Droplet(nx, ny, nz, xm, ym, zm, H.default_params(core_otc=1), mode=gpu.MODE_TDDFT)
    o = O.Oracle(nx, ny, nz, xm, ym, zm, O.default_params(core_otc=1), mode=O.MODE_TDDFT)
    dt = 5e-4 / I.PSTOK
    errs = []
    for x in (d, o):
        x.set("DEN", I.woods_saxon(nx, ny, nz, xm, ym, zm, 30.0, seed=13, center=(0.3, -0.2, 4.0)))
        potion, delinter, rmax = I.impurity_table(nz, zm)
        x.set_impurities(potion[None, :], delinter, rmax, [108.0 * I.MP_U])
        x.set_impurity_state(np.array([[0.0, 0.4, -2.0]]), np.array([[0.0, 0.0, 3.0]]) * I.PSTOK, np.zeros((1, 3)))
        x.tddft_init()
        errs.append(np.array([x.tddft_iteration(dt, it, 3) for it in range(1, 7)]))
    assert relerr(errs[0][3:, 0], errs[1][3:, 0]) < RTOL_TRAJ
    for name in ("PSI", "DEN", "POT4", "PSIOLD1", "HPSIOLD1"):
        assert relerr(d.field(name), o.field(name)) < RTOL_TRAJ, name
    sg, so = d.impurity_state(), o.impurity_state()
    for k in ("rimp", "force"):
        assert relerr(sg[k], so[k]) < RTOL_TRAJ, k
    d.close()


def test_graph_replay_matches_eager(gpu):
    """hedft_tddft_run_pc replays a captured 6-step graph; results equal eager steps."""
    n, xmax, n4 = 32, 6.4, 60.0
    _say("graph vs eager")
    import he_dft_2024_amd as H
    p = H.default_params(core_otc=1)
    dt = 5e-4 / I.PSTOK
    outs = []
    for use_graph in (False, True):
        d = H.Droplet(n, n, n, xmax, xmax, xmax, p)
        I.tddft_setup(d, n, xmax, n4, seed=3, r0=(0.0, 0.0, -4.0))
        d.tddft_init()
        for it in range(1, 4):
            d.tddft_iteration(dt, it, 3)
        d.run_pc(dt, 13, use_graph)
        d.run_pc(dt, 5, use_graph)
        outs.append((d.field("PSI"), d.impurity_state()["rimp"]))
        d.close()
    assert relerr(outs[1][0], outs[0][0]) < 1e-13
    assert relerr(outs[1][1], outs[0][1]) < 1e-13


def test_graph_remainder_replays_match_eager(gpu):
    """Runs of q periods + r < 6 iterations replay the period graph q times and
    a graph of the r remaining iterations captured from the same slot phase
    (no eager steps inside the timed region); every field matches eager steps
    bit for bit, also as the phase moves between runs and a captured
    remainder graph is reused or recaptured, and when the graphs of a run were
    captured ahead of the iterations before it (hedft_tddft_prepare_pc)."""
    n, xmax, n4 = 32, 6.4, 60.0
    _say("graph remainder vs eager")
    import he_dft_2024_amd as H
    p = H.default_params(core_otc=1)
    dt = 5e-4 / I.PSTOK
    names = ("PSI", "DEN", "POT4", "PC", "PSIOLD1", "PSIOLD2", "PSIOLD3", "HPSIOLD1", "HPSIOLD2")
    outs = []
    for use_graph, prepare in ((False, False), (True, False), (True, True)):
        d = H.Droplet(n, n, n, xmax, xmax, xmax, p)
        I.tddft_setup(d, n, xmax, n4, seed=3, r0=(0.0, 0.0, -4.0))
        d.tddft_init()
        for it in range(1, 4):
            d.tddft_iteration(dt, it, 3)
        if prepare:
            d.prepare_pc(dt, 3, 20)
        snaps = []
        for k in (3, 20, 20, 8, 14, 5, 11):
            d.run_pc(dt, k, use_graph)
            snaps.append([d.field(x) for x in names] + [d.impurity_state()["rimp"]])
        outs.append(snaps)
        d.close()
    for v in (1, 2):
        for i, (a, b) in enumerate(zip(outs[v], outs[0])):
            for name, x, y in zip(names + ("rimp",), a, b):
                assert np.array_equal(x, y), (v, i, name)


def test_uimp_on_grid_line_vs_oracle(gpu):
    """An impurity on a grid line puts grid points at integral r/DelInter
    (r = 4 A exactly here), where the reference's int()/mod() interpolation
    (potenimp.f90:101-111) jumps by one table node; both sides take the same
    branch, and one RK + one PC step stay in parity."""
    n, xmax, n4 = 32, 6.4, 60.0
    _say("uimp on grid line")
    d, o = _pair(n, xmax, gpu.MODE_TDDFT, core_otc=1)
    dt = 5e-4 / I.PSTOK
    for x in (d, o):
        I.tddft_setup(x, n, xmax, n4, seed=0, r0=(0.0, 0.0, -4.0))
        x.tddft_init()
    assert relerr(d.field("UIMP"), o.field("UIMP")) < 1e-14
    for x in (d, o):
        x.tddft_iteration(dt, 1, 1)
        x.tddft_iteration(dt, 2, 1)
    for name in ("PSI", "DEN", "POT4", "HPSIOLD1", "HPSIOLD2", "PC"):
        assert relerr(d.field(name), o.field(name)) < RTOL_TRAJ, name
    d.close()


def test_statics_steepest_descent_vs_oracle(gpu):
    """evolo + ironing driver iterations (the statics default when Lrkpc is
    off) against the oracle, whose evolo is pinned to the reference's own."""
    n, xmax, n4 = 32, 6.4, 50.0
    _say("statics steepest descent 10 iterations")
    d, o = _pair(n, xmax, gpu.MODE_STATICS, core_otc=1)
    den = I.woods_saxon(n, n, n, xmax, xmax, xmax, n4, seed=8)
    dt = 0.015 * 0.4 ** 2 / (4 * 6.05969638298)
    hist = []
    for x in (d, o):
        x.set("DEN", den)
        mu = x.statics_init(n4)
        mu, _ = x.evolo(0.0, n4, mu)  # first chemical potential (BCN4HeDFT.f90:1050)
        h = [mu]
        for it in range(10):
            mu, err = x.statics_iteration_sd(dt, n4, mu)
            h.append(mu)
        hist.append(h)
    assert relerr(hist[0], hist[1]) < RTOL_TRAJ
    for name in ("PSI", "DEN", "POT4", "DENCG", "UALPHAS"):
        assert relerr(d.field(name), o.field(name)) < RTOL_TRAJ, name
    eg, eo = d.energy(), o.energy()
    for k in ("ekin4", "elj4", "ecor4", "ealphas", "etot", "norm"):
        assert abs(eg[k] - eo[k]) <= 1e-10 * abs(eo[k]) + 1e-12, k
    d.close()


@pytest.mark.parametrize("icon", [13, 0, 8, 18, 24])
def test_xy_stencil_paths_all_icons_vs_oracle(gpu, icon):
    """The x+y stencil kernels (grad den, alpha_s divergence, Laplacian of psi,
    |grad psi|^2) on a non-cubic grid with every boundary mode: poten, one RK
    step and the energy against the oracle."""
    nx, ny, nz, h, n4 = 64, 96, 32, 0.4, 40.0
    _say("xy stencils icon", icon, (nx, ny, nz))
    import he_dft_2024_amd as H
    xm, ym, zm = 0.5 * nx * h, 0.5 * ny * h, 0.5 * nz * h
    d = H.Droplet(nx, ny, nz, xm, ym, zm, H.default_params(core_otc=1, icon=icon))
    o = O.Oracle(nx, ny, nz, xm, ym, zm, O.default_params(core_otc=1, icon=icon))
    den = I.woods_saxon(nx, ny, nz, xm, ym, zm, n4, seed=9, center=(0.8, -1.2, 0.4))
    dt = 5e-4 / I.PSTOK
    for x in (d, o):
        x.set("DEN", den)
        x.tddft_init()
        x.poten()
    for name in ("DXDEN", "DYDEN", "DZDEN", "UALPHAS", "POT4"):
        assert relerr(d.field(name), o.field(name)) < RTOL_OP, name
    for x in (d, o):
        x.steprk(dt)
    assert relerr(d.field("PSI"), o.field("PSI")) < RTOL_TRAJ
    eg, eo = d.energy(), o.energy()
    for k in ("ekin4", "ealphas", "etot"):
        assert abs(eg[k] - eo[k]) <= 1e-10 * abs(eo[k]), k
    d.close()


def test_tddft_three_impurities_with_pair_terms_vs_oracle(gpu):
    """Three classical impurities with different tables, masses and pair
    potentials: the force pass's run-time impurity loop, the pair force F_ij
    and the pair energy (potenimp.f90:40-58, energy.f90:168-177), eager steps
    and graph replay, and the standalone forceimp."""
    n, xmax, n4 = 32, 6.4, 60.0
    _say("tddft three impurities with pair terms")
    d, o = _pair(n, xmax, gpu.MODE_TDDFT, core_otc=1)
    dt = 5e-4 / I.PSTOK
    potion, delinter, rmax = I.impurity_table(n, xmax)
    potion2, _, _ = I.impurity_table(n, xmax, eps=12.0, sigma=3.5)
    potion3, _, _ = I.impurity_table(n, xmax, eps=30.0, sigma=4.4)
    v, dv, dl = I.pair_tables(3, 4 * xmax, npair=20001,
                              pairs=[{}, {"eps": 12.0, "sigma": 3.5}, {"eps": 30.0, "sigma": 4.4}])
    den = I.woods_saxon(n, n, n, xmax, xmax, xmax, n4, seed=8)
    r0 = np.array([[0.0, 0.5, -4.0], [1.0, -0.5, 4.5], [-3.0, 2.0, 0.5]])
    v0 = np.array([[0.0, 0.0, 3.0], [0.5, 0.0, -2.0], [1.0, -1.0, 0.0]]) * I.PSTOK
    for x in (d, o):
        x.set("DEN", den)
        x.set_impurities(np.stack([potion, potion2, potion3]), delinter, rmax,
                         [108.0 * I.MP_U, 40.0 * I.MP_U, 20.0 * I.MP_U])
        x.set_impurity_pairs(v, dv, dl)
        x.set_impurity_state(r0, v0, np.zeros((3, 3)))
        x.tddft_init()
        for it in range(1, 10):
            x.tddft_iteration(dt, it, 3)
    d.run_pc(dt, 6, True)
    for it in range(10, 16):
        o.tddft_iteration(dt, it, 3)
    assert relerr(d.field("PSI"), o.field("PSI")) < RTOL_TRAJ
    sg, so = d.impurity_state(), o.impurity_state()
    for k in ("rimp", "vimp", "aimp", "force"):
        assert relerr(sg[k], so[k]) < RTOL_TRAJ, k
    eg, eo = d.energy(), o.energy()
    assert abs(eo["eimpu_impu"]) > 1e-3
    for k in ("eimpu_impu", "ekinx", "eHeX", "eimpu", "etot"):
        assert abs(eg[k] - eo[k]) <= 1e-10 * abs(eo[k]) + 1e-12, k
    for x in (d, o):
        x.forceimp()
    assert relerr(d.impurity_state()["force"], o.impurity_state()["force"]) < RTOL_OP
    d.close()


def test_run_entry_points_reject_bad_arguments(gpu):
    """hedft_statics_run and hedft_tddft_prepare_pc: wrong mode, counts out of
    range and PC iterations before the RK start are errors, not launches."""
    import ctypes as C
    import he_dft_2024_amd as H
    _say("statics_run / prepare_pc argument checks")
    L = H.lib()
    d0 = C.c_double(0)
    s = H.Droplet(32, 32, 32, 6.4, 6.4, 6.4, H.default_params(core_otc=1), mode=H.MODE_STATICS)
    s.set("DEN", I.woods_saxon(32, 32, 32, 6.4, 6.4, 6.4, 50.0, seed=0))
    s.statics_init(50.0)
    args = lambda it0, nit, n4: (s._c, 1e-5, it0, nit, n4, 1, C.byref(d0), C.byref(d0), C.byref(d0),  # noqa: E731
                                 C.byref(d0))
    assert L.hedft_statics_run(*args(0, 5, 50.0)) == -1      # iterations count from 1
    assert L.hedft_statics_run(*args(1, 0, 50.0)) == -1      # nothing to run
    assert L.hedft_statics_run(*args(1, 5, 0.0)) == -1       # no atoms
    assert L.hedft_statics_run(*args(5, 5, 50.0)) != 0       # PC before the RK start
    assert L.hedft_tddft_prepare_pc(s._c, 1e-5, 0, 12) == -1  # a statics context
    s.close()
    t = H.Droplet(32, 32, 32, 6.4, 6.4, 6.4, H.default_params(core_otc=1), mode=H.MODE_TDDFT)
    assert L.hedft_statics_run(t._c, 1e-5, 1, 5, 50.0, 1, None, None, None, None) == -1  # a TDDFT context
    assert L.hedft_tddft_prepare_pc(t._c, 1e-5, -1, 12) == -1  # negative skip
    t.close()
