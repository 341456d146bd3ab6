"""The 2.5-D stepper kernel (x, y and z second derivatives of psi and the
RK-Gill / Hamming update in one pass; hedft_tuning.lap25 > 0) against the
separate x+y Laplacian pass followed by the z pass (lap25 = 0): both sum the
same stencil terms in the same order, so every field must agree bit for bit
(the corrector's sum |pc|, errHe, is summed over the other kernel's blocks
and agrees to rounding).
Covered: the three boundary modes (one-sided rows icon 13, periodic icon 8,
antiperiodic icon 18), both tile shapes (32 x 8 on 32-wide planes, 16 x 16 on
48-wide ones), chunk lengths that do and do not divide nz, chunks shorter
than the 6-plane boundary band, classical impurities, the absorbing layer, the
quantum impurity's psix, mixed precision, and graph replay.  The oracle
comparison of the default (lap25) path is in the trajectory tests of
test_gpu_parity.py, test_gpu_headline.py and test_gpu_quantum.py."""
import numpy as np
import pytest

from he_dft_2024_amd import inputs as I
from he_dft_2024_amd import quantum as Q

pytestmark = pytest.mark.gpu

DT = 5e-4 / I.PSTOK
SLOTS = ("PSI", "DEN", "PC", "PSIOLD1", "PSIOLD2", "PSIOLD3", "HPSIOLD1", "HPSIOLD2", "POT4")


def _say(*a):
    print("[case]", *a, flush=True)


def _droplet(gpu, n, lap25, icon=13, absorb=False, precision=0):
    xmax = 0.2 * n
    kw = dict(core_otc=1, icon=icon, precision=precision)
    if absorb:
        kw.update(lambdah=2.0, tmean=(xmax - 2.0,) * 3, tsurf=(2.0,) * 3)
    d = gpu.Droplet(n, n, n, xmax, xmax, xmax, gpu.default_params(**kw), mode=gpu.MODE_TDDFT)
    t = d.get_tuning()
    t["lap25"] = lap25
    d.set_tuning(t)
    assert d.get_tuning()["lap25"] == lap25
    I.tddft_setup(d, n, xmax, 0.45 * n ** 3 * 0.0022, seed=5, r0=(0.3, -0.2, -2.0), v_aps=(0.0, 1.0, 3.0))
    return d


def _run(d, nrk=3, npc=5, graph_steps=0):
    d.tddft_init()
    errs = [d.tddft_iteration(DT, it, nrk) for it in range(1, nrk + npc + 1)]
    if graph_steps:
        d.run_pc(DT, graph_steps, True)
    return np.array(errs)


def _same(a, b):
    for name in SLOTS:
        assert np.array_equal(a.field(name), b.field(name)), name
    ea, eb = a.energy(), b.energy()
    for key in ("etot", "ekin4", "elj4", "ecor4", "norm"):
        assert ea[key] == eb[key], key


@pytest.mark.parametrize("n,icon,tz,absorb", [(32, 13, 16, False), (32, 13, 5, True), (32, 8, 16, False),
                                              (32, 18, 7, True), (48, 13, 16, True), (48, 8, 4, False),
                                              (32, 13, 32, False)])
def test_lap25_equals_separate_pass(gpu, n, icon, tz, absorb):
    _say("lap25 vs separate pass", n, "icon", icon, "tz", tz, "absorber", absorb)
    a = _droplet(gpu, n, tz, icon, absorb)
    b = _droplet(gpu, n, 0, icon, absorb)
    ea, eb = _run(a), _run(b)
    assert np.array_equal(ea[:, 1:], eb[:, 1:])
    assert np.abs(ea[:, 0] - eb[:, 0]).max() <= 1e-13 * np.abs(eb[:, 0]).max()
    _same(a, b)
    ia, ib = a.impurity_state(), b.impurity_state()
    for key in ("rimp", "vimp", "aimp"):
        assert np.array_equal(ia[key], ib[key]), key
    a.close()
    b.close()


def test_lap25_graph_and_mixed_equal_separate_pass(gpu):
    _say("lap25 vs separate pass: graph replay, mixed precision")
    for prec in (0, 1):
        a = _droplet(gpu, 32, 16, precision=prec)
        b = _droplet(gpu, 32, 0, precision=prec)
        _run(a, graph_steps=13)
        _run(b, graph_steps=13)
        _same(a, b)
        a.close()
        b.close()


def test_lap25_quantum_impurity_equals_separate_pass(gpu):
    _say("lap25 vs separate pass: quantum impurity")
    n, xmax = 32, 6.4
    v = Q.lennard_jones(10.22, 2.556)
    uk = Q.uimp_k_table(n, n, n, xmax, xmax, xmax, v, 2.0, float(v(2.0)))
    ds = []
    for tz in (16, 0):
        d = gpu.Droplet(n, n, n, xmax, xmax, xmax, gpu.default_params(core_otc=1), mode=gpu.MODE_TDDFT)
        t = d.get_tuning()
        t["lap25"] = tz
        d.set_tuning(t)
        I.quantum_setup(d, n, xmax, 60.0, uk, seed=3, m_imp_u=2 * 4.002603, mu4=-1.5, r0=(0.4, -0.8, 1.2),
                        width=1.3, k=(0.6, 0.0, -0.3))
        _run(d, graph_steps=6)
        ds.append(d)
    for name in ("PSI", "PSIX", "PSIXOLD1", "HPSIXOLD2", "PCX", "POTX4", "UPOTX"):
        assert np.array_equal(ds[0].field(name), ds[1].field(name)), name
    for d in ds:
        d.close()


def test_lap25_tuning_checks(gpu):
    """Every supported grid edge is a multiple of 16, so TDDFT contexts take
    the 2.5-D kernel by default; statics contexts and negative values are
    rejected."""
    _say("lap25 tuning checks")
    d = gpu.Droplet(48, 32, 64, 9.6, 6.4, 12.8, gpu.default_params(core_otc=1), mode=gpu.MODE_TDDFT)
    assert d.get_tuning()["lap25"] == 16
    t = d.get_tuning()
    t["lap25"] = -1
    with pytest.raises(gpu.HedftError):
        d.set_tuning(t)
    d.close()
    s = gpu.Droplet(32, 32, 32, 6.4, 6.4, 6.4, mode=gpu.MODE_STATICS)
    assert s.get_tuning()["lap25"] == 0
    t = s.get_tuning()
    t["lap25"] = 16
    with pytest.raises(gpu.HedftError):
        s.set_tuning(t)
    s.close()
