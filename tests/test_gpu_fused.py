"""The single-stream stage C (hedft_tuning.fused = 1) against the two-stream
one (fused = 0, the default).  Fused, the alpha_s divergence d_x(f int_x) +
d_y(f int_y) + d_z(f int_z) and the pot4 assembly are one 2.5-D pass after
stage C's convolution, where the two-stream mean field runs the x+y part on
the side stream beside that convolution and the z part with the assembly
afterwards.  Both sum the same stencil terms in the same order and round the
x+y part to the storage type before adding the z part, so every field must
agree bit for bit.
Covered: the three boundary modes (icon 13, 8, 18), both tile shapes (32 x 8
on 32-wide planes, 16 x 16 on 48-wide ones), the absorbing layer, one and two
classical impurities, the quantum impurity, mixed precision, graph replay and
a poten call on its own.  The oracle comparison of the default path is in the
trajectory tests of test_gpu_parity.py, test_gpu_headline.py and
test_gpu_quantum.py."""
import numpy as np
import pytest

from he_dft_2024_amd import inputs as I
from he_dft_2024_amd import quantum as Q

pytestmark = pytest.mark.gpu

DT = 5e-4 / I.PSTOK
FIELDS = ("PSI", "DEN", "PC", "PSIOLD1", "HPSIOLD1", "HPSIOLD2", "POT4", "UALPHAS", "DENALF", "DENCG",
          "DXDEN", "DZDEN")


def _say(*a):
    print("[case]", *a, flush=True)


def _fused(d, on):
    t = d.get_tuning()
    t["fused"] = on
    d.set_tuning(t)
    assert d.get_tuning()["fused"] == on


def _droplet(gpu, n, fused, icon=13, absorb=False, precision=0, two=False):
    xmax = 0.2 * n
    kw = dict(core_otc=1, icon=icon, precision=precision)
    if absorb:
        kw.update(lambdah=2.0, tmean=(xmax - 2.0,) * 3, tsurf=(2.0,) * 3)
    d = gpu.Droplet(n, n, n, xmax, xmax, xmax, gpu.default_params(**kw), mode=gpu.MODE_TDDFT)
    _fused(d, fused)
    n4 = 0.45 * n ** 3 * 0.0022
    if two:
        potion, delinter, rmax = I.impurity_table(n, xmax)
        potion2, _, _ = I.impurity_table(n, xmax, eps=12.0, sigma=3.5)
        v, dv, dl = I.pair_tables(2, 4 * xmax, npair=20001)
        d.set("DEN", I.woods_saxon(n, n, n, xmax, xmax, xmax, n4, seed=5))
        d.set_impurities(np.stack([potion, potion2]), delinter, rmax, [108.0 * I.MP_U, 40.0 * I.MP_U])
        d.set_impurity_pairs(v, dv, dl)
        d.set_impurity_state(np.array([[0.3, -0.2, -2.0], [-1.0, 0.5, 2.5]]),
                             np.array([[0.0, 1.0, 3.0], [0.5, 0.0, -2.0]]) * I.PSTOK, np.zeros((2, 3)))
    else:
        I.tddft_setup(d, n, xmax, n4, seed=5, r0=(0.3, -0.2, -2.0), v_aps=(0.0, 1.0, 3.0))
    return d


def _run(d, nrk=3, npc=5, graph_steps=0):
    d.tddft_init()
    errs = [d.tddft_iteration(DT, it, nrk) for it in range(1, nrk + npc + 1)]
    if graph_steps:
        d.run_pc(DT, graph_steps, True)
    return np.array(errs)


def _same(a, b, fields=FIELDS):
    for name in fields:
        assert np.array_equal(a.field(name), b.field(name)), name
    ea, eb = a.energy(), b.energy()
    for key in ("etot", "ekin4", "elj4", "ecor4", "ealphas", "eHeX", "norm"):
        assert ea[key] == eb[key], key


@pytest.mark.parametrize("n,icon,absorb,two", [(32, 13, False, False), (32, 8, True, False),
                                               (32, 18, True, True), (48, 13, True, False),
                                               (48, 8, False, True)])
def test_fused_mean_field_equals_two_stream(gpu, n, icon, absorb, two):
    _say("fused vs two-stream mean field", n, "icon", icon, "absorber", absorb, "two impurities", two)
    a = _droplet(gpu, n, 1, icon, absorb, two=two)
    b = _droplet(gpu, n, 0, icon, absorb, two=two)
    ea, eb = _run(a), _run(b)
    assert np.array_equal(ea, eb)
    _same(a, b)
    ia, ib = a.impurity_state(), b.impurity_state()
    for key in ("rimp", "vimp", "aimp", "force"):
        assert np.array_equal(ia[key], ib[key]), key
    a.close()
    b.close()


def test_fused_graph_and_mixed_equal_two_stream(gpu):
    _say("fused vs two-stream mean field: graph replay, FP64 and mixed precision")
    for prec in (0, 1):
        a = _droplet(gpu, 32, 1, precision=prec, absorb=True)
        b = _droplet(gpu, 32, 0, precision=prec, absorb=True)
        _run(a, graph_steps=13)
        _run(b, graph_steps=13)
        _same(a, b)
        a.close()
        b.close()


def test_fused_poten_alone_equals_two_stream(gpu):
    _say("fused vs two-stream: poten on its own")
    ds = [_droplet(gpu, 32, f, absorb=True) for f in (1, 0)]
    for d in ds:
        d.tddft_init()
        d.poten()
    _same(ds[0], ds[1], ("POT4", "UALPHAS", "DENALF", "DENCG", "DELJ4", "INTXALF", "INTZALF"))
    for d in ds:
        d.close()


def test_fused_quantum_impurity_equals_two_stream(gpu):
    _say("fused vs two-stream mean field: quantum impurity")
    n, xmax = 32, 6.4
    v = Q.lennard_jones(10.22, 2.556)
    uk = Q.uimp_k_table(n, n, n, xmax, xmax, xmax, v, 2.0, float(v(2.0)))
    ds = []
    for f in (1, 0):
        d = gpu.Droplet(n, n, n, xmax, xmax, xmax, gpu.default_params(core_otc=1), mode=gpu.MODE_TDDFT)
        _fused(d, f)
        I.quantum_setup(d, n, xmax, 60.0, uk, seed=3, m_imp_u=2 * 4.002603, mu4=-1.5, r0=(0.4, -0.8, 1.2),
                        width=1.3, k=(0.6, 0.0, -0.3))
        _run(d, graph_steps=6)
        ds.append(d)
    for name in ("PSI", "PSIX", "PCX", "POT4", "POTX4", "UPOTX", "UALPHAS"):
        assert np.array_equal(ds[0].field(name), ds[1].field(name)), name
    for d in ds:
        d.close()


def test_fused_tuning_checks(gpu):
    """fused is 0 (the default) or 1; it is kept by set_tuning calls that leave
    it out."""
    _say("fused tuning checks")
    d = gpu.Droplet(32, 32, 32, 6.4, 6.4, 6.4, gpu.default_params(core_otc=1), mode=gpu.MODE_TDDFT)
    assert d.get_tuning()["fused"] == 0
    _fused(d, 1)
    t = d.get_tuning()
    t["fused"] = 2
    with pytest.raises(gpu.HedftError):
        d.set_tuning(t)
    d.set_tuning({"lap_late": 4})
    assert d.get_tuning()["fused"] == 1
    d.close()
