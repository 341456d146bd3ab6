"""Classical-impurity integrator and pair terms on the GPU: exactness under a
constant acceleration, the two-body harmonic problem against the oracle (the
pair force through eager steps and graph replay), the state guards of the
C-ABI (pair tables required for nimp >= 2; a new impurity set invalidates the
captured graph)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import relerr
from he_dft_2024_amd import inputs as I

pytestmark = pytest.mark.gpu

N_BOX, XM_BOX = 32, 12.8


def _say(*a):
    print("[case]", *a, flush=True)


def _empty(x, nimp, masses):
    x.set("DEN", np.zeros((N_BOX, N_BOX, N_BOX)))
    x.set_impurities(np.zeros((nimp, 101)), 0.5, 50.0, [m * I.MP_U for m in masses])


def test_impurity_pc_exact_under_constant_acceleration(gpu):
    """RK-Gill + Hamming PC integrate r(t) = r0 + v0 t + a t^2/2 exactly
    (Dynamics_TDDFT/steprk.f90:93-117, steppc.f90:80-97,149-174): a wrong slot
    rotation or coefficient misses at O(1)."""
    _say("impurity PC, constant acceleration")
    d = gpu.Droplet(N_BOX, N_BOX, N_BOX, XM_BOX, XM_BOX, XM_BOX, gpu.default_params(), mode=gpu.MODE_TDDFT)
    _empty(d, 1, [40.0])
    R0 = np.array([[1.0, 0.0, -2.0]])
    V0 = np.array([[0.3, -0.2, 0.5]]) * I.PSTOK
    A = np.array([[2.0, -1.0, 0.5]]) * I.PSTOK ** 2
    d.set_impurity_state(R0, V0, np.zeros((1, 3)))
    d.tddft_init()
    d.set_impurity_state(None, None, A)
    dt = 0.01
    for _ in range(3):
        d.steprk(dt)
    for _ in range(20):
        d.steppc(dt)
    T = 23 * dt
    s = d.impurity_state()
    np.testing.assert_allclose(s["rimp"], R0 + V0 * T + 0.5 * A * T * T, rtol=0, atol=1e-13)
    np.testing.assert_allclose(s["vimp"], V0 + A * T, rtol=0, atol=1e-12)
    d.close()


def test_two_body_pair_force_vs_oracle(gpu):
    """Two impurities bound by a linear pair force over an empty box: eager RK
    start, then graph-replayed PC iterations, against the oracle (whose
    first-order convergence to the harmonic solution is pinned on the CPU)."""
    _say("two-body pair force, eager + graph")
    d = gpu.Droplet(N_BOX, N_BOX, N_BOX, XM_BOX, XM_BOX, XM_BOX, gpu.default_params(), mode=gpu.MODE_TDDFT)
    o = O.Oracle(N_BOX, N_BOX, N_BOX, XM_BOX, XM_BOX, XM_BOX, O.default_params(), O.MODE_TDDFT)
    dl = 0.05
    r = dl * np.arange(1001)
    R0 = np.array([[1.0, 0.0, -2.0], [-1.0, 0.5, 1.5]])
    V0 = np.array([[0.0, 0.3, 0.1], [0.2, -0.1, 0.0]]) * I.PSTOK
    dt = 0.01
    for x in (d, o):
        _empty(x, 2, [40.0, 80.0])
        x.set_impurity_pairs((2.5 * r * r)[None, :], (5.0 * r)[None, :], dl)
        x.set_impurity_state(R0, V0, np.zeros((2, 3)))
        x.tddft_init()
        for it in range(1, 4):
            x.tddft_iteration(dt, it, 3)
    d.run_pc(dt, 12, True)
    for it in range(4, 16):
        o.tddft_iteration(dt, it, 3)
    sg, so = d.impurity_state(), o.impurity_state()
    for k in ("rimp", "vimp", "aimp", "force"):
        assert relerr(sg[k], so[k]) < 1e-13, k
    assert abs(d.energy()["eimpu_impu"] - o.energy()["eimpu_impu"]) <= 1e-13 * abs(o.energy()["eimpu_impu"])
    d.close()


def test_several_impurities_need_pair_tables(gpu):
    _say("nimp=2 without pair tables")
    d = gpu.Droplet(N_BOX, N_BOX, N_BOX, XM_BOX, XM_BOX, XM_BOX, gpu.default_params(), mode=gpu.MODE_TDDFT)
    _empty(d, 2, [40.0, 80.0])
    d.set_impurity_state(np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]]), np.zeros((2, 3)), np.zeros((2, 3)))
    with pytest.raises(gpu.HedftError, match="ESTATE|order|state"):
        d.tddft_init()
    with pytest.raises(ValueError):
        d.set_impurity_pairs(np.zeros((2, 10)), np.zeros((2, 10)), 0.1)   # one pair, not two
    d.close()


def test_removing_impurities_invalidates_graph(gpu):
    """A graph captured with an impurity must not be replayed after the
    impurity set changes: graph replay after set_impurities(0) equals eager
    steps of the same state."""
    n, xmax, n4 = 32, 6.4, 60.0
    _say("graph after set_impurities(0)")
    dt = 5e-4 / I.PSTOK
    outs = []
    for use_graph in (True, False):
        d = gpu.Droplet(n, n, n, xmax, xmax, xmax, gpu.default_params(core_otc=1), mode=gpu.MODE_TDDFT)
        I.tddft_setup(d, n, xmax, n4, seed=3, r0=(0.0, 0.0, -4.0))
        d.tddft_init()
        for it in range(1, 4):
            d.tddft_iteration(dt, it, 3)
        d.run_pc(dt, 6, use_graph)
        d.set_impurities(np.zeros((0, 10)), 1.0, 10.0, np.zeros(0))
        d.run_pc(dt, 6, use_graph)
        outs.append(d.field("PSI"))
        assert np.abs(d.field("UIMP")).max() == 0.0
        d.close()
    assert relerr(outs[0], outs[1]) < 1e-13
