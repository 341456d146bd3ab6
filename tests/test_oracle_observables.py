"""Pins of the oracle's observables block (Dynamics_TDDFT/DFT4He3d.f90:709-836,
r_cm.f90) against closed forms, on the CPU: states of angular momentum +N
about each axis, a boosted packet's centre-of-mass velocity k/m, Gaussian second
moments, and r_cm's division by the nominal atom number."""
import numpy as np

from oracle import oracle as O
from he_dft_2024_amd import inputs as I

N, XMAX = 32, 6.4


def _oracle(psi, den=None):
    o = O.Oracle(N, N, N, XMAX, XMAX, XMAX, O.default_params())
    o.set("PSI", psi)
    o.set("DEN", np.abs(psi) ** 2 if den is None else den)
    return o


def _coords(c=(0.0, 0.0, 0.0)):
    x = I.grid_axes(N, XMAX)
    return (x[None, None, :] - c[0], x[None, :, None] - c[1], x[:, None, None] - c[2])


def _normalise(psi, n4):
    dv = (2 * XMAX / N) ** 3
    return psi * np.sqrt(n4 / ((np.abs(psi) ** 2).sum() * dv))


def test_angular_momentum_of_rotating_states():
    h = 2 * XMAX / N
    c = (-6.4 + 16 * h, -6.4 + 15 * h, -6.4 + 17 * h)  # a grid point near the centre
    X, Y, Z = _coords(c)
    g = np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / (2 * 1.2 ** 2))
    n4 = 20
    for axis, psi in ((2, (X + 1j * Y) * g), (0, (Y + 1j * Z) * g), (1, (Z + 1j * X) * g)):
        o = _oracle(_normalise(psi, n4))
        ob = o.observables(n4)
        assert abs(ob["norm"] - n4) < 1e-10
        assert np.abs(ob["rcm"] - np.array(c)).max() < 1e-8  # box truncation of the tails
        want = np.zeros(3)
        want[axis] = n4
        assert np.abs(ob["l"] - want).max() < 2e-5 * n4, (axis, ob["l"])
        assert np.abs(ob["vcom"] * O.MHE).max() < 1e-10


def test_com_velocity_of_boosted_packet():
    X, Y, Z = _coords()
    k = np.array([0.7, -0.3, 0.45])
    psi = np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / (2 * 1.5 ** 2) + 1j * (k[0] * X + k[1] * Y + k[2] * Z))
    o = _oracle(_normalise(psi, 10))
    ob = o.observables(10)
    assert np.abs(ob["vcom"] * O.MHE - k).max() < 1e-6


def test_second_moments_and_nominal_n4():
    h = 2 * XMAX / N
    s = 1.3
    c = (0.5 * h, -h, 0.0)
    X, Y, Z = _coords(c)
    den = np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / s ** 2)
    psi = np.sqrt(den).astype(np.complex128)
    o = _oracle(psi, den)
    norm = np.pi ** 1.5 * s ** 3
    ob = o.observables(7)
    assert abs(ob["norm"] - norm) < 1e-10 * norm
    # r_cm divides by the nominal n4, not by the integral
    assert np.abs(ob["rcm"] - np.array(c) * norm / 7).max() < 1e-8  # box truncation of the tails
    want = norm * (s ** 2 / 2 + np.array(c) ** 2)
    assert np.abs(ob["r2"] - want).max() < 1e-8 * want.max()
