"""Writes the committed golden fixtures under tests/golden/.

stencil_ref.npz : effective derivnD matrices (n = 26 points along x) for
                  id = 1, 2 and icon in {0, 8, 13, 18, 20, 24}, produced by the
                  reference's own Init_deriv_p/Deriv3D_p (init_deriv_parallel.f90,
                  newder.f, redef.f and the mates.f coefficient table) compiled
                  into oracle/_ref from /root/reference.  Column j = derivnD of
                  the unit vector e_j, h = 1.
fvlj_exact.json : the Orsay-Trento LJ form factor of fforma.f90 evaluated from
                  its defining integral with mpmath at 40 digits, at |p| values
                  of a 128^3 grid with h = 0.4 A (including the zone corner).
Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))


def stencil_fixture():
    from oracle import oracle as O
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref not built (make -C oracle ref)")
    n = 26
    out = {}
    for icon in (0, 8, 13, 18, 20, 24):
        for id_ in (1, 2):
            m = np.zeros((n, n))
            for j in range(n):
                e = np.zeros((1, 1, n))
                e[0, 0, j] = 1.0
                m[:, j] = O.ref_derivnd(id_, e, 1.0, 1, icon)[0, 0, :]
            out[f"icon{icon}_id{id_}"] = m
    np.savez_compressed(os.path.join(HERE, "stencil_ref.npz"), **out)


def fvlj_fixture():
    import mpmath as mp
    mp.mp.dps = 40
    eps, sig, h = 10.22, 2.556, 2.190323
    hp = 1.0 / (128 * 0.4)
    ps = sorted({hp * np.sqrt(i * i + j * j + k * k)
                 for (i, j, k) in [(1, 0, 0), (3, 4, 0), (10, 0, 0), (20, 5, 3), (40, 40, 0),
                                   (64, 0, 0), (50, 60, 30), (64, 64, 64), (33, 17, 64)]})

    def exact(p):
        a = 2 * mp.pi * p
        s11 = mp.quadosc(lambda t: mp.sin(a * t) / t ** 11, [h, mp.inf], omega=a)
        s5 = mp.quadosc(lambda t: mp.sin(a * t) / t ** 5, [h, mp.inf], omega=a)
        return 8 * eps * sig ** 6 * (sig ** 6 * s11 - s5) / p
    rows = [{"p": float(p), "fvlj": float(exact(mp.mpf(p)))} for p in ps]
    with open(os.path.join(HERE, "fvlj_exact.json"), "w") as f:
        json.dump({"eps": eps, "sigma": sig, "h": h, "values": rows}, f, indent=1)


if __name__ == "__main__":
    stencil_fixture()
    fvlj_fixture()
    print("golden fixtures written")
