"""hedft_forceimp_noalloc (no context, no allocation, per-GPU static scratch)
against hedft_forceimp on a context holding the same fields: bit-identical
forces, on two calls in a row, the second with a smaller grid and fewer
impurities (the scratch keeps the first call's contents).  Device buffers are
allocated through the HIP runtime the library itself links (libamdhip64); the
library call goes through the C-ABI."""
import ctypes as C

import numpy as np
import pytest

from he_dft_2024_amd import inputs as I

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DevBuf:
    """Device copy of a numpy array (hipMalloc + hipMemcpy)."""
    def __init__(self, a):
        self.a = np.ascontiguousarray(a)
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), self.a.nbytes) == 0
        assert hip().hipMemcpy(self.ptr, self.a.ctypes.data_as(C.c_void_p), self.a.nbytes, 1) == 0  # H2D

    def get(self):
        out = np.empty_like(self.a)
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes, 2) == 0  # D2H
        return out

    def free(self):
        hip().hipFree(self.ptr)


def _case(H, n, xmax, nimp, seed):
    print("[case] forceimp_noalloc n", n, "nimp", nimp, flush=True)
    d = H.Droplet(n, n, n, xmax, xmax, xmax, H.default_params(core_otc=1), mode=H.MODE_TDDFT)
    rng = np.random.default_rng(seed)
    d.set("DEN", I.woods_saxon(n, n, n, xmax, xmax, xmax, 60.0, seed=seed))
    d.derden()
    lj = [{}, {"eps": 12.0, "sigma": 3.5}, {"eps": 30.0, "sigma": 4.4}, {"eps": 5.0, "sigma": 3.0}][:nimp]
    tabs = []
    for kw in lj:
        potion, delinter, rmax = I.impurity_table(n, xmax, **kw)
        tabs.append(potion)
    potion = np.ascontiguousarray(np.stack(tabs))
    d.set_impurities(potion, delinter, rmax, [40.0 * I.MP_U] * nimp)
    v = dv = None
    npair, dl = 0, 0.0
    if nimp >= 2:
        v, dv, dl = I.pair_tables(nimp, 4 * xmax, npair=20001, pairs=lj[:nimp * (nimp - 1) // 2] if nimp == 3 else None)
        npair = v.shape[1]
        d.set_impurity_pairs(v, dv, dl)
    R = rng.uniform(-0.5 * xmax, 0.5 * xmax, (nimp, 3))
    d.set_impurity_state(R, np.zeros((nimp, 3)), np.zeros((nimp, 3)))
    d.forceimp()
    f_ctx = d.impurity_state()["force"]
    bufs = [DevBuf(d.field(nm)) for nm in ("DXDEN", "DYDEN", "DZDEN")]
    tpot = DevBuf(potion)
    trimp = DevBuf(R.ravel(order="F"))
    tdv = DevBuf(dv) if dv is not None else None
    tforce = DevBuf(np.full(3 * nimp, np.nan))
    tflag = DevBuf(np.zeros(1, dtype=np.int32))
    assert hip().hipDeviceSynchronize() == 0
    rc = H.lib().hedft_forceimp_noalloc(C.byref(d.grid), 0, None, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, nimp,
                                        potion.shape[1], delinter, rmax, tpot.ptr, trimp.ptr, npair, dl,
                                        tdv.ptr if tdv else None, tforce.ptr, tflag.ptr)
    assert rc == 0, H.lib().hedft_strerror(rc)
    assert hip().hipDeviceSynchronize() == 0
    f = tforce.get().reshape(3, nimp).T
    assert np.abs(f).max() > 0.0
    assert np.array_equal(f, f_ctx), (f, f_ctx)
    assert int(tflag.get()[0]) == 0
    for b in bufs + [tpot, trimp, tforce, tflag] + ([tdv] if tdv else []):
        b.free()
    d.close()


def test_forceimp_noalloc_matches_context_path(gpu):
    _case(gpu, 48, 9.6, 3, seed=1)
    _case(gpu, 32, 6.4, 1, seed=2)   # smaller grid, one impurity: stale scratch must not leak in


def test_forceimp_noalloc_rejects_too_many_impurities(gpu):
    from he_dft_2024_amd.hedft import Grid
    g = Grid(32, 32, 32, 0, 6.4, 6.4, 6.4)
    z = C.c_void_p(16)   # never dereferenced: the size check comes first
    rc = gpu.lib().hedft_forceimp_noalloc(C.byref(g), 0, None, z, z, z, 5, 10, 0.1, 1.0, z, z, 10, 0.1, z, z,
                                          None)
    assert rc == -1
