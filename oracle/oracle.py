"""ctypes binding of the CPU FP64 oracle (oracle/hedft_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker.  The product library never
imports it.  It also binds oracle/_ref/libhedft_ref.so (the reference's own
stencil and kernel routines compiled from its Fortran sources) when that
library has been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libhedft_ref.so")
REFQ_PATH = os.path.join(HERE, "_ref", "libhedft_refq.so")

MODE_STATICS, MODE_TDDFT = 0, 1

# field ids (same numbering as enum hedft_field in include/hedft.h)
FIELDS = [
    "DEN", "PSI", "POT4", "DENCG", "DELJ4", "DENALF", "FALFS", "INTXALF", "INTYALF", "INTZALF",
    "UALPHAS", "DXDEN", "DYDEN", "DZDEN", "UIMP", "UEXT", "TIMEC", "PSIOLD1", "PSIOLD2",
    "PSIOLD3", "HPSIOLD1", "HPSIOLD2", "PC", "Q", "FDEN", "FVLJ4", "WCGK", "KALFS", "WORK1",
    "WORK2", "WORKC1", "WORKK1",
    "PSIX", "POTX4", "UPOTX", "DENX", "UIMPK", "PSIXOLD1", "PSIXOLD2", "PSIXOLD3", "HPSIXOLD1",
    "HPSIXOLD2", "PCX", "QX", "PSIL",
]
FIELD = {name: i for i, name in enumerate(FIELDS)}
MHE = 4.002572164 * 0.02061484  # helium mass in K^-1 A^-2 (Dynamics_TDDFT/DFT4He3d.f90:103,315)
ENERGY_NAMES = ["ekin4", "elj4", "ecor4", "ealphas", "esolid", "eext", "etot4", "ekinx", "eHeX",
                "eimpu_impu", "eimpu", "etot", "norm"]


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("pad_", C.c_int32),
                ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double)]


class Params(C.Structure):
    _fields_ = [
        ("h2o2m4", C.c_double), ("cp4", C.c_double), ("cpp4", C.c_double),
        ("alphas", C.c_double), ("l", C.c_double), ("den0s", C.c_double),
        ("csolid", C.c_double), ("beta_solid", C.c_double), ("den_m", C.c_double),
        ("h4", C.c_double), ("eps4", C.c_double), ("sigma4", C.c_double),
        ("core_otc", C.c_int32), ("lsolid", C.c_int32), ("npd", C.c_int32), ("icon", C.c_int32),
        ("dtmin", C.c_double), ("sdmu", C.c_double), ("dmumax", C.c_double), ("psimin", C.c_double),
        ("lambdah", C.c_double), ("tmean", C.c_double * 3), ("tsurf", C.c_double * 3),
        ("precision", C.c_int32), ("lexternal", C.c_int32),
    ]


def default_params(**kw) -> Params:
    """Reference module defaults (He_statics/modules.f90:15-19,109-133,214-226,
    297-299; Dynamics_TDDFT/modules.f90 he4; DFT4He3d.f90:136-138)."""
    p = Params()
    p.h2o2m4, p.cp4, p.cpp4 = 6.05969638298, -2.41186e4, 1.85850e6
    p.alphas, p.l, p.den0s = 54.31, 1.0, 0.04
    p.csolid, p.beta_solid, p.den_m = 3.1577504e4, 40.0, 0.37
    p.h4, p.eps4, p.sigma4 = 2.190323, 10.22, 2.556
    p.core_otc, p.lsolid, p.npd, p.icon = 0, 0, 13, 13
    p.dtmin, p.sdmu, p.dmumax, p.psimin = 1.0e-10, 0.1, 100.0, -5.0
    p.lambdah = 2.0
    for a in range(3):
        p.tmean[a] = 1000.0
        p.tsurf[a] = 2.0
    p.precision, p.lexternal = 0, 0
    for k, v in kw.items():
        if k in ("tmean", "tsurf"):
            for a in range(3):
                getattr(p, k)[a] = v[a]
        else:
            setattr(p, k, v)
    return p


_lib = None
_ref = None
_refq = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_int
        L.oracle_create.restype = vp
        L.oracle_create.argtypes = [C.POINTER(Grid), C.POINTER(Params), ip]
        L.oracle_destroy.argtypes = [vp]
        L.oracle_field.restype = dp
        L.oracle_field.argtypes = [vp, ip, C.POINTER(C.c_size_t)]
        L.oracle_derivnd.argtypes = [ip, ip, ip, ip, C.c_double, ip, vp, vp, ip, ip]
        L.oracle_stencil_matrix.argtypes = [ip, ip, ip, C.c_double, vp]
        L.oracle_fft_r2c.argtypes = [ip, ip, ip, vp, vp]
        L.oracle_fft_c2r.argtypes = [ip, ip, ip, vp, vp]
        L.oracle_sine_integral.restype = C.c_double
        L.oracle_sine_integral.argtypes = [C.c_double]
        L.oracle_fvlj_radial.restype = C.c_double
        L.oracle_fvlj_radial.argtypes = [C.c_double] * 4
        for name in ("oracle_derden", "oracle_term_alfa", "oracle_poten", "oracle_statics_refresh",
                     "oracle_potenimp", "oracle_forceimp", "oracle_update_aimp", "oracle_tddft_init"):
            getattr(L, name).argtypes = [vp]
        L.oracle_steprkr.argtypes = [vp, C.c_double, dp, dp, C.c_double]
        L.oracle_steppcr.argtypes = [vp, C.c_double, dp, dp, C.c_double, dp]
        L.oracle_steprk.argtypes = [vp, C.c_double]
        L.oracle_steppc.argtypes = [vp, C.c_double, dp, dp, dp]
        L.oracle_set_impurities.argtypes = [vp, ip, ip, C.c_double, C.c_double, vp, vp]
        L.oracle_set_impurity_state.argtypes = [vp, vp, vp, vp]
        L.oracle_set_impurity_pairs.argtypes = [vp, ip, C.c_double, vp, vp]
        L.oracle_get_impurity_state.argtypes = [vp, vp, vp, vp, vp, C.POINTER(C.c_int)]
        L.oracle_energy.argtypes = [vp, vp]
        L.oracle_statics_init.argtypes = [vp, C.c_double, dp]
        L.oracle_statics_iteration.argtypes = [vp, C.c_double, ip, C.c_double, dp, dp, dp]
        L.oracle_statics_iteration_log.argtypes = [vp, C.c_double, ip, C.c_double, dp, dp, dp]
        L.oracle_steprkrl.argtypes = [vp, C.c_double, dp, dp, C.c_double]
        L.oracle_steppcrl.argtypes = [vp, C.c_double, dp, dp, C.c_double, dp]
        L.oracle_statics_log_start.argtypes = [vp]
        L.oracle_evolox.argtypes = [vp, C.c_double, C.c_double, C.c_double, ip, dp, dp]
        L.oracle_evolo.argtypes = [vp, C.c_double, C.c_double, C.c_double, ip, C.c_double, dp, dp]
        L.oracle_statics_iteration_sd.argtypes = [vp, C.c_double, C.c_double, C.c_double, ip, C.c_double, dp, dp]
        L.oracle_ironing.argtypes = [ip, ip, ip, vp]
        L.oracle_tddft_iteration.argtypes = [vp, C.c_double, ip, ip, dp, dp, dp]
        L.oracle_set_quantum_impurity.argtypes = [vp, vp, C.c_double, C.c_double]
        L.oracle_set_quantum_impurities.argtypes = [vp, ip, vp, vp, vp, vp, C.c_double]
        L.oracle_select_quantum.argtypes = [vp, ip]
        L.oracle_quantum_centroid.argtypes = [vp, dp, vp]
        L.oracle_observables.argtypes = [vp, ip, C.c_double, vp]
        L.oracle_num_threads.restype = ip
        _lib = L
    return _lib


def ref():
    """The reference's own routines (oracle/_ref), or None when not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = C.CDLL(REF_PATH)
        ip, vp, d = C.c_int, C.c_void_p, C.c_double
        R.ref_init_deriv.argtypes = [ip, ip, ip]
        R.ref_deriv3d.argtypes = [ip, ip, ip, ip, d, ip, vp, vp, ip, ip]
        R.ref_fforma.argtypes = [ip, ip, ip, vp, vp, d, d, d, d]
        R.ref_initcg.argtypes = [ip, ip, ip, vp, vp, d]
        R.ref_si.restype = d
        R.ref_si.argtypes = [d]
        R.ref_evolo.argtypes = [ip, ip, ip, d, d, d, ip, vp, vp, vp, vp, d, d, d, ip, d, C.POINTER(d),
                                C.POINTER(d), vp, vp]
        R.ref_read_density.argtypes = [C.c_char_p, ip, C.c_char, vp, vp, ip, C.POINTER(C.c_int)]
        R.ref_read_wavefunction.argtypes = [C.c_char_p, ip, C.c_char, vp, C.POINTER(C.c_int), vp, vp, ip, vp, ip,
                                            C.POINTER(C.c_int)]
        R.ref_init_deriv(13, 2, 1)
        _ref = R
    return _ref


def ref_q():
    """The reference's impurity-potential library and FT_V_spline
    (Dynamics_quantum_impurities, oracle/_ref), or None when not built."""
    global _refq
    if _refq is None and os.path.exists(REFQ_PATH):
        R = C.CDLL(REFQ_PATH)
        d, ip = C.c_double, C.c_int
        R.refq_select_pot.restype = d
        R.refq_select_pot.argtypes = [C.c_char_p, ip, d, d, d]
        R.refq_ft_v_spline.restype = d
        R.refq_ft_v_spline.argtypes = [d, d, d, C.c_char_p, ip, d, d, ip]
        _refq = R
    return _refq


def ref_select_pot(selec: str, r, r_cutoff: float, umax: float) -> np.ndarray:
    """Select_Pot(selec, r, r_cutoff, umax) of the reference at each r."""
    R = ref_q()
    b = selec.encode()
    return np.array([R.refq_select_pot(b, len(b), float(x), r_cutoff, umax) for x in np.atleast_1d(r)])


def ref_ft_v_spline(selec: str, p, pmax: float, hq: float, r_cutoff: float, umax: float, parabolic=False):
    """FT_V_spline(p, pmax, hq, selec, r_cutoff, umax) of the reference at each p."""
    R = ref_q()
    b = selec.encode()
    return np.array([R.refq_ft_v_spline(float(x), pmax, hq, b, len(b), r_cutoff, umax, int(parabolic))
                     for x in np.atleast_1d(p)])


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def derivnd(id_, f: np.ndarray, h: float, iv: int, icon: int) -> np.ndarray:
    """derivnD on an array shaped (nz, ny, nx) (C order == Fortran (nx,ny,nz))."""
    f = np.ascontiguousarray(f)
    nz, ny, nx = f.shape
    ncomp = 2 if np.iscomplexobj(f) else 1
    df = np.zeros_like(f)
    lib().oracle_derivnd(id_, nx, ny, nz, h, iv, _ptr(f), _ptr(df), icon, ncomp)
    return df


def ref_derivnd(id_, f: np.ndarray, h: float, iv: int, icon: int) -> np.ndarray:
    R = ref()
    f = np.ascontiguousarray(f)
    nz, ny, nx = f.shape
    df = np.zeros_like(f)
    R.ref_deriv3d(id_, nx, ny, nz, h, iv, _ptr(f), _ptr(df), icon, 1 if np.iscomplexobj(f) else 0)
    return df


def ironing(f: np.ndarray) -> np.ndarray:
    """One smoothing pass of ironing.f90 (oracle restatement) on a copy."""
    g = np.ascontiguousarray(f, dtype=np.float64).copy()
    nz, ny, nx = g.shape
    lib().oracle_ironing(nx, ny, nz, _ptr(g))
    return g


def ref_evolo(psi, pot4, psi1, psi2, h, icon, deltat, n4, mu4, vdt=(0.6, 0.1), iron=1):
    """The reference's evolo (oracle/_ref) on copies of the given arrays:
    returns (psi, psi1, psi2, den, hpsi, mu4, mu4err)."""
    R = ref()
    arrs = [np.ascontiguousarray(a, dtype=np.float64).copy() for a in (psi, pot4, psi1, psi2)]
    nz, ny, nx = arrs[0].shape
    den = np.zeros_like(arrs[0])
    hp = np.zeros_like(arrs[0])
    m, e = C.c_double(mu4), C.c_double(0)
    R.ref_evolo(nx, ny, nz, h[0], h[1], h[2], icon, _ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]), _ptr(arrs[3]),
                deltat, vdt[0], vdt[1], iron, n4, C.byref(m), C.byref(e), _ptr(den), _ptr(hp))
    return arrs[0], arrs[2], arrs[3], den, hp, m.value, e.value


def ref_read_density(path, n, cchar=b"#"):
    """The reference's readen.f90 records read by the Fortran runtime
    (oracle/_ref): returns (header[13], den[n])."""
    R = ref()
    hdr = np.zeros(13)
    den = np.zeros(n)
    err = C.c_int(0)
    b = path.encode()
    R.ref_read_density(b, len(b), cchar, _ptr(hdr), _ptr(den), n, C.byref(err))
    if err.value != 0:
        raise IOError(f"Fortran read failed (iostat {err.value})")
    return hdr, den


def ref_read_wavefunction(path, n, maximp=4, cchar=b"#"):
    """readenc.f90 mode 2 records via the Fortran runtime: (header[9], nimp,
    rimp[maximp,3] Fortran order, vimp, psi[n])."""
    R = ref()
    hdr = np.zeros(9)
    ni = C.c_int(0)
    rimp = np.zeros((3, maximp))
    vimp = np.zeros((3, maximp))
    psi = np.zeros(n, dtype=np.complex128)
    err = C.c_int(0)
    b = path.encode()
    R.ref_read_wavefunction(b, len(b), cchar, _ptr(hdr), C.byref(ni), _ptr(rimp), _ptr(vimp), maximp, _ptr(psi),
                            n, C.byref(err))
    if err.value != 0:
        raise IOError(f"Fortran read failed (iostat {err.value})")
    return hdr, ni.value, rimp, vimp, psi


def stencil_matrix(id_, icon, h=1.0, K=13):
    m = np.zeros(K * K)
    lib().oracle_stencil_matrix(id_, K, icon, h, _ptr(m))
    return m.reshape(K, K)


def fft_r2c(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    nz, ny, nx = a.shape
    out = np.zeros((nz, ny, nx // 2 + 1), dtype=np.complex128)
    lib().oracle_fft_r2c(nx, ny, nz, _ptr(a), _ptr(out))
    return out


def fft_c2r(s: np.ndarray, nx: int) -> np.ndarray:
    s = np.array(s, dtype=np.complex128, order="C", copy=True)
    nz, ny, _ = s.shape
    out = np.zeros((nz, ny, nx), dtype=np.float64)
    lib().oracle_fft_c2r(nx, ny, nz, _ptr(s), _ptr(out))
    return out


class Oracle:
    """One droplet: the reference's module state and its hot-path routines."""

    def __init__(self, nx, ny, nz, xmax, ymax, zmax, params: Params | None = None,
                 mode: int = MODE_TDDFT):
        self.grid = Grid(nx, ny, nz, 0, xmax, ymax, zmax)
        self.params = params if params is not None else default_params()
        self.mode = mode
        self.shape = (nz, ny, nx)
        self.hshape = (nz, ny, nx // 2 + 1)
        self.h = (2 * xmax / nx, 2 * ymax / ny, 2 * zmax / nz)
        self.dxyz = self.h[0] * self.h[1] * self.h[2]
        self._c = lib().oracle_create(C.byref(self.grid), C.byref(self.params), mode)
        self.nimp = 0

    def close(self):
        if self._c:
            lib().oracle_destroy(self._c)
            self._c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- field access (zero-copy numpy views) -----------------------------
    def field(self, name: str) -> np.ndarray:
        cnt = C.c_size_t()
        p = lib().oracle_field(self._c, FIELD[name], C.byref(cnt))
        flat = np.ctypeslib.as_array(p, shape=(cnt.value,))  # a view of the state: copy to keep a snapshot
        n = self.shape[0] * self.shape[1] * self.shape[2]
        nh = self.hshape[0] * self.hshape[1] * self.hshape[2]
        if cnt.value == n:
            return flat.reshape(self.shape)
        if cnt.value == 2 * n:
            return flat.view(np.complex128).reshape(self.shape)
        if cnt.value == nh:
            return flat.reshape(self.hshape)
        return flat.view(np.complex128).reshape(self.hshape)

    def set(self, name: str, value):
        self.field(name)[...] = value

    # ---- routines ------------------------------------------------------------
    def derden(self): lib().oracle_derden(self._c)
    def term_alfa(self): lib().oracle_term_alfa(self._c)
    def poten(self): lib().oracle_poten(self._c)
    def statics_refresh(self): lib().oracle_statics_refresh(self._c)
    def potenimp(self): lib().oracle_potenimp(self._c)
    def forceimp(self): lib().oracle_forceimp(self._c)
    def update_aimp(self): lib().oracle_update_aimp(self._c)
    def tddft_init(self): lib().oracle_tddft_init(self._c)

    def steprkr(self, dt, mu4, n4):
        m, e = C.c_double(mu4), C.c_double(0)
        lib().oracle_steprkr(self._c, dt, C.byref(m), C.byref(e), n4)
        return m.value, e.value

    def steppcr(self, dt, mu4, n4):
        m, e, pc = C.c_double(mu4), C.c_double(0), C.c_double(0)
        lib().oracle_steppcr(self._c, dt, C.byref(m), C.byref(e), n4, C.byref(pc))
        return m.value, e.value, pc.value

    def steprk(self, dt): lib().oracle_steprk(self._c, dt)

    def steppc(self, dt):
        a, b, c = C.c_double(0), C.c_double(0), C.c_double(0)
        lib().oracle_steppc(self._c, dt, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def set_impurities(self, potion: np.ndarray, delinter: float, rmax: float, mass):
        potion = np.ascontiguousarray(potion, dtype=np.float64)
        if potion.ndim == 1:
            potion = potion[None, :]
        mass = np.ascontiguousarray(np.atleast_1d(mass), dtype=np.float64)
        self.nimp = potion.shape[0]
        lib().oracle_set_impurities(self._c, self.nimp, potion.shape[1], delinter, rmax,
                                    _ptr(potion), _ptr(mass))

    def set_impurity_pairs(self, vpair, dvpair, delpair: float):
        """Pair tables (npairs, npair): energy (selec_gs_k_k) and force
        (drselec_gs_k_k) functions of |R_k - R_m| on r_i = i*delpair."""
        v = np.ascontiguousarray(np.atleast_2d(vpair), dtype=np.float64)
        dv = np.ascontiguousarray(np.atleast_2d(dvpair), dtype=np.float64)
        assert v.shape == dv.shape and v.shape[0] == self.nimp * (self.nimp - 1) // 2
        lib().oracle_set_impurity_pairs(self._c, v.shape[1], delpair, _ptr(v), _ptr(dv))

    def set_impurity_state(self, rimp=None, vimp=None, aimp=None):
        def conv(a):
            if a is None:
                return None, None
            a = np.asfortranarray(np.asarray(a, dtype=np.float64).reshape(self.nimp, 3))
            flat = np.ascontiguousarray(a.ravel(order="F"))
            return flat, _ptr(flat)
        r, rp = conv(rimp); v, vp = conv(vimp); a, ap = conv(aimp)
        lib().oracle_set_impurity_state(self._c, rp, vp, ap)

    def impurity_state(self):
        m = self.nimp * 3
        r, v, a, f = (np.zeros(m) for _ in range(4))
        ls = C.c_int(0)
        lib().oracle_get_impurity_state(self._c, _ptr(r), _ptr(v), _ptr(a), _ptr(f), C.byref(ls))
        shp = lambda x: x.reshape(3, self.nimp).T.copy()
        return {"rimp": shp(r), "vimp": shp(v), "aimp": shp(a), "force": shp(f), "lstopimp": ls.value}

    def energy(self) -> dict:
        out = np.zeros(len(ENERGY_NAMES))
        lib().oracle_energy(self._c, _ptr(out))
        return dict(zip(ENERGY_NAMES, out.tolist()))

    def statics_init(self, n4):
        m = C.c_double(0)
        lib().oracle_statics_init(self._c, n4, C.byref(m))
        return m.value

    def statics_iteration(self, dt, iterrk, n4, mu4):
        m, e, pc = C.c_double(mu4), C.c_double(0), C.c_double(0)
        lib().oracle_statics_iteration(self._c, dt, iterrk, n4, C.byref(m), C.byref(e), C.byref(pc))
        return m.value, e.value, pc.value

    def statics_refresh_poten(self):
        """the driver's refresh and poten after the steps (BCN4HeDFT.f90:1125-1160)"""
        lib().oracle_statics_refresh(self._c)
        lib().oracle_poten(self._c)

    def evolox(self, dtx, epsx, vdt=(0.6, 0.1), ironx=1):
        e, r = C.c_double(epsx), C.c_double(0)
        lib().oracle_evolox(self._c, dtx, vdt[0], vdt[1], ironx, C.byref(e), C.byref(r))
        return e.value, r.value

    def statics_log_start(self):
        lib().oracle_statics_log_start(self._c)

    def steprkrl(self, dt, mu4, n4):
        m, e = C.c_double(mu4), C.c_double(0)
        lib().oracle_steprkrl(self._c, dt, C.byref(m), C.byref(e), n4)
        return m.value, e.value

    def steppcrl(self, dt, mu4, n4):
        m, e, pc = C.c_double(mu4), C.c_double(0), C.c_double(0)
        lib().oracle_steppcrl(self._c, dt, C.byref(m), C.byref(e), n4, C.byref(pc))
        return m.value, e.value, pc.value

    def statics_iteration_log(self, dt, iterrk, n4, mu4):
        m, e, pc = C.c_double(mu4), C.c_double(0), C.c_double(0)
        lib().oracle_statics_iteration_log(self._c, dt, iterrk, n4, C.byref(m), C.byref(e), C.byref(pc))
        return m.value, e.value, pc.value

    def evolo(self, dt, n4, mu4, vdt=(0.6, 0.1), iron=1):
        m, e = C.c_double(mu4), C.c_double(0)
        lib().oracle_evolo(self._c, dt, vdt[0], vdt[1], iron, n4, C.byref(m), C.byref(e))
        return m.value, e.value

    def statics_iteration_sd(self, dt, n4, mu4, vdt=(0.6, 0.1), iron=1):
        m, e = C.c_double(mu4), C.c_double(0)
        lib().oracle_statics_iteration_sd(self._c, dt, vdt[0], vdt[1], iron, n4, C.byref(m), C.byref(e))
        return m.value, e.value

    def tddft_iteration(self, dt, it, nrk):
        a, b, c = C.c_double(0), C.c_double(0), C.c_double(0)
        lib().oracle_tddft_iteration(self._c, dt, it, nrk, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def observables(self, n4, mhe=MHE):
        out = np.zeros(13)
        lib().oracle_observables(self._c, int(n4), float(mhe), _ptr(out))
        return {"norm": out[0], "rcm": out[1:4].copy(), "l": out[4:7].copy(), "vcom": out[7:10].copy(),
                "r2": out[10:13].copy()}

    def set_quantum_impurity(self, uimp_k, m_imp_u, mu4=0.0):
        k = np.ascontiguousarray(np.broadcast_to(np.asarray(uimp_k, dtype=np.float64), self.hshape))
        lib().oracle_set_quantum_impurity(self._c, _ptr(k), float(m_imp_u), float(mu4))

    def set_quantum_impurities(self, uimp_ks, masses_u, mu4=0.0, uimp_kk=None):
        """N_imp quantum impurities: uimp_ks[i] = uimp_k(:,:,:,i) (a real half
        spectrum each), masses_u[i] in u, uimp_kk = {(i, j): uimp_k_k(:,:,:,i,j)}
        for the pairs whose Selec_gs_k_k is not 'null' (i != j; row i is the
        impurity that feels the potential)."""
        nq = len(masses_u)
        uk = np.ascontiguousarray(np.stack([np.broadcast_to(np.asarray(u, dtype=np.float64), self.hshape)
                                            for u in uimp_ks]))
        ukk = np.zeros((nq, nq) + tuple(self.hshape))
        pair = np.zeros(nq * nq, dtype=np.int32)
        for (i, j), u in (uimp_kk or {}).items():
            if i == j:
                raise ValueError("uimp_k_k(i,i) is not used")
            ukk[j, i] = np.broadcast_to(np.asarray(u, dtype=np.float64), self.hshape)
            pair[i + j * nq] = 1  # uimp_k_k(:,:,:,i,j) column-major
        m = np.ascontiguousarray(masses_u, dtype=np.float64)
        lib().oracle_set_quantum_impurities(self._c, nq, _ptr(uk), _ptr(ukk), pair.ctypes.data_as(C.c_void_p),
                                            _ptr(m), float(mu4))

    def select_quantum(self, k):
        """The impurity the PSIX* / UPOTX / UIMPK fields and quantum_centroid
        refer to."""
        lib().oracle_select_quantum(self._c, int(k))

    def quantum_centroid(self):
        nrm = C.c_double(0)
        r = np.zeros(3)
        lib().oracle_quantum_centroid(self._c, C.byref(nrm), _ptr(r))
        return nrm.value, r


def num_threads() -> int:
    return lib().oracle_num_threads()
