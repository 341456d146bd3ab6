"""The Fortran side of the drop-in boundary: module hedft_gpu (ISO_C_BINDING)
driving the library the way the TDDFT driver loop does.  The check program
(he_dft_2024_amd/fortran/hedft_fortran_check, built by build()) and the Python
binding run the same RK start + PC iterations on the same inputs; the results
must be identical, since both call the same C entry points in the same order."""
import os
import subprocess

import numpy as np
import pytest

from he_dft_2024_amd import inputs as I

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "he_dft_2024_amd", "fortran")
EXE = os.path.join(FDIR, "_build", "hedft_fortran_check")
QEXE = os.path.join(FDIR, "_build", "hedft_fortran_qcheck")


def test_fortran_sources_present():
    for f in ("hedft_gpu.f90", "hedft_fortran_check.f90", "hedft_fortran_qcheck.f90", "hedft_fortran_scheck.f90",
              "hedft_dropin_quantum_run.f90", "hedft_dropin_tddft.f90",
              "hedft_dropin_quantum.f90", "hedft_dropin_statics.f90", "hedft_dropin_tddft_run.f90",
              "hedft_dropin_statics_run.f90", "Makefile"):
        assert os.path.exists(os.path.join(FDIR, f)), f


@pytest.mark.gpu
def test_fortran_driver_matches_python_binding(gpu, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("Fortran check program not built (amdflang absent at build time)")
    n, xmax, n4, niter, nrk = 32, 6.4, 60.0, 6, 3
    dt = 5e-4 / I.PSTOK
    print("[case] fortran drop-in 32^3 3 RK + 3 PC", flush=True)
    den = I.woods_saxon(n, n, n, xmax, xmax, xmax, n4, seed=4)
    potion, delinter, rmax = I.impurity_table(n, xmax)
    mass = 108.0 * I.MP_U
    r0 = np.array([0.0, 0.0, -4.0])
    v0 = np.array([0.0, 0.0, 3.0]) * I.PSTOK
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([n, niter, nrk], dtype=np.int32).tofile(f)
        np.array([xmax, n4, dt, delinter, rmax, mass], dtype=np.float64).tofile(f)
        np.array([potion.size], dtype=np.int32).tofile(f)
        np.asarray(potion, dtype=np.float64).tofile(f)
        r0.tofile(f)
        v0.tofile(f)
        np.ascontiguousarray(den, dtype=np.float64).tofile(f)
    r = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(fout, dtype=np.float64)
    err_f, etot_f, ekin_f, norm_f = raw[:4]
    psi_f = raw[4:].view(np.complex128).reshape(n, n, n)

    d = gpu.Droplet(n, n, n, xmax, xmax, xmax, gpu.default_params(core_otc=1), mode=gpu.MODE_TDDFT)
    d.set("DEN", den)
    d.set_impurities(potion[None, :], delinter, rmax, [mass])
    d.set_impurity_state(r0[None, :], v0[None, :], np.zeros((1, 3)))
    d.tddft_init()
    for it in range(1, niter + 1):
        e = d.tddft_iteration(dt, it, nrk)
    en = d.energy()
    psi = d.field("PSI")
    d.close()
    assert err_f == e[0]
    assert etot_f == en["etot"] and ekin_f == en["ekin4"] and norm_f == en["norm"]
    assert np.array_equal(psi_f, psi)


@pytest.mark.gpu
def test_fortran_quantum_loop_matches_python_binding(gpu, tmp_path):
    """The quantum-impurity loop driven from Fortran equals the Python binding's,
    bit for bit (same entry points, same order)."""
    if not os.path.exists(QEXE):
        pytest.skip("Fortran check program not built (amdflang absent at build time)")
    from he_dft_2024_amd import quantum as Q
    n, xmax, n4, niter, nrk = 32, 6.4, 60.0, 6, 3
    dt, m_imp, mu4 = 5e-4 / I.PSTOK, 2 * 4.002603, -0.7
    print("[case] fortran quantum loop 32^3 3 RK + 3 PC", flush=True)
    v = Q.lennard_jones(10.22, 2.556)
    uk = Q.uimp_k_table(n, n, n, xmax, xmax, xmax, v, 2.0, float(v(2.0)))
    den = I.woods_saxon(n, n, n, xmax, xmax, xmax, n4, seed=6)
    psix = I.gaussian_packet(n, n, n, xmax, xmax, xmax, r0=(0.5, -0.4, 1.0), width=1.3, k=(0.4, 0.0, 0.2))
    fin, fout = tmp_path / "qin.bin", tmp_path / "qout.bin"
    with open(fin, "wb") as f:
        np.array([n, niter, nrk], dtype=np.int32).tofile(f)
        np.array([xmax, dt, m_imp, mu4], dtype=np.float64).tofile(f)
        np.ascontiguousarray(uk, dtype=np.float64).tofile(f)
        np.ascontiguousarray(den, dtype=np.float64).tofile(f)
        np.ascontiguousarray(psix, dtype=np.complex128).tofile(f)
    r = subprocess.run([QEXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(fout, dtype=np.float64)
    scal = raw[:9]
    psix_f = raw[9:].view(np.complex128).reshape(n, n, n)

    d = gpu.Droplet(n, n, n, xmax, xmax, xmax, gpu.default_params(), mode=gpu.MODE_TDDFT)
    d.set_quantum_impurity(uk, m_imp, mu4)
    d.set("DEN", den)
    d.set("PSIX", psix)
    d.tddft_init()
    for it in range(1, niter + 1):
        e = d.tddft_iteration(dt, it, nrk)
    en = d.energy()
    nrm, rc = d.quantum_centroid()
    px = d.field("PSIX")
    d.close()
    assert scal[0] == e[0] and scal[1] == e[1]
    assert scal[2] == en["etot"] and scal[3] == en["ekinx"] and scal[4] == en["eHeX"]
    assert scal[5] == nrm and np.array_equal(scal[6:9], rc)
    assert np.array_equal(psix_f, px)


DEXE = os.path.join(FDIR, "_build", "hedft_dropin_tddft_run")
SEXE = os.path.join(FDIR, "_build", "hedft_dropin_statics_run")


@pytest.mark.gpu
def test_dropin_tddft_reference_names_match_python_sequence(gpu, tmp_path):
    """A driver-shaped Fortran program linked with hedft_dropin_tddft.o and the
    reference's own Dynamics_TDDFT/modules.f90 calls STEPRK / STEPPC / potenimp
    / poten / forceimp by their reference names in the order of
    DFT4He3d.f90:671-696 (aimp = F/m on the host).  The same entry points
    called from Python in that order give bit-identical results, and the
    graph-ready driver entry (hedft_tddft_iteration, force fused into the
    gradient pass) agrees to rounding."""
    if not os.path.exists(DEXE):
        pytest.skip("drop-in driver program not built (reference tree or amdflang absent at build time)")
    n, xmax, n4, niter = 32, 6.4, 60.0, 6
    dt = 5e-4 / I.PSTOK
    print("[case] fortran drop-in by reference names, 32^3 3 RK + 3 PC", flush=True)
    den = I.woods_saxon(n, n, n, xmax, xmax, xmax, n4, seed=4)
    potion, delinter, rmax = I.impurity_table(n, xmax)
    mass = 108.0 * I.MP_U
    r0 = np.array([0.0, 0.5, -4.0])
    v0 = np.array([0.0, 0.0, 3.0]) * I.PSTOK
    fin, fout = tmp_path / "din.bin", tmp_path / "dout.bin"
    with open(fin, "wb") as f:
        np.array([n, niter], dtype=np.int32).tofile(f)
        np.array([xmax, dt, delinter, rmax, mass], dtype=np.float64).tofile(f)
        np.array([potion.size], dtype=np.int32).tofile(f)
        np.asarray(potion, dtype=np.float64).tofile(f)
        r0.tofile(f)
        v0.tofile(f)
        np.ascontiguousarray(den, dtype=np.float64).tofile(f)
    r = subprocess.run([DEXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(fout, dtype=np.float64)
    err_f, etot_f, ekin_f, ehex_f = raw[:4]
    rimp_f, vimp_f, f_f = raw[4:7], raw[7:10], raw[10:13]
    psi_f = raw[13:].view(np.complex128).reshape(n, n, n)

    # the same calls from Python
    psi0 = np.sqrt(den).astype(np.complex128)
    d = gpu.Droplet(n, n, n, xmax, xmax, xmax, gpu.default_params(core_otc=1), mode=gpu.MODE_TDDFT)
    d.set("DEN", np.abs(psi0) ** 2)
    d.set("PSI", psi0)
    d.set_impurities(potion[None, :], delinter, rmax, [mass])
    d.set_impurity_state(r0[None, :], v0[None, :], np.zeros((1, 3)))

    def after_step():
        d.potenimp()
        d.poten()
        d.forceimp()
        return d.impurity_state()["force"] / mass
    aimp = after_step()
    err = 0.0
    for it in range(1, niter + 1):
        d.set_impurity_state(aimp=aimp)
        if it <= 3:
            d.steprk(dt)
        else:
            err = d.steppc(dt)[0]
        aimp = after_step()
    en = d.energy()
    st = d.impurity_state()
    assert err_f == err
    assert etot_f == en["etot"] and ekin_f == en["ekin4"] and ehex_f == en["eHeX"]
    assert np.array_equal(rimp_f, st["rimp"][0]) and np.array_equal(vimp_f, st["vimp"][0])
    assert np.array_equal(f_f, st["force"][0])
    assert np.array_equal(psi_f, d.field("PSI"))
    d.close()
    # the driver entry point with the fused force reduction
    e = gpu.Droplet(n, n, n, xmax, xmax, xmax, gpu.default_params(core_otc=1), mode=gpu.MODE_TDDFT)
    e.set("DEN", den)
    e.set_impurities(potion[None, :], delinter, rmax, [mass])
    e.set_impurity_state(r0[None, :], v0[None, :], np.zeros((1, 3)))
    e.tddft_init()
    for it in range(1, niter + 1):
        e.tddft_iteration(dt, it, 3)
    psi_e = e.field("PSI")
    assert np.abs(psi_e - psi_f).max() < 1e-12 * np.abs(psi_f).max()
    assert abs(e.energy()["etot"] - etot_f) < 1e-12 * abs(etot_f)
    e.close()


@pytest.mark.gpu
def test_dropin_statics_reference_names_match_python_sequence(gpu, tmp_path):
    """The He_statics drop-in (hedft_dropin_statics.o + the reference's own
    He_statics/modules.f90): STEPRKR / STEPPCR by name with the driver's refresh
    (BCN4HeDFT.f90:1113-1160), bit-identical to the same entry points from
    Python and to hedft_statics_iteration."""
    if not os.path.exists(SEXE):
        pytest.skip("drop-in driver program not built (reference tree or amdflang absent at build time)")
    n, xmax, n4, niter = 32, 6.4, 50, 10
    dt = 0.01 * 0.4 ** 2 / (4 * 6.05969638298)
    print("[case] fortran statics drop-in by reference names, 32^3 4 RK + 6 PC", flush=True)
    den = I.woods_saxon(n, n, n, xmax, xmax, xmax, float(n4), seed=3)
    fin, fout = tmp_path / "sin.bin", tmp_path / "sout.bin"
    with open(fin, "wb") as f:
        np.array([n, niter, n4], dtype=np.int32).tofile(f)
        np.array([xmax, dt], dtype=np.float64).tofile(f)
        np.ascontiguousarray(den, dtype=np.float64).tofile(f)
    r = subprocess.run([SEXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(fout, dtype=np.float64)
    mus_f = raw[:niter]
    errpc_f, etot_f, ekin_f = raw[niter:niter + 3]
    psi_f = raw[niter + 3:].reshape(n, n, n)

    d = gpu.Droplet(n, n, n, xmax, xmax, xmax, gpu.default_params(core_otc=1), mode=gpu.MODE_STATICS)
    d.set("DEN", den)
    mu = d.statics_init(float(n4))
    mus, epc = [], 0.0
    for it in range(1, niter + 1):
        if it <= 4:
            mu, _ = d.steprkr(dt, mu, float(n4))
        else:
            mu, _, epc = d.steppcr(dt, mu, float(n4))
        d.statics_refresh()
        mus.append(mu)
    en = d.energy()
    assert np.array_equal(mus_f, np.array(mus))
    assert errpc_f == epc and etot_f == en["etot"] and ekin_f == en["ekin4"]
    assert np.array_equal(psi_f, d.field("PSI"))
    d.close()
    s = gpu.Droplet(n, n, n, xmax, xmax, xmax, gpu.default_params(core_otc=1), mode=gpu.MODE_STATICS)
    s.set("DEN", den)
    mu = s.statics_init(float(n4))
    for it in range(1, niter + 1):
        mu, _ = s.statics_iteration(dt, it, float(n4), mu)
    assert mu == mus_f[-1]
    assert np.array_equal(psi_f, s.field("PSI"))
    s.close()


QDEXE = os.path.join(FDIR, "_build", "hedft_dropin_quantum_run")


@pytest.mark.gpu
def test_dropin_quantum_reference_names_match_python_sequence(gpu, tmp_path):
    """The quantum-impurity drop-in (hedft_dropin_quantum.o + the reference's own
    Dynamics_quantum_impurities/modules.f90) with N_imp = 2: uimp_k_k(:,:,:,1,2)
    set, Selec_gs_k_k(2,1) = 'null'.  STEPRK / STEPPC / potenimp / poten by
    their reference names in the order of DFT4He3d.f90:667-677 are bit-identical
    to the same entry points from Python."""
    if not os.path.exists(QDEXE):
        pytest.skip("drop-in driver program not built (reference tree or amdflang absent at build time)")
    from he_dft_2024_amd import quantum as Q
    n, xmax, n4, niter, nq, mu4 = 32, 6.4, 60.0, 6, 2, -1.5
    dt = 5e-4 / I.PSTOK
    print("[case] fortran quantum drop-in by reference names, 2 impurities, 32^3 3 RK + 3 PC", flush=True)
    den = I.woods_saxon(n, n, n, xmax, xmax, xmax, n4, seed=4)
    v, vxx = Q.lennard_jones(10.22, 2.556), Q.lennard_jones(20.0, 3.0)
    uk = Q.uimp_k_table(n, n, n, xmax, xmax, xmax, v, 2.0, float(v(2.0)))
    ukk = Q.uimp_k_table(n, n, n, xmax, xmax, xmax, vxx, 2.5, float(vxx(2.5)))
    masses = np.array([2 * 4.002603, 20.18])
    psix = [I.gaussian_packet(n, n, n, xmax, xmax, xmax, r0=(0.4, -0.8, 1.2), width=1.3, k=(0.6, 0.0, -0.3)),
            I.gaussian_packet(n, n, n, xmax, xmax, xmax, r0=(-1.5, 0.9, -0.6), width=1.1, k=(0.0, -0.4, 0.5))]
    pair = np.zeros((nq, nq), dtype=np.int32)       # pair[j, i] = Fortran pair(i, j)
    pair[1, 0] = 1                                    # uimp_k_k(:,:,:,1,2): impurity 2 acts on impurity 1
    ukk_f = np.zeros((nq, nq) + uk.shape)             # [j, i] = uimp_k_k(:,:,:,i,j)
    ukk_f[1, 0] = ukk
    fin, fout = tmp_path / "qin.bin", tmp_path / "qout.bin"
    with open(fin, "wb") as f:
        np.array([n, niter, nq], dtype=np.int32).tofile(f)
        np.array([xmax, dt, mu4], dtype=np.float64).tofile(f)
        masses.tofile(f)
        pair.tofile(f)
        np.stack([uk, 0.7 * uk]).tofile(f)
        ukk_f.tofile(f)
        np.ascontiguousarray(den, dtype=np.float64).tofile(f)
        np.stack(psix).astype(np.complex128).tofile(f)
    r = subprocess.run([QDEXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(fout, dtype=np.float64)
    scal = raw[:6]
    off, nn = 6, n ** 3
    psi_f = raw[off:off + 2 * nn].view(np.complex128).reshape(n, n, n)
    off += 2 * nn
    psix_f = raw[off:off + 2 * nq * nn].view(np.complex128).reshape(nq, n, n, n)
    off += 2 * nq * nn
    upotx_f = raw[off:off + nq * nn].reshape(nq, n, n, n)
    potx4_f = raw[off + nq * nn:].reshape(n, n, n)

    psi0 = np.sqrt(den).astype(np.complex128)
    d = gpu.Droplet(n, n, n, xmax, xmax, xmax, gpu.default_params(core_otc=1), mode=gpu.MODE_TDDFT)
    d.set_quantum_impurities([uk, 0.7 * uk], masses, mu4=mu4, uimp_kk={(0, 1): ukk})
    d.set("DEN", np.abs(psi0) ** 2)
    d.set("PSI", psi0)
    for k in range(nq):
        d.select_quantum(k)
        d.set("PSIX", psix[k])
    d.select_quantum(0)
    d.potenimp()
    d.poten()
    err = (0.0, 0.0, 0.0)
    for it in range(1, niter + 1):
        if it <= 3:
            d.steprk(dt)
        else:
            err = d.steppc(dt)
        d.potenimp()
        d.poten()
    en = d.energy()
    assert scal[0] == err[0] and scal[1] == err[1]
    assert abs(en["eimpu_impu"]) > 0
    for i, k in enumerate(("etot", "ekinx", "eHeX", "eimpu_impu")):
        assert scal[2 + i] == en[k], k
    assert np.array_equal(psi_f, d.field("PSI"))
    assert np.array_equal(potx4_f, d.field("POTX4"))
    for k in range(nq):
        d.select_quantum(k)
        assert np.array_equal(psix_f[k], d.field("PSIX")), k
        assert np.array_equal(upotx_f[k], d.field("UPOTX")), k
    d.close()


SEXE2 = os.path.join(FDIR, "_build", "hedft_fortran_scheck")


@pytest.mark.gpu
def test_fortran_statics_runs_match_python_binding(gpu, tmp_path):
    """Module hedft_gpu's hedft_statics_run / hedft_get_tuning /
    hedft_last_elapsed from Fortran: batches of k statics iterations give the
    same mu4, mu4err, errpc and psi as Droplet.statics_run, bit for bit."""
    if not os.path.exists(SEXE2):
        pytest.skip("Fortran check program not built (amdflang absent at build time)")
    n, xmax, n4, nbatch, k = 32, 6.4, 50.0, 4, 7
    dt = 0.01 * 0.4 ** 2 / (4 * 6.05969638298)
    print("[case] fortran statics runs, 32^3 4 x 7 iterations", flush=True)
    den = I.woods_saxon(n, n, n, xmax, xmax, xmax, n4, seed=0)
    fin, fout = tmp_path / "sin.bin", tmp_path / "sout.bin"
    with open(fin, "wb") as f:
        np.array([n, nbatch, k], dtype=np.int32).tofile(f)
        np.array([xmax, dt, n4], dtype=np.float64).tofile(f)
        np.ascontiguousarray(den, dtype=np.float64).tofile(f)
    r = subprocess.run([SEXE2, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(fout, dtype=np.float64)
    hist_f, ok, psi_f = raw[:3 * nbatch].reshape(nbatch, 3), raw[3 * nbatch], raw[3 * nbatch + 1:]
    assert ok == 1.0
    d = gpu.Droplet(n, n, n, xmax, xmax, xmax, gpu.default_params(core_otc=1), mode=gpu.MODE_STATICS)
    d.set("DEN", den)
    d.statics_init(n4)
    hist = []
    for b in range(nbatch):
        m, e, pc, _ = d.statics_run(dt, 1 + b * k, k, n4)
        hist.append((m, e, pc))
    assert np.array_equal(hist_f, np.array(hist))
    assert np.array_equal(psi_f, d.field("PSI").ravel())
    d.close()
