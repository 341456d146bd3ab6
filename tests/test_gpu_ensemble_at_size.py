"""The batched and serving patterns at the sizes of their workloads:

* the 48-droplet seeded ensemble (seed 3, N log-uniform 200-2000, impurity at
  1-8 A/ps, impact parameter 0-R) on 128^3 grids, mixed precision: after the RK
  start and 6 batched PC iterations every droplet equals its run alone, bit for
  bit; one droplet is compared with the FP64 oracle per energy term;
* the ragged variant (each droplet on the smallest of 128^3 / 192^3 that holds
  it): batch equals alone;
* the serving loop with 24 jobs drawn like the ensemble (seed 4), 10 % of them
  quantum-impurity jobs, Poisson arrivals at 20 jobs/s, 128^3 mixed: each
  result equals the job run alone."""
import numpy as np
import pytest

from oracle import oracle as O
from he_dft_2024_amd import inputs as I

pytestmark = pytest.mark.gpu

DT = 5e-4 / I.PSTOK


def _droplet(gpu, e, prec):
    g, xm = e["grid"], e["xmax"]
    d = gpu.Droplet(g, g, g, xm, xm, xm, gpu.default_params(core_otc=1, precision=prec))
    I.tddft_setup(d, g, xm, e["n4"], seed=e["seed"], r0=(e["b"], 0.0, -20.0), v_aps=(0.0, 0.0, e["v"]))
    d.tddft_init()
    for it in range(1, 5):
        d.tddft_iteration(DT, it, 4)
    return d


def _batch_vs_alone(gpu, ens, nsteps=6):
    alone = []
    for k, e in enumerate(ens):
        d = _droplet(gpu, e, gpu.PREC_MIXED)
        d.run_pc(DT, nsteps, True)
        alone.append((d.field("PSI"), d.energy()["etot"]))
        d.close()
    ds = [_droplet(gpu, e, gpu.PREC_MIXED) for e in ens]
    gpu.batch_run_pc(ds, DT, nsteps, chunk=6)
    for k, (d, (psi, etot)) in enumerate(zip(ds, alone)):
        assert np.array_equal(d.field("PSI"), psi), k
        assert d.energy()["etot"] == etot, k
    return ds


def test_c3_ensemble_48_at_128_batch_equals_alone_and_oracle(gpu):
    print("[case] 48-droplet 128^3 mixed ensemble: batch vs alone", flush=True)
    ens = I.ensemble(count=48, seed=3, grid=128)
    ds = _batch_vs_alone(gpu, ens)
    # one droplet of the ensemble against the FP64 oracle (same 4 RK + 6 PC)
    k = int(np.argmin([abs(e["n4"] - 800.0) for e in ens]))
    e = ens[k]
    print("[case] ensemble droplet", k, "N =", round(e["n4"], 1), "vs oracle", flush=True)
    g, xm = e["grid"], e["xmax"]
    o = O.Oracle(g, g, g, xm, xm, xm, O.default_params(core_otc=1), O.MODE_TDDFT)
    I.tddft_setup(o, g, xm, e["n4"], seed=e["seed"], r0=(e["b"], 0.0, -20.0), v_aps=(0.0, 0.0, e["v"]))
    o.tddft_init()
    for it in range(1, 11):
        o.tddft_iteration(DT, it, 4)
    eg, eo = ds[k].energy(), o.energy()
    for key in ("ekin4", "elj4", "ecor4", "ealphas", "eHeX", "etot", "norm"):
        assert abs(eg[key] - eo[key]) <= 2e-6 * abs(eo[key]) + 1e-9, (key, eg[key], eo[key])
    for d in ds:
        d.close()


def test_ragged_ensemble_128_192_batch_equals_alone(gpu):
    print("[case] ragged 128/192 ensemble (every 4th droplet): batch vs alone", flush=True)
    ens = I.ensemble(count=48, seed=3, grids=(128, 192))[::4]
    assert {e["grid"] for e in ens} == {128, 192}
    for d in _batch_vs_alone(gpu, ens):
        d.close()


def test_c4_serving_24_jobs_with_quantum_mix_equals_alone(gpu):
    print("[case] serving 24 jobs at 128^3 (10% quantum): mix vs alone", flush=True)
    n = 24
    rng = np.random.default_rng(4)
    arr = np.cumsum(rng.exponential(1.0 / 20.0, size=n))
    quantum = np.zeros(n, dtype=bool)
    quantum[rng.choice(n, size=max(1, round(0.1 * n)), replace=False)] = True
    ens = I.ensemble(count=n, seed=4, grid=128)
    jobs = [gpu.make_job(arrival_s=float(t), n=128, h=0.4, n4=e["n4"], nsteps=12, nrk=4,
                         dt=(5e-5 / I.PSTOK) if q else DT, precision=gpu.PREC_MIXED, r0=(e["b"], 0.0, -20.0),
                         v0=(0.0, 0.0, e["v"]), seed=e["seed"], quantum=bool(q), mass_amu=4.0 if q else 108.0,
                         width=1.5)
            for t, e, q in zip(arr, ens, quantum)]
    mix, st = gpu.serve(jobs, hbm_budget=64e9, chunk=6)
    assert st["completed"] == n and st["max_concurrent"] >= 2
    for k, j in enumerate(jobs):
        j.arrival_s = 0.0
        alone, _ = gpu.serve([j], hbm_budget=64e9, chunk=6)
        a, m = alone[0], mix[k]
        assert a["status"] == 0 and m["status"] == 0, k
        for key in ("etot", "etot4", "ekin4", "norm"):
            assert a[key] == m[key], (k, key)
        assert a["rimp"] == m["rimp"], k
