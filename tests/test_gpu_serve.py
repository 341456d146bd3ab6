"""Serving loop (hedft_serve): jobs arriving over time are admitted by an HBM
budget, stay resident, overlap on the device, and each result is identical to
the same job run alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _jobs(gpu, count=6, seed=11):
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / 50.0, size=count))  # Poisson, 50 jobs/s
    jobs = []
    for k in range(count):
        n = 32 if k % 2 == 0 else 48
        jobs.append(gpu.make_job(arrival_s=float(arrivals[k]), n=n, n4=float(rng.uniform(40, 120)),
                                 nsteps=int(rng.choice([12, 18, 30])), nrk=3,
                                 precision=gpu.PREC_MIXED if k % 3 == 0 else gpu.PREC_FP64,
                                 r0=(float(rng.uniform(0, 1)), 0.0, -4.0),
                                 v0=(0.0, 0.0, float(rng.uniform(1, 8))), seed=100 + k))
    return jobs


def test_serve_mix_equals_alone(gpu):
    print("[case] serve 6 jobs mixed vs alone", flush=True)
    jobs = _jobs(gpu)
    big = sum(gpu.job_bytes(j) for j in jobs) * 2
    mix, st = gpu.serve(jobs, hbm_budget=big, chunk=6)
    assert st["completed"] == len(jobs)
    assert st["max_concurrent"] >= 2
    assert st["p99_ms_per_1000"] >= st["p50_ms_per_1000"] > 0
    for k, j in enumerate(jobs):
        j.arrival_s = 0.0
        alone, _ = gpu.serve([j], hbm_budget=big, chunk=6)
        a, m = alone[0], mix[k]
        assert a["status"] == 0 and m["status"] == 0
        for key in ("etot", "etot4", "ekin4", "norm"):
            assert a[key] == m[key], (k, key, a[key], m[key])
        assert a["rimp"] == m["rimp"]
        assert np.isfinite(m["etot"])


def test_serve_respects_hbm_budget(gpu):
    print("[case] serve budget", flush=True)
    jobs = _jobs(gpu, count=5, seed=12)
    for j in jobs:
        j.arrival_s = 0.0
        j.n = 32  # equal footprints: the budget holds two at a time
    per = max(gpu.job_bytes(j) for j in jobs)
    res, st = gpu.serve(jobs, hbm_budget=2.2 * per, chunk=6)
    assert st["completed"] == len(jobs)
    assert st["max_concurrent"] <= 2
    assert st["peak_bytes"] <= 2.2 * per
    # admission is in arrival order
    admits = [r["admit_s"] for r in res]
    assert admits == sorted(admits)


def test_serve_rejects_job_larger_than_budget(gpu):
    j = gpu.make_job(n=64)
    with pytest.raises(gpu.HedftError):
        gpu.serve([j], hbm_budget=gpu.job_bytes(j) / 2)
