#!/usr/bin/env python3
"""bench.py — real-time TDDFT Hamming-PC steps/s of one 4He droplet on one MI355X.

Workload (the headline configuration): N=1000 droplet on a 128^3 grid, h = 0.4 A
(box +-25.6 A), Orsay-Trento functional with the alpha_s term (core 'OTC',
lsolid off), 13-point stencils with icon=13, one classical impurity of 108 amu
with a Lennard-Jones He-X potential (eps 20 K, sigma 4 A) starting 20 A from the
centre at 300 m/s toward it, dt = 0.5 fs, 4 RK-Gill start steps, then timed
Hamming PC iterations (each = STEPPC + potenimp/poten/forceimp/aimp of the
driver, i.e. two mean-field evaluations).  FP64 throughout (the drop-in path).

A "step" is one PC iteration of one droplet; inputs are resident in HBM when
the timed region starts; time is measured with HIP events on the library's
stream around K graph-replayed iterations.  The graphs the timed run replays
are captured (host only) before the W warmup iterations, so that the warmup
iterations run on the GPU right before the timed region: a GPU left idle for
the capture starts the timed region at lower clocks (0.3 s idle cost ~4 % over
18 iterations).

Prints ONE JSON line (rank 0).  --gpus N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json

import numpy as np
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=24)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--n4", type=float, default=1000.0)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--mixed-steps", type=int, default=600,
                    help="PC iterations timed on the mixed-precision path (0 skips it)")
    ap.add_argument("--batch", type=int, default=48,
                    help="droplets in the mixed-precision batched ensemble (0 skips it)")
    ap.add_argument("--batch-steps", type=int, default=60, help="PC iterations per droplet in the batched leg")
    ap.add_argument("--serve-jobs", type=int, default=24, help="jobs in the serving leg (0 skips it)")
    ap.add_argument("--large-steps", type=int, default=48,
                    help="PC iterations of the 256^3 N=5000 mixed-precision droplet (0 skips it)")
    ap.add_argument("--statics-steps", type=int, default=200,
                    help="imaginary-time iterations of the 64^3 N=300 relaxation (0 skips it)")
    ap.add_argument("--serve-steps", type=int, default=1000, help="PC iterations per serving job")
    ap.add_argument("--ragged-steps", type=int, default=24,
                    help="PC iterations per droplet of the ragged (128/192) batched ensemble (0 skips it)")
    ap.add_argument("--accuracy-steps", type=int, default=1002,
                    help="PC iterations after which the mixed path is compared with the FP64 path")
    ap.add_argument("--cpu-statics-steps", type=int, default=40,
                    help="statics iterations of the 64^3 CPU sample (PC iterations 11.. are timed)")
    ap.add_argument("--quantum-steps", type=int, default=48,
                    help="PC iterations in the 256^3 quantum-impurity leg (0 skips it)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-steps", type=int, default=12, help="PC iterations in the CPU sample")
    ap.add_argument("--probe-steps", type=int, default=24,
                    help="eager PC iterations timed per kernel for the roofline object")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a plain 512 MiB device copy (counter calibration) and exit")
    return ap.parse_args()


def algorithmic_bytes_per_pc_step(n: int, w: int = 8) -> float:
    """Algorithmic bytes of one Hamming PC iteration of one droplet (OTC):
    2 mean-field evaluations of 14 3-D FFTs each + predictor + corrector.
    C_fft = w N + 3 (2w) Nh; B_poten = 14 C_fft + 8 w Nh + 20 w N;
    B_pred = B_corr = 8 (2w) N + 2 w N."""
    N = n ** 3
    Nh = (n // 2 + 1) * n * n
    cfft = w * N + 3 * (2 * w) * Nh
    b_poten = 14 * cfft + 8 * w * Nh + 20 * w * N
    b_pc = 8 * (2 * w) * N + 2 * w * N
    return 2 * b_poten + 2 * b_pc


def bytes_fused_step_kernel(n: int, w: int = 8) -> float:
    """Algorithmic bytes of one fused H psi + predictor (or corrector) launch:
    B_pred = B_corr = 8 (2w) N + 2 w N (complex psi history reads/writes, pot4
    in, den out).  The kernel is the 2.5-D pass (x, y and z stencils of psi
    and the update together); its tile halos are re-reads, not algorithmic
    traffic."""
    N = n ** 3
    return 8 * (2 * w) * N + 2 * w * N


def load_traffic(kernel: str, n: int):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc
    (FETCH_SIZE and WRITE_SIZE, scaled by a plain copy's counters), as committed
    next to the profiles; None when no such measurement exists."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t.get(f"{kernel}@{n}")
    except (OSError, ValueError):
        return None


def copy_calibration():
    import torch
    x = torch.ones(64 << 20, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": x.numel() * 8, "copy_bytes_written": y.numel() * 8}), flush=True)


def main():
    a = parse()
    if a.copy_calib:
        copy_calibration()
        return
    if a.gpus != 1:
        print("bench.py: this build runs one droplet per GPU on a single GPU; "
              f"--gpus {a.gpus} is not supported (the multi-GPU design is documented, not built).",
              file=sys.stderr)
        sys.exit(2)
    import he_dft_2024_amd as H
    from he_dft_2024_amd import inputs as I

    n = a.grid
    xmax = 0.2 * n          # h = 0.4 A
    dt = 5e-4 / I.PSTOK     # 0.5 fs
    p = H.default_params(core_otc=1, lsolid=0)
    drop = H.Droplet(n, n, n, xmax, xmax, xmax, p, mode=H.MODE_TDDFT)
    I.tddft_setup(drop, n, xmax, a.n4, seed=1, r0=(0.0, 0.0, -20.0), v_aps=(0.0, 0.0, 3.0))
    drop.tddft_init()
    nrk = 4
    for it in range(1, nrk + 1):
        drop.tddft_iteration(dt, it, nrk)
    use_graph = not a.no_graph
    if use_graph:
        drop.prepare_pc(dt, a.warmup, a.steps)
    if a.warmup:
        drop.run_pc(dt, a.warmup, use_graph)
    ms = drop.run_pc(dt, a.steps, use_graph)
    steps_per_s = a.steps / (ms / 1000.0)
    e = drop.energy()
    if not np.isfinite(e["etot"]):
        print("bench.py: non-finite energy after the timed steps", file=sys.stderr)
        sys.exit(3)

    bytes_step = algorithmic_bytes_per_pc_step(n)
    step_gbs = bytes_step * steps_per_s / 1e9

    # dominant kernel: the fused H psi + predictor / corrector pass, timed
    # live with HIP events on the library's main stream over real PC steps
    roofline = None
    if a.probe_steps > 0:
        ms_pred, ms_corr = drop.probe_pc(dt, a.probe_steps)
        kname, kms = ("fused_hpsi_corrector", ms_corr) if ms_corr >= ms_pred else ("fused_hpsi_predictor", ms_pred)
        kb = bytes_fused_step_kernel(n)
        ach = kb / (kms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": load_traffic(kname, n),
                    "kernel": kname, "avg_ms": round(kms, 5), "algorithmic_bytes": kb,
                    "avg_ms_pred": round(ms_pred, 5), "avg_ms_corr": round(ms_corr, 5),
                    "launches": a.probe_steps}
        e = drop.energy()

    # mixed-precision path on the same workload: same RK start, warmup, timed
    # and probe iteration counts, so its energy per atom is compared with the
    # FP64 path's after the same number of steps
    mixed = None
    if a.mixed_steps > 0:
        pm = H.default_params(core_otc=1, lsolid=0, precision=H.PREC_MIXED)
        dm = H.Droplet(n, n, n, xmax, xmax, xmax, pm, mode=H.MODE_TDDFT)
        I.tddft_setup(dm, n, xmax, a.n4, seed=1, r0=(0.0, 0.0, -20.0), v_aps=(0.0, 0.0, 3.0))
        dm.tddft_init()
        for it in range(1, nrk + 1):
            dm.tddft_iteration(dt, it, nrk)
        if use_graph:
            dm.prepare_pc(dt, a.warmup, a.mixed_steps)
        if a.warmup:
            dm.run_pc(dt, a.warmup, use_graph)
        msm = dm.run_pc(dt, a.mixed_steps, use_graph)
        # accuracy: both paths advanced to the same fixed number of PC
        # iterations (at least --accuracy-steps), then compared per energy term
        # and by the density's relative L2 error
        done_fp = a.warmup + a.steps + a.probe_steps
        done_mx = a.warmup + a.mixed_steps
        horizon = max(a.accuracy_steps, done_fp, done_mx)
        if horizon > done_fp:
            drop.run_pc(dt, horizon - done_fp, use_graph)
        if horizon > done_mx:
            dm.run_pc(dt, horizon - done_mx, use_graph)
        e = drop.energy()
        em = dm.energy()
        den_fp, den_mx = drop.field("DEN"), dm.field("DEN")
        dm.close()
        terms = {k: abs(em[k] - e[k]) / abs(e[k]) for k in ("ekin4", "elj4", "ecor4", "ealphas", "eHeX", "ekinx",
                                                            "etot4", "etot", "norm") if abs(e[k]) > 0}
        mixed = {"value": round(a.mixed_steps / (msm / 1000.0), 3), "unit": "PC steps/s",
                 "ms_per_step": round(msm / a.mixed_steps, 4), "steps": a.mixed_steps, "dtype": "f32 field / f64 psi",
                 "speedup_vs_fp64": round((a.mixed_steps / msm) / (a.steps / ms), 3),
                 "pc_steps_compared": horizon,
                 "energy_per_atom_rel_err_vs_fp64": abs(em["etot"] - e["etot"]) / abs(e["etot"]),
                 "term_rel_err_vs_fp64": {k: float("%.3e" % v) for k, v in terms.items()},
                 "density_l2_rel_err_vs_fp64": float(np.linalg.norm(den_mx - den_fp) / np.linalg.norm(den_fp))}

    batched = None
    if a.batch > 0:
        batched = batched_leg(H, I, a, dt, steps_per_s)
    ragged = None
    if a.batch > 0 and a.ragged_steps > 0:
        ragged = ragged_leg(H, I, a, dt)

    serving = None
    if a.serve_jobs > 0:
        serving = serving_leg(H, I, a, dt)

    large = None
    if a.large_steps > 0:
        large = large_leg(H, I, a, dt)

    statics = None
    if a.statics_steps > 0:
        statics = statics_leg(H, I, a)

    quantum = None
    if a.quantum_steps > 0:
        quantum = quantum_leg(H, I, a)

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(n, xmax, a.n4, dt, a.cpu_steps)
        if a.cpu_statics_steps > 10:
            cpu["statics_64"] = cpu_statics_baseline(a.cpu_statics_steps)
            if statics is not None:
                statics["speedup_vs_cpu"] = round(statics["value"] / cpu["statics_64"]["value"], 1)

    out = {
        "metric": "TDDFT steps/s at 128^3 (FP64 plain, one droplet)",
        "value": round(steps_per_s, 3),
        "unit": "PC steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms / a.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded Woods-Saxon droplet (not relaxed) + LJ impurity table",
        "config": {"workload": f"TDDFT Hamming-PC iteration, N={int(a.n4)} droplet, {n}^3 grid, h=0.4 A, "
                               "OTC, 1 classical LJ impurity 108 amu at 300 m/s",
                   "grid": n, "n4": a.n4, "dt_fs": 0.5, "graph": use_graph},
        "step_algorithmic_GBps": round(step_gbs, 1),
        "step_roofline_frac": round(step_gbs / HBM_PEAK_GBS, 4),
        "step_algorithmic_bytes": bytes_step,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "energy_per_atom_K": e["etot4"] / a.n4,
        "mixed": mixed,
        "mixed_batched": batched,
        "mixed_batched_ragged": ragged,
        "serving": serving,
        "mixed_256": large,
        "statics_64": statics,
        "quantum_256": quantum,
    }
    print(json.dumps(out), flush=True)


def batched_leg(H, I, a, dt, plain_steps_per_s):
    """Seeded ensemble of a.batch droplets (N log-uniform 200-2000, impurity at
    1-8 A/ps with impact parameter 0-R, 128^3 grid), mixed precision, advanced
    together by hedft_batch_run_pc; aggregate PC steps/s over all droplets."""
    ens = I.ensemble(count=a.batch, seed=3, grid=a.grid)
    drops = []
    for k, e in enumerate(ens):
        g, xm = e["grid"], e["xmax"]
        d = H.Droplet(g, g, g, xm, xm, xm, H.default_params(core_otc=1, precision=H.PREC_MIXED))
        I.tddft_setup(d, g, xm, e["n4"], seed=e["seed"], r0=(e["b"], 0.0, -20.0), v_aps=(0.0, 0.0, e["v"]))
        d.tddft_init()
        for it in range(1, 5):
            d.tddft_iteration(dt, it, 4)
        drops.append(d)
    # autotune the launch structure on a scratch droplet of the same grid and
    # precision, then apply the choice to the whole ensemble
    g0 = ens[0]["grid"]
    scratch = H.Droplet(g0, g0, g0, ens[0]["xmax"], ens[0]["xmax"], ens[0]["xmax"],
                        H.default_params(core_otc=1, precision=H.PREC_MIXED))
    I.tddft_setup(scratch, g0, ens[0]["xmax"], ens[0]["n4"], seed=ens[0]["seed"], r0=(ens[0]["b"], 0.0, -20.0),
                  v_aps=(0.0, 0.0, ens[0]["v"]))
    scratch.tddft_init()
    for it in range(1, 5):
        scratch.tddft_iteration(dt, it, 4)
    tuning = scratch.autotune(dt, 12)
    scratch.close()
    for d in drops:
        d.set_tuning(tuning)
    H.batch_run_pc(drops, dt, 6, chunk=6)  # warmup (graph capture, caches)
    ms = H.batch_run_pc(drops, dt, a.batch_steps, chunk=6)
    agg = len(drops) * a.batch_steps / (ms / 1000.0)
    bad = sum(1 for d in drops if not np.isfinite(d.energy()["etot"]))
    for d in drops:
        d.close()
    return {"value": round(agg, 2), "unit": "PC steps/s (aggregate)", "n_droplets": len(drops),
            "steps_each": a.batch_steps, "ms": round(ms, 3), "dtype": "f32 field / f64 psi",
            "speedup_vs_fp64_plain": round(agg / plain_steps_per_s, 3),
            "clamped_droplets": sum(1 for e in ens if e["clamped"]), "nonfinite": bad,
            "autotuned": {k: tuning[k] for k in ("xy_split_fwd", "xy_split_inv", "no_side", "prefer_eager",
                                                 "lap_late")}}


def ragged_leg(H, I, a, dt):
    """The ensemble with uneven grids: the same seeded draws, each droplet on the
    smallest of {128, 192} points that holds 2R + 16 A (19 at 128^3, 29 at 192^3,
    none truncated), mixed precision, advanced together by hedft_batch_run_pc.
    Reports droplet PC steps/s and the work-weighted rate in 128^3-droplet
    steps/s (each droplet weighted by its grid volume / 128^3)."""
    ens = I.ensemble(count=a.batch, seed=3, grids=(128, 192))
    drops = []
    for e in ens:
        g, xm = e["grid"], e["xmax"]
        d = H.Droplet(g, g, g, xm, xm, xm, H.default_params(core_otc=1, precision=H.PREC_MIXED))
        I.tddft_setup(d, g, xm, e["n4"], seed=e["seed"], r0=(e["b"], 0.0, -20.0), v_aps=(0.0, 0.0, e["v"]))
        d.tddft_init()
        for it in range(1, 5):
            d.tddft_iteration(dt, it, 4)
        drops.append(d)
    H.batch_run_pc(drops, dt, 6, chunk=6)
    ms = H.batch_run_pc(drops, dt, a.ragged_steps, chunk=6)
    agg = len(drops) * a.ragged_steps / (ms / 1000.0)
    w = sum((e["grid"] / 128.0) ** 3 for e in ens) / len(ens)
    bad = sum(1 for d in drops if not np.isfinite(d.energy()["etot"]))
    for d in drops:
        d.close()
    return {"value": round(agg, 2), "unit": "PC steps/s (aggregate, droplets of 128^3 and 192^3)",
            "n_droplets": len(drops), "grids": {str(g): sum(1 for e in ens if e["grid"] == g) for g in (128, 192)},
            "steps_each": a.ragged_steps, "ms": round(ms, 3),
            "work_weighted_128_equiv_steps_per_s": round(agg * w, 2), "clamped_droplets": 0, "nonfinite": bad}


def serving_leg(H, I, a, dt):
    """Continuous-batching loop: Poisson arrivals at 20 jobs/s (seed 4) of
    droplets drawn like the ensemble (N 200-2000, 1-8 A/ps, impact parameter
    0-R, 128^3, mixed precision), 10% of them with a quantum impurity,
    admitted by a 64 GB HBM budget, each run for a.serve_steps PC iterations;
    latency percentiles per 1000 steps."""
    rng = np.random.default_rng(4)
    arr = np.cumsum(rng.exponential(1.0 / 20.0, size=a.serve_jobs))
    quantum = np.zeros(a.serve_jobs, dtype=bool)  # 10% of the jobs, chosen by the same seeded stream
    quantum[rng.choice(a.serve_jobs, size=max(1, round(0.1 * a.serve_jobs)), replace=False)] = True
    ens = I.ensemble(count=a.serve_jobs, seed=4, grid=a.grid)
    # quantum-impurity jobs: a 4 u wave packet (width 1.5 A) through the same LJ pair
    # potential, at the reference quantum example's step of 5e-5 ps
    jobs = [H.make_job(arrival_s=float(t), n=e["grid"], h=0.4, n4=e["n4"], nsteps=a.serve_steps, nrk=4,
                       dt=(5e-5 / I.PSTOK) if q else dt, precision=H.PREC_MIXED, r0=(e["b"], 0.0, -20.0),
                       v0=(0.0, 0.0, e["v"]), seed=e["seed"], quantum=bool(q), mass_amu=4.0 if q else 108.0,
                       width=1.5)
            for t, e, q in zip(arr, ens, quantum)]
    res, st = H.serve(jobs, hbm_budget=64e9, chunk=12)
    ok = sum(1 for r in res if r["status"] == 0 and np.isfinite(r["etot"]))
    # time from admission to completion (context creation, RK start and the PC
    # iterations), per 1000 PC iterations, beside the arrival-to-completion latency
    svc = np.array([(r["done_s"] - r["admit_s"]) * 1000.0 / a.serve_steps * 1000.0 for r in res
                    if r["status"] == 0])
    first, last = min(r["admit_s"] for r in res), max(r["done_s"] for r in res)
    return {"jobs": a.serve_jobs, "quantum_jobs": int(quantum.sum()), "steps_each": a.serve_steps,
            "arrival_rate_per_s": 20.0, "completed_ok": ok,
            "p50_ms_per_1000_steps": round(st["p50_ms_per_1000"], 2),
            "p99_ms_per_1000_steps": round(st["p99_ms_per_1000"], 2),
            "p50_service_ms_per_1000_steps": round(float(np.percentile(svc, 50)), 2),
            "p99_service_ms_per_1000_steps": round(float(np.percentile(svc, 99)), 2),
            "aggregate_steps_per_s": round(st["steps_per_s"], 2),
            "aggregate_steps_per_s_first_admit_to_last_done": round(ok * a.serve_steps / (last - first), 2),
            "max_concurrent": st["max_concurrent"], "wall_s": round(st["wall_s"], 3)}


def large_leg(H, I, a, dt):
    """Single large droplet: N=5000 on 256^3 (h=0.4 A), impurity at 5 A/ps,
    mixed precision; PC steps/s and its algorithmic bandwidth."""
    n = 256
    xm = 0.5 * n * 0.4
    res = {}
    for prec in (H.PREC_MIXED, H.PREC_FP64):
        d = H.Droplet(n, n, n, xm, xm, xm, H.default_params(core_otc=1, precision=prec))
        I.tddft_setup(d, n, xm, 5000.0, seed=7, r0=(0.0, 0.0, -20.0), v_aps=(0.0, 0.0, 5.0))
        d.tddft_init()
        for it in range(1, 5):
            d.tddft_iteration(dt, it, 4)
        d.run_pc(dt, 6, True)
        ms = d.run_pc(dt, a.large_steps, True)
        res[prec] = (ms, d.energy(), d.field("DEN"))
        d.close()
    ms, e, den = res[H.PREC_MIXED]
    ms64, e64, den64 = res[H.PREC_FP64]
    sps = a.large_steps / (ms / 1000.0)
    N, Nh = n ** 3, (n // 2 + 1) * n * n
    b = 544 * N + 736 * Nh  # mixed: FP32 field, FP64 psi (same counting convention)
    terms = {k: float("%.3e" % (abs(e[k] - e64[k]) / abs(e64[k])))
             for k in ("ekin4", "elj4", "ecor4", "ealphas", "eHeX", "ekinx", "etot", "norm") if abs(e64[k]) > 0}
    return {"value": round(sps, 3), "unit": "PC steps/s", "grid": n, "n4": 5000.0, "steps": a.large_steps,
            "algorithmic_GBps": round(b * sps / 1e9, 1), "energy_per_atom_K": e["etot4"] / 5000.0,
            "norm": e["norm"], "fp64_steps_per_s": round(a.large_steps / (ms64 / 1000.0), 3),
            "term_rel_err_vs_fp64": terms,
            "density_l2_rel_err_vs_fp64": float(np.linalg.norm(den - den64) / np.linalg.norm(den64)),
            "pc_steps_compared": 6 + a.large_steps}


def statics_leg(H, I, a):
    """Imaginary-time relaxation: N=300 on 64^3 (h=0.4 A), OTC, 4 RK-Gill +
    Hamming PC iterations with the local time step, dtau = 1e-4 (FP64)."""
    n, xm, n4 = 64, 12.8, 300.0
    d = H.Droplet(n, n, n, xm, xm, xm, H.default_params(core_otc=1), mode=H.MODE_STATICS)
    d.set("DEN", I.woods_saxon(n, n, n, xm, xm, xm, n4, seed=0))
    mu = d.statics_init(n4)
    t0 = time.perf_counter()
    for it in range(1, a.statics_steps + 1):
        mu = d.statics_iteration(1e-4, it, n4, mu)[0]
    el = time.perf_counter() - t0
    e = d.energy()
    d.close()
    out = {"value": round(a.statics_steps / el, 2), "unit": "iterations/s (host-timed, syncs each step)",
           "grid": n, "n4": n4, "steps": a.statics_steps, "E_per_N_K": e["etot"] / n4, "mu_K": mu,
           "norm": e["norm"]}
    # the same iterations queued without host syncs, PC iterations replayed
    # from graphs (hedft_statics_run): RK start and 6 warmup iterations, then
    # the rest timed with HIP events; the end state equals the loop above
    if a.statics_steps > 16:
        d = H.Droplet(n, n, n, xm, xm, xm, H.default_params(core_otc=1), mode=H.MODE_STATICS)
        d.set("DEN", I.woods_saxon(n, n, n, xm, xm, xm, n4, seed=0))
        d.statics_init(n4)
        d.statics_run(1e-4, 1, 10, n4)
        k = a.statics_steps - 10
        mu_g, _, _, ms = d.statics_run(1e-4, 11, k, n4)
        eg = d.energy()
        d.close()
        out["graph"] = {"value": round(k / (ms / 1000.0), 2), "unit": "iterations/s (graph replay, HIP events)",
                        "iterations_timed": k, "same_end_state": bool(eg["etot"] == e["etot"] and mu_g == mu)}
    return out


def quantum_leg(H, I, a):
    """Quantum-impurity TDDFT at the size of the reference's own example
    (Dynamics_quantum_impurities/input_files/input.0: 256^3, +-32 A, N=1000,
    Lsolid so core OT, icon 24, t*mean 30 A, deltat 5e-5 ps, impurity mass
    4.002603 u), FP64.  The X-He kernel is a Lennard-Jones stand-in (eps 1 K,
    sigma 5.3 A, capped at 3.17502 A) transformed by he_dft_2024_amd.quantum; the
    reference's He2S_fci table is not shipped.  PC steps/s over graph replays."""
    from he_dft_2024_amd import quantum as Q
    n, xm, n4 = 256, 32.0, 1000.0
    dt = 5e-5 / I.PSTOK
    p = H.default_params(core_otc=0, lsolid=1, icon=24, tmean=(30.0, 30.0, 30.0))
    d = H.Droplet(n, n, n, xm, xm, xm, p, mode=H.MODE_TDDFT)
    v = Q.lennard_jones(1.0, 5.3)
    t0 = time.perf_counter()
    uk = Q.uimp_k_table(n, n, n, xm, xm, xm, v, 3.17502, float(v(3.17502)))
    t_table = time.perf_counter() - t0
    I.quantum_setup(d, n, xm, n4, uk, seed=11, m_imp_u=4.002603, width=2.0)
    d.tddft_init()
    for it in range(1, 4):
        d.tddft_iteration(dt, it, 3)
    d.run_pc(dt, 6, True)
    ms = d.run_pc(dt, a.quantum_steps, True)
    e = d.energy()
    nrm, rc = d.quantum_centroid()
    d.close()
    sps = a.quantum_steps / (ms / 1000.0)
    return {"value": round(sps, 3), "unit": "PC steps/s", "grid": n, "n4": n4, "steps": a.quantum_steps,
            "ms_per_step": round(ms / a.quantum_steps, 3), "dtype": "f64", "uimp_k_table_s": round(t_table, 2),
            "etot_K": e["etot"], "ekinx_K": e["ekinx"], "eHeX_K": e["eHeX"], "psix_norm": nrm,
            "psix_centre_A": [round(x, 6) for x in rc]}


def cpu_baseline(n, xmax, n4, dt, nsteps):
    """The FP64 oracle (C restatement, OpenMP) timed on the host on a bounded
    sample: nsteps PC iterations of the same workload after its RK start."""
    from oracle import oracle as O
    from he_dft_2024_amd import inputs as I
    o = O.Oracle(n, n, n, xmax, xmax, xmax, O.default_params(core_otc=1), O.MODE_TDDFT)
    I.tddft_setup(o, n, xmax, n4, seed=1, r0=(0.0, 0.0, -20.0), v_aps=(0.0, 0.0, 3.0))
    o.tddft_init()
    for it in range(1, 5):
        o.tddft_iteration(dt, it, 4)
    t0 = time.perf_counter()
    for it in range(5, 5 + nsteps):
        o.tddft_iteration(dt, it, 4)
    el = time.perf_counter() - t0
    return {"value": round(nsteps / el, 5), "unit": "PC steps/s", "cores": O.num_threads(),
            "kind": "port", "sample": f"{nsteps} PC iterations of the same {n}^3 workload after a 4-step "
                                      f"RK start (oracle, OpenMP, own mixed-radix FFT)"}


def cpu_statics_baseline(nsteps):
    """The FP64 oracle on the 64^3 N=300 imaginary-time relaxation (4 RK-Gill +
    Hamming PC, dtau = 1e-4), timed over PC iterations 11..nsteps on the host."""
    from oracle import oracle as O
    from he_dft_2024_amd import inputs as I
    n, xm, n4 = 64, 12.8, 300.0
    o = O.Oracle(n, n, n, xm, xm, xm, O.default_params(core_otc=1), O.MODE_STATICS)
    o.set("DEN", I.woods_saxon(n, n, n, xm, xm, xm, n4, seed=0))
    mu = o.statics_init(n4)
    for it in range(1, 11):
        mu = o.statics_iteration(1e-4, it, n4, mu)[0]
    t0 = time.perf_counter()
    for it in range(11, nsteps + 1):
        mu = o.statics_iteration(1e-4, it, n4, mu)[0]
    el = time.perf_counter() - t0
    return {"value": round((nsteps - 10) / el, 3), "unit": "iterations/s", "cores": O.num_threads(),
            "kind": "port", "sample": f"statics iterations 11..{nsteps} of the 64^3 N=300 relaxation (oracle)"}


if __name__ == "__main__":
    main()
