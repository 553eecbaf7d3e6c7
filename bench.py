"""Benchmark of the RRGalaxSim hot path on one MI355X.

Headline workload (one JSON line on stdout): the plain PM path -- exponential
disk N = 1,000,000 (scale 1, sech^2 h = 0.1, v_c from the enclosed mass + 10%
dispersion, seed 1, R < 7.5), FP32 particles and grid, 128^3 cells over
[-8, 8]^3, CIC deposit, 200 damped-wave iterations per step (tolerance 0, so
no early exit), -grad phi, trilinear gather, dt = 5e-3; `--steps` steps timed
(50 by default).  value = N * steps / seconds, in Mparticle-steps/s.

Also reported (same line): the mixed-precision direct sum on the disk+bulge
workload (N = 262,144, eps = 0.005) and the plain FP64 direct-sum kernel at
the same N, as interactions/s; the dominant kernel's roofline; and the CPU
oracle timed on a bounded sample of the headline workload.

python bench.py [--gpus 1] [--steps 50] [--warmup 3] [--no-direct] [--no-cpu] [--skip legs]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "N-body Ginteractions/s, PM Mparticle-steps/s at stated err; adjoint/fwd time"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
C = 4574.337022617616
PROFILES = os.path.join(ROOT, "profiles")
# rocprofv3 --pmc summaries holding the measured HBM bytes per launch of the
# headline stencil launches, by iterations per launch (FETCH_SIZE and
# WRITE_SIZE per dispatch, each scaled by what the counter reads for a 128 MiB
# device-to-device copy); each file states the commands that made it.
TRAFFIC_FILES = {4: "wave_fused_pm1e6_128_hbm_counters.txt", 1: "wave_pm1e6_128_hbm_counters.txt"}


def profiled_traffic(levels):
    """(bytes per launch, file) from the summary's `traffic_bytes_per_launch = B` line, or (None, None)."""
    name = TRAFFIC_FILES.get(levels)
    if name is None:
        return None, None
    try:
        with open(os.path.join(PROFILES, name)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    return int(line.replace(":", "=", 1).split("=")[1].split()[0]), "profiles/" + name
    except OSError:
        pass
    return None, None


def wave_schedule(rrg, iters, nx, real="fp32", fused=True):
    """Launches of one unchecked solve of `iters` iterations on a periodic
    grid, by iterations per launch, as the library splits them."""
    return rrg.wave_schedule(nx, iters, real=real, fused=fused)

PM = dict(n=1_000_000, nx=128, dx=0.125, dt=5e-3, iterations=200, setup_iterations=2000, diss_base=0.785,
          diss_extra=0.01)
RESORT_EVERY = 48  # the library's default steps between PM particle re-sorts


def _device_sync():
    try:
        import torch

        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def pm_bench(rrg, ics, steps, warmup, real="fp32", force_err=True):
    pos, vel, m = ics.exp_disk(PM["n"], seed=1)
    wave = dict(diss_base=PM["diss_base"], diss_extra=PM["diss_extra"], wave_speed=C,
                setup_iterations=PM["setup_iterations"], iterations=PM["iterations"], tolerance=0.0)
    p = rrg.make_params(dt=PM["dt"], grid=dict(nx=PM["nx"], dx=PM["dx"], bc_frac=0.0), wave=wave, real=real)
    sim = rrg.Simulation(p, pos, vel, m)
    sim.precalc()
    sim.sync_particles_and_grid()
    acc0 = sim.state()["acc"]
    # throughput: steps as a user runs them (no stage events: one graph
    # launch per step); the warmup captures that graph
    sim.set_timing(False)
    sim.calc_step(max(warmup, 1))
    sim.synchronize()
    _device_sync()
    t0 = time.perf_counter()
    sim.calc_step(steps)
    sim.synchronize()
    _device_sync()
    t1 = time.perf_counter()
    # stage breakdown and the stencil's launch time: at least one whole
    # re-sort cycle of steps again with HIP events between the stages (one
    # graph per stage, captured by the first of these steps), so the
    # per-step stage times carry the re-sort's share
    sim.set_timing(True)
    sim.calc_step(1)
    sim.synchronize()
    sim.set_timing(True)
    estep = max(steps, RESORT_EVERY)
    te0 = time.perf_counter()
    sim.calc_step(estep)
    sim.synchronize()
    te1 = time.perf_counter()
    tm = sim.timing()
    stats = sim.stats()
    sim.close()
    # the stencil launch time with every kernel launched on its own (no
    # graph), the form a kernel tracer times (bench.py --profile-pass --eager),
    # and each stage's eager time over steps 16-19 after the re-sort of
    # precalc (the deposit's time grows as particles leave their sorted runs)
    pe = rrg.make_params(dt=PM["dt"], grid=dict(nx=PM["nx"], dx=PM["dx"], bc_frac=0.0),
                         wave=dict(wave, setup_iterations=200), real=real, graphs=False)
    with rrg.Simulation(pe, pos, vel, m) as se:
        se.precalc()
        se.calc_step(16)
        se.set_timing(True)
        se.calc_step(3)
        eager_tm = se.timing()
        wave_eager_ms, wave_eager_launches = eager_tm["wave"]
    ncell = PM["nx"] ** 3
    rb = 4 if real == "fp32" else 8
    wave_ms, wave_launches = tm["wave"]
    avg_wave_s = wave_ms / 1e3 / max(wave_launches, 1)
    iters = PM["iterations"]
    sched = wave_schedule(rrg, iters, PM["nx"], real=real)
    assert sum(sched.values()) == round(wave_launches / estep), (sched, wave_launches, estep)
    fused = sched[4] + sched[3] + sched[2]
    n = PM["n"]
    # Effective bytes, fp32, per unit of work of the PM step (what the step
    # would move with one stencil iteration per launch): the stencil 16 B
    # per cell-iteration (phi^n, phi^(n-1), source read, phi^(n+1)
    # written), gradient 16 B per cell, deposit 16 B per particle + 8 B per
    # cell, fixed point to source 12 B per cell, gather 24 B per particle,
    # kick/drift 60 B per particle.
    step_bytes = iters * 16 * ncell + 16 * ncell + (16 * n + 8 * ncell) + 12 * ncell + 24 * n + 60 * n
    # Compulsory HBM bytes of the kernels that run, each grid and particle
    # array moved once: a fused stencil launch reads phi^n, phi^(n-1) and the
    # source and writes its last two levels (20 B per cell), a single launch
    # 16 B per cell; predict + deposit x, v, a, m in and x, v* out (64 B per
    # particle) and the int64 grid read and written (16 B per cell); fixed
    # point to source 8 B in and the source out (12 B per cell; it also
    # clears the cells that hold mass, a few percent, not counted); gather +
    # kick (with -grad phi computed from phi on the way)
    # x, v, v* in, a, v out (60 B per particle) and phi read (4 B per cell).
    # The re-sort (every 48 steps) is not counted.
    own_bytes = (fused * 20 * ncell + sched[1] * 16 * ncell + (16 + 12 + 4) * ncell + (64 + 60) * n)
    secs = t1 - t0
    out = {}
    if force_err:
        # PM force vs an FP64 direct sum over all particles (softening = dx):
        # the method error of the PM model (periodic source, cell-size
        # resolution)
        idx, ref = _fp64_sample_forces(pos, m, np.full(n, PM["dx"] ** 2), nsample=4096)
        err = np.linalg.norm(acc0[idx] - ref, axis=1) / np.linalg.norm(ref, axis=1)
        out["pm_force_err_vs_fp64_direct"] = {"median": float(np.median(err)), "p99": float(np.percentile(err, 99)),
                                              "sample": "4096 particles (seed 5), t=0, direct sum softened by dx"}
    return out | {
        "value": n * steps / secs / 1e6,
        "ms_per_step": secs / steps * 1e3,
        "ms_per_step_with_stage_events": (te1 - te0) / estep * 1e3,
        "stage_ms_per_step": {k: v[0] / estep for k, v in tm.items() if v[1]},
        "stage_event_steps": estep,
        "resort_every_steps": RESORT_EVERY,
        "resorts_in_stage_steps": int(tm.get("sort", (0, 0))[1] // 3),
        # the timed steps may fall between re-sorts; this adds the re-sort's
        # per-step share measured over the stage-event steps
        "ms_per_step_resort_amortized": secs / steps * 1e3 + tm.get("sort", (0.0, 0))[0] / estep,
        "wave_avg_launch_us": avg_wave_s * 1e6,
        "wave_avg_launch_us_eager": wave_eager_ms * 1e3 / max(wave_eager_launches, 1),
        "stage_us_eager_16_after_sort": {k: v[0] * 1e3 / 3 for k, v in eager_tm.items()
                                         if k in ("deposit", "source", "gather") and v[1]},
        # the other stages' compulsory bytes over their eager stage times (the
        # event bracket includes the launch, so these are lower bounds)
        "stage_rates_eager": {
            k: {"compulsory_bytes": b, "gbs": b / (eager_tm[k][0] * 1e-3 / 3) / 1e9,
                "frac": b / (eager_tm[k][0] * 1e-3 / 3) / 1e9 / HBM_PEAK_GBS, "per_unit": u}
            for k, b, u in (("deposit", 16 * rb * n + 16 * ncell,
                             f"{16 * rb} B per particle (x, v, a, m in; x, v* out) + 16 B per cell (int64 grid)"),
                            ("source", (8 + rb) * ncell,
                             f"{8 + rb} B per cell (int64 grid read, source written; the cells holding mass, a "
                             "few percent, are also cleared for the next deposit, not counted)"),
                            ("gather", 15 * rb * n + rb * ncell,
                             f"{15 * rb} B per particle (x, v, v* in; a, v out) + phi read once"))
            if eager_tm.get(k, (0, 0))[1]},
        "wave_ms_per_step": wave_ms / estep,
        "wave_us_per_iteration": wave_ms * 1e3 / (iters * estep),
        "wave_launches_per_step": {"four_iterations": sched[4], "three_iterations": sched[3],
                                   "two_iterations": sched[2], "single": sched[1]},
        "wave_iterations_per_launch": max(k for k, v in sched.items() if v),
        "wave_fused_launch_bytes": 20 * ncell,
        "step_effective_bytes": step_bytes,
        "step_compulsory_bytes": own_bytes,
        "wave_iterations_per_step": stats["last_wave_iterations"],
    }


def pm_fp64_bench(rrg, ics, steps=10, warmup=2):
    """The headline PM workload with FP64 particles and grid (the reference's
    default precision): throughput, step time and stage breakdown."""
    r = pm_bench(rrg, ics, steps, warmup, real="fp64", force_err=False)
    return {"value": r["value"], "unit": "Mparticle-steps/s", "ms_per_step": r["ms_per_step"],
            "ms_per_step_with_stage_events": r["ms_per_step_with_stage_events"],
            "stage_ms_per_step": r["stage_ms_per_step"], "wave_us_per_iteration": r["wave_us_per_iteration"],
            "wave_launches_per_step": r["wave_launches_per_step"],
            "config": {"workload": "as the headline, FP64 state", "timed_steps": steps}}


def _fp64_sample_forces(pos, m, s2, nsample=512):
    """Host FP64 direct sum a_i = sum_j m_j d_ij (r^2 + (s2_i + s2_j)/2)^(-3/2), d_ij = x_j - x_i,
    for a seeded sample of targets (the reference kernel's formula, nbody.cu:19-65); one target at
    a time over all sources, targets spread over host threads (numpy releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor

    idx = np.sort(np.random.default_rng(5).choice(len(m), nsample, replace=False))
    ref = np.empty((nsample, 3))
    x, y, z = (np.ascontiguousarray(pos[:, c]) for c in range(3))

    def one(a):
        t = idx[a]
        dx, dy, dz = x - x[t], y - y[t], z - z[t]
        q = dx * dx + dy * dy + dz * dz + 0.5 * (s2[t] + s2)
        w = m / (q * np.sqrt(q))
        ref[a] = ((w * dx).sum(), (w * dy).sum(), (w * dz).sum())

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        list(ex.map(one, range(nsample)))
    return idx, ref


def direct_bench(rrg, ics, steps=20):
    """Config 3: `steps` steps of each force kind; the first captures the
    stage graphs, the others are timed (with stage events, for the kernel
    time); force errors at t=0 and the energy drift over the run."""
    pos, vel, m = ics.disk_bulge()
    n = len(m)
    s2 = np.full(n, 0.005 ** 2)
    out, acc0, energy0, energy1 = {}, {}, {}, {}
    k = steps - 1
    for force in ("mixed", "fp32", "fp64", "fp64_exact"):
        p = rrg.make_params(dt=1e-3, force=force)
        sim = rrg.Simulation(p, pos, vel, m, s2)
        sim.precalc()
        sim.sync_particles_and_grid()
        acc0[force] = sim.state()["acc"]
        energy0[force] = sim.conservation()
        sim.set_timing(True)  # warm up with stage events, as timed
        sim.calc_step(1)
        sim.synchronize()
        sim.set_timing(True)
        t0 = time.perf_counter()
        sim.calc_step(k)
        sim.synchronize()
        t1 = time.perf_counter()
        tm = sim.timing()
        energy1[force] = sim.conservation()
        sim.close()
        kern_s = tm["direct"][0] / 1e3 / k
        inter = float(n) * n
        out[force] = {
            "ginteractions_per_s": inter * k / (t1 - t0) / 1e9,
            "kernel_ginteractions_per_s": inter / kern_s / 1e9,
            "kernel_ms": kern_s * 1e3,
            "tflops_20_per_interaction": 20 * inter / kern_s / 1e12,
        }
    idx, ref = _fp64_sample_forces(pos, m, s2, nsample=4096)
    e_ref = energy0["fp64_exact"]["Ek"] + energy0["fp64_exact"]["Ep"]
    for force in acc0:
        err = np.linalg.norm(acc0[force][idx] - ref, axis=1) / np.linalg.norm(ref, axis=1)
        e = energy0[force]["Ek"] + energy0[force]["Ep"]
        e1 = energy1[force]["Ek"] + energy1[force]["Ep"]
        out[force]["force_rel_err_median"] = float(np.median(err))
        out[force]["force_rel_err_p99"] = float(np.percentile(err, 99))
        out[force]["force_rel_err_max"] = float(np.max(err))
        out[force]["energy_rel_err_vs_fp64_exact"] = abs(e - e_ref) / abs(e_ref)
        out[force]["energy_drift"] = abs(e1 / e - 1.0)
    for f32 in ("mixed", "fp32"):
        for f64 in ("fp64", "fp64_exact"):
            out[f"{f32}_over_{f64}"] = (out[f32]["kernel_ginteractions_per_s"] /
                                        out[f64]["kernel_ginteractions_per_s"])
    out["config"] = {"workload": "disk N=196608 + Plummer bulge N=65536 (M=0.2, a=0.2), eps=0.005, dt=1e-3",
                     "n": n, "steps": steps, "timed_steps": k,
                     "kinds": {"fp64": "FP64 pairs, device-library rsqrt", "fp64_exact": "FP64 pairs, IEEE 1./sqrt "
                               "(the reference kernel's arithmetic)", "mixed": "FP32 pairs on hi/lo positions + "
                               "two-sum accumulation", "fp32": "FP32 pairs on FP32 positions + two-sum accumulation"},
                     "force_error": f"|a - a_ref| / |a_ref| at t=0 on {len(idx)} sampled particles (seed 5); a_ref = "
                                    "FP64 direct sum over all N sources on the host (numpy)",
                     "energy_error": "|E - E_ref| / |E_ref| of Ek + Ep at t=0, E_ref from the fp64_exact kernel",
                     "energy_drift": f"|E(t_{steps}) / E(0) - 1| over the {steps} steps"}
    return out


def _ensemble_force_err(pos, m, s2, off, acc, per_system=32):
    """Median over each system's seeded sample of targets of |a - a_ref| / |a_ref|,
    a_ref from a host FP64 direct sum over the system's own sources (the
    reference kernel's formula); returns the worst system's median and the
    largest error seen."""
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(9)
    jobs = []
    for g in range(len(off) - 1):
        a, b = int(off[g]), int(off[g + 1])
        for t in np.sort(rng.choice(b - a, min(per_system, b - a), replace=False)):
            jobs.append((g, a, b, a + int(t)))

    def one(job):
        g, a, b, t = job
        d = pos[a:b] - pos[t]
        q = np.einsum("ij,ij->i", d, d) + 0.5 * (s2[t] + s2[a:b])
        ref = (m[a:b] / (q * np.sqrt(q))) @ d
        return g, float(np.linalg.norm(acc[t] - ref) / np.linalg.norm(ref))

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        res = list(ex.map(one, jobs))
    per = {}
    for g, e in res:
        per.setdefault(g, []).append(e)
    return max(float(np.median(v)) for v in per.values()), max(e for _, e in res)


def ensemble_bench(rrg, ics, steps=100):
    """64 systems, N_g = round(2^U(12,16)) (seed 2), Plummer for even g and
    exponential disk for odd g, eps = 0.01, dt = 1e-3, FP32 pair forces on
    FP64 state: the mixed kind (hi/lo positions) at the top level and the fp32
    kind beside it; one persistent work-queue launch per force evaluation,
    hipGraph-replayed steps, autotuned launch shape; force errors at t = 0
    against a host FP64 direct sum per system."""
    sizes = ics.ensemble_sizes(64, seed=2)
    rng = np.random.default_rng(2)
    parts = [ics.plummer(int(nn), rng=rng) if g % 2 == 0 else ics.exp_disk(int(nn), rng=rng)
             for g, nn in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    pos = np.concatenate([q[0] for q in parts]); vel = np.concatenate([q[1] for q in parts])
    m = np.concatenate([q[2] for q in parts])
    s2 = np.full(len(m), 0.01 ** 2)
    inter = float(np.sum(sizes.astype(np.float64) ** 2))

    def run(force):
        p = rrg.make_params(dt=1e-3, force=force, autotune=True)
        sim = rrg.Simulation(p, pos, vel, m, s2, offsets=off)
        sim.precalc()
        acc0 = sim.state()["acc"]
        sim.calc_step(9)  # captures the eight-step and the one-step graph
        sim.synchronize()
        g0 = sim.stats()["graph_launches"]
        t0 = time.perf_counter()
        sim.calc_step(steps)
        sim.synchronize()
        t1 = time.perf_counter()
        stats = sim.stats()
        # kernel time of the force launch, from events on the context stream
        sim.set_timing(True)
        sim.calc_step(1)  # captures the per-stage graphs
        sim.set_timing(True)
        sim.calc_step(3)
        tm = sim.timing()
        sim.close()
        kern_s = tm["direct"][0] / 1e3 / 3
        worst_median, worst = _ensemble_force_err(pos, m, s2, off, acc0)
        return {
            "ginteractions_per_s": inter * steps / (t1 - t0) / 1e9,
            "kernel_ginteractions_per_s": inter / kern_s / 1e9,
            "ms_per_step": (t1 - t0) / steps * 1e3,
            "force_kernel_ms": kern_s * 1e3,
            "fp32_vector_frac": 20 * inter / kern_s / 1e12 / FP32_PEAK_TFLOPS,
            "kernel_launches_per_step": 3,
            "graph_launches_per_step": (stats["graph_launches"] - g0) / steps,
            "tuned_shape": {"ib": stats["tile_ib"], "sources_per_item": stats["slice_len"],
                            "workgroups": stats["workgroups"], "items": stats["work_items"],
                            "candidates": stats["tune_candidates"]},
            "force_rel_err_worst_system_median": worst_median,
            "force_rel_err_max": worst,
        }

    out = run("mixed")
    out["fp32"] = run("fp32")
    out["config"] = {"workload": "64 galaxies, N log-uniform in [2^12, 2^16] (seed 2), Plummer/disk alternating, "
                                 "eps=0.01, dt=1e-3, FP32 pair forces / FP64 state (top level: the mixed kind; "
                                 "`fp32`: plain FP32 positions)", "systems": 64,
                     "n_total": int(off[-1]), "sum_n2": inter, "timed_steps": steps,
                     "force_error": "|a - a_ref| / |a_ref| at t=0 on 32 sampled targets per system (seed 9); "
                                    "a_ref = FP64 direct sum over the system's sources on the host (numpy)"}
    return out


def adjoint_bench(rrg, ics, n=65_536, steps=200, every=20):
    """Plummer N=65,536 (seed 0), eps=0.01, dt=1e-3, 200 steps; target = the
    same ICs with masses x exp(0.05 N(0,1)) (seed 3) run forward; loss = mean
    |x_T - x_T^target|^2; gradients w.r.t. x0, v0 and m with checkpoints every
    20 steps.  t_forward = the recorded forward run, t_adjoint = rrg_adjoint.
    Gradient error: the mixed-kind gradients against the FP64-kind gradients
    of the same run (the FP64 path is pinned to the CPU FP64 reverse-mode
    gradient in tests/test_gpu_adjoint.py)."""
    pos, vel, m = ics.plummer(n, seed=0)
    s2 = np.full(n, 0.01 ** 2)
    mt = m * np.exp(0.05 * np.random.default_rng(3).standard_normal(n))
    p = rrg.make_params(dt=1e-3, force="mixed")
    with rrg.Simulation(p, pos, vel, mt, s2) as sim:
        sim.precalc()
        sim.calc_step(steps)
        target = sim.state()["pos"]
    out = {}
    for force in ("mixed", "fp64"):
        p = rrg.make_params(dt=1e-3, force=force)
        with rrg.Simulation(p, pos, vel, m, s2) as sim:
            sim.precalc()
            sim.adjoint_begin(every)
            sim.synchronize()
            t0 = time.perf_counter()
            sim.calc_step(steps)
            sim.synchronize()
            t1 = time.perf_counter()
            xT = sim.state()["pos"]
            loss = float(np.mean(np.sum((xT - target) ** 2, axis=1)))
            t2 = time.perf_counter()
            g = sim.adjoint(2 * (xT - target) / n)
            t3 = time.perf_counter()
        out[force] = {"t_forward_s": t1 - t0, "t_adjoint_s": t3 - t2, "adjoint_over_forward": (t3 - t2) / (t1 - t0),
                      "reverse_ginteractions_per_s": float(n) * n * steps / (t3 - t2) / 1e9, "loss": loss,
                      "grads": g}
    ref = out["fp64"]["grads"]
    errs = {}
    for name, a, b in zip(("x0", "v0", "m"), out["mixed"]["grads"], ref):
        errs[name] = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    for f in out.values():
        del f["grads"]
    return {"mixed": out["mixed"], "fp64": out["fp64"], "mixed_grad_rel_err_vs_fp64": errs,
            "config": {"workload": "Plummer N=65536, eps=0.01, dt=1e-3, 200 steps, checkpoint every 20, "
                                   "loss = mean |x_T - x_T(target masses)|^2", "n": n, "steps": steps}}


def wave_scaling(rrg, ics, sizes=(64, 128, 256, 512), steps=2, real="fp32"):
    """The stencil across grid sizes: the headline disk on 64^3 - 512^3 grids
    over the same [-8,8]^3 box (20 setup iterations, 200 per step, `real`
    state), with several iterations per launch where the library has a launch
    shape and with one per launch; time per iteration and per launch from the
    stage events; `frac` = the launches' compulsory bytes (fused: 5 reals per
    cell, single: 4) over the stage time against HBM peak, and
    `effective_gbs` = 4 reals per cell-iteration (16 B in FP32, 32 B in FP64)
    over the same time."""
    rb = 4 if real == "fp32" else 8
    pos, vel, m = ics.exp_disk(PM["n"], seed=1)
    out = {}
    for nx in sizes:
        row = {"cells": nx ** 3}
        for fused in (True, False):
            wave = dict(diss_base=PM["diss_base"], diss_extra=PM["diss_extra"], wave_speed=C, setup_iterations=20,
                        iterations=PM["iterations"], tolerance=0.0)
            p = rrg.make_params(dt=PM["dt"], grid=dict(nx=nx, dx=16.0 / nx), wave=wave, real=real, wave_fused=fused)
            with rrg.Simulation(p, pos, vel, m) as sim:
                sim.precalc()
                sim.set_timing(True)
                sim.calc_step(3)  # captures the stage graphs of every buffer arrangement
                sim.synchronize()
                sim.set_timing(True)
                sim.calc_step(steps)
                tm = sim.timing()
            ms, launches = tm["wave"]
            sched = wave_schedule(rrg, PM["iterations"], nx, real=real, fused=fused)
            assert sum(sched.values()) * steps == launches, (nx, fused, sched, launches)
            own = ((sched[4] + sched[3] + sched[2]) * 5.0 + sched[1] * 4.0) * rb * nx ** 3 * steps
            alg = 4.0 * rb * nx ** 3 * PM["iterations"] * steps
            row["fused" if fused else "single"] = {
                "launches_per_step": {str(k): v for k, v in sched.items() if v}, "us_per_launch": ms * 1e3 / launches,
                "us_per_iteration": ms * 1e3 / (PM["iterations"] * steps),
                "compulsory_gbs": own / (ms * 1e-3) / 1e9, "frac": own / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "effective_gbs": alg / (ms * 1e-3) / 1e9}
        out[str(nx)] = row
    return out


def _sample_setup(rrg, ics):
    """The sample run's particles and its params (a function of `graphs`)."""
    pos, vel, m = ics.exp_disk(262_144, seed=6, r_max=8.0)
    pos, vel = pos * 0.25, vel * 2.0
    wave = dict(diss_base=0.1, diss_extra=0.01, wave_speed=C, setup_iterations=30000, iterations=3000,
                tolerance=1e-6, steady_check_step=30, steady_hold_steps=3)
    return (pos, vel, m), lambda graphs: rrg.make_params(dt=0.002, grid=dict(nx=512, dx=0.025, bc_frac=0.2),
                                                         wave=wave, real="fp64", graphs=graphs)


def profile_pass_sample(rrg, ics, steps):
    """The sample run for a kernel trace: precalc and `steps` steps with every
    kernel launched on its own and the check read back by the host after each
    checking launch (the device-checked form replays graphs)."""
    (pos, vel, m), params = _sample_setup(rrg, ics)
    with rrg.Simulation(params(False), pos, vel, m) as sim:
        sim.precalc()
        setup_iters = sim.stats()["last_wave_iterations"]
        sim.sync_particles_and_grid()
        iters = []
        for _ in range(steps):
            sim.calc_step(1)
            iters.append(sim.stats()["last_wave_iterations"])
        sim.synchronize()
    print(json.dumps({"profile_pass": "sample", "setup_iterations": setup_iters, "iterations_each": iters}), flush=True)


def sample_bench(rrg, ics, steps=3):
    """The reference's sample run (sample/params.json): 512^3 cells of
    dx = 0.025 (box [-6.4, 6.4]^3), FP64 (the reference's default real),
    diss_base 0.1, up to 30000 setup iterations and up to 3000 per step with
    the steady check (tolerance 1e-6 every 30 iterations, hold 3),
    dt = 0.002.  Particles: N = 262,144 in a synthetic disk shaped like
    sample/ini/S.txt (exponential, scale 0.25, truncated at R = 2; S.txt
    itself is not shipped).  Run twice: the check decided on the device with
    graph-replayed steps, and decided on the host after each check (a
    blocking read-back per check, as the reference does).  After one warm-up
    step, `steps` steps are timed one at a time, each with its iteration
    count."""
    (pos, vel, m), params = _sample_setup(rrg, ics)
    out = {}
    for mode, graphs in (("device_check", True), ("host_check", False)):
        with rrg.Simulation(params(graphs), pos, vel, m) as sim:
            t0 = time.perf_counter()
            sim.precalc()
            sim.synchronize()
            t1 = time.perf_counter()
            setup_iters = sim.stats()["last_wave_iterations"]
            sim.sync_particles_and_grid()
            sim.calc_step(1)
            sim.synchronize()
            ms, iters = [], []
            for _ in range(steps):
                a = time.perf_counter()
                sim.calc_step(1)
                sim.synchronize()
                ms.append((time.perf_counter() - a) * 1e3)
                iters.append(sim.stats()["last_wave_iterations"])
            graph_launches = sim.stats()["graph_launches"]
        out[mode] = {"ms_per_step": float(np.mean(ms)), "ms_each": ms, "iterations_each": iters,
                     "us_per_iteration": float(np.sum(ms) * 1e3 / max(1, np.sum(iters))),
                     "setup_s": t1 - t0, "setup_iterations": setup_iters, "graph_launches": graph_launches}
    out["config"] = {"workload": "sample/params.json wave settings: 512^3, dx=0.025, fp64, diss_base=0.1, tolerance "
                                 "1e-6 every 30 iterations (hold 3), <= 3000 iterations per step, dt=0.002; synthetic "
                                 "disk N=262144 shaped like sample/ini/S.txt", "timed_steps": steps}
    return out


def c1_bench(rrg, ics, n=4096, steps=100):
    """Plain FP64 direct sum: Plummer N=4096 (seed 0), eps=0.01, dt=1e-3, 100
    steps; interactions/s over the steps and the relative energy drift."""
    pos, vel, m = ics.plummer(n, seed=0)
    s2 = np.full(n, 0.01 ** 2)
    out = {}
    for force in ("fp64_exact", "fp64"):
        p = rrg.make_params(dt=1e-3, force=force)
        with rrg.Simulation(p, pos, vel, m, s2) as sim:
            sim.precalc()
            c0 = sim.conservation()
            # the first 9 steps capture the eight-step and the one-step graph;
            # the remaining steps are timed
            sim.calc_step(8)
            sim.calc_step(1)
            sim.synchronize()
            timed = steps - 9
            t0 = time.perf_counter()
            sim.calc_step(timed)
            sim.synchronize()
            t1 = time.perf_counter()
            c1 = sim.conservation()
        e0, e1 = c0["Ek"] + c0["Ep"], c1["Ek"] + c1["Ep"]
        out[force] = {"ginteractions_per_s": float(n) * n * timed / (t1 - t0) / 1e9, "ms_per_step": (t1 - t0) / timed * 1e3,
                      "energy_drift": abs(e1 / e0 - 1.0)}
    out["config"] = {"workload": "Plummer N=4096 (seed 0), eps=0.01, dt=1e-3, 100 steps (steps 10-100 timed), FP64",
                     "n": n,
                     "steps": steps}
    return out


def pm_adjoint_bench(rrg, ics, n=262_144, steps=200, every=20):
    """PM variant of the differentiable run: exponential disk N=262,144
    (seed 1), 64^3 cells over [-8,8]^3, 200 damped-wave iterations per step
    (2000 at setup), dt=5e-3, 200 steps, checkpoints every 20 steps; target =
    the same ICs with masses x exp(0.05 N(0,1)) (seed 3) run forward; loss =
    mean |x_T - x_T^target|^2.  t_forward = precalc + recorded steps (the whole
    differentiated map), t_adjoint = rrg_adjoint.  Gradient error: FP32 state
    against the FP64 state's gradients of the same run (the FP64 path is
    pinned to the CPU FP64 reverse-mode gradient in tests/test_gpu_pm_adjoint.py)."""
    pos, vel, m = ics.exp_disk(n, seed=1)
    mt = m * np.exp(0.05 * np.random.default_rng(3).standard_normal(n))
    wave = dict(diss_base=PM["diss_base"], diss_extra=PM["diss_extra"], wave_speed=C, setup_iterations=2000,
                iterations=200, tolerance=0.0)
    grid = dict(nx=64, dx=0.25)
    p = rrg.make_params(dt=5e-3, grid=grid, wave=wave, real="fp32")
    with rrg.Simulation(p, pos, vel, mt) as sim:
        sim.precalc()
        sim.sync_particles_and_grid()
        sim.calc_step(steps)
        target = sim.state()["pos"]
    out = {}
    seed64 = None
    for real in ("fp64", "fp32"):
        p = rrg.make_params(dt=5e-3, grid=grid, wave=wave, real=real)
        with rrg.Simulation(p, pos, vel, m) as sim:
            sim.synchronize()
            t0 = time.perf_counter()
            sim.precalc()
            sim.adjoint_begin(every)
            sim.synchronize()
            t1 = time.perf_counter()
            sim.calc_step(steps)
            sim.synchronize()
            t2 = time.perf_counter()
            xT = sim.state()["pos"]
            loss = float(np.mean(np.sum((xT - target) ** 2, axis=1)))
            seed = 2 * (xT - target) / n
            t3 = time.perf_counter()
            g = sim.adjoint(seed)
            t4 = time.perf_counter()
            r = {"t_forward_s": t2 - t0, "t_forward_steps_s": t2 - t1, "t_adjoint_s": t4 - t3,
                 "adjoint_over_forward": (t4 - t3) / (t2 - t0), "reversed_steps_per_s": steps / (t4 - t3),
                 "loss": loss, "grads": g}
            if real == "fp64":
                seed64 = seed
            else:
                r["grads_fp64_seed"] = sim.adjoint(seed64)
                sim.set_timing(True)  # stage breakdown from a second, event-timed sweep
                sim.adjoint(seed)
                r["adjoint_stage_ms"] = {k: round(v[0], 3) for k, v in sim.timing().items() if v[1]}
                sim.set_timing(False)
        out[real] = r

    def rel(a, b):
        return {name: float(np.max(np.abs(x - y)) / np.max(np.abs(y))) for name, x, y in zip(("x0", "v0", "m"), a, b)}

    ref = out["fp64"]["grads"]
    same_seed = rel(out["fp32"].pop("grads_fp64_seed"), ref)
    own_seed = rel(out["fp32"]["grads"], ref)
    seed_diff = float(np.max(np.abs(seed - seed64)) / np.max(np.abs(seed64)))  # `seed` is the FP32 run's
    for f in out.values():
        del f["grads"]
    return {"fp32": out["fp32"], "fp64": out["fp64"],
            "fp32_grad_rel_err_vs_fp64": same_seed,
            "fp32_grad_rel_err_vs_fp64_own_loss": own_seed,
            "fp32_loss_seed_rel_diff": seed_diff,
            "grad_error_note": "same dL/dx_T seed (the FP64 run's) through both sweeps; the own-loss figure also "
                               "carries the FP32 forward's divergence from the FP64 trajectory in dL/dx_T",
            "config": {"workload": "PM: exponential disk N=262144, 64^3 grid on [-8,8]^3, 200 wave iterations/step "
                                   "(2000 setup), dt=5e-3, 200 steps, checkpoint every 20, loss = mean |x_T - "
                                   "x_T(target masses)|^2", "n": n, "steps": steps}}


def _cores():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def _oracle():
    import oracle

    oracle.lib()
    return oracle


def cpu_baseline(ics, target_s=10.0, real="fp32", one_core=False):
    """The CPU oracle on a bounded sample of the headline workload: the full
    1e6-particle / 128^3 problem, K PM steps (each with the 200 wave
    iterations; K sized for about `target_s` of CPU work), timed as
    (precalc + K steps) - (precalc only)."""
    oracle = _oracle()
    pos, vel, m = ics.exp_disk(PM["n"], seed=1)
    dtype = np.float32 if real == "fp32" else np.float64
    kw = dict(nx=PM["nx"], dx=PM["dx"], dt=PM["dt"], diss_base=PM["diss_base"], wave_speed=C, setup_iterations=0,
              iterations=PM["iterations"], tolerance=0.0, fixed_e=60, dtype=dtype)

    def per_step(k):
        t0 = time.perf_counter()
        oracle.pm_run(pos, vel, m, 0, **kw)
        t1 = time.perf_counter()
        oracle.pm_run(pos, vel, m, k, **kw)
        t2 = time.perf_counter()
        return ((t2 - t1) - (t1 - t0)) / k

    k = int(min(100, max(2, target_s / max(per_step(2), 1e-3))))
    t = per_step(k)
    out = {"value": PM["n"] / t / 1e6, "unit": "Mparticle-steps/s", "cores": _cores(), "kind": "port",
           "sample": f"oracle/rrg_oracle.cpp ({real}, OpenMP over cells and particles, deposit into per-thread "
                     f"grids) on the full workload, {k} PM steps of 200 wave iterations ({k * t:.1f} s), minus "
                     "precalc"}
    if one_core:
        prev = oracle.set_threads(1)
        try:
            t1 = per_step(1)
        finally:
            oracle.set_threads(prev)
        out["value_1_core"] = PM["n"] / t1 / 1e6
        out["sample_1_core"] = f"the same on one thread, 1 PM step ({t1:.1f} s), minus precalc"
    return out


def cpu_baseline_stencil(nx, real, iters, dx=None):
    """The oracle's wave_diss_iteration (wave.cu:106-193, periodic, OpenMP
    over cells) on an nx^3 grid: microseconds per iteration over `iters`
    iterations (the values do not change the work)."""
    oracle = _oracle()
    dtype = np.float32 if real == "fp32" else np.float64
    dx = 16.0 / nx if dx is None else dx
    ncell = nx ** 3
    phi, old, mass = np.zeros(ncell, dtype), np.zeros(ncell, dtype), np.zeros(ncell, dtype)
    dt = 0.5 * dx / (C * np.sqrt(3.0))
    oracle.wave_iteration(nx, dx, PM["diss_base"], C, dt, True, 1.0 / (nx * dx) ** 3, phi, old, mass, dtype=dtype)
    t0 = time.perf_counter()
    for _ in range(iters):
        phi, old = oracle.wave_iteration(nx, dx, PM["diss_base"], C, dt, True, 1.0 / (nx * dx) ** 3, phi, old, mass,
                                         dtype=dtype), phi
    t = time.perf_counter() - t0
    return {"value": t / iters * 1e6, "unit": f"us per stencil iteration ({nx}^3 {real}, lower is better)",
            "cores": _cores(), "kind": "port",
            "sample": f"oracle wave_iteration, {nx}^3 {real}, {iters} iterations after one untimed, {t:.2f} s"}


def cpu_baseline_c1(ics):
    """The oracle (FP64, IEEE 1/sqrt as nbody.cu:50-53) on the whole C1 run:
    precalc + 100 steps of Plummer N=4096, N^2 interactions per evaluation."""
    oracle = _oracle()
    n, steps = 4096, 100
    pos, vel, m = ics.plummer(n, seed=0)
    t0 = time.perf_counter()
    oracle.nbody_run(pos, vel, m, np.full(n, 0.01 ** 2), 1e-3, steps)
    t = time.perf_counter() - t0
    return {"value": float(n) * n * (steps + 1) / t / 1e9, "unit": "Ginteractions/s", "cores": _cores(),
            "kind": "port", "sample": f"oracle nbody_run, the whole workload (101 force evaluations), {t:.2f} s"}


def cpu_baseline_direct(ics, ntarget=4096):
    """The oracle's FP64 pair sum (nbody.cu:19-65 arithmetic) for a seeded
    sample of targets against all N sources of the disk + bulge workload."""
    oracle = _oracle()
    pos, _, m = ics.disk_bulge()
    n = len(m)
    idx = np.sort(np.random.default_rng(5).choice(n, ntarget, replace=False))
    t0 = time.perf_counter()
    oracle.accel_subset(pos, m, np.full(n, 0.005 ** 2), idx)
    t = time.perf_counter() - t0
    return {"value": float(ntarget) * n / t / 1e9, "unit": "Ginteractions/s", "cores": _cores(), "kind": "port",
            "sample": f"oracle accel_subset, {ntarget} targets (seed 5) x all {n} sources, FP64, {t:.2f} s"}


def cpu_baseline_ensemble(ics, per_system=256):
    """The oracle's FP64 pair sum over the 64-galaxy ensemble: a seeded sample
    of targets in every system against all of that system's sources."""
    oracle = _oracle()
    sizes = ics.ensemble_sizes(64, seed=2)
    rng = np.random.default_rng(2)
    parts = [ics.plummer(int(nn), rng=rng) if g % 2 == 0 else ics.exp_disk(int(nn), rng=rng)
             for g, nn in enumerate(sizes)]
    srng = np.random.default_rng(9)
    inter = 0.0
    t = 0.0
    for (pos, _, m) in parts:
        idx = np.sort(srng.choice(len(m), per_system, replace=False))
        t0 = time.perf_counter()
        oracle.accel_subset(pos, m, np.full(len(m), 0.01 ** 2), idx)
        t += time.perf_counter() - t0
        inter += float(per_system) * len(m)
    return {"value": inter / t / 1e9, "unit": "Ginteractions/s", "cores": _cores(), "kind": "port",
            "sample": f"oracle accel_subset, {per_system} targets per system (seed 9) x all its sources, 64 systems, "
                      f"FP64, {t:.2f} s"}


def cpu_baseline_adjoint(ics, n=4096, steps=20):
    """The oracle's FP64 reverse-mode gradient of the direct-sum run (forward
    run with stored positions, then the reversed steps), Plummer N=4096, 20
    steps; the rate counts N^2 pair interactions per reversed step."""
    oracle = _oracle()
    pos, vel, m = ics.plummer(n, seed=0)
    g = np.random.default_rng(3).standard_normal((n, 3)) / n
    t0 = time.perf_counter()
    oracle.nbody_adjoint(pos, vel, m, np.full(n, 0.01 ** 2), 1e-3, steps, g)
    t = time.perf_counter() - t0
    return {"value": float(n) * n * steps / t / 1e9, "unit": "Ginteractions/s of the reverse sweep (N^2 per "
            "reversed step, forward recomputation included)", "cores": _cores(), "kind": "port",
            "sample": f"oracle nbody_adjoint, Plummer N={n}, {steps} steps, FP64, {t:.2f} s"}


def cpu_baseline_adjoint_pm(ics, n=262_144):
    """The oracle's FP64 reverse-mode gradient of the PM run at the workload's
    size (64^3 cells, N=262,144, 200 wave iterations per step): the time per
    reversed step from runs of 3 and 1 steps (200 setup iterations)."""
    oracle = _oracle()
    pos, vel, m = ics.exp_disk(n, seed=1)
    g = np.random.default_rng(3).standard_normal((n, 3)) / n
    kw = dict(nx=64, dx=0.25, dt=5e-3, diss_base=PM["diss_base"], wave_speed=C, setup_iterations=200,
              iterations=200, fixed_e=60)
    t0 = time.perf_counter()
    oracle.pm_adjoint(pos, vel, m, 1, g, **kw)
    t1 = time.perf_counter()
    oracle.pm_adjoint(pos, vel, m, 3, g, **kw)
    t2 = time.perf_counter()
    per = ((t2 - t1) - (t1 - t0)) / 2
    return {"value": 1.0 / per, "unit": "reversed PM steps/s (forward recomputation included)", "cores": _cores(),
            "kind": "port", "sample": f"oracle pm_adjoint, 64^3, N={n}, (t(3 steps) - t(1 step)) / 2 = {per:.2f} s"}


def profile_pass_direct(rrg, ics, steps, graphs=True):
    """Short run for counter passes over the direct-sum kernels: the disk +
    bulge workload (N = 262,144, eps = 0.005), precalc + `steps` steps for
    each force kind (one acceleration launch per step, plus the potential
    and acceleration at precalc)."""
    pos, vel, m = ics.disk_bulge()
    s2 = np.full(len(m), 0.005 ** 2)
    for force in ("mixed", "fp32", "fp64", "fp64_exact"):
        p = rrg.make_params(dt=1e-3, force=force, graphs=graphs)
        with rrg.Simulation(p, pos, vel, m, s2) as sim:
            sim.precalc()
            sim.calc_step(steps)
            sim.synchronize()
        print(json.dumps({"profile_pass": "direct", "force": force, "steps": steps}), flush=True)


def profile_pass(rrg, ics, steps, graphs=True):
    """Short run for hardware-counter passes: `steps` PM steps of the headline
    workload (setup iterations cut to 200), then one device-to-device copy of
    exactly 128 MiB through the HIP runtime (its copy kernel reads and writes
    128 MiB), whose FETCH_SIZE / WRITE_SIZE per dispatch calibrate the
    counters' units.  graphs=False launches every kernel eagerly, for kernel
    tracers that do not follow graph launches."""
    import ctypes

    pos, vel, m = ics.exp_disk(PM["n"], seed=1)
    wave = dict(diss_base=PM["diss_base"], diss_extra=PM["diss_extra"], wave_speed=C, setup_iterations=200,
                iterations=PM["iterations"], tolerance=0.0)
    p = rrg.make_params(dt=PM["dt"], grid=dict(nx=PM["nx"], dx=PM["dx"], bc_frac=0.0), wave=wave, real="fp32",
                        graphs=graphs)
    with rrg.Simulation(p, pos, vel, m) as sim:
        sim.precalc()
        sim.calc_step(steps)
        sim.synchronize()
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    nbytes = 2 ** 27
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"profile_pass": "done", "pm_steps": steps, "copy_bytes_each_way": nbytes}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-direct", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-pass", action="store_true", help="short PM run + calibration copy for rocprofv3 --pmc")
    ap.add_argument("--eager", action="store_true", help="with --profile-pass: launch kernels eagerly, not from graphs")
    ap.add_argument("--workload", default="pm", choices=("pm", "direct", "sample"),
                    help="with --profile-pass: the headline PM run (pm), the direct-sum kinds on the disk + bulge "
                         "workload (direct) or the reference's sample run (sample, with --eager)")
    ap.add_argument("--skip", default="", help="comma-separated legs to leave out: direct, c1, scaling, ensemble, "
                                              "adjoint, adjoint_pm, pm_fp64, sample, cpu")
    args = ap.parse_args()
    skip = set(x for x in args.skip.split(",") if x)
    t_start = time.perf_counter()

    def done(leg):  # progress on stderr; stdout carries only the JSON line
        print(f"bench: {leg} done at {time.perf_counter() - t_start:.1f} s", file=sys.stderr, flush=True)
    if args.profile_pass:
        import rrgalaxsim_amd as rrg
        from rrgalaxsim_amd import ics

        rrg.lib()
        if args.workload == "direct":
            profile_pass_direct(rrg, ics, args.steps, graphs=not args.eager)
        elif args.workload == "sample":
            if not args.eager:
                sys.exit("bench.py: the sample profile pass runs eagerly (--eager)")
            profile_pass_sample(rrg, ics, args.steps)
        else:
            profile_pass(rrg, ics, args.steps, graphs=not args.eager)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); the hot path has no "
              "multi-GPU implementation", file=sys.stderr)
        sys.exit(2)
    import rrgalaxsim_amd as rrg
    from rrgalaxsim_amd import ics

    rrg.lib()
    pm = pm_bench(rrg, ics, args.steps, args.warmup)
    done("pm")
    # Dominant kernel: the stencil launch that runs `per_launch` iterations.
    # Its compulsory HBM bytes per launch: phi^n, phi^(n-1) and the source
    # read once (12 B per cell in FP32) and its last two levels written
    # (8 B per cell) for a fused launch; 16 B per cell for a single
    # iteration.  Launch time: the wave stage's HIP-event time on the
    # library's stream over its launches.
    cells = PM["nx"] ** 3
    per_launch = pm["wave_iterations_per_launch"]
    launch_s = pm["wave_avg_launch_us"] * 1e-6
    compulsory = (20 if per_launch > 1 else 16) * cells
    effective = 16 * cells * per_launch
    traffic, traffic_src = profiled_traffic(per_launch) if PM["nx"] == 128 else (None, None)
    step_s = pm["ms_per_step"] * 1e-3
    line = {
        "metric": METRIC,
        "value": pm["value"],
        "unit": "Mparticle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": pm["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "PM: exponential disk N=1e6 (seed 1), 128^3 grid on [-8,8]^3, CIC + 200 damped-wave "
                               "iterations + gradient + trilinear gather, dt=5e-3, fp32",
                   "n": PM["n"], "nx": PM["nx"], "wave_iterations": PM["iterations"]},
        "roofline": {"bound": "hbm",
                     "kernel": f"k_wave_fused<float, {per_launch}>" if per_launch > 1 else "k_wave<float>",
                     "achieved": compulsory / launch_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": compulsory / launch_s / 1e9 / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes": compulsory, "launch_us": launch_s * 1e6,
                     "launch_us_eager": pm["wave_avg_launch_us_eager"],
                     "launch_note": "launch_us: HIP events on the library's stream around the stencil launches of "
                                    "graph-replayed steps (the timed form); launch_us_eager: the same with every "
                                    "kernel launched on its own, the form the committed kernel trace times",
                     "per_unit": (f"{compulsory // cells} B per cell per launch ({cells} cells, {per_launch} "
                                  "iterations): phi^n, phi^(n-1) and the source read once, the last two levels "
                                  "written" if per_launch > 1 else f"16 B per cell ({cells} cells): phi^n, "
                                  "phi^(n-1), source read, phi^(n+1) written"),
                     "effective_gbs": effective / launch_s / 1e9,
                     "effective_note": "16 B per cell-iteration x iterations per launch over the launch time: the "
                                       "rate a one-iteration-per-launch stencil would need; not an HBM rate, since "
                                       "the fused launch keeps its intermediate levels on chip"},
        "step_roofline": {"compulsory_bytes": pm["step_compulsory_bytes"],
                          "achieved_gbs": pm["step_compulsory_bytes"] / step_s / 1e9,
                          "frac": pm["step_compulsory_bytes"] / step_s / 1e9 / HBM_PEAK_GBS,
                          "effective_bytes": pm["step_effective_bytes"],
                          "effective_gbs": pm["step_effective_bytes"] / step_s / 1e9},
        "ms_per_step_with_stage_events": pm["ms_per_step_with_stage_events"],
        "stage_ms_per_step": pm["stage_ms_per_step"],
        "stage_us_eager_16_after_sort": pm["stage_us_eager_16_after_sort"],
        "stage_rates_eager": pm["stage_rates_eager"],
        "stage_event_steps": pm["stage_event_steps"],
        "resort_every_steps": pm["resort_every_steps"],
        "resorts_in_stage_steps": pm["resorts_in_stage_steps"],
        "ms_per_step_resort_amortized": pm["ms_per_step_resort_amortized"],
        "wave_avg_launch_us": pm["wave_avg_launch_us"],
        "wave_us_per_iteration": pm["wave_us_per_iteration"],
        "wave_launches_per_step": pm["wave_launches_per_step"],
        "pm_force_err_vs_fp64_direct": pm["pm_force_err_vs_fp64_direct"],
    }
    if not args.no_direct:
        legs = [("direct", "direct_sum", lambda: direct_bench(rrg, ics)), ("c1", "direct_sum_c1", lambda: c1_bench(rrg, ics)),
                ("scaling", "wave_scaling", lambda: wave_scaling(rrg, ics)),
                ("scaling", "wave_scaling_fp64", lambda: wave_scaling(rrg, ics, real="fp64")),
                ("ensemble", "ensemble", lambda: ensemble_bench(rrg, ics)),
                ("adjoint", "adjoint", lambda: adjoint_bench(rrg, ics)),
                ("adjoint_pm", "adjoint_pm", lambda: pm_adjoint_bench(rrg, ics)),
                ("pm_fp64", "pm_fp64", lambda: pm_fp64_bench(rrg, ics)),
                ("sample", "sample_workload", lambda: sample_bench(rrg, ics))]
        for leg, key, fn in legs:
            if leg in skip:
                continue
            line[key] = fn()
            done(leg)
        if "direct_sum" in line:
            d = line["direct_sum"]
            for k in ("mixed", "fp32"):
                d[k]["fp32_vector_frac"] = d[k]["tflops_20_per_interaction"] / FP32_PEAK_TFLOPS
            for k in ("fp64", "fp64_exact"):
                d[k]["fp64_vector_frac"] = d[k]["tflops_20_per_interaction"] / FP64_PEAK_TFLOPS
    if "wave_scaling" in line:
        ws = line["wave_scaling"]
        line["roofline"]["hbm_bound_grids"] = {
            nx: {"fused_frac": ws[nx]["fused"]["frac"], "single_frac": ws[nx]["single"]["frac"],
                 "fused_us_per_iteration": ws[nx]["fused"]["us_per_iteration"]} for nx in ("256", "512") if nx in ws}
    if not args.no_cpu and "cpu" not in skip:
        line["cpu_baseline"] = cpu_baseline(ics, one_core=True)
        per_leg = [("direct_sum", cpu_baseline_direct), ("direct_sum_c1", cpu_baseline_c1),
                   ("ensemble", cpu_baseline_ensemble), ("adjoint", cpu_baseline_adjoint),
                   ("adjoint_pm", cpu_baseline_adjoint_pm),
                   ("pm_fp64", lambda i: cpu_baseline(i, target_s=5.0, real="fp64")),
                   ("wave_scaling", lambda i: {"128": cpu_baseline_stencil(128, "fp32", 20)}),
                   ("wave_scaling_fp64", lambda i: {"128": cpu_baseline_stencil(128, "fp64", 20)}),
                   ("sample_workload", lambda i: cpu_baseline_stencil(512, "fp64", 3, dx=0.025))]
        for key, fn in per_leg:
            if key in line:
                line[key]["cpu_baseline"] = fn(ics)
        done("cpu")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
