"""Benchmark: particle-steps/s of the PCISPH dam break on one MI355X, plain and
with the learned pressure warm start.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu] [--mode M]

Workload (default): the 524,288-particle PCISPH dam break (64x128x64 block at
0.00625 m in a 1.6x1.0x0.8 m box, dt 0.001 s, 1 % density-error tolerance,
3..50 corrections, +-1e-5 m seeded jitter).  A step is one PhysicSolver::update
of the whole scene; the state stays resident in HBM during the timed region.
Rate = particles * K / wall seconds of the K timed steps (HIP events on the
library's stream, bracketed by device synchronisation).

Three modes run the same K steps from the same initial state: plain (zero
pressure start, as the reference), learned (the committed network seeds the
first correction; same loop, same tolerance, the library's fallback on) and
learned_always (the seed on every step, no fallback).  A learned mode's rate
counts only when its run is at least as accurate as the plain one (no more
steps missing the tolerance, no larger worst step error, no higher energy
peak); `value` is the fastest mode that counts (`value_mode`, `value_reason`
say which and why).

Without --steps the run times steps 5-235 of the scene's configured 2000: in
every mode (and on the CPU oracle) the reference's update rule and clamping
box collider leave particles pressed against the walls above 1 % from step
23 on, and the scene turns unstable once the spreading front reaches the far
z wall (about step 233), after which the run measures an exploding state.
Each mode reports its qualified-step fraction; --steps 2000 times the whole
configured run.

Prints ONE JSON line.  Extra keys: per-step PCISPH iteration statistics,
convergence qualification, the dominant kernel's roofline (algorithmic bytes
per launch over its event-timed average duration) and the CPU baseline (the
oracle's OpenMP restatement on this host, over the same steps as the GPU's
timed run up to 40 of them, and the GPU's rate over exactly those steps).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic (compulsory) bytes per particle for each kernel: every particle's
# own state read or written once, FP32 SoA (vec3 = 12 B, scalar = 4 B);
# neighbour re-reads not counted.  Per PCISPH correction 108 B = predicted
# density + pressure update + error max (24 B) + pressure force + re-prediction
# (84 B); fixed part 348 B at three radix passes.
ALGO_BYTES = {
    "pci_density": 24.0,
    "pci_pressure": 84.0,
    "density": 16.0,
    "forces": 40.0 + 52.0,   # non-pressure forces + first prediction
    "pressure_sph": 36.0,
    "integrate": 88.0,
}


def step_bytes(k: float, passes: int = 3) -> float:
    """Compulsory bytes per particle-step: 300 + 16P fixed + 108 per correction."""
    return 300.0 + 16.0 * passes + 108.0 * k


KERNELS = ("grid_setup", "keys", "radix_pass", "reorder", "cell_table", "density", "forces", "pressure_sph",
           "pci_density", "mlp_seed", "pci_pressure", "integrate")


def _solver(scene, weights):
    from liquidsimulation_amd.solver import solver_for
    from liquidsimulation_amd import warm_start
    s = solver_for(scene.params)
    if weights is not None:
        warm_start.apply(s, weights)
    s.set_colliders(scene.colliders)
    s.set_particles(scene.x, scene.v)
    return s


def run_gpu(scene, steps: int, warmup: int, profile: bool, weights=None, marks=(40, 230), fixed_tuning=None):
    """Timed run of `steps` steps after `warmup`, with HIP events on the
    library's stream after the first marks[0], marks[1], ... timed steps (no
    synchronisation between them), then (profile) the same warm-up + steps
    again from the same initial state with per-launch HIP events: the run is
    deterministic, so the per-kernel averages are those of the timed steps;
    that rerun also samples the total energy every 50 steps."""
    import torch

    s = _solver(scene, weights)
    if warmup:
        s.update(scene.dt, warmup)
    if fixed_tuning is not None:
        s.set_tuning(*fixed_tuning)
        tuning = dict(zip(("use_graphs", "lookahead", "correction_threads", "lanes_per_particle"), fixed_tuning),
                      fixed=True)
    else:
        tuning = s.autotune(scene.dt, 10) if scene.params.method == 1 else None
    stream = torch.cuda.ExternalStream(s.stream())
    bounds = [0] + sorted({m for m in marks if 0 < m < steps}) + [steps]
    evs = [torch.cuda.Event(enable_timing=True) for _ in bounds]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record(stream)
    stats = []
    for k in range(1, len(bounds)):
        stats += s.update(scene.dt, bounds[k] - bounds[k - 1])
        evs[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = evs[0].elapsed_time(evs[-1])
    windows = {f"{warmup + bounds[0]}-{warmup + b}": evs[0].elapsed_time(evs[k])
               for k, b in enumerate(bounds) if k and b < steps}
    its = np.array([x.iterations for x in stats], np.int64)
    errs = np.array([x.max_density_error_ratio for x in stats], np.float64)
    conv = np.array([x.converged for x in stats], np.int64)
    used = np.array([x.warm_start_used for x in stats], np.int64)
    st = s.get_particles()
    m = s.getParticleMass()
    g = np.asarray(scene.params.gravity, np.float64)
    energy = {"kinetic_J": float(0.5 * m * np.sum(st["v"].astype(np.float64) ** 2)),
              "potential_J": float(-m * np.sum(st["x"].astype(np.float64) @ g))}
    s.close()
    kernels = {}
    same = None
    e_samples = []
    if profile:
        s = _solver(scene, weights)
        if warmup:
            s.update(scene.dt, warmup)
        if tuning:
            s.set_tuning(tuning["use_graphs"], tuning["lookahead"], tuning["correction_threads"],
                         tuning.get("lanes_per_particle", 0))
        ke, pe = s.get_energy()
        e_samples.append(float(ke.sum() + pe.sum()))
        s.profile(True)
        s.profile_reset()
        rerun = []
        for k in range(0, steps, 50):
            rerun += s.update(scene.dt, min(50, steps - k))
            ke, pe = s.get_energy()
            e_samples.append(float(ke.sum() + pe.sum()))
        for name in KERNELS:
            n, tot = s.profile_query(name)
            if n:
                kernels[name] = {"launches": n, "total_ms": tot, "avg_us": 1000.0 * tot / n}
        s.profile(False)
        same = bool(np.array_equal(np.array([x.iterations for x in rerun], np.int64), its))
        s.close()
    return {"ms": ms, "wall": wall, "its": its, "errs": errs, "conv": conv, "used": used, "kernels": kernels,
            "energy": energy, "energy_samples": e_samples, "rerun_identical": same, "tuning": tuning,
            "windows_ms": windows, "bounds": bounds, "warmup": warmup}


def run_ensemble(ens, steps: int, warmup: int, weights=None, profile: bool = False, always: bool = False):
    """All scenes of an ensemble in one context: warm-up, then `steps` timed
    steps (HIP events on the library stream).  With weights the learned seed
    runs with the library's fallback, or on every step (always)."""
    import torch
    from liquidsimulation_amd.solver import Ensemble
    from liquidsimulation_amd.abi import WARM_LEARNED, WARM_LEARNED_ALWAYS, WARM_NONE
    from liquidsimulation_amd import warm_start
    mode = WARM_NONE if weights is None else (WARM_LEARNED_ALWAYS if always else WARM_LEARNED)
    e = Ensemble(ens.params.replace(warm_start=mode))
    if weights is not None:
        warm_start.apply(e, weights)
    e.set_scenes([m.x for m in ens.members], [m.v for m in ens.members], gravity=ens.gravity)
    for k, m in enumerate(ens.members):
        e.set_scene_colliders(k, m.colliders)
    if warmup:
        e.update(ens.dt, warmup)
    tuning = e.autotune(ens.dt, 5)
    stream = torch.cuda.ExternalStream(e.stream())
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    stats = e.update(ens.dt, steps)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1)
    last = e.scene_stats()
    kernels = {}
    if profile:
        e.profile(True)
        e.profile_reset()
        e.update(ens.dt, min(steps, 50))
        for name in KERNELS:
            n, tot = e.profile_query(name)
            if n:
                kernels[name] = 1000.0 * tot / n
        e.profile(False)
    e.close()
    its = np.array([x.iterations for x in stats], np.int64)
    errs = np.array([x.max_density_error_ratio for x in stats], np.float64)
    return {"value": ens.n * steps / (ms / 1000.0), "ms_per_step": ms / steps,
            "iterations_max_over_scenes": {"mean": float(its.mean()), "max": int(its.max())},
            # a step counts as converged only when every scene met the tolerance
            "nonconverged_steps": int(sum(1 - x.converged for x in stats)),
            "worst_density_error": float(errs.max()) if errs.size else 0.0,
            "energy_J": None,
            "seeded_steps": int(sum(x.warm_start_used for x in stats)),
            "last_step_scene_iterations": [x.iterations for x in last],
            "nonconverged_scenes_last_step": int(sum(1 - x.converged for x in last)),
            "kernels_avg_us": kernels, "tuning": tuning}


def run_members_alone(ens, steps: int, warmup: int):
    """The same scenes one context at a time (rate over `steps` steps each)."""
    import torch
    total_ms = 0.0
    for k, m in enumerate(ens.members):
        sc = m.with_params(gravity=ens.gravity[k])
        s = _solver(sc, None)
        if warmup:
            s.update(sc.dt, warmup)
        stream = torch.cuda.ExternalStream(s.stream())
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ev0.record(stream)
        s.update(sc.dt, steps)
        ev1.record(stream)
        torch.cuda.synchronize()
        total_ms += ev0.elapsed_time(ev1)
        s.close()
    return {"value": ens.n * steps / (total_ms / 1000.0), "ms_per_step": total_ms / steps,
            "steps_per_scene": steps}


def bench_ensemble(args):
    from liquidsimulation_amd import scenes, warm_start
    ens = scenes.WORKLOADS[args.workload]()
    steps = args.steps if args.steps_given else ens.steps
    w = warm_start.load() if (args.mode == "both" and os.path.exists(warm_start.WEIGHTS)) else None
    plain = run_ensemble(ens, steps, args.warmup, None, not args.no_profile)
    learned = run_ensemble(ens, steps, args.warmup, w, False) if w is not None else None
    learned_always = run_ensemble(ens, steps, args.warmup, w, False, always=True) if w is not None else None
    alone = run_members_alone(ens, max(20, steps // 10), args.warmup)
    modes = {"learned": learned, "learned_always": learned_always}
    value_mode, value_reason = choose_value(plain, modes)
    head = plain if value_mode == "plain" else modes[value_mode]
    out = {
        "metric": "particle-steps/s of an ensemble of independent PCISPH dam breaks in one context, 1 GPU",
        "value": head["value"], "unit": "particle-steps/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic (seeded dam-break lattices, per-scene block dims, gravity, sphere)",
        "config": {"workload": ens.name, "scenes": len(ens.members), "particles": ens.n,
                   "particles_min": int(min(m.n for m in ens.members)),
                   "particles_max": int(max(m.n for m in ens.members)), "dt": ens.dt,
                   "warm_start": "learned" if learned is not None else "none"},
        "value_mode": value_mode, "value_reason": value_reason,
        "plain": plain, "learned": learned, "learned_always": learned_always,
        "scenes_one_at_a_time": alone,
        "speedup_batched_vs_one_at_a_time": plain["value"] / alone["value"],
    }
    print(json.dumps(out), flush=True)


def _window_stats(res, scene, lo, hi):
    its, errs, conv = res["its"][lo:hi], res["errs"][lo:hi], res["conv"][lo:hi]
    return {"iterations_mean": float(np.mean(its)) if its.size else 0.0,
            "nonconverged_steps": int(np.sum(conv == 0)) if scene.params.method == 1 else 0,
            "worst_density_error": float(errs.max()) if errs.size else 0.0,
            "mean_final_density_error": float(errs.mean()) if errs.size else 0.0}


def mode_summary(res, scene, steps):
    its, errs = res["its"], res["errs"]
    nonconv = int(np.sum(res["conv"] == 0)) if scene.params.method == 1 else 0
    es = res["energy_samples"]
    windows = {}
    for name, ms in res["windows_ms"].items():
        a, b = (int(v) for v in name.split("-"))
        lo, hi = a - res["warmup"], b - res["warmup"]
        w = {"value": scene.n * (hi - lo) / (ms / 1000.0), "steps": hi - lo}
        w.update(_window_stats(res, scene, lo, hi))
        windows[f"steps {a}-{b}"] = w
    return {
        "value": scene.n * steps / (res["ms"] / 1000.0),
        "ms_per_step": res["ms"] / steps,
        "iterations": {"mean": float(np.mean(its)) if its.size else 0.0,
                       "p50": float(np.percentile(its, 50)) if its.size else 0,
                       "p99": float(np.percentile(its, 99)) if its.size else 0,
                       "max": int(its.max()) if its.size else 0},
        "qualified": nonconv == 0,
        "qualified_step_fraction": 1.0 - nonconv / max(len(its), 1),
        "nonconverged_steps": nonconv,
        "worst_density_error": float(errs.max()) if errs.size else 0.0,
        "mean_final_density_error": float(errs.mean()) if errs.size else 0.0,
        "seeded_steps": int(res["used"].sum()),
        "windows": windows,
        "end_energy": res["energy"],
        "energy_J": ({"start": es[0], "max": max(es), "end": es[-1], "samples_every_steps": 50} if es else None),
        "kernels_avg_us": {k: v["avg_us"] for k, v in res["kernels"].items()},
        "profiled_rerun_identical": res["rerun_identical"],
        "tuning": res["tuning"],
    }


def _accuracy_reasons(plain, learned):
    reasons = []
    if learned["nonconverged_steps"] > plain["nonconverged_steps"]:
        reasons.append("more steps miss the tolerance")
    if learned["worst_density_error"] > plain["worst_density_error"]:
        reasons.append("larger worst step error")
    el, ep = learned["energy_J"], plain["energy_J"]
    if el and ep and el["max"] > 1.05 * max(ep["max"], ep["start"]):
        reasons.append("higher energy peak")
    return reasons


def choose_value(plain, learned):
    """A learned mode's rate counts only when its run is at least as accurate
    as the plain one over the same steps: no more steps that miss the
    tolerance, no larger worst step error, and no higher energy peak (5 %
    slack on the energy).  `value` is the fastest of the modes that count;
    plain always counts.  `learned` maps mode name -> summary (or None)."""
    best, why = "plain", []
    for name, lm in learned.items():
        if lm is None:
            continue
        reasons = _accuracy_reasons(plain, lm)
        if reasons:
            why.append(f"{name}: " + ", ".join(reasons))
        elif lm["value"] < (plain if best == "plain" else learned[best])["value"]:
            why.append(f"{name}: at least as accurate, slower")
        else:
            why.append(f"{name}: at least as accurate, faster")
            best = name
    return best, "; ".join(why) if why else "no weights"


def cpu_baseline(scene, warm: int, timed: int):
    from oracle import pyoracle as O
    nthreads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    nthreads = max(1, min(nthreads, 64))
    st = O.State.from_scene(scene.x, scene.v)
    if warm:
        O.run_ocl(scene.params, scene.colliders, scene.dt, warm, st, nthreads)
    t0 = time.perf_counter()
    status, its, errs = O.run_ocl(scene.params, scene.colliders, scene.dt, timed, st, nthreads)
    dt = time.perf_counter() - t0
    return {
        "value": scene.n * timed / dt,
        "unit": "particle-steps/s",
        "cores": nthreads,
        "kind": "port",
        "sample": f"{warm} warm-up + {timed} timed steps of {scene.name} from its initial state "
                  f"(OpenMP restatement of the OpenCL-kernel path, {nthreads} threads, "
                  f"mean {float(np.mean(its)):.2f} corrections/step)",
        "seconds": dt,
        "status": int(status),
    }


def copy_calibration():
    """A device-to-device copy of known size: rocprofv3's FETCH_SIZE and
    WRITE_SIZE read on it give the factors that turn those counters into bytes."""
    import torch
    n = 64 * 1024 * 1024
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4, "copies": 5}), flush=True)


TRAFFIC = os.path.join(ROOT, "profiles", "pcisph_dam_break_524k_traffic.json")
ISSUE = os.path.join(ROOT, "profiles", "pcisph_dam_break_524k_issue.json")
# VALU issue peak: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD every 4 cycles at 2.4 GHz
VALU_PEAK = 256 * 4 * 2.4e9 / 4


def issue_roofline(kernel: str):
    """The stencil kernels are bound by instruction issue, not HBM: VALU issue
    of `kernel` from the committed rocprofv3 counter summary (SQ_INSTS_VALU and
    SQ_INSTS_VALU_TRANS_F32 per launch over that run's own launch time; a
    transcendental occupies its SIMD four issue slots), or None."""
    try:
        with open(ISSUE) as f:
            k = json.load(f)["kernels"][kernel]
    except (OSError, KeyError, ValueError):
        return None
    slots = k["valu"] + 3.0 * k["valu_trans"]
    achieved = slots / (k["avg_us"] * 1e-6)
    return {"bound": "valu", "kernel": kernel, "achieved": achieved, "peak": VALU_PEAK, "unit": "wave-instr/s",
            "frac": achieved / VALU_PEAK, "valu_per_launch": k["valu"], "transcendental_per_launch": k["valu_trans"],
            "counter_run_avg_us": k["avg_us"], "source": os.path.relpath(ISSUE, ROOT)}


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    try:
        with open(TRAFFIC) as f:
            t = json.load(f)
        k = t["kernels"][kernel]
        return k["fetch_bytes"] + k["write_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=230)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="pcisph_dam_break_524k")
    ap.add_argument("--no-profile", action="store_true",
                    help="skip the event-timed rerun that gives per-kernel averages")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-warm", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="timed CPU steps (default: the GPU's timed steps, at most 40)")
    ap.add_argument("--mode", choices=("both", "plain"), default="both")
    ap.add_argument("--tuning", default=None,
                    help="graphs,lookahead,threads[,lanes] (e.g. 0,2,256): fixed launch settings, no "
                         "autotuning (counter runs); lanes per particle 0 (automatic), 1 or 4")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy five times (scales the HBM byte counters)")
    args = ap.parse_args()
    args.steps_given = "--steps" in sys.argv
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (the SPH step has no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from liquidsimulation_amd import scenes, warm_start
    from liquidsimulation_amd.abi import WARM_LEARNED, WARM_LEARNED_ALWAYS, WARM_NONE
    if args.workload.startswith("ensemble"):
        bench_ensemble(args)
        return
    scene = scenes.WORKLOADS[args.workload]()
    fixed = None
    if args.tuning:
        vals = [int(v) for v in args.tuning.split(",")]
        fixed = (bool(vals[0]), vals[1], vals[2], vals[3] if len(vals) > 3 else 0)
    cpu_steps = args.cpu_steps if args.cpu_steps > 0 else min(args.steps, 40)
    marks = (cpu_steps, 230)
    summaries = {"learned": None, "learned_always": None}
    runs = {}
    if args.mode == "both" and scene.params.method == 1 and os.path.exists(warm_start.WEIGHTS):
        w = warm_start.load()
        for name, mode in (("learned", WARM_LEARNED), ("learned_always", WARM_LEARNED_ALWAYS)):
            sl = scene.with_params(warm_start=mode)
            runs[name] = run_gpu(sl, args.steps, args.warmup, not args.no_profile, weights=w, marks=marks,
                                 fixed_tuning=fixed)
            summaries[name] = mode_summary(runs[name], sl, args.steps)
    scene = scene.with_params(warm_start=WARM_NONE)
    runs["plain"] = res = run_gpu(scene, args.steps, args.warmup, not args.no_profile, marks=marks,
                                  fixed_tuning=fixed)
    plain = mode_summary(res, scene, args.steps)
    learned = summaries["learned"]
    n = scene.n
    value_mode, value_reason = choose_value(plain, summaries)
    head = plain if value_mode == "plain" else summaries[value_mode]
    value = head["value"]
    tol = scene.params.max_density_error_ratio
    # the roofline is that of the run `value` comes from
    kernels = runs[value_mode]["kernels"]
    dom = max(kernels.items(), key=lambda kv: kv[1]["total_ms"])[0] if kernels else None
    roof = None
    if dom:
        per = ALGO_BYTES.get(dom)
        avg_s = kernels[dom]["avg_us"] * 1e-6
        if per:
            achieved = per * n / avg_s / 1e9
            roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom), "kernel": dom,
                    "algo_bytes_per_launch": per * n, "avg_launch_us": kernels[dom]["avg_us"]}
    mean_k = head["iterations"]["mean"]
    step_roof = n * step_bytes(mean_k) * args.steps / (head["ms_per_step"] * args.steps / 1000.0) / 1e9
    out = {
        "metric": "particle-steps/s at <=1% density error (plain vs learned warm start), 1 GPU",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded dam-break lattice with jitter)",
        "config": {"workload": scene.name, "particles": n, "dt": scene.dt,
                   "method": "PCISPH" if scene.params.method == 1 else "SPH",
                   "tolerance": tol, "min_iterations": scene.params.min_iterations,
                   "max_iterations": scene.params.max_iterations,
                   "stall_iterations": scene.params.stall_iterations,
                   "warm_start": "learned" if learned is not None else "none"},
        "value_mode": value_mode,
        "value_reason": value_reason,
        "iterations": head["iterations"],
        "qualified": head["qualified"],
        "qualified_step_fraction": head["qualified_step_fraction"],
        "nonconverged_steps": head["nonconverged_steps"],
        "worst_density_error": head["worst_density_error"],
        "plain": plain,
        "learned": learned,
        "learned_always": summaries["learned_always"],
        "speedup_learned_vs_plain": (learned["value"] / plain["value"]) if learned is not None else None,
        "speedup_learned_always_vs_plain": (summaries["learned_always"]["value"] / plain["value"])
        if summaries["learned_always"] is not None else None,
        "roofline": roof,
        "issue_roofline": issue_roofline(dom) if dom else None,
        "step_roofline": {"achieved_GBps": step_roof, "frac": step_roof / HBM_PEAK_GBS,
                          "bytes_per_particle_step": step_bytes(mean_k)},
        "kernels": kernels,
        "wall_s": res["wall"],
    }
    if not args.no_cpu:
        cb = cpu_baseline(scene, args.cpu_warm, cpu_steps)
        # the GPU over the same steps (the whole timed run, or its first window)
        if args.cpu_warm == args.warmup:
            if cpu_steps == args.steps:
                same = {"window": f"steps {args.warmup}-{args.warmup + args.steps}", "value": head["value"],
                        "iterations_mean": head["iterations"]["mean"]}
            else:
                win = f"steps {args.warmup}-{args.warmup + cpu_steps}"
                same = dict(window=win, **{k: head["windows"][win][k] for k in ("value", "iterations_mean")}) \
                    if win in head["windows"] else None
            if same:
                same["gpu_over_cpu"] = same["value"] / cb["value"]
                cb["gpu_same_steps"] = same
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
