"""Throughput of the checkpointed forward + adjoint stirring rollout on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload single_n64]

One "step" is one gradient evaluation of every instance of the workload: reset to the step-0
state, the full forward rollout (recording the stick force, checkpoints every --ckpt substeps),
the force-MSE loss against a target profile, and the full checkpointed adjoint giving
dL/dparams.  Inputs are resident on the GPU before timing; the loss runs on the host over
the T x 3 force series.

Prints one JSON line:
  value      fwd+adjoint particle-steps/s (liquid particles x substeps x evaluations / s)
  extra      gradient evals/s, forward-only and adjoint split, the batched workload figure
  roofline   the dominant kernel's algorithmic HBM bytes / its average launch time
  cpu_baseline  the reference's structure on the host (FP32, per-particle parallel loops with atomic
             grid adds, OpenMP on the host's cores; and one core) on a bounded sample of the same
             workload; the FP64 oracle checks the GPU on that sample
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from typing import Optional

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from diffstir_amd import scene as S  # noqa: E402
from diffstir_amd.engine import Engine  # noqa: E402
from diffstir_amd.loss import batch_loss  # noqa: E402

HBM_PEAK_GBS = 8000.0
# A HIP event pair on a launch (hipExtLaunchKernel) reads the kernel's execution plus the gap with
# which the dispatch follows the packet that records the start event (untimed launches follow each
# other with no gap): 4.0 us on MI355X, the median over the seven tiled kernels of (event sample -
# traced duration of the same dispatches) in profiles/event_timing_single_n64.md.
DISPATCH_GAP_US = 4.0
# the autotuner's choices per workload signature (diffstir_amd.autotune.signature), written by
# `bench.py --retune` on an MI355X and read by default runs, so that runs and profiles of the
# same command use the same kernels
TUNED_FILE = os.path.join(ROOT, "diffstir_amd", "tuned_options.json")
FWD_BYTES = 192.0            # FP32 x, v, C, F read once and written once per particle-step

# Algorithmic HBM bytes per particle (all species) of one launch of each tiled kernel: the
# particle data the kernel must read and write (FP32 state, FP64 J and SVD factor M).
KERNEL_BYTES_PER_PARTICLE = {
    "k_p2g_tiled": 96.0 + 8.0 + 36.0,                     # read x, v, C, F, J; write F
    "k_g2p_tiled": 12.0 + 36.0 + 60.0 + 8.0,              # read x, F; write x, v, C, J
    "k_g2p_tiled_factor": 12.0 + 36.0 + 60.0 + 8.0 + 80.0,  # ... and write J, M (the recompute)
    "k_g2p_adj_tiled": 12.0 + 60.0,                       # read x and the adjoints of x, v, C
    "k_p2g_adj_stress": 72.0 + 72.0 + 36.0 + 80.0 + 96.0,  # read C, F, gathered, adj F, J/M; write adj
    "k_g2p_p2g_tiled": 12.0 + 36.0 + 60.0 + 36.0,         # read x, F; write x, v, C, F (G2P + next P2G)
}
# the adjoint P2G gathers: read x, v, A and the adjoints of x, v, C; then either the stress half
# too (fused: read C, F, adj F, J/M, write the adjoint) or the 18 gathered floats (split)
ADJ_P2G_BYTES = {"fused": 120.0 + 188.0 + 96.0, "split": 120.0 + 72.0}


def fwd_adj_bytes(K: int) -> float:
    """Algorithmic bytes per fwd+adjoint particle-step with checkpoints every K substeps:
    forward 192 + segment recompute 192 + adjoint (read primal 96, read/write adjoint 192)
    + checkpoint store 96/K."""
    return 672.0 + 96.0 / K


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_targets(eng: Engine, scenes):
    """Target force profiles from the same instances at the ground-truth parameters."""
    gt = np.stack([S.hb_params_draw(np.random.default_rng(S.GROUND_TRUTH_SEED), s.params[3]) for s in scenes])
    eng.set_params(gt)
    eng.reset()
    eng.forward()
    tgt = eng.forces()
    return tgt


def candidate_params(scenes):
    return np.stack([S.hb_params_draw(np.random.default_rng(7 + b), s.params[3]) for b, s in enumerate(scenes)])


def grad_eval(eng: Engine, targets, kind):
    eng.reset()
    eng.forward()
    F = eng.forces()
    losses, g = batch_loss(F, targets, kind)
    d = eng.backward(g)
    return losses, d


def device_free_bytes() -> int:
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    hip.hipMemGetInfo(C.byref(free), C.byref(total))
    return int(free.value)


def run_gpu(workload: str, steps: int, warmup: int, ckpt: int, kernels: str = "tiled", timing_every: int = 10,
            autotune: bool = True, graph: bool = True, kernel_sample: bool = True, substeps: int = 0,
            overlap: bool = True, sort_every: Optional[int] = None, reduce_parts: Optional[int] = None,
            options: Optional[dict] = None, retune: bool = False, warm_start: int = 0):
    scenes = S.workload(workload)
    if warm_start:
        # counter runs: start from the positions the rollout reaches after `warm_start` substeps
        # (the stirred state the autotuner also times), velocities from rest
        from diffstir_amd.autotune import warmed_scenes
        with Engine(scenes, ckpt_every=ckpt) as tmp:
            scenes = warmed_scenes(scenes, tmp, warm_start)
        log(f"[bench] warm start: positions after {warm_start} substeps")
    if substeps:
        # a shortened rollout (profiling runs whose every launch is serialised by the profiler)
        scenes = [s.truncated(min(substeps + 1, s.max_steps)) for s in scenes]
    free0 = device_free_bytes()
    n_liq = sum(s.n_liquid for s in scenes)
    n_all = sum(s.n_liquid + s.n_stick + s.n_container for s in scenes)
    T = scenes[0].max_steps
    log(f"[bench] workload {workload}: {len(scenes)} instance(s), {n_liq} liquid particles, {T - 1} substeps, "
        f"checkpoint every {ckpt}")
    eng = Engine(scenes, ckpt_every=ckpt, kernels=kernels, overlap=overlap)
    tuned = None
    if options and kernels == "tiled":
        # fixed execution options (profiling runs: no tuner candidates among the traced launches)
        eng.configure(**options)
        log(f"[bench] fixed options: {options}")
    elif autotune and kernels == "tiled":
        from diffstir_amd.autotune import tune
        t0 = time.perf_counter()
        if retune:
            # a fresh choice, stored in the options file the default runs read
            tuned = tune(scenes, ckpt_every=ckpt, engine=eng)
            disk = {}
            if os.path.exists(TUNED_FILE):
                with open(TUNED_FILE) as f:
                    disk = json.load(f)
            disk[tuned["signature"]] = tuned
            with open(TUNED_FILE, "w") as f:
                json.dump(disk, f, indent=1)
        else:
            # the committed choice for this workload when there is one, else a fresh one
            tuned = tune(scenes, ckpt_every=ckpt, engine=eng, cache_file=TUNED_FILE)
            eng.configure(**tuned["options"])
        log(f"[bench] autotuned options ({time.perf_counter() - t0:.1f} s): {tuned['options']}")
    if sort_every is not None:
        eng.configure(sort_every=sort_every)
    if reduce_parts is not None:
        eng.configure(reduce_parts=reduce_parts)
    graph = graph and kernels == "tiled"
    eng.set_graph(graph)
    targets = make_targets(eng, scenes)
    eng.set_params(candidate_params(scenes))
    kind = scenes[0].cfg.loss_kind
    for i in range(warmup):
        grad_eval(eng, targets, kind)
        log(f"[bench] warmup {i + 1}/{warmup} done")
    eng.synchronize()
    fwd_ms = bwd_ms = 0.0
    t0 = time.perf_counter()
    for i in range(steps):
        losses, d = grad_eval(eng, targets, kind)
        f, b = eng.timing()
        fwd_ms += f
        bwd_ms += b
        log(f"[bench] step {i + 1}/{steps}: fwd {f:.1f} ms, adjoint {b:.1f} ms, loss[0] {losses[0]:.6g}")
    eng.synchronize()
    elapsed = time.perf_counter() - t0
    mem_used = free0 - device_free_bytes()
    if not np.all(np.isfinite(d)):
        raise RuntimeError("non-finite gradient")
    kt = {}
    if kernel_sample and kernels == "tiled":
        # one more evaluation, outside the timed region, launched kernel by kernel with every
        # timing_every-th substep's kernels timed by HIP events (graph replays are not sampled)
        eng.set_graph(False)
        eng.set_kernel_timing(timing_every)
        eng.kernel_timing(reset=True)
        grad_eval(eng, targets, kind)
        eng.synchronize()
        kt = eng.kernel_timing()
        eng.set_kernel_timing(0)
    psteps = float(n_liq) * (T - 1) * steps
    res = dict(workload=workload, scenes=scenes, n_liq=n_liq, n_all=n_all, T=T, elapsed=elapsed, psteps=psteps,
               fwd_ms=fwd_ms / steps, bwd_ms=bwd_ms / steps, B=len(scenes), kernels=kt,
               options=eng.options() if kernels == "tiled" else None, tuned=tuned, mem_used=mem_used, graph=graph,
               timing_every=timing_every)
    eng.close()
    return res


def plain_single(workload: str, sample_steps: int, ckpt: int) -> dict:
    """The plain one-instance GPU form: one thread per particle with global float atomics into a
    dense grid (the per-particle kernel family), timed on the first `sample_steps` substeps of
    the workload's first instance and scaled to its whole rollout."""
    sc = S.workload(workload)[0]
    T = sc.max_steps
    short = sc.truncated(sample_steps + 1)
    with Engine([short], ckpt_every=ckpt, kernels="per_particle") as e:
        for _ in range(2):   # the first run warms up
            e.reset()
            t0 = time.perf_counter()
            e.forward()
            F = e.forces()
            e.backward(np.ones_like(F))
            e.synchronize()
            el = time.perf_counter() - t0
    per_eval = el * (T - 1) / sample_steps
    return {"kernels": "per_particle", "workload": workload, "sample_substeps": sample_steps,
            "gradient_evals_per_s": 1.0 / per_eval,
            "fwd_adjoint_particle_steps_per_s": sc.n_liquid * sample_steps / el}


def cpu_baseline(workload: str, sample_steps: int, ckpt: int, one_core_steps: int = 4, check_workloads=()):
    """The reference's execution structure on the host: the FP32 restatement of the substep with
    one parallel loop over the particles of each species and atomic adds into a dense grid (as
    Taichi runs simulation_functions.py:104-290 on ti.cpu; oracle/mpm_oracle.c built with
    ORACLE_ATOMIC), OpenMP on the host's cores, on the first `sample_steps` substeps of the
    workload's first instance: the checkpointed forward + adjoint (the forward with checkpoints,
    each segment's recompute and the step adjoints), timed alone.  The same on one core over
    `one_core_steps` substeps.  The FP64 oracle then checks the GPU on the same sample (not timed):
    the tiled kernels run it and their relative errors are reported, for `workload` and for
    instance 0 of each (workload, substeps) of `check_workloads`."""
    from oracle.oracle import Oracle, loss_and_grad
    ref = Oracle("f32at")
    sc = S.workload(workload)[0].truncated(sample_steps + 1)
    threads = ref.set_threads(0)
    # dL/dforces of the sample's loss (target: half the forces), from an untimed forward
    F = ref.rollout(sc, nsteps=sample_steps)
    _, g = loss_and_grad(F, 0.5 * F, sc.cfg.loss_kind)
    t0 = time.perf_counter()
    ref.rollout_grad(sc, g, nsteps=sample_steps, ckpt=min(ckpt, sample_steps))
    el = time.perf_counter() - t0
    sc1 = S.workload(workload)[0].truncated(one_core_steps + 1)
    F1 = ref.rollout(sc1, nsteps=one_core_steps)
    _, g1 = loss_and_grad(F1, 0.5 * F1, sc1.cfg.loss_kind)
    ref.set_threads(1)
    t0 = time.perf_counter()
    ref.rollout_grad(sc1, g1, nsteps=one_core_steps, ckpt=min(ckpt, one_core_steps))
    el1 = time.perf_counter() - t0
    ref.set_threads(threads)
    out = dict(value=sc.n_liquid * sample_steps / el, unit="fwd+adjoint particle-steps/s", cores=threads,
               kind="port",
               sample=f"{workload} instance 0 ({sc.n_liquid} liquid particles), first {sample_steps} substeps, "
                      f"checkpointed forward + adjoint, FP32 with the reference's structure (per-particle "
                      f"parallel loops, atomic grid adds), {threads} OpenMP threads, {el:.1f} s",
               host_cpus=os.cpu_count(),
               one_core={"value": sc1.n_liquid * one_core_steps / el1, "substeps": one_core_steps,
                         "seconds": el1})
    o = Oracle("f64")
    Fo = o.rollout(sc, nsteps=sample_steps)
    _, go = loss_and_grad(Fo, 0.5 * Fo, sc.cfg.loss_kind)
    _, dth = o.rollout_grad(sc, go, nsteps=sample_steps, ckpt=min(ckpt, sample_steps))
    checks = {workload: gpu_vs_oracle(sc, Fo, dth, ckpt)}
    for wl, n in check_workloads:
        sb = S.workload(wl)[0].truncated(n + 1)
        Fb = o.rollout(sb, nsteps=n)
        _, gb = loss_and_grad(Fb, 0.5 * Fb, sb.cfg.loss_kind)
        _, db = o.rollout_grad(sb, gb, nsteps=n, ckpt=min(ckpt, n))
        checks[wl] = gpu_vs_oracle(sb, Fb, db, ckpt)
    out["gpu_vs_oracle"] = checks
    return out


def gpu_vs_oracle(sc, F_o, dth_o, ckpt) -> dict:
    """Relative errors of the GPU's stick forces and parameter gradient against the FP64 oracle's
    on the same truncated scene, with the same loss (target: half the oracle's forces)."""
    from diffstir_amd.loss import compute_loss
    T = sc.max_steps
    with Engine([sc], ckpt_every=min(ckpt, T - 1)) as e:
        e.reset()
        e.forward()
        Fg = e.forces()[0]
        _, gg = compute_loss(Fg, 0.5 * F_o, sc.cfg.loss_kind)
        dg = e.backward(gg[None])[0]
    fe = float(np.abs(Fg[2:] - F_o[2:]).max() / np.abs(F_o[2:]).max())
    ge = float(np.abs(dg - dth_o).max() / np.abs(dth_o).max())
    return {"substeps": T - 1, "liquid_particles": sc.n_liquid, "force_rel_err": fe, "grad_rel_err": ge,
            "tolerance": {"force": 1e-4, "grad": 1e-3}}


def roofline(r) -> dict:
    """The dominant tiled kernel (largest sampled GPU time): its algorithmic bytes per launch
    divided by its average launch time, both from this run.  The samples are HIP event pairs on
    the kernel's own launch (hipExtLaunchKernel on its stream), every `timing_every`-th substep of
    one launched evaluation after the timed region, each sampled substep starting behind a short
    hold of the stream so that the GPU never waits for the host.  A sample is the kernel's
    execution plus the dispatch gap that the event-recording packet causes, DISPATCH_GAP_US
    (measured on MI355X against the kernel trace: profiles/event_timing_single_n64.md); the
    kernel's time is the sample minus that gap."""
    bpp = dict(KERNEL_BYTES_PER_PARTICLE)
    if r.get("options"):
        bpp["k_p2g_adj_tiled"] = ADJ_P2G_BYTES[r["options"]["adjoint_p2g"]]
    kt = {k: v for k, v in r["kernels"].items() if k in bpp and v[1] > 0}
    if not kt:
        return {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None, "traffic": None}
    name = max(kt, key=lambda k: kt[k][0])
    tot_ms, n = kt[name]
    raw_us = tot_ms / n * 1e3
    avg_s = (raw_us - DISPATCH_GAP_US) / 1e6
    per_launch = bpp[name] * r["n_all"]
    achieved = per_launch / avg_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{r['workload']}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            t = json.load(f)
        if t.get("kernel") == name and t.get("options") == r.get("options"):
            traffic = t["bytes_per_launch"]
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": name,
            "avg_launch_us": avg_s * 1e6, "event_sample_us": raw_us, "dispatch_gap_us": DISPATCH_GAP_US,
            "algorithmic_bytes_per_launch": per_launch, "sampled_launches": n,
            "kernel_avg_us": {k: (v[0] / v[1] * 1e3 - DISPATCH_GAP_US if v[1] else None)
                              for k, v in r["kernels"].items()}}


def trace_summary(path: str):
    """Per-kernel average durations from a rocprofv3 kernel-trace CSV of this script: over all
    dispatches, and over the dispatches the kernel sampling timed with HIP events (the first
    dispatch of each kernel after each k_hold on the same queue), in microseconds."""
    import csv
    import re
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    short = lambda n: re.sub(r"^void |\(.*$", "", n).replace("ds::", "")
    allk, sampled = {}, {}
    seen = {}
    for r in rows:
        name, q = short(r["Kernel_Name"]), r["Queue_Id"]
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        if name == "k_hold":
            seen[q] = set()
            continue
        allk.setdefault(name, []).append(d)
        if q in seen and name not in seen[q]:
            seen[q].add(name)
            sampled.setdefault(name, []).append(d)
    out = {}
    for name, v in sorted(allk.items(), key=lambda kv: -sum(kv[1])):
        sv = sampled.get(name, [])
        out[name] = {"dispatches": len(v), "avg_us": sum(v) / len(v), "total_ms": sum(v) / 1e3,
                     "sampled": len(sv), "sampled_avg_us": (sum(sv) / len(sv)) if sv else None}
    print(json.dumps(out, indent=1))


def calibrate_copy(mib: int):
    """A plain device-to-device copy of `mib` MiB (three times), for scaling the FETCH_SIZE /
    WRITE_SIZE counters of a profiled run."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    hip.hipMemset(a, 1, C.c_size_t(n))
    for _ in range(3):
        hip.hipMemcpy(b, a, C.c_size_t(n), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": n, "copies": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="single_n64")
    ap.add_argument("--ckpt", type=int, default=50)
    ap.add_argument("--cpu-sample-steps", type=int, default=100)
    ap.add_argument("--cpu-one-core-steps", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernels", default="tiled", choices=["tiled", "per_particle"])
    ap.add_argument("--batch-workload", default="batch32_n64")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0)
    ap.add_argument("--trace-summary", default=None,
                    help="summarise a rocprofv3 kernel-trace CSV of this script (no GPU work) and exit")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-overlap", action="store_true",
                    help="run the adjoint's segment recompute on the main stream (no second stream)")
    ap.add_argument("--substeps", type=int, default=0,
                    help="shorten the rollout to this many substeps (profiling only; 0 = the workload's own)")
    ap.add_argument("--warm-start", type=int, default=0,
                    help="start from the positions the rollout reaches after this many substeps (counter runs "
                         "of a shortened rollout in the stirred state; 0 = the initial fill)")
    ap.add_argument("--sort-every", type=int, default=None,
                    help="sort the particles every this many substeps (a divisor of --ckpt; 0 = once per "
                         "checkpoint interval); overrides the autotuner's choice")
    ap.add_argument("--reduce-parts", type=int, default=None, choices=[1, 2, 4],
                    help="wavefronts per node block of the grid reductions; overrides the autotuner's choice")
    ap.add_argument("--retune", action="store_true",
                    help="run the autotuner even where diffstir_amd/tuned_options.json holds a choice, and store it")
    ap.add_argument("--options", default=None,
                    help="JSON of fixed execution options (Engine.configure keywords) instead of the autotuner")
    a = ap.parse_args()
    if a.trace_summary:
        trace_summary(a.trace_summary)
        return
    if a.calibrate_copy_mib:
        calibrate_copy(a.calibrate_copy_mib)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}): the instances of a batch share one device "
              "and nothing here uses a second GPU", file=sys.stderr)
        sys.exit(2)
    fixed = json.loads(a.options) if a.options else None
    r = run_gpu(a.workload, a.steps, a.warmup, a.ckpt, a.kernels, autotune=not a.no_autotune, graph=not a.no_graph,
                substeps=a.substeps, overlap=not a.no_overlap, sort_every=a.sort_every, reduce_parts=a.reduce_parts,
                options=fixed, retune=a.retune, warm_start=a.warm_start)
    value = r["psteps"] / r["elapsed"]
    ms_per_step = 1000.0 * r["elapsed"] / a.steps
    bpps = fwd_adj_bytes(a.ckpt)
    achieved = value * bpps / 1e9
    line = {
        "metric": "fwd+adjoint particle-steps/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 SVD, SVD backward, loss and gradient accumulation)",
        "data": "synthetic (seeded cylinder fill, circular stick path)",
        "config": {"workload": a.workload, "instances": r["B"], "liquid_particles": r["n_liq"],
                   "substeps": r["T"] - 1, "n_grid": r["scenes"][0].cfg.n_grid, "ckpt_every": a.ckpt,
                   "density": r["scenes"][0].cfg.m_liq / r["scenes"][0].cfg.vol_liq,
                   "kernels": a.kernels, "options": r["options"], "autotuned": r["tuned"] is not None,
                   "hip_graph": r["graph"]},
        "extra": {"gradient_evals_per_s": r["B"] * a.steps / r["elapsed"], "fwd_ms": r["fwd_ms"],
                  "adjoint_ms": r["bwd_ms"], "device_memory_GB": r["mem_used"] / 1e9,
                  "whole_evaluation_GBps": achieved,
                  "whole_evaluation_basis": f"{bpps:.1f} algorithmic B per fwd+adjoint particle-step"},
        "roofline": roofline(r),
    }
    if not a.no_batch:
        rb = run_gpu(a.batch_workload, 1, 1, a.ckpt, a.kernels, autotune=not a.no_autotune, graph=not a.no_graph,
                     kernel_sample=False, overlap=not a.no_overlap, retune=a.retune)
        line["extra"]["batched"] = {
            "workload": a.batch_workload, "instances": rb["B"], "liquid_particles": rb["n_liq"],
            "options": rb["options"],
            "gradient_evals_per_s": rb["B"] / rb["elapsed"],
            "fwd_adjoint_particle_steps_per_s": rb["psteps"] / rb["elapsed"],
            "ratio_to_single_instance_evals": (rb["B"] / rb["elapsed"]) / line["extra"]["gradient_evals_per_s"],
        }
        plain = plain_single(a.workload, 200, a.ckpt)
        line["extra"]["plain_single_instance"] = plain
        line["extra"]["batched"]["ratio_to_plain_single_instance_evals"] = (
            line["extra"]["batched"]["gradient_evals_per_s"] / plain["gradient_evals_per_s"])
    if not a.no_cpu_baseline:
        # the other workloads' first instances on short samples: the batched sweep, the
        # identification scene (Herschel-Bulkley, 3000 substeps) and the n_grid 128 instance
        checks = () if a.no_batch else ((a.batch_workload, a.cpu_sample_steps), ("ident16_n64", a.cpu_sample_steps),
                                        ("large_n128", 4))
        line["cpu_baseline"] = cpu_baseline(a.workload, a.cpu_sample_steps, a.ckpt, a.cpu_one_core_steps,
                                            check_workloads=checks)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
