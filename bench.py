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
  cpu_baseline  the FP64 CPU oracle (single thread) on a bounded sample of the same workload
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

from diffstir_amd import scene as S  # noqa: E402
from diffstir_amd.engine import Engine  # noqa: E402
from diffstir_amd.loss import batch_loss  # noqa: E402

HBM_PEAK_GBS = 8000.0
FWD_BYTES = 192.0            # FP32 x, v, C, F read once and written once per particle-step


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


def run_gpu(workload: str, steps: int, warmup: int, ckpt: int, kernels: str = "tiled"):
    scenes = S.workload(workload)
    n_liq = sum(s.n_liquid for s in scenes)
    T = scenes[0].max_steps
    log(f"[bench] workload {workload}: {len(scenes)} instance(s), {n_liq} liquid particles, {T - 1} substeps, "
        f"checkpoint every {ckpt}")
    eng = Engine(scenes, ckpt_every=ckpt, kernels=kernels)
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
    if not np.all(np.isfinite(d)):
        raise RuntimeError("non-finite gradient")
    psteps = float(n_liq) * (T - 1) * steps
    res = dict(scenes=scenes, n_liq=n_liq, T=T, elapsed=elapsed, psteps=psteps, fwd_ms=fwd_ms / steps,
               bwd_ms=bwd_ms / steps, B=len(scenes))
    eng.close()
    return res


def cpu_baseline(workload: str, sample_steps: int, ckpt: int):
    """The FP64 oracle, one thread, on the first `sample_steps` substeps of the workload's first
    instance (forward + checkpointed adjoint)."""
    from oracle.oracle import Oracle, loss_and_grad
    o = Oracle("f64")
    sc = S.workload(workload)[0].truncated(sample_steps + 1)
    t0 = time.perf_counter()
    F = o.rollout(sc, nsteps=sample_steps)
    L, g = loss_and_grad(F, 0.5 * F, sc.cfg.loss_kind)
    o.rollout_grad(sc, g, nsteps=sample_steps, ckpt=min(ckpt, sample_steps))
    el = time.perf_counter() - t0
    # the timed region ran the forward twice (plain + inside rollout_grad); count one fwd+adj
    return dict(value=sc.n_liquid * sample_steps / el, unit="fwd+adjoint particle-steps/s", cores=1, kind="port",
                sample=f"{workload} instance 0, first {sample_steps} substeps, forward + checkpointed adjoint "
                       f"(plus one extra forward), FP64 oracle, {el:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="single_n64")
    ap.add_argument("--ckpt", type=int, default=50)
    ap.add_argument("--cpu-sample-steps", type=int, default=12)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernels", default="tiled", choices=["tiled", "per_particle"])
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}): the instances of a batch share one device "
              "and nothing here uses a second GPU", file=sys.stderr)
        sys.exit(2)
    r = run_gpu(a.workload, a.steps, a.warmup, a.ckpt, a.kernels)
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
                   "kernels": a.kernels},
        "extra": {"gradient_evals_per_s": r["B"] * a.steps / r["elapsed"], "fwd_ms": r["fwd_ms"],
                  "adjoint_ms": r["bwd_ms"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "basis": f"whole evaluation, {bpps:.1f} algorithmic B per fwd+adjoint particle-step"},
    }
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(a.workload, a.cpu_sample_steps, a.ckpt)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
