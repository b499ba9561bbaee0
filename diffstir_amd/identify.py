"""Batched parameter identification: Adam in log-parameter space over B random restarts.

Every restart is one instance of a single Engine, so one forward + adjoint launch sequence
serves all restarts.  The target is a synthetic stick-force profile from the forward model at
ground-truth parameters, optionally with Gaussian noise of a given relative size.  The loss
per restart is the force MSE over x and z; the update is Adam on phi = log(theta) with
dL/dphi = theta * dL/dtheta.  The reference fits with its own heuristic step rule instead
(src/gradient_descent_module.py:11-42, see gradient_descent.parameters_fitting).
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from .engine import Engine
from .loss import batch_loss
from .scene import Scene


def synthetic_target(scene: Scene, truth: Sequence[float], noise: float = 0.01, seed: int = 0,
                     ckpt_every: int = 50, return_clean: bool = False):
    """Forces of `scene` at parameters `truth` plus Gaussian noise of `noise` x the RMS force
    (and the noiseless forces with return_clean)."""
    with Engine([scene], ckpt_every=ckpt_every) as e:
        e.set_params(np.asarray(truth, dtype=np.float64)[None])
        e.reset()
        e.forward()
        F = e.forces()[0]
    rms = np.sqrt(np.mean(F[2:] ** 2)) if F.shape[0] > 2 else 1.0
    rng = np.random.default_rng(seed)
    tgt = F + noise * rms * rng.normal(size=F.shape)
    return (tgt, F) if return_clean else tgt


def identify(scenes: Sequence[Scene], target: np.ndarray, init_params: np.ndarray, iterations: int = 100,
             lr: float = 1e-2, fit_mask: Sequence[bool] = (True, True, True, False), betas=(0.9, 0.999),
             eps: float = 1e-8, ckpt_every: int = 50, callback=None, engine: Optional[Engine] = None,
             lr_final: Optional[float] = None) -> dict:
    """Adam on log-parameters for B restarts (one instance each) against one target profile.
    With `lr_final` the step size follows a cosine from `lr` at the first iteration to `lr_final`
    at the last; otherwise it stays `lr`.

    Returns the per-iteration losses [iters, B], parameters [iters+1, B, 4] and the final
    parameters."""
    B = len(scenes)
    theta = np.asarray(init_params, dtype=np.float64).reshape(B, 4).copy()
    if np.any(theta[:, np.asarray(fit_mask)] <= 0):
        raise ValueError("log-parameter fitting needs positive initial values for the fitted parameters")
    mask = np.asarray(fit_mask, dtype=bool)
    phi = np.where(mask, np.log(np.where(theta > 0, theta, 1.0)), 0.0)
    m = np.zeros_like(phi)
    v = np.zeros_like(phi)
    targets = np.broadcast_to(np.asarray(target, dtype=np.float64), (B,) + np.asarray(target).shape)
    own = engine is None
    eng = engine if engine is not None else Engine(list(scenes), ckpt_every=ckpt_every)
    losses, hist = [], [theta.copy()]
    try:
        for it in range(iterations):
            eng.set_params(theta)
            eng.reset()
            eng.forward()
            F = eng.forces()
            L, g = batch_loss(F, targets, 1)
            d = eng.backward(g)
            gphi = np.where(mask, theta * d, 0.0)
            m = betas[0] * m + (1 - betas[0]) * gphi
            v = betas[1] * v + (1 - betas[1]) * gphi * gphi
            mh = m / (1 - betas[0] ** (it + 1))
            vh = v / (1 - betas[1] ** (it + 1))
            lr_it = lr
            if lr_final is not None and iterations > 1:
                lr_it = lr_final + 0.5 * (lr - lr_final) * (1.0 + np.cos(np.pi * it / (iterations - 1)))
            phi = phi - lr_it * mh / (np.sqrt(vh) + eps)
            theta = np.where(mask, np.exp(phi), theta)
            losses.append(L)
            hist.append(theta.copy())
            if callback is not None:
                callback(it, L, theta)
    finally:
        if own:
            eng.close()
    return {"losses": np.array(losses), "params": np.array(hist), "final": theta}


def main(argv=None):
    """python -m diffstir_amd.identify [--workload ident16_n64] [--iterations 100] [--lr 1e-2]
    Recover the ground-truth Herschel-Bulkley parameters of the workload's scene from its
    synthetic force profile (1% noise) with B random restarts batched; prints one JSON line."""
    import argparse
    import json
    import time
    from . import scene as S
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="ident16_n64")
    ap.add_argument("--iterations", type=int, default=100)
    ap.add_argument("--lr", type=float, default=1e-2)
    ap.add_argument("--lr-final", type=float, default=None,
                    help="cosine decay of the step size from --lr to this value over the iterations")
    ap.add_argument("--noise", type=float, default=0.01)
    ap.add_argument("--ckpt", type=int, default=50)
    ap.add_argument("--out", default="", help="also write the loss curve and parameter history (JSON)")
    a = ap.parse_args(argv)
    scenes = S.workload(a.workload)
    truth = np.asarray(scenes[0].params, dtype=np.float64)
    tgt, clean = synthetic_target(scenes[0], truth, noise=a.noise, ckpt_every=a.ckpt, return_clean=True)
    from .loss import compute_loss
    floor, _ = compute_loss(clean, tgt, scenes[0].cfg.loss_kind)
    rng = np.random.default_rng(11)
    inits = np.stack([S.hb_params_draw(rng, truth[3]) for _ in scenes])
    inits[:, :3] = np.maximum(inits[:, :3], 1e-2)
    t0 = time.perf_counter()
    res = identify(scenes, tgt, inits, iterations=a.iterations, lr=a.lr, ckpt_every=a.ckpt, lr_final=a.lr_final,
                   callback=lambda it, L, th: print(f"[identify] iteration {it}: loss min {np.min(L):.4g} "
                                                    f"median {np.median(L):.4g}", flush=True))
    el = time.perf_counter() - t0
    best = int(np.argmin(res["losses"][-1]))
    fin = res["params"][-1][best]
    summary = {"workload": a.workload, "restarts": len(scenes), "iterations": a.iterations,
               "substeps": scenes[0].max_steps - 1, "s_per_iteration": el / a.iterations,
               "gradient_evals_per_s": len(scenes) * a.iterations / el,
               "lr": a.lr, "lr_final": a.lr_final, "noise_floor_loss": float(floor),
               "loss_first_min": float(np.min(res["losses"][0])), "loss_last_min": float(np.min(res["losses"][-1])),
               "loss_first_median": float(np.median(res["losses"][0])),
               "loss_last_median": float(np.median(res["losses"][-1])),
               "truth": truth[:3].tolist(), "best_restart": fin[:3].tolist(),
               "relative_error": (np.abs(fin[:3] - truth[:3]) / truth[:3]).tolist()}
    if a.out:
        with open(a.out, "w") as f:
            json.dump(dict(summary, losses=res["losses"].tolist(), params=res["params"].tolist(),
                           inits=inits.tolist()), f)
    print(json.dumps(summary))


if __name__ == "__main__":
    main()
