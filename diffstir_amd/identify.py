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


def _xz_entries(F: np.ndarray) -> np.ndarray:
    """The force entries the x/z force MSE compares: x over rows 0..T-2, then z over the same."""
    T = F.shape[0]
    return np.concatenate([F[:T - 1, 0], F[:T - 1, 2]])


def identify_lm(scene: Scene, target: np.ndarray, init_params: np.ndarray, iterations: int = 20,
                fit_mask: Sequence[bool] = (True, True, True, False), rel_step: float = 3e-2,
                damping: float = 1e-3, max_step: float = 0.5, ckpt_every: int = 50, callback=None) -> dict:
    """Levenberg-Marquardt on log-parameters for R restarts of `scene` against one target profile
    (x/z force MSE, the loss of identify): every iteration runs, per restart, the trial point and
    its central-difference neighbours (step `rel_step` per fitted log-parameter) as instances of
    one Engine, so one batched forward gives each restart's residual and Jacobian.  A trial whose
    loss is below the restart's best is accepted (damping / 3), otherwise the damping grows x4 and
    the next trial is taken from the best point again; the step solves
    (J^T J + lambda diag(J^T J)) d = -J^T r and is scaled down to at most `max_step` in every
    log-parameter.  A batch in which some trial leaves the stable range (DS_ENONFINITE from the
    device) counts as rejected for every restart, and their step bounds halve (an accepted trial
    doubles its restart's bound again, up to `max_step`).

    Returns the per-iteration best losses [iterations+1, R], the best parameters
    [iterations+1, R, 4] and the final best parameters."""
    theta0 = np.asarray(init_params, dtype=np.float64).reshape(-1, 4)
    R = theta0.shape[0]
    idx = [i for i in range(4) if fit_mask[i]]
    m = len(idx)
    if np.any(theta0[:, idx] <= 0):
        raise ValueError("log-parameter fitting needs positive initial values for the fitted parameters")
    per = 1 + 2 * m
    tgt = _xz_entries(np.asarray(target, dtype=np.float64))
    trial = theta0.copy()
    best = [None] * R
    lam = np.full(R, float(damping))
    cap = np.full(R, float(max_step))
    losses, hist = [], []

    def propose(b):
        th, _, rb, Jb = best[b]
        A = Jb.T @ Jb
        d = -np.linalg.solve(A + lam[b] * np.diag(np.diag(A)), Jb.T @ rb)
        d *= min(1.0, cap[b] / max(np.max(np.abs(d)), 1e-300))
        nxt = th.copy()
        nxt[idx] = th[idx] * np.exp(d)
        trial[b] = nxt

    from ._lib import DS_ENONFINITE, DiffStirError
    with Engine([scene] * (R * per), ckpt_every=ckpt_every) as e:
        for it in range(iterations + 1):
            pts = []
            for b in range(R):
                pts.append(trial[b])
                for i in idx:
                    for sgn in (1.0, -1.0):
                        p = trial[b].copy()
                        p[i] *= np.exp(sgn * rel_step)
                        pts.append(p)
            e.set_params(np.stack(pts))
            e.reset()
            try:
                e.forward()
                F = e.forces()
            except DiffStirError as err:
                if err.code != DS_ENONFINITE or any(bb is None for bb in best):
                    raise
                F = None
            for b in range(R):
                if F is None:
                    lam[b] *= 4.0
                    cap[b] *= 0.5
                    propose(b)
                    continue
                T = F.shape[1]
                base = b * per
                r = _xz_entries(F[base]) - tgt
                L = float(np.sum(r * r) / (2 * (T - 1)))
                J = np.stack([(_xz_entries(F[base + 1 + 2 * k]) - _xz_entries(F[base + 2 + 2 * k])) / (2 * rel_step)
                              for k in range(m)], axis=1)
                if best[b] is None or L < best[b][1]:
                    best[b] = (trial[b].copy(), L, r, J)
                    lam[b] /= 3.0
                    cap[b] = min(max_step, 2.0 * cap[b])
                else:
                    lam[b] *= 4.0
                propose(b)
            losses.append([bb[1] for bb in best])
            hist.append(np.stack([bb[0] for bb in best]))
            if callback is not None:
                callback(it, np.array(losses[-1]), hist[-1])
    return {"losses": np.array(losses), "params": np.array(hist), "final": hist[-1]}


def sensitivity(scene: Scene, theta: Sequence[float], rel_step: float = 3e-2,
                fit_mask: Sequence[bool] = (True, True, True, False), noise: float = 0.01,
                ckpt_every: int = 50, scan: Optional[Sequence[float]] = None,
                points: Optional[Sequence[Sequence[float]]] = None) -> dict:
    """How well the x/z stick-force profile of `scene` determines the fitted parameters near
    `theta`: central differences (step `rel_step` in log-parameter space) of the force profile,
    all perturbed runs batched as instances of one Engine.  The step is wide because each force
    entry carries FP32 rounding of about 1e-7 of the largest force, which a step of 1e-3 turns into
    errors of several percent in single entries of J; at 3e-2 the truncation error of the central
    difference stays near 1e-3.

    With r the x/z force entries the loss compares (rows 0..T-2, as in loss.compute_loss for the
    force MSE) and J = dr/dphi (phi = log theta over the fitted parameters):

    * ``jacobian``: J, [2 (T-1), fitted] (x entries, then z entries);
    * ``gauss_newton``: J^T J / (T-1), the Hessian of the force-MSE loss at a zero residual;
    * ``eigvals``/``eigvecs``: its eigen-decomposition (ascending), so the last column is the best
      determined parameter combination and the first the worst;
    * ``std_log``: the least-squares estimate's standard deviation per fitted log-parameter when
      the target carries Gaussian noise of `noise` x the RMS force (synthetic_target's noise),
      sqrt(diag(sigma^2 (J^T J)^-1)); ``std_along_eig``: sigma / sqrt(eig(J^T J));
    * ``correlation``: the estimate's correlation matrix;
    * with `scan` (offsets t): ``scan_loss``, the loss against the noiseless profile at
      phi + t * (weakest eigenvector), with its quadratic prediction ``scan_quadratic``;
    * with `points` (fitted parameter triples): ``point_loss``, the loss of each against the
      noiseless profile;
    * ``scan_chi2`` / ``point_chi2``: those losses as sum((r - r0)^2) / sigma^2, the chi-square at
      which the noisy target tells the point from `theta` (about 1 or less: not at all).
    """
    theta = np.asarray(theta, dtype=np.float64).reshape(4)
    idx = [i for i in range(4) if fit_mask[i]]
    m = len(idx)
    pts = [theta.copy()]
    for i in idx:
        for sgn in (1.0, -1.0):
            p = theta.copy()
            p[i] *= np.exp(sgn * rel_step)
            pts.append(p)
    with Engine([scene] * len(pts), ckpt_every=ckpt_every) as e:
        e.set_params(np.stack(pts))
        e.reset()
        e.forward()
        F = e.forces()
        T = F.shape[1]

        def entries(f):
            return np.concatenate([f[:T - 1, 0], f[:T - 1, 2]])

        r0 = entries(F[0])
        J = np.stack([(entries(F[1 + 2 * k]) - entries(F[2 + 2 * k])) / (2 * rel_step) for k in range(m)], axis=1)
        JtJ = J.T @ J
        w, V = np.linalg.eigh(JtJ)
        rms = np.sqrt(np.mean(F[0][2:] ** 2)) if T > 2 else 1.0
        sigma = noise * rms
        cov = sigma ** 2 * np.linalg.inv(JtJ)
        sd = np.sqrt(np.diag(cov))
        out = {"theta": theta.tolist(), "fitted": idx, "rel_step": rel_step, "entries": int(r0.size),
               "rms_force": float(rms), "noise_sigma": float(sigma),
               "jacobian": J, "jacobian_norms": np.linalg.norm(J, axis=0).tolist(),
               "gauss_newton": (JtJ / (T - 1)).tolist(), "eigvals": (w / (T - 1)).tolist(),
               "eigvecs": V.tolist(), "condition": float(w[-1] / w[0]),
               "std_log": sd.tolist(), "std_along_eig": (sigma / np.sqrt(w)).tolist(),
               "correlation": (cov / np.outer(sd, sd)).tolist()}

        def losses_at(ps):
            # rerun in the same context, one instance per parameter set (padding with theta)
            out_l = []
            for c0 in range(0, len(ps), len(pts)):
                blk = list(ps[c0:c0 + len(pts)])
                e.set_params(np.stack(blk + [theta] * (len(pts) - len(blk))))
                e.reset()
                e.forward()
                Fs = e.forces()
                for b in range(len(blk)):
                    d = entries(Fs[b]) - r0
                    out_l.append(float(np.sum(d * d) / (2 * (T - 1))))
            return np.array(out_l)

        chi2 = 2 * (T - 1) / sigma ** 2
        if scan is not None:
            scan = np.asarray(scan, dtype=np.float64)
            sp = []
            for t in scan:
                p = theta.copy()
                p[idx] *= np.exp(t * V[:, 0])
                sp.append(p)
            L = losses_at(sp)
            out["scan_t"] = scan.tolist()
            out["scan_loss"] = L.tolist()
            out["scan_chi2"] = (chi2 * L).tolist()
            out["scan_quadratic"] = (0.5 * scan ** 2 * w[0] / (T - 1)).tolist()
        if points is not None:
            pp = []
            for q in points:
                p = theta.copy()
                p[idx] = np.asarray(q, dtype=np.float64)
                pp.append(p)
            L = losses_at(pp)
            out["points"] = [np.asarray(q, dtype=np.float64).tolist() for q in points]
            out["point_loss"] = L.tolist()
            out["point_chi2"] = (chi2 * L).tolist()
    return out


def main(argv=None):
    """python -m diffstir_amd.identify [--workload ident16_n64] [--iterations 100] [--lr 1e-2]
    Recover the ground-truth Herschel-Bulkley parameters of the workload's scene from its
    synthetic force profile (1% noise) with B random restarts batched; prints one JSON line.
    --method lm runs identify_lm instead of Adam, from the random restarts, from the ground truth
    (--start truth: where the noisy target's least-squares minimum lies) or from the final
    parameters of an earlier run's --out file (--start PATH)."""
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
    ap.add_argument("--sensitivity", action="store_true",
                    help="instead of fitting: the force profile's sensitivity to the parameters at the "
                         "ground truth, the predicted parameter error at --noise, and a loss scan along "
                         "the worst-determined parameter combination")
    ap.add_argument("--rel-step", type=float, default=3e-2, help="central-difference step of --sensitivity")
    ap.add_argument("--at", action="append", default=[],
                    help="with --sensitivity: also the loss of the noiseless profile at these fitted "
                         "parameters (comma-separated, one triple per --at)")
    ap.add_argument("--method", choices=("adam", "lm"), default="adam")
    ap.add_argument("--start", default="random", help="random, truth, or an earlier --out JSON file")
    a = ap.parse_args(argv)
    scenes = S.workload(a.workload)
    truth = np.asarray(scenes[0].params, dtype=np.float64)
    if a.sensitivity:
        t0 = time.perf_counter()
        res = sensitivity(scenes[0], truth, rel_step=a.rel_step, noise=a.noise, ckpt_every=a.ckpt,
                          scan=np.linspace(-0.6, 0.6, 13),
                          points=[[float(v) for v in q.split(",")] for q in a.at] or None)
        res["workload"] = a.workload
        res["seconds"] = time.perf_counter() - t0
        res.pop("jacobian")
        if a.out:
            with open(a.out, "w") as f:
                json.dump(res, f)
        print(json.dumps(res))
        return
    tgt, clean = synthetic_target(scenes[0], truth, noise=a.noise, ckpt_every=a.ckpt, return_clean=True)
    from .loss import compute_loss
    floor, _ = compute_loss(clean, tgt, scenes[0].cfg.loss_kind)
    rng = np.random.default_rng(11)
    inits = np.stack([S.hb_params_draw(rng, truth[3]) for _ in scenes])
    inits[:, :3] = np.maximum(inits[:, :3], 1e-2)
    if a.start == "truth":
        inits = truth[None].copy()
    elif a.start != "random":
        with open(a.start) as f:
            inits = np.asarray(json.load(f)["params"][-1], dtype=np.float64)

    def progress(it, L, th):
        print(f"[identify] iteration {it}: loss min {np.min(L):.4g} median {np.median(L):.4g}", flush=True)

    t0 = time.perf_counter()
    if a.method == "lm":
        # 4 restarts (28 instances) per Engine keep the checkpoint and segment buffers of a
        # 148k-particle, 3000-substep scene near 70 GB
        parts = [identify_lm(scenes[0], tgt, inits[g:g + 4], iterations=a.iterations, rel_step=a.rel_step,
                             ckpt_every=a.ckpt, callback=progress) for g in range(0, len(inits), 4)]
        res = {"losses": np.concatenate([r["losses"] for r in parts], axis=1),
               "params": np.concatenate([r["params"] for r in parts], axis=1)}
    else:
        res = identify(scenes[:len(inits)], tgt, inits, iterations=a.iterations, lr=a.lr, ckpt_every=a.ckpt,
                       lr_final=a.lr_final, callback=progress)
    el = time.perf_counter() - t0
    best = int(np.argmin(res["losses"][-1]))
    fin = res["params"][-1][best]
    R = len(inits)
    rate = ({"forward_evals_per_s": R * (1 + 2 * 3) * (a.iterations + 1) / el} if a.method == "lm"
            else {"gradient_evals_per_s": R * a.iterations / el, "lr": a.lr, "lr_final": a.lr_final})
    summary = {"workload": a.workload, "method": a.method, "start": a.start, "restarts": R,
               "iterations": a.iterations, "substeps": scenes[0].max_steps - 1,
               "s_per_iteration": el / a.iterations, **rate, "noise": a.noise, "noise_floor_loss": float(floor),
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
