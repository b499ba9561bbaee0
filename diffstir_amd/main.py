"""Model selection over the three rheologies, as in the reference's main.py.

    python -m diffstir_amd.main --traj data/traj.npy [--config config/constitutive_parameters_config.npy]
                                [--arguments config/arguments.json] [--epochs 35] [--frames N]
                                [--out output/]

For each model (Carreau, Cross, Herschel-Bulkley) the parameters start from the configuration
array (shape (3, 3, 4): initial values, lower bounds, upper bounds per model, main.py:5-14) and
are fitted with `parameters_fitting`; the fitted parameters and losses go to
parameters_and_losses_result.json (main.py:19-24); the best model is replayed and its force
profile exported (main.py:26-30).

The reference's loop assigns `constitutive_parameters[i]` with the model index i inside the
per-parameter loop (main.py:11-14), so only one entry per model is set; here entry j gets
config[0, i, j] as the loop evidently intends.
"""
from __future__ import annotations

import argparse
import json
import os

import numpy as np

from . import scene as S
from .force_calculation import force_calculation_and_animation
from .gradient_descent import parameters_fitting
from .simulator import Simulator


def model_selection(traj: np.ndarray, config: np.ndarray, epochs: int = 35, frames=None, arguments=None,
                    out_dir=None, seed: int = 0, ckpt_every: int = 50, verbose: bool = False) -> dict:
    config = np.asarray(config, dtype=np.float64).reshape(3, 3, 4)
    params_out, losses = [], []
    for i in range(3):
        sc = S.reference_scene(traj, seed=seed, arguments=arguments, frames=frames, params=config[0, i], model=i)
        sim = Simulator(sc, ckpt_every=ckpt_every)
        try:
            best, hist = parameters_fitting(sim, epochs, config[1, i], config[2, i], verbose=verbose)
        finally:
            sim.close()
        params_out.append([float(v) for v in best])
        losses.append(float(min(hist["losses"])))
    result = {"parameters": params_out, "losses": losses}
    best_model = int(np.argmin(losses))
    sc = S.reference_scene(traj, seed=seed, arguments=arguments, frames=frames, params=params_out[best_model],
                           model=best_model)
    sim = Simulator(sc, ckpt_every=ckpt_every)
    try:
        replay = force_calculation_and_animation(sim, (sc.max_steps - 1) * sc.cfg.dt, out_dir=out_dir)
    finally:
        sim.close()
    result["best_model"] = best_model
    if out_dir is not None:
        os.makedirs(out_dir, exist_ok=True)
        with open(os.path.join(out_dir, "parameters_and_losses_result.json"), "w") as f:
            json.dump({"parameters": params_out, "losses": losses}, f, indent=4)
    result["replay"] = replay
    return result


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--traj", required=True)
    ap.add_argument("--config", default=None)
    ap.add_argument("--arguments", default=None)
    ap.add_argument("--epochs", type=int, default=35)
    ap.add_argument("--frames", type=int, default=None)
    ap.add_argument("--out", default="output")
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    traj = np.load(a.traj, allow_pickle=False)
    cfg = np.load(a.config, allow_pickle=False) if a.config else np.ones((3, 3, 4))
    args = S.load_arguments(a.arguments) if a.arguments else None
    r = model_selection(traj, cfg, epochs=a.epochs, frames=a.frames, arguments=args, out_dir=a.out, seed=a.seed,
                        verbose=True)
    print(json.dumps({"parameters": r["parameters"], "losses": r["losses"], "best_model": r["best_model"]}))


if __name__ == "__main__":
    main()
