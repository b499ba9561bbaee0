"""Writes the committed fixtures of tests/golden/.

    python tests/golden/make_fixtures.py [--reference /root/reference]

Two kinds of fixture:

* reference data (copied from the reference's data/ and output/ when the reference is
  present; these are data files, inputs and outputs of the reference, no source):
    ref_traj.npy               data/traj.npy, the raw stick recording (8000 x 3 float64)
    ref_stick_positions.npz    every 5th row of output/stick_positions.npy (the reference's
                               stick-centre series, force_calculation_and_animation_module.py:54-56)
    ref_force_profile_ends.npz first/last 16 entries of output/force_profile_{x,z}.npy
* oracle vectors (FP64 CPU oracle, oracle/mpm_oracle.c, on seeded scenes): gold_*.npz with the
  inputs needed to rerun them, the forces, a loss target, the loss and dL/dparams.
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from diffstir_amd import scene as S  # noqa: E402
from oracle.oracle import Oracle, loss_and_grad  # noqa: E402


def copy_reference(ref: str):
    traj = np.load(os.path.join(ref, "data", "traj.npy"))
    np.save(os.path.join(HERE, "ref_traj.npy"), traj)
    sp = np.load(os.path.join(ref, "output", "stick_positions.npy"))
    np.savez_compressed(os.path.join(HERE, "ref_stick_positions.npz"), rows=np.arange(0, sp.shape[0], 5),
                        pos=sp[::5], n_rows=sp.shape[0])
    fx = np.load(os.path.join(ref, "output", "force_profile_x.npy"))
    fz = np.load(os.path.join(ref, "output", "force_profile_z.npy"))
    np.savez_compressed(os.path.join(HERE, "ref_force_profile_ends.npz"), x_head=fx[:16], x_tail=fx[-16:],
                        z_head=fz[:16], z_tail=fz[-16:], n=fx.shape[0])


def target_for(F: np.ndarray, seed: int) -> np.ndarray:
    """A loss target near the forces: 0.8 F plus noise of 10% of the force scale."""
    rng = np.random.default_rng(seed)
    scale = np.abs(F[2:]).mean() if F.shape[0] > 2 else 1.0
    return 0.8 * F + rng.normal(size=F.shape) * 0.1 * scale


def gold(o: Oracle, name: str, sc: S.Scene, nsteps: int, ckpt: int, seed: int, save_inputs: bool):
    F = o.rollout(sc, nsteps=nsteps)
    tgt = target_for(F, seed)
    L, g = loss_and_grad(F, tgt, sc.cfg.loss_kind)
    F2, thb = o.rollout_grad(sc, g, nsteps=nsteps, ckpt=ckpt)
    assert np.array_equal(F, F2)
    out = dict(forces=F, target=tgt, loss=L, dl_dparams=thb, params=sc.params, nsteps=nsteps,
               loss_kind=sc.cfg.loss_kind, model=sc.cfg.model,
               checksum=np.array([sc.x_liquid.astype(np.float64).sum(), sc.x_stick.astype(np.float64).sum(),
                                  sc.x_container.astype(np.float64).sum(), sc.traj.astype(np.float64).sum()]))
    if save_inputs:
        out.update(x_liquid=sc.x_liquid, x_stick=sc.x_stick, x_container=sc.x_container, traj=sc.traj)
    np.savez_compressed(os.path.join(HERE, f"gold_{name}.npz"), **out)
    print(f"gold_{name}: loss {L:.6g} grad {thb}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    a = ap.parse_args()
    if os.path.isdir(a.reference):
        copy_reference(a.reference)
    o = Oracle("f64")
    gold(o, "tiny_hb", S.tiny_scene(16, 40, seed=3), 40, 8, 11, True)
    traj = np.load(os.path.join(HERE, "ref_traj.npy"))
    for model, mname in [(0, "carreau"), (1, "cross"), (2, "hb")]:
        params = {0: (2.0, 0.5, 0.7, 3.0), 1: (2.0, 0.5, 0.6, 2.0), 2: (1.5, 2.0, 0.8, 1.0)}[model]
        sc = S.reference_scene(traj, seed=0, frames=31, params=params, model=model)
        gold(o, f"ref_{mname}", sc, 30, 10, 12 + model, model == 2)
    sc = S.workload("plain_n32")[0].truncated(21)
    gold(o, "plain_n32_20", sc, 20, 10, 15, False)


if __name__ == "__main__":
    main()
