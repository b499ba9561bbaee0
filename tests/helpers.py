"""Shared helpers of the test suite: fixtures on disk, error metrics, scene rebuilds."""
from __future__ import annotations

import dataclasses
import os

import numpy as np

from diffstir_amd import scene as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name: str):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def ref_traj() -> np.ndarray:
    return np.load(os.path.join(GOLDEN, "ref_traj.npy"), allow_pickle=False)


def force_err(F: np.ndarray, Fref: np.ndarray, skip: int = 2) -> float:
    """max_s |F(s) - Fref(s)| / max_s |Fref(s)| over rows s >= skip (row 1 carries the stick's
    start-up spike, which would make any relative figure look small)."""
    d = np.abs(np.asarray(F)[skip:] - np.asarray(Fref)[skip:]).max()
    return float(d / max(np.abs(np.asarray(Fref)[skip:]).max(), 1e-300))


def rel_err(a, b) -> np.ndarray:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


def grad_err(g, gref) -> float:
    """Largest |g - gref| relative to the largest |gref| component (one scale per vector:
    components that are structurally tiny, like the bulk-modulus derivative, are judged on
    the vector's scale)."""
    g = np.asarray(g, dtype=np.float64)
    gref = np.asarray(gref, dtype=np.float64)
    return float(np.abs(g - gref).max() / max(np.abs(gref).max(), 1e-300))


def gold_scene(name: str) -> tuple:
    """Rebuild the scene of a gold_*.npz fixture and return (scene, fixture)."""
    z = golden(f"gold_{name}.npz")
    if name == "tiny_hb":
        sc = S.tiny_scene(16, 40, seed=3)
    elif name.startswith("ref_"):
        model = {"carreau": 0, "cross": 1, "hb": 2}[name[4:]]
        sc = S.reference_scene(ref_traj(), seed=0, frames=31, params=tuple(z["params"]), model=model)
    elif name == "plain_n32_20":
        sc = S.workload("plain_n32")[0].truncated(21)
    else:
        raise KeyError(name)
    if "x_liquid" in z.files:
        sc = dataclasses.replace(sc, x_liquid=z["x_liquid"], x_stick=z["x_stick"], x_container=z["x_container"],
                                 traj=z["traj"])
    sc = dataclasses.replace(sc, params=np.asarray(z["params"], dtype=np.float64))
    chk = np.array([sc.x_liquid.astype(np.float64).sum(), sc.x_stick.astype(np.float64).sum(),
                    sc.x_container.astype(np.float64).sum(), sc.traj.astype(np.float64).sum()])
    assert np.allclose(chk, z["checksum"], rtol=0, atol=1e-6 * np.abs(z["checksum"]).max()), \
        f"scene generator drifted from the {name} fixture"
    return sc, z


def reference_layout(root, frames: int, frames_per_clip: int = 10, force=None):
    """A directory in the reference repository's layout (config/arguments.json, data/traj.npy,
    data/force.npy) holding the first `frames` frames of the reference's data/traj.npy, for
    diffstir_amd.reference_api.load()."""
    import json
    a = dict(S.REFERENCE_ARGUMENTS)
    sim_keys = ("dt", "n_grid", "uniform_scaling_factor", "n_model_particles_generated",
                "n_operator_model_particles_generated", "n_container_model_particles_generated")
    doc = {"simulation_setting": {k: a[k] for k in sim_keys},
           "scene_information": {k: v for k, v in a.items() if k not in sim_keys}}
    doc["simulation_setting"]["frames_per_clip"] = frames_per_clip
    os.makedirs(os.path.join(root, "config"), exist_ok=True)
    os.makedirs(os.path.join(root, "data"), exist_ok=True)
    with open(os.path.join(root, "config", "arguments.json"), "w") as f:
        json.dump(doc, f)
    np.save(os.path.join(root, "data", "traj.npy"), ref_traj()[:frames])
    f = np.zeros(frames) if force is None else np.asarray(force, dtype=np.float64)[:frames]
    np.save(os.path.join(root, "data", "force.npy"), f)
    return str(root)
