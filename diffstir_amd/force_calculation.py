"""Forward replay and force export (src/force_calculation_and_animation_module.py:3-21, 54-56).

`force_calculation_and_animation` runs int(presentation_time / dt - 1) substeps with the
current parameters (no tape), computes the loss, and writes the reference's output files:

    force_profile_x.npy, force_profile_z.npy   float32 (max_steps,): the stick force per step;
                                               rows 0 and max_steps-1 are 0, as in the
                                               reference (fcam:16-21, sf:330-336)
    stick_positions.npy                        float32 (max_steps - 1, 3): the mean stick position
                                               of frames 0..max_steps-2 (fcam:54,56 saves a
                                               max_steps-frame field without its last frame)
    liquid_point_clouds.npy                    float32 (max_steps, N, 3), the whole position field
                                               (fcam:13,55), only with save_point_clouds=True

Frames the replay does not reach stay 0, as in a freshly allocated field.  The reference's own
call, main.py:30 with presentation_time = (max_steps - 1) dt, runs max_steps - 2 substeps and so
fills every row of stick_positions.npy (its output/stick_positions.npy has 7999 rows for 8000
frames) and every frame of the point clouds but the last.

The interactive Open3D animation and the blocking input() loop of the reference (fcam:22-78)
are not part of this package.
"""
from __future__ import annotations

import os
from typing import Optional

import numpy as np

from .simulator import Simulator


def replay_steps(presentation_time: float, dt: float, max_steps: int) -> int:
    """Substeps of the replay loop, fcam:8: int(presentation_time / dt - 1), within the
    max_steps - 1 substeps the fields hold."""
    return max(0, min(int(presentation_time / dt - 1), max_steps - 1))


def export_frames(frames: np.ndarray, max_steps: int) -> np.ndarray:
    """The filled frames (nsteps + 1 of them) placed into a max_steps-frame field whose other
    frames are 0 (the shape of x.to_numpy() / operator_x.to_numpy(), fcam:13-14)."""
    frames = np.asarray(frames, dtype=np.float32)
    if frames.shape[0] > max_steps:
        raise ValueError(f"{frames.shape[0]} frames do not fit a {max_steps}-frame field")
    out = np.zeros((max_steps,) + frames.shape[1:], np.float32)
    out[:frames.shape[0]] = frames
    return out


def stick_positions(stick_centres: np.ndarray, max_steps: int) -> np.ndarray:
    """fcam:54,56: the per-frame mean stick position of a max_steps-frame field without its last
    frame, (max_steps - 1, 3)."""
    return export_frames(stick_centres, max_steps)[:-1]


def force_calculation_and_animation(sim: Simulator, presentation_time: float, out_dir: Optional[str] = None,
                                    save_point_clouds: bool = False) -> dict:
    nsteps = replay_steps(presentation_time, sim.dt, sim.max_steps)
    sim.initialize_objects()
    stick = np.zeros((nsteps + 1, 3), np.float32)
    clouds = [] if save_point_clouds else None
    stick[0] = sim.engine.read_stick(0)["x"].mean(0)
    if clouds is not None:
        clouds.append(sim.engine.read_liquid(0)["x"])
    for f in range(nsteps):
        sim.forward(f)
        stick[f + 1] = sim.engine.read_stick(0)["x"].mean(0)
        if clouds is not None:
            clouds.append(sim.engine.read_liquid(0)["x"])
    loss = sim.compute_loss()
    F = sim.total_operator_force.copy()
    # compute_loss accumulates steps 0..max_steps-2 only: the last row stays 0 (sf:332)
    F[sim.max_steps - 1:] = 0.0
    force_x = F[:, 0].astype(np.float32)
    force_z = F[:, 2].astype(np.float32)
    sp = stick_positions(stick, sim.max_steps)
    out = {"loss": loss, "force_x": force_x, "force_z": force_z, "stick_positions": sp, "steps": nsteps}
    if clouds is not None:
        out["liquid_point_clouds"] = export_frames(np.stack(clouds), sim.max_steps)
    if out_dir is not None:
        os.makedirs(out_dir, exist_ok=True)
        np.save(os.path.join(out_dir, "force_profile_x.npy"), force_x)
        np.save(os.path.join(out_dir, "force_profile_z.npy"), force_z)
        np.save(os.path.join(out_dir, "stick_positions.npy"), sp)
        if clouds is not None:
            np.save(os.path.join(out_dir, "liquid_point_clouds.npy"), out["liquid_point_clouds"])
    return out
