"""Forward replay and force export (src/force_calculation_and_animation_module.py:3-21, 54-56).

`force_calculation_and_animation` runs int(presentation_time / dt - 1) substeps with the
current parameters (no tape), computes the loss, and writes the reference's output files:

    force_profile_x.npy, force_profile_z.npy   float32 (max_steps,): the stick force per step;
                                               rows 0 and max_steps-1 are 0, as in the
                                               reference (fcam:16-21, sf:330-336)
    stick_positions.npy                        float32 (steps, 3): mean stick position per step,
                                               without the last one (fcam:54-56)
    liquid_point_clouds.npy                    float32 (steps+1, N, 3), only with
                                               save_point_clouds=True (fcam:55)

The interactive Open3D animation and the blocking input() loop of the reference (fcam:22-78)
are not part of this package.
"""
from __future__ import annotations

import os
from typing import Optional

import numpy as np

from .simulator import Simulator


def force_calculation_and_animation(sim: Simulator, presentation_time: float, out_dir: Optional[str] = None,
                                    save_point_clouds: bool = False) -> dict:
    nsteps = int(presentation_time / sim.dt - 1)
    nsteps = max(0, min(nsteps, sim.max_steps - 1))
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
    out = {"loss": loss, "force_x": force_x, "force_z": force_z, "stick_positions": stick[:-1]}
    if clouds is not None:
        out["liquid_point_clouds"] = np.stack(clouds)
    if out_dir is not None:
        os.makedirs(out_dir, exist_ok=True)
        np.save(os.path.join(out_dir, "force_profile_x.npy"), force_x)
        np.save(os.path.join(out_dir, "force_profile_z.npy"), force_z)
        np.save(os.path.join(out_dir, "stick_positions.npy"), stick[:-1])
        if clouds is not None:
            np.save(os.path.join(out_dir, "liquid_point_clouds.npy"), out["liquid_point_clouds"])
    return out
