"""Parameter fitting driven by the adjoint (src/gradient_descent_module.py).

`parameters_fitting` runs the reference's epoch loop: a taped rollout per epoch (forward of
every substep, compute_loss, adjoint), then each parameter is updated by the reference's
magnitude-matching step rule `single_parameter_fitting1` (gdm:11-42); the best epoch's
parameters are kept (gdm:124-130).  The reference returns from inside its clip loop after the
first clip (gdm:131-139), so only the first clip is fitted, with three times the epochs
(gdm:97-98); `parameters_fitting` does the same.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .simulator import Simulator


def magnitude_level(a: float) -> int:
    """gdm:3-9: floor(log10(max(|a|, 1e-11))), 0 for a NaN argument."""
    abs_a = max(abs(a), 1e-11)
    lg = math.log10(abs_a) if not math.isnan(abs_a) else float("nan")
    if math.isnan(lg):
        lg = 0
    return math.floor(lg)


def single_parameter_fitting1(parameter: float, parameter_grad: float, lr: float, lower_bound: float,
                              upper_bound: float, epochs: int, epoch_index: int) -> Tuple[float, float]:
    """gdm:11-42, with the parameter, its bounds and learning rate passed in (the reference reads
    them from module globals).  Returns (new parameter, new learning rate).

    A step of lr * grad is rescaled by powers of 3 so that its decimal magnitude matches the
    distance to the bound it heads for divided by the remaining epochs; out-of-bounds steps are
    halved until they land inside.  A zero or NaN gradient leaves both unchanged
    (abs(nan) > 0 is False, gdm:18)."""
    period = epochs - epoch_index
    base = 3
    if abs(parameter_grad) > 0:
        delta = lr * parameter_grad
        new_magnitude = 0
        if parameter_grad < 0:
            new_magnitude = magnitude_level((upper_bound - parameter) / period)
        if parameter_grad > 0:
            new_magnitude = magnitude_level((lower_bound - parameter) / period)
        if epoch_index < 1:
            lr *= pow(base, new_magnitude - magnitude_level(delta))
            delta = lr * parameter_grad
        parameter_update = parameter - delta
        if lower_bound < parameter_update < upper_bound:
            parameter = parameter_update
            lr *= pow(base, new_magnitude - magnitude_level(delta))
        else:
            if parameter_update < lower_bound:
                new_magnitude = magnitude_level((upper_bound - parameter) / period)
            if parameter_update > upper_bound:
                new_magnitude = magnitude_level((lower_bound - parameter) / period)
            lr *= pow(base, new_magnitude - magnitude_level(delta))
            parameter_update = parameter - lr * parameter_grad
            for _ in range(4096):
                if not (parameter_update > upper_bound or parameter_update < lower_bound):
                    break
                lr *= 0.5
                parameter_update = parameter - lr * parameter_grad
            parameter = parameter_update
    return parameter, lr


def parameters_fitting(sim: Simulator, epoch_setting: int, lower_bound: Sequence[float],
                       upper_bound: Sequence[float], verbose: bool = False,
                       n_fit: Optional[int] = None, first_clip: bool = True) -> Tuple[List[float], dict]:
    """gdm:84-139 for one clip: taped epochs with the step rule above, 3 * epoch_setting of them
    for the first clip (gdm:96-98), epoch_setting for later ones.

    Returns (best parameters, history with losses, parameters and gradients per epoch).
    `n_fit` limits the fitted parameters to the first n (Herschel-Bulkley leaves params[3]
    unused when the bulk modulus is fixed)."""
    if first_clip:
        sim.initialize_objects()
    n_par = 4 if n_fit is None else n_fit
    lr = [1.0] * 4
    grad_list = [0.0] * 4
    grad_shock = [0] * 4
    losses, params_hist, grads_hist = [], [], []
    epochs = epoch_setting * 3 if first_clip else epoch_setting
    lo = [float(v) for v in lower_bound]
    hi = [float(v) for v in upper_bound]
    for i in range(epochs):
        with sim.tape():
            sim.initialize_ti_field()
            sim.forward_steps()
            sim.compute_loss()
        losses.append(sim.loss)
        params_hist.append(sim.constitutive_parameters.copy())
        g = sim.constitutive_parameters_grad.copy()
        grads_hist.append(g)
        if verbose:
            print(f"epoch {i}: loss {sim.loss:.6g} params {sim.constitutive_parameters} grad {g}")
        for j in range(4):
            if grad_list[j] * g[j] < 0:
                grad_shock[j] = 1
            grad_list[j] = g[j]
        for j in range(n_par):
            p, lr[j] = single_parameter_fitting1(float(sim.constitutive_parameters[j]), float(grad_list[j]), lr[j],
                                                 lo[j], hi[j], epochs, i)
            sim.constitutive_parameters[j] = p
    best = int(np.argmin(losses))
    sim.constitutive_parameters[:] = params_hist[best]
    sim.loss = losses[best]
    return list(params_hist[best]), {"losses": losses, "params": params_hist, "grads": grads_hist, "best": best}


def fit_clips(sim: Simulator, trajectory: np.ndarray, forces: np.ndarray, n_clips: int, epoch_setting: int,
              lower_bound: Sequence[float], upper_bound: Sequence[float], verbose: bool = False,
              n_fit: Optional[int] = None) -> Tuple[List[float], List[dict]]:
    """The clip loop of gdm:86-139 as it evidently intends: clip c fits against frames
    [c F, (c + 1) F) of the trajectory and target (F = the simulator's max_steps), starting from
    the particle positions the previous clip's last rollout ended with (gdm:131-138), and hands
    its best parameters on.  The reference returns inside the loop after clip 0 (gdm:139) and
    slices the same frames for the next clip as for the current one (gdm:134-135); neither is
    reproduced here.  Returns (the last clip's best parameters, per-clip histories)."""
    F = sim.max_steps
    trajectory = np.asarray(trajectory)
    forces = np.asarray(forces, dtype=np.float64)
    if trajectory.shape[0] < n_clips * F or forces.shape[0] < n_clips * F:
        raise ValueError(f"{n_clips} clips of {F} frames need {n_clips * F} trajectory and target frames")
    hists = []
    best = list(sim.constitutive_parameters)
    for clip in range(n_clips):
        if clip > 0:
            sim.hand_off_clip(trajectory[clip * F:(clip + 1) * F], forces[clip * F:(clip + 1) * F])
        else:
            sim.operator_forces_seq = forces[:F]
        best, hist = parameters_fitting(sim, epoch_setting, lower_bound, upper_bound, verbose=verbose, n_fit=n_fit,
                                        first_clip=clip == 0)
        hists.append(hist)
    return best, hists
