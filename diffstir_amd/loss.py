"""Losses over the stick-force series (host, FP64; the series is T x 3 numbers per instance).

kind 0 (LOSS_REF_L1_X) is the reference's loss (simulation_functions.py:330-339):
    loss = sum_{s=0}^{T-2} |F_x(s) - target(s)| / (T - 1)
with d|u|/du = sign(u) (0 at 0).  kind 1 (LOSS_MSE_XZ) is the mean squared error over x and z:
    loss = sum_{s=0}^{T-2} [(F_x - t_x)^2 + (F_z - t_z)^2] / (2 (T - 1)).
The last row (s = T-1) is outside both sums, as in the reference's loop bound.
"""
from __future__ import annotations

import numpy as np

from .scene import LOSS_MSE_XZ, LOSS_REF_L1_X


def compute_loss(forces: np.ndarray, target: np.ndarray, kind: int):
    """Loss and dL/dforces for one instance ([T, 3] forces; target [T, 3] or [T] = x only)."""
    F = np.asarray(forces, dtype=np.float64)
    T = F.shape[0]
    tgt = np.asarray(target, dtype=np.float64)
    if tgt.ndim == 1:
        t3 = np.zeros((T, 3))
        t3[:, 0] = tgt
        tgt = t3
    d = F[:T - 1] - tgt[:T - 1]
    g = np.zeros_like(F)
    if kind == LOSS_REF_L1_X:
        loss = np.abs(d[:, 0]).sum() / (T - 1)
        g[:T - 1, 0] = np.sign(d[:, 0]) / (T - 1)
    elif kind == LOSS_MSE_XZ:
        loss = ((d[:, 0] ** 2).sum() + (d[:, 2] ** 2).sum()) / (2 * (T - 1))
        g[:T - 1, 0] = d[:, 0] / (T - 1)
        g[:T - 1, 2] = d[:, 2] / (T - 1)
    else:
        raise ValueError(f"unknown loss kind {kind}")
    return float(loss), g


def batch_loss(forces: np.ndarray, targets: np.ndarray, kind: int):
    """Per-instance losses [B] and dL/dforces [B, T, 3] (the batch loss is their sum)."""
    B = forces.shape[0]
    losses = np.zeros(B)
    grads = np.zeros_like(np.asarray(forces, dtype=np.float64))
    for b in range(B):
        losses[b], grads[b] = compute_loss(forces[b], targets[b], kind)
    return losses, grads
