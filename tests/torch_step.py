"""An independent PyTorch FP64 restatement of one substep, used only to check the oracle's
hand-written adjoint with autograd (simulation_functions.py:39-42, 104-290).

Vectorised over particles with index_add scatters; the branches of the grid update are
torch.where selections on the computed values, so autograd differentiates the branch taken.
"""
from __future__ import annotations

import torch


def _stencil(x, inv_dx):
    X = x * inv_dx
    b = (X.detach() - 0.5).to(torch.int64)          # truncation (positions are > dx/2)
    fx = X - b.to(x.dtype)
    w = [0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1.0) ** 2, 0.5 * (fx - 0.5) ** 2]
    return b, fx, w


def _offsets():
    for i in range(3):
        for j in range(3):
            for k in range(3):
                yield i, j, k


def _viscosity(model, th, g):
    if model == 0:
        return th[1] + (th[0] - th[1]) * (1 + (th[3] * g) ** 2) ** ((th[2] - 1) / 2)
    if model == 1:
        return th[1] + (th[0] - th[1]) * (1 / (1 + (th[3] * g) ** (th[2] - 1)))
    de = th[0] * g ** (th[2] - 1)
    return th[1] / g + torch.where(de > 0, de, torch.zeros_like(de))


def step(cfg, th, st, traj_f, traj_f1):
    """One substep.  st: dict of tensors lx, lv, lC(n,3,3), lF, sx, sv, sC, cx, cv, cC.
    Returns (new state dict, sum of the stick's dv)."""
    n = cfg.n_grid
    dx = 1.0 / n
    inv_dx = float(n)
    dt = cfg.dt
    dtype = st["lx"].dtype
    P = torch.zeros(n * n * n, 4, dtype=dtype)
    eye = torch.eye(3, dtype=dtype)

    def scatter(P, x, m, u, A):
        b, fx, w = _stencil(x, inv_dx)
        for i, j, k in _offsets():
            off = torch.tensor([i, j, k], dtype=dtype)
            dpos = (off - fx) * dx
            W = w[i][:, 0] * w[j][:, 1] * w[k][:, 2]
            mom = W[:, None] * (m * u + (A @ dpos[:, :, None])[:, :, 0])
            idx = ((b[:, 0] + i) * n + (b[:, 1] + j)) * n + (b[:, 2] + k)
            P = P.index_add(0, idx, torch.cat([mom, (W * m)[:, None]], 1))
        return P

    C = st["lC"]
    Fn = (eye + dt * C) @ st["lF"]
    J = torch.linalg.det(Fn)
    gamma = torch.sqrt((C ** 2).sum((1, 2))) + 1e-6
    eta = _viscosity(cfg.model, th, gamma)
    k = th[3] if cfg.k_from_param else cfg.k_bulk
    stress = eta[:, None, None] * 0.5 * (C + C.transpose(1, 2)) \
        + eye * ((k / 2) * (J - 1 / J - 1) * J)[:, None, None]
    alpha = -dt * cfg.vol_liq * 4 * inv_dx * inv_dx
    A = alpha * stress + cfg.m_liq * C
    u = st["lv"] + torch.tensor([0.0, dt * cfg.gravity, 0.0], dtype=dtype)
    P = scatter(P, st["lx"], cfg.m_liq, u, A)
    if st["sx"].shape[0]:
        P = scatter(P, st["sx"], cfg.m_stick, st["sv"], cfg.m_stick * st["sC"])
    if st["cx"].shape[0]:
        P = scatter(P, st["cx"], cfg.m_cont, st["cv"], cfg.m_cont * st["cC"])

    # grid update
    ar = torch.arange(n)
    I, Jn, K = torch.meshgrid(ar, ar, ar, indexing="ij")
    I, Jn, K = I.reshape(-1), Jn.reshape(-1), K.reshape(-1)
    If, Kf = I.to(dtype), K.to(dtype)
    v = P[:, :3] / (P[:, 3:] + 1e-10)
    d = torch.stack([If * dx - cfg.wall_cx, torch.zeros(I.shape, dtype=dtype), Kf * dx - cfg.wall_cz], 1)
    dd = (d ** 2).sum(1)
    wall = dd > cfg.wall_r ** 2
    nrm = d / torch.sqrt(torch.where(wall, dd, torch.ones_like(dd)))[:, None]
    v = torch.where(wall[:, None], v - nrm * (v * nrm).sum(1, keepdim=True), v)
    zero = torch.zeros_like(v[:, 0])
    bd = cfg.bound
    v0 = torch.where((I < bd) & (v[:, 0] < 0), zero, v[:, 0])
    v0 = torch.where((I > n - bd) & (v0 > 0), zero, v0)
    v2 = torch.where((K < bd) & (v[:, 2] < 0), zero, v[:, 2])
    v2 = torch.where((K > n - bd) & (v2 > 0), zero, v2)
    v1 = torch.where((Jn < bd) & (v[:, 1] < 0), zero, v[:, 1])
    top = (Jn > n - bd) & (v1 > 0)
    V = torch.stack([torch.where(top, zero, v0), torch.where(top, zero, v1), torch.where(top, zero, v2)], 1)

    def gather(x):
        b, fx, w = _stencil(x, inv_dx)
        nv = torch.zeros(x.shape[0], 3, dtype=dtype)
        nC = torch.zeros(x.shape[0], 3, 3, dtype=dtype)
        for i, j, k in _offsets():
            off = torch.tensor([i, j, k], dtype=dtype)
            dp = off - fx
            W = w[i][:, 0] * w[j][:, 1] * w[k][:, 2]
            idx = ((b[:, 0] + i) * n + (b[:, 1] + j)) * n + (b[:, 2] + k)
            g = V[idx]
            nv = nv + W[:, None] * g
            nC = nC + 4 * inv_dx * W[:, None, None] * g[:, :, None] * dp[:, None, :]
        return nv, nC

    out = {}
    nv, nC = gather(st["lx"])
    out["lv"] = nv
    out["lx"] = st["lx"] + dt * nv
    out["lC"] = nC
    out["lF"] = Fn
    vk = (torch.as_tensor(traj_f1, dtype=dtype) - torch.as_tensor(traj_f, dtype=dtype)) / dt
    dvsum = torch.zeros(3, dtype=dtype)
    if st["sx"].shape[0]:
        sv, sC = gather(st["sx"])
        dvsum = (vk - sv).sum(0)
        out["sC"] = sC
        out["sx"] = st["sx"] + vk * dt
        out["sv"] = vk.expand_as(sv)
    if st["cx"].shape[0]:
        cv, cC = gather(st["cx"])
        out["cC"] = cC
        out["cx"] = st["cx"]
        out["cv"] = torch.zeros_like(cv)
    return out, dvsum
