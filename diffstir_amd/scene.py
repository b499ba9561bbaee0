"""Scenes for the stirring simulator: physical constants, particle fills and stick paths.

A `Scene` is one simulation instance: a `SceneConfig` (grid, time step, rheology, wall,
species masses) plus the initial particle positions of the three species and the stick
trajectory.  Two families are provided:

* `reference_scene` restates the reference's scene set-up (src/simulator/initialization.py):
  constants from config/arguments.json (init:14-54), the trajectory transform of
  data/traj.npy (init:74-78) and rejection-sampled container, liquid and stick particles
  (init:79-117) -- seeded here, where the reference draws from the unseeded global RNG.
* `stirring_scene` builds the synthetic workloads: a jittered 2x2x2-per-cell cylinder fill,
  a particle-sampled vertical stick on a circle / ellipse path with optional vertical
  oscillation, no container particles (the cylindrical slip wall of the grid update holds
  the liquid), Herschel-Bulkley rheology with the bulk modulus as the fourth parameter.

Arrays are float32 in the layout the HIP library takes (particles x 3, row-major).
"""
from __future__ import annotations

import dataclasses
import json
import math
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

MODEL_CARREAU = 0
MODEL_CROSS = 1
MODEL_HERSCHEL_BULKLEY = 2
MODEL_NAMES = {MODEL_CARREAU: "carreau", MODEL_CROSS: "cross", MODEL_HERSCHEL_BULKLEY: "herschel_bulkley"}

LOSS_REF_L1_X = 0   # sum_s |F_x(s) - target_x(s)| / (T-1), s = 0..T-2     (sf:330-339)
LOSS_MSE_XZ = 1     # sum_s [(F_x - t_x)^2 + (F_z - t_z)^2] / (2 (T-1)), s = 0..T-2


@dataclass
class SceneConfig:
    """Constants of one simulation instance (the reference bakes these in at import, init:14-54)."""

    n_grid: int = 32
    dt: float = 2e-4
    bound: int = 3                       # init:15
    gravity: float = -9.81               # sf:152 (times uniform_scaling_factor)
    model: int = MODEL_HERSCHEL_BULKLEY
    k_bulk: float = 20000.0              # sf:112 hard-codes k = 20000
    k_from_param: bool = False           # True: k = params[3] (Herschel-Bulkley leaves it unused)
    wall_r: float = 0.0415 - 0.5 / 32    # sf:205: r = 0.0415 - 0.5*dx
    wall_cx: float = 0.5                 # sf:206
    wall_cz: float = 0.5
    m_liq: float = 1.0
    vol_liq: float = 1.0
    m_stick: float = 1.0
    m_cont: float = 1.0
    loss_kind: int = LOSS_MSE_XZ

    @property
    def dx(self) -> float:
        return 1.0 / self.n_grid

    @property
    def inv_dx(self) -> float:
        return float(self.n_grid)

    def as_dict(self) -> dict:
        return dataclasses.asdict(self)


@dataclass
class Scene:
    cfg: SceneConfig
    x_liquid: np.ndarray                 # (N, 3) float32
    x_stick: np.ndarray                  # (Ns, 3) float32
    x_container: np.ndarray              # (Nc, 3) float32
    traj: np.ndarray                     # (T, 3) float32: stick reference point per frame
    params: np.ndarray = field(default_factory=lambda: np.zeros(4))   # (4,) float64
    name: str = "scene"

    @property
    def n_liquid(self) -> int:
        return int(self.x_liquid.shape[0])

    @property
    def n_stick(self) -> int:
        return int(self.x_stick.shape[0])

    @property
    def n_container(self) -> int:
        return int(self.x_container.shape[0])

    @property
    def max_steps(self) -> int:
        """Frames in the trajectory; a rollout advances max_steps - 1 substeps (gdm:105)."""
        return int(self.traj.shape[0])

    def truncated(self, frames: int) -> "Scene":
        """The same scene with only the first `frames` trajectory frames."""
        return dataclasses.replace(self, traj=np.ascontiguousarray(self.traj[:frames]))


# ----------------------------------------------------------------------------- reference scene

REFERENCE_ARGUMENTS = {
    # config/arguments.json:2-23 (the learning_setting block is never read by the reference)
    "dt": 1e-3, "n_grid": 71, "uniform_scaling_factor": 1,
    "n_model_particles_generated": 6500,
    "n_operator_model_particles_generated": 500,
    "n_container_model_particles_generated": 500,
    "container_diameter": 0.095, "container_height": 0.18, "container_wall_thickness": 0.003,
    "liquid_height": 0.125, "operator_bottom_height": 0.02, "operator_radius": 0.0125,
    "operator_length": 0.2, "liquid_density": 1650, "operator_density": 770, "container_density": 1000,
}


def load_arguments(path: Optional[str]) -> dict:
    """Flatten an arguments.json (simulation_setting + scene_information) or return the defaults."""
    if path is None:
        return dict(REFERENCE_ARGUMENTS)
    with open(path) as f:
        a = json.load(f)
    out = dict(a["simulation_setting"])
    out.update(a["scene_information"])
    return out


def transform_trajectory(raw: np.ndarray):
    """init:74-78: swap y and z, centre on the bounding-box midpoint.

    Returns (trajectory, initial_pos_related_to_center).  The reference subtracts the centre
    a second time from the already-centred first frame (init:78); that offset places the
    stick particles (init:117) and is kept as is."""
    t = np.asarray(raw, dtype=np.float64) @ np.array([[1, 0, 0], [0, 0, 1], [0, 1, 0]], dtype=np.float64)
    center = (np.amax(t, axis=0) - np.amin(t, axis=0)) / 2
    center = center + np.amin(t, axis=0)
    t = t - center
    init_rel = t[0] - center
    return t, init_rel


def reference_scene(traj: np.ndarray, seed: int = 0, arguments: Optional[dict] = None,
                    frames: Optional[int] = None, params: Sequence[float] = (1.0, 1.0, 1.0, 1.0),
                    model: int = MODEL_HERSCHEL_BULKLEY) -> Scene:
    """The reference's scene (init:14-117) with a seeded generator.

    `traj` is the raw (T, 3) recording (data/traj.npy).  The loss defaults to the reference's
    L1 on the x force (sf:330-339)."""
    a = dict(REFERENCE_ARGUMENTS if arguments is None else arguments)
    rng = np.random.default_rng(seed)
    dt = float(a["dt"]); n_grid = int(a["n_grid"])
    dx = 1.0 / n_grid
    bound = 3
    n_l = int(a["n_model_particles_generated"])
    n_c = int(a["n_container_model_particles_generated"])
    n_o = int(a["n_operator_model_particles_generated"])
    scale = float(a["uniform_scaling_factor"])
    c_rad = a["container_diameter"] / 2
    c_h = a["container_height"]; wt = a["container_wall_thickness"]
    l_h = a["liquid_height"]; o_r = a["operator_radius"]; o_len = a["operator_length"]
    o_bot = a["operator_bottom_height"]
    l_rad = c_rad - wt                                                   # init:41
    o_vol = o_r ** 2 * o_len * np.pi / n_o                              # init:42
    l_vol = l_rad ** 2 * l_h * np.pi / n_l                              # init:43
    o_mass = a["operator_density"] * o_vol                              # init:45
    l_mass = a["liquid_density"] * l_vol                                # init:46
    area_bottom = np.pi * c_rad ** 2                                    # init:47-51
    area_wall = a["container_diameter"] * 2 * np.pi * c_h
    c_area = area_wall + area_bottom
    n_wall = int(n_c * (area_wall / c_area))
    n_bottom = n_c - n_wall
    c_vol = (c_area * wt) / n_c                                         # init:52-53
    c_mass = c_vol * a["container_density"]

    t, init_rel = transform_trajectory(traj)

    def sample(n, rmin2, rmax2, radius, ylo, yhi):
        # init:80-112: uniform rejection sampling in an annulus/disc times a y-interval
        out = np.empty((n, 3))
        i = 0
        while i < n:
            m = max(64, 2 * (n - i))
            xz = rng.uniform(-radius, radius, size=(m, 2))
            y = rng.uniform(ylo, yhi, size=m)
            r2 = (xz ** 2).sum(1)
            ok = (r2 >= rmin2) & (r2 <= rmax2)
            k = min(int(ok.sum()), n - i)
            sel = np.nonzero(ok)[0][:k]
            out[i:i + k, 0] = xz[sel, 0]
            out[i:i + k, 1] = y[sel]
            out[i:i + k, 2] = xz[sel, 1]
            i += k
        return out

    cont = np.concatenate([
        sample(n_wall, (c_rad - wt) ** 2, c_rad ** 2, c_rad, dx + wt, dx + c_h + wt),
        sample(n_bottom, 0.0, c_rad ** 2, c_rad, dx, dx + wt),
    ]) if n_c > 0 else np.zeros((0, 3))
    liq = sample(n_l, 0.0, l_rad ** 2, l_rad, 2 * dx + wt, l_h + wt + 2 * dx)
    op = sample(n_o, 0.0, o_r ** 2, o_r, dx + o_bot + wt, o_len + o_bot + dx + wt)
    place = np.array([0.5, (bound - 1) * dx, 0.5])                      # init:113-117
    cont = cont + place
    liq = liq + place
    op = op + place + init_rel
    if frames is not None:
        t = t[:frames]
    cfg = SceneConfig(n_grid=n_grid, dt=dt, bound=bound, gravity=-9.81 * scale, model=model,
                      k_bulk=20000.0, k_from_param=False, wall_r=0.0415 - 0.5 * dx,
                      wall_cx=0.5, wall_cz=0.5, m_liq=l_mass, vol_liq=l_vol,
                      m_stick=o_mass, m_cont=c_mass, loss_kind=LOSS_REF_L1_X)
    return Scene(cfg=cfg, x_liquid=liq.astype(np.float32), x_stick=op.astype(np.float32),
                 x_container=cont.astype(np.float32), traj=t.astype(np.float32),
                 params=np.asarray(params, dtype=np.float64), name="reference")


# ----------------------------------------------------------------------------- synthetic stirring

def jittered_cylinder(rng: np.random.Generator, n_grid: int, cx: float, cz: float, radius: float,
                      y0: float, y1: float) -> np.ndarray:
    """Particles on a 2x2x2-per-cell lattice, each jittered uniformly inside its sub-cell,
    kept where the point lies in the vertical cylinder (cx, cz, radius) x [y0, y1).
    Returned in lattice order (x-major), which is also grid-block order."""
    dx = 1.0 / n_grid
    h = 0.5 * dx
    i0 = max(int(math.floor((cx - radius) / h)) - 1, 0)
    i1 = int(math.ceil((cx + radius) / h)) + 1
    k0 = max(int(math.floor((cz - radius) / h)) - 1, 0)
    k1 = int(math.ceil((cz + radius) / h)) + 1
    j0 = max(int(math.floor(y0 / h)) - 1, 0)
    j1 = int(math.ceil(y1 / h)) + 1
    I, J, K = np.meshgrid(np.arange(i0, i1), np.arange(j0, j1), np.arange(k0, k1), indexing="ij")
    sub = np.stack([I.ravel(), J.ravel(), K.ravel()], axis=1).astype(np.float64)
    pts = (sub + rng.uniform(0.0, 1.0, size=sub.shape)) * h
    r2 = (pts[:, 0] - cx) ** 2 + (pts[:, 2] - cz) ** 2
    keep = (r2 <= radius * radius) & (pts[:, 1] >= y0) & (pts[:, 1] < y1)
    return pts[keep]


def stick_path(frames: int, dt: float, radius_x: float, radius_z: float, omega: float,
               y_amp: float = 0.0, y_freq: float = 0.0, cx: float = 0.5, cz: float = 0.5) -> np.ndarray:
    """Stick reference point per frame: an ellipse in the xz-plane (a circle when the radii are
    equal) traversed at angular speed omega, plus an optional vertical oscillation."""
    t = np.arange(frames, dtype=np.float64) * dt
    out = np.zeros((frames, 3))
    out[:, 0] = cx + radius_x * np.cos(omega * t)
    out[:, 1] = y_amp * np.sin(2 * np.pi * y_freq * t)
    out[:, 2] = cz + radius_z * np.sin(omega * t)
    return out


def hb_params_draw(rng: np.random.Generator, k_bulk: float = 1e3) -> np.ndarray:
    """Herschel-Bulkley [K, tau_y, n, k_bulk]: tau_y ~ U[0,20], K ~ U[0.1,5], n ~ U[0.5,1.5]."""
    tau_y = rng.uniform(0.0, 20.0)
    K = rng.uniform(0.1, 5.0)
    n = rng.uniform(0.5, 1.5)
    return np.array([K, tau_y, n, k_bulk], dtype=np.float64)


GROUND_TRUTH_SEED = 1234


def stirring_scene(n_grid: int = 32, substeps: int = 1000, dt: float = 2e-4, fill_radius: float = 0.3,
                   fill_height: float = 0.25, stick_radius: float = 0.03, path_rx: float = 0.15,
                   path_rz: float = 0.15, omega: float = 2 * math.pi, y_amp: float = 0.0,
                   y_freq: float = 0.0, density: float = 1000.0, stick_density: float = 1000.0,
                   k_bulk: float = 1e3, params: Optional[Sequence[float]] = None, seed: int = 0,
                   name: str = "stirring") -> Scene:
    """A cylinder of liquid (8 particles per cell) stirred by a particle-sampled vertical stick.

    The fill's bottom sits at y = bound*dx; the cylindrical slip wall is 2 cells outside the
    fill; the stick spans from 2 cells above the floor to 4 cells above the fill's top, raised by
    the vertical oscillation's amplitude so that its lowest point stays 2 cells above the floor."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / n_grid
    bound = 3
    y0 = bound * dx
    liq = jittered_cylinder(rng, n_grid, 0.5, 0.5, fill_radius, y0, y0 + fill_height)
    frames = substeps + 1
    traj = stick_path(frames, dt, path_rx, path_rz, omega, y_amp, y_freq)
    s0 = traj[0]
    lift = abs(y_amp)
    stick = jittered_cylinder(rng, n_grid, s0[0], s0[2], stick_radius, y0 + 2 * dx + lift,
                              y0 + fill_height + 4 * dx + lift)
    stick[:, 1] += s0[1]
    p_vol = dx ** 3 / 8.0
    if params is None:
        params = hb_params_draw(np.random.default_rng(GROUND_TRUTH_SEED), k_bulk)
    cfg = SceneConfig(n_grid=n_grid, dt=dt, bound=bound, gravity=-9.81, model=MODEL_HERSCHEL_BULKLEY,
                      k_bulk=k_bulk, k_from_param=True, wall_r=fill_radius + 2 * dx, wall_cx=0.5,
                      wall_cz=0.5, m_liq=density * p_vol, vol_liq=p_vol, m_stick=stick_density * p_vol,
                      m_cont=0.0, loss_kind=LOSS_MSE_XZ)
    return Scene(cfg=cfg, x_liquid=liq.astype(np.float32), x_stick=stick.astype(np.float32),
                 x_container=np.zeros((0, 3), np.float32), traj=traj.astype(np.float32),
                 params=np.asarray(params, dtype=np.float64), name=name)


def workload(name: str, seed: int = 0) -> List[Scene]:
    """The named workloads (one Scene per batched instance).

    plain_n32      n_grid 32, r 0.3, h 0.25, about 18.5k particles, 1000 substeps
    single_n64     n_grid 64, about 148k particles, 2000 substeps
    batch32_n64    32 instances at n_grid 64, h ~ U[0.15,0.30], omega ~ U[pi,4pi], own parameters
    ident16_n64    16 restarts at n_grid 64, 3000 substeps
    large_n128     n_grid 128, r 0.35, h 0.3, stick r 0.02 on a 0.2/0.1 ellipse with a
                   0.05 vertical oscillation at 1 Hz, 4000 substeps
    """
    if name == "plain_n32":
        return [stirring_scene(32, 1000, seed=seed, name=name)]
    if name == "single_n64":
        return [stirring_scene(64, 2000, seed=seed, name=name)]
    if name == "batch32_n64":
        out = []
        for b in range(32):
            r = np.random.default_rng(seed * 1000 + b)
            h = r.uniform(0.15, 0.30)
            om = r.uniform(math.pi, 4 * math.pi)
            out.append(stirring_scene(64, 2000, fill_height=h, omega=om, params=hb_params_draw(r),
                                      seed=seed * 1000 + b, name=f"{name}[{b}]"))
        return out
    if name == "ident16_n64":
        return [stirring_scene(64, 3000, seed=seed, name=f"{name}[{b}]") for b in range(16)]
    if name == "large_n128":
        return [stirring_scene(128, 4000, fill_radius=0.35, fill_height=0.3, stick_radius=0.02,
                               path_rx=0.2, path_rz=0.1, y_amp=0.05, y_freq=1.0, seed=seed, name=name)]
    raise KeyError(name)


def tiny_scene(n_grid: int = 16, substeps: int = 20, seed: int = 0, fill_radius: float = 0.2,
               fill_height: float = 0.2, params: Optional[Sequence[float]] = None) -> Scene:
    """A small stirring scene for tests (a few thousand particles)."""
    return stirring_scene(n_grid, substeps, fill_radius=fill_radius, fill_height=fill_height,
                          stick_radius=0.06, path_rx=0.1, path_rz=0.1, seed=seed, params=params,
                          name="tiny")
