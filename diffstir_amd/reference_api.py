"""The reference's simulation module on the HIP library: a drop-in for the driver's import.

The reference driver (src/gradient_descent_module.py, src/force_calculation_and_animation_module.py
and main.py through them) reaches the simulator with `from src.simulator.simulation_functions
import *`: module-global Taichi fields read and written from Python (`constitutive_parameters[j]`,
`constitutive_parameters.grad[j]`, `loss[None]`, `model_index[None]`, `initial_x.from_numpy(...)`,
`x.to_numpy()`, ...; initialization.py:123-165), the operator pair forward(s)/backward(s) with
compute_loss, initialize_objects and initialize_ti_field (simulation_functions.py:293-365), the
`ti.ad.Tape` that records them (gradient_descent_module.py:103-107), and the constants of
initialization.py:14-78.  `from diffstir_amd.reference_api import *` provides the same names, so
those files run with only that import changed.

The scene is built at import, as the reference builds it, from the reference's data layout under
the directory DIFFSTIR_REFERENCE_ROOT (default: the current directory): config/arguments.json,
data/traj.npy and data/force.npy.  Without them the import succeeds, the constants are None and
every call raises until `load(root)` is called (names already star-imported keep their values,
so call `load` before the star import in that case).

Differences a caller can see, all where the checkpointed design keeps less state than Taichi's
full per-step fields:
  * x.to_numpy() / operator_x.to_numpy() / container_x.to_numpy() return a (max_steps, n, 3)
    ndarray when every frame up to the current step was recorded (substeps taken outside a tape,
    as in the force replay); otherwise a frame series whose element [k] is available for the
    current step and for the last step of the previous rollout (what the clip hand-off reads,
    gradient_descent_module.py:131-132), any other frame raising IndexError;
  * operator_kinematic_seq.from_numpy / operator_forces_seq.from_numpy accept fewer rows than
    max_steps and overwrite those rows (Taichi raises on the clip-long slice of
    gradient_descent_module.py:136-137);
  * the fields hold float32 values like the reference's ti.f32 fields, the device computes with
    the float64 values of those float32 numbers.
"""
from __future__ import annotations

import dataclasses
import json  # noqa: F401  (the reference's star import provides it to main.py)
import math  # noqa: F401  (used by gradient_descent_module.py through the star import)
import os
from typing import Optional

import numpy as np

from . import scene as S
from .simulator import Simulator

__all__ = [
    "ti", "np", "math", "os", "json", "load",
    # constants of initialization.py
    "dim", "bound", "dt", "n_grid", "frames_per_clip", "dx", "inv_dx", "n_particles", "n_container_particles",
    "n_operator_particles", "uniform_scaling_factor", "operator_p_vol", "liquid_p_vol", "operator_p_mass",
    "liquid_p_mass", "container_p_vol", "container_p_mass", "operator_density", "liquid_density",
    "container_density", "operator_mass", "operator_trajectory", "force_magnitudes", "total_frames", "n_clips",
    "max_steps", "n_simulation_frames", "n_constitutive_parameters", "liquid_particles", "operator_particles",
    "container_particles",
    # fields
    "model_index", "learning_rate", "constitutive_parameters", "parameters_lower_bound", "parameters_upper_bound",
    "loss", "initial_x", "initial_operator_x", "initial_container_x", "operator_kinematic_seq",
    "operator_forces_seq", "total_operator_dv", "x", "operator_x", "container_x",
    # operators
    "initialize_objects", "initialize_ti_field", "forward", "backward", "compute_loss",
]

_f32 = np.float32


# ----------------------------------------------------------------------------- state

class _State:
    """The simulator behind the fields, and what Taichi would keep in them between calls."""

    def __init__(self):
        self.sim: Optional[Simulator] = None
        self.scene: Optional[S.Scene] = None
        self.model = 0
        self.params = np.zeros(4, dtype=np.float32)
        self.lr = np.zeros(4, dtype=np.float32)
        self.loss = _f32(0.0)
        self.target = None             # operator_forces_seq, (max_steps,)
        self.scene_dirty = False       # positions or trajectory changed since the upload
        self.frames = {}               # frame index -> (liquid, stick, container) positions
        self.frames_complete = False   # every frame 0..step recorded (substeps outside a tape)
        self.in_tape = False
        self.ckpt_every = 50

    def need(self) -> Simulator:
        if self.scene is None:
            raise RuntimeError("no reference scene: set DIFFSTIR_REFERENCE_ROOT to a directory with "
                               "config/arguments.json and data/{traj,force}.npy, or call load(root)")
        if self.sim is None or self.sim.scene.cfg.model != self.model:
            self._build()
        return self.sim

    def _build(self):
        if self.sim is not None:
            self.sim.close()
        cfg = dataclasses.replace(self.scene.cfg, model=int(self.model))
        self.scene = dataclasses.replace(self.scene, cfg=cfg)
        self.sim = Simulator(self.scene, ckpt_every=self.ckpt_every)
        self.sim.model_index = int(self.model)
        self.scene_dirty = False

    def positions(self):
        sim = self.sim
        xl = sim.engine.read_liquid(0)["x"]
        xs = sim.engine.read_stick(0)["x"] if self.scene.n_stick else np.zeros((0, 3), np.float32)
        return xl, xs, np.asarray(self.scene.x_container, np.float32)


_S = _State()


# ----------------------------------------------------------------------------- ti stand-in

class _Tape:
    """ti.ad.Tape(loss=loss) (gradient_descent_module.py:103): clears the parameter gradient,
    records the forward substeps of the body, and on exit runs the checkpointed adjoint of the
    recorded rollout for the loss compute_loss() accumulated."""

    def __init__(self, loss=None, clear_gradients=True):
        self._cm = None

    def __enter__(self):
        sim = _S.need()
        self._cm = sim.tape()
        self._cm.__enter__()
        _S.in_tape = True
        return self

    def __exit__(self, *exc):
        _S.in_tape = False
        return self._cm.__exit__(*exc)


class _Ad:
    Tape = _Tape


class _Ti:
    """The parts of the taichi module the driver files use: ti.ad.Tape, ti.max (gdm:4) and the
    type names of their annotations (ti.i32, gdm:84; ti.f32, fcam:3)."""
    ad = _Ad()
    f32 = float
    i32 = int

    @staticmethod
    def max(*a):
        return max(*a)

    @staticmethod
    def min(*a):
        return min(*a)


ti = _Ti()


# ----------------------------------------------------------------------------- fields

def _none_key(key):
    if key is not None and key != ():
        raise IndexError("a 0-d field is indexed with [None]")


class _ScalarField:
    """A 0-d field accessed as f[None] (loss, model_index)."""

    def __init__(self, get, put):
        self._get, self._put = get, put

    def __getitem__(self, key):
        _none_key(key)
        return self._get()

    def __setitem__(self, key, value):
        _none_key(key)
        self._put(value)

    def to_numpy(self):
        return np.array(self._get())


class _Vec32Field:
    """A float32 vector field of fixed length backed by host memory (learning_rate)."""

    def __init__(self, store: str):
        self._store = store

    def _arr(self) -> np.ndarray:
        return getattr(_S, self._store)

    def __getitem__(self, j):
        return float(self._arr()[j])

    def __setitem__(self, j, value):
        self._arr()[j] = _f32(value)

    def from_numpy(self, a):
        a = np.asarray(a, dtype=np.float32).reshape(self._arr().shape)
        self._arr()[:] = a

    def to_numpy(self):
        return self._arr().copy()

    def fill(self, value):
        self._arr()[:] = _f32(value)


class _GradView:
    """constitutive_parameters.grad: dL/dparams of the last taped rollout (float32 values)."""

    def __getitem__(self, j):
        sim = _S.sim
        return 0.0 if sim is None else float(_f32(sim.constitutive_parameters_grad[j]))

    def __setitem__(self, j, value):
        _S.need().constitutive_parameters_grad[j] = float(_f32(value))

    def to_numpy(self):
        sim = _S.sim
        return np.zeros(4, np.float32) if sim is None else sim.constitutive_parameters_grad.astype(np.float32)

    def fill(self, value):
        _S.need().constitutive_parameters_grad[:] = float(_f32(value))


class _ParameterField(_Vec32Field):
    """constitutive_parameters (initialization.py:134) with its gradient field."""

    def __init__(self):
        super().__init__("params")
        self.grad = _GradView()


class _PositionsField:
    """initial_x / initial_operator_x / initial_container_x (initialization.py:148-150): the
    positions initialize_ti_field starts from."""

    def __init__(self, species: str):
        self._sp = species

    def from_numpy(self, a):
        _S.need()
        old = getattr(_S.scene, self._sp)
        a = np.asarray(a, dtype=np.float32).reshape(old.shape)
        _S.scene = dataclasses.replace(_S.scene, **{self._sp: a.copy()})
        _S.scene_dirty = True

    def to_numpy(self):
        _S.need()
        return np.asarray(getattr(_S.scene, self._sp), np.float32).copy()


class _KinematicSeqField:
    """operator_kinematic_seq (initialization.py:151): the stick trajectory, max_steps frames."""

    def from_numpy(self, a):
        _S.need()
        a = np.asarray(a, dtype=np.float32).reshape(-1, 3)
        t = np.array(_S.scene.traj, dtype=np.float32)
        if a.shape[0] > t.shape[0]:
            raise ValueError(f"{a.shape[0]} rows do not fit a {t.shape[0]}-frame field")
        t[:a.shape[0]] = a
        _S.scene = dataclasses.replace(_S.scene, traj=t)
        _S.scene_dirty = True

    def to_numpy(self):
        _S.need()
        return np.asarray(_S.scene.traj, np.float32).copy()


class _ForcesSeqField:
    """operator_forces_seq (initialization.py:152): the target x force per step."""

    def from_numpy(self, a):
        _S.need()
        a = np.asarray(a, dtype=np.float32).reshape(-1)
        if a.shape[0] > _S.target.shape[0]:
            raise ValueError(f"{a.shape[0]} rows do not fit a {_S.target.shape[0]}-frame field")
        _S.target[:a.shape[0]] = a

    def to_numpy(self):
        _S.need()
        return _S.target.copy()

    def __getitem__(self, s):
        _S.need()
        return float(_S.target[s])


class _TotalDvField:
    """total_operator_dv (initialization.py:153): sum_p dv[s, p] per step, filled by compute_loss
    (simulation_functions.py:332-333) for steps 0..max_steps-2."""

    def to_numpy(self):
        sim = _S.need()
        F = np.asarray(sim.total_operator_force, dtype=np.float64).reshape(-1, 3)
        out = (F * (sim.dt / sim.scene.cfg.m_stick)).astype(np.float32)
        out[sim.max_steps - 1:] = 0.0
        return out


class FrameSeries:
    """The frames of a per-step position field that the checkpointed rollout can return: the
    current step, and frames recorded or kept from the previous rollout; [k] of any other frame
    raises IndexError."""

    def __init__(self, species: int, n: int):
        self._sp, self._n = species, n
        self.shape = (_S.sim.max_steps, n, 3)

    def __len__(self):
        return self.shape[0]

    def __getitem__(self, k):
        if not isinstance(k, (int, np.integer)):
            raise IndexError("index the frames one at a time: the rollout keeps checkpoints, not every frame")
        T = self.shape[0]
        k = int(k) + T if k < 0 else int(k)
        if k == _S.sim.step:
            return _S.positions()[self._sp]
        if k in _S.frames:
            return _S.frames[k][self._sp].copy()
        raise IndexError(f"frame {k} is not kept: the checkpointed rollout returns the current step "
                         f"({_S.sim.step}), the frames of a rollout taken outside a tape and the last step "
                         "of the previous rollout")

    def __array__(self, dtype=None):
        raise TypeError("not every frame is kept (substeps ran inside a tape): index single frames")


class _FrameField:
    """x / operator_x / container_x (initialization.py:155-158), positions per step."""

    def __init__(self, species: int, count_attr: str):
        self._sp, self._count = species, count_attr

    def to_numpy(self):
        sim = _S.need()
        n = getattr(_S.scene, self._count)
        if _S.frames_complete:
            out = np.zeros((sim.max_steps, n, 3), np.float32)
            for k, fr in _S.frames.items():
                out[k] = fr[self._sp]
            out[sim.step] = _S.positions()[self._sp]
            return out
        return FrameSeries(self._sp, n)


def _set_model(v):
    _S.model = int(v)


model_index = _ScalarField(lambda: _S.model, _set_model)
learning_rate = _Vec32Field("lr")
constitutive_parameters = _ParameterField()
parameters_lower_bound = [0] * 4
parameters_upper_bound = [0] * 4
loss = _ScalarField(lambda: float(_S.loss), lambda v: setattr(_S, "loss", _f32(v)))
initial_x = _PositionsField("x_liquid")
initial_operator_x = _PositionsField("x_stick")
initial_container_x = _PositionsField("x_container")
operator_kinematic_seq = _KinematicSeqField()
operator_forces_seq = _ForcesSeqField()
total_operator_dv = _TotalDvField()
x = _FrameField(0, "n_liquid")
operator_x = _FrameField(1, "n_stick")
container_x = _FrameField(2, "n_container")


# ----------------------------------------------------------------------------- operators

def initialize_objects():
    """simulation_functions.py:293-298: the scene's initial positions and trajectory into the
    fields, then initialize_ti_field()."""
    _S.need()
    _S.scene = dataclasses.replace(_S.scene, x_liquid=np.asarray(liquid_particles, np.float32),
                                   x_stick=np.asarray(operator_particles, np.float32),
                                   x_container=np.asarray(container_particles, np.float32),
                                   traj=np.asarray(operator_trajectory[:max_steps], np.float32))
    _S.scene_dirty = True
    initialize_ti_field()


def initialize_ti_field():
    """simulation_functions.py:301-319: step-0 state from the initial positions, force
    accumulators cleared.  The previous rollout's last frame stays readable (Taichi keeps the
    frames a reset does not overwrite)."""
    sim = _S.need()
    kept = {}
    if sim.step > 0:
        kept[sim.step] = _S.positions()
    if _S.scene_dirty:
        sim.scene = _S.scene
        sim.engine.scenes[0] = _S.scene
        sim.engine.upload_scene(0, _S.scene)
        _S.scene_dirty = False
    sim.constitutive_parameters[:] = _S.params.astype(np.float64)
    sim.initialize_ti_field()
    _S.frames = kept
    _S.frames_complete = False


def forward(s: int):
    """One substep s -> s+1 (simulation_functions.py:341-348).  Outside a tape every frame is kept
    for x.to_numpy() (the force replay, force_calculation_and_animation_module.py:8-15)."""
    sim = _S.need()
    if s == 0:
        sim.constitutive_parameters[:] = _S.params.astype(np.float64)
        if not _S.in_tape:
            _S.frames = {0: _S.positions()}
            _S.frames_complete = True
    sim.forward(s)
    if _S.frames_complete and not _S.in_tape:
        _S.frames[sim.step] = _S.positions()


def backward(s: int):
    _S.need().backward(s)


def compute_loss():
    """simulation_functions.py:330-339: adds the L1 loss of the x force against
    operator_forces_seq to loss[None] and fills total_operator_dv."""
    sim = _S.need()
    sim.operator_forces_seq = _S.target.astype(np.float64)
    L = sim.compute_loss()
    _S.loss = _f32(float(_S.loss) + L)


# ----------------------------------------------------------------------------- loading

def load(root: Optional[str] = None, seed: int = 0, ckpt_every: int = 50):
    """Build the reference scene from root/config/arguments.json and root/data/{traj,force}.npy
    (initialization.py:14-117, seeded) and set this module's constants."""
    g = globals()
    root = root or os.environ.get("DIFFSTIR_REFERENCE_ROOT") or os.getcwd()
    args_path = os.path.join(root, "config", "arguments.json")
    traj_path = os.path.join(root, "data", "traj.npy")
    force_path = os.path.join(root, "data", "force.npy")
    if not (os.path.exists(args_path) and os.path.exists(traj_path) and os.path.exists(force_path)):
        return False
    args = S.load_arguments(args_path)
    raw = np.load(traj_path, allow_pickle=False)
    forces = np.load(force_path, allow_pickle=False)
    sc = S.reference_scene(raw, seed=seed, arguments=args, params=np.zeros(4), model=0)
    c = sc.cfg
    T = raw.shape[0]
    fpc = int(args["frames_per_clip"])
    if fpc >= T:
        fpc = T - 1
    o_vol = args["operator_radius"] ** 2 * args["operator_length"] * np.pi / args["n_operator_model_particles_generated"]
    if _S.sim is not None:
        _S.sim.close()
    _S.__init__()
    _S.scene, _S.ckpt_every = sc, ckpt_every
    _S.target = np.zeros(T, np.float32)
    t, _ = S.transform_trajectory(raw)
    g.update(dim=3, bound=c.bound, dt=c.dt, n_grid=c.n_grid, frames_per_clip=fpc, dx=c.dx, inv_dx=c.inv_dx,
             n_particles=sc.n_liquid, n_container_particles=sc.n_container, n_operator_particles=sc.n_stick,
             uniform_scaling_factor=args["uniform_scaling_factor"], operator_p_vol=o_vol, liquid_p_vol=c.vol_liq,
             operator_p_mass=c.m_stick, liquid_p_mass=c.m_liq,
             container_p_vol=(c.m_cont / args["container_density"]), container_p_mass=c.m_cont,
             operator_density=args["operator_density"], liquid_density=args["liquid_density"],
             container_density=args["container_density"], operator_mass=c.m_stick * sc.n_stick,
             operator_trajectory=t, force_magnitudes=forces, total_frames=T, n_clips=int(T / fpc), max_steps=T,
             n_simulation_frames=T, n_constitutive_parameters=4, liquid_particles=np.asarray(sc.x_liquid),
             operator_particles=np.asarray(sc.x_stick), container_particles=np.asarray(sc.x_container))
    return True


for _name in __all__:
    globals().setdefault(_name, None)
load()
