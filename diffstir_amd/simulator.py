"""The reference's operator interface for the substep, on the HIP library.

The reference (src/simulator/simulation_functions.py) exposes module-level functions over
module-global Taichi fields:

    initialize_objects()       upload particles and trajectory, then initialize_ti_field()  (sf:293-298)
    initialize_ti_field()      step-0 state, cleared force accumulators                    (sf:301-319)
    forward(s)                 one substep s -> s+1                                         (sf:341-348)
    backward(s)                its adjoint, replayed by ti.ad.Tape                          (sf:351-365)
    compute_loss()             stick force per step -> loss                                 (sf:330-339)
    with ti.ad.Tape(loss=loss): ...                                                        (gdm:103-107)

`Simulator` holds the same state (constitutive_parameters, model_index, loss, the target
operator_forces_seq, the per-step force series) for one scene and implements the same calls on
an `Engine`.  `Simulator.tape()` mirrors ti.ad.Tape: inside it, forward(s) calls are recorded;
on exit the loss's gradient is propagated by the checkpointed adjoint and lands in
`constitutive_parameters_grad`.  The module-level functions act on the simulator installed
with `setup(scene)`, like the reference's module-global fields.
"""
from __future__ import annotations

import contextlib
from typing import Optional

import numpy as np

from .engine import Engine
from .loss import compute_loss as _loss
from .scene import Scene


class Simulator:
    def __init__(self, scene: Scene, ckpt_every: int = 50, kernels: str = "tiled", particles_per_thread: int = 2):
        self.scene = scene
        self.dt = scene.cfg.dt
        self.max_steps = scene.max_steps
        self.engine = Engine([scene], ckpt_every=ckpt_every, kernels=kernels,
                             particles_per_thread=particles_per_thread)
        self.constitutive_parameters = np.asarray(scene.params, dtype=np.float64).copy()
        self.constitutive_parameters_grad = np.zeros(4)
        self.model_index = int(scene.cfg.model)
        self.loss = 0.0
        # the target force series (operator_forces_seq, initialization.py:152): x force per step
        # for the reference loss; [T, 3] for the x/z MSE
        self.operator_forces_seq = np.zeros((self.max_steps, 3))
        self.total_operator_force = np.zeros((self.max_steps, 3))
        self._recording: Optional[list] = None
        self._params_on_device = None

    # ------------------------------------------------------------------ state
    def close(self):
        self.engine.close()

    def _sync_params(self):
        if self._params_on_device is None or not np.array_equal(self._params_on_device, self.constitutive_parameters):
            self.engine.set_params(self.constitutive_parameters[None])
            self._params_on_device = self.constitutive_parameters.copy()

    def initialize_objects(self):
        """sf:293-298: particles and trajectory are uploaded when the Engine is built; reset."""
        self.initialize_ti_field()

    def initialize_ti_field(self):
        """sf:301-319: step-0 state and cleared force accumulators."""
        if self.model_index != self.scene.cfg.model:
            raise ValueError("model_index differs from the scene's model; build a Simulator per model")
        self._sync_params()
        self.engine.reset()
        self.total_operator_force[:] = 0.0

    @property
    def step(self) -> int:
        return self.engine.step

    def hand_off_clip(self, trajectory: np.ndarray, target: Optional[np.ndarray] = None):
        """The clip hand-off of gdm:131-138: the next clip starts from the particle positions of
        the last rollout's final step (initial_x, initial_operator_x), with the next trajectory
        frames and target; then initialize_ti_field.  The reference writes a clip-long slice
        into max_steps-long fields there (a shape mismatch); here the slice must have the
        simulator's max_steps frames."""
        import dataclasses
        trajectory = np.asarray(trajectory, dtype=np.float32)
        if trajectory.shape != (self.max_steps, 3):
            raise ValueError(f"the next clip's trajectory must have {self.max_steps} frames, got {trajectory.shape}")
        xl = self.engine.read_liquid(0)["x"]
        xs = self.engine.read_stick(0)["x"] if self.scene.n_stick else self.scene.x_stick
        self.scene = dataclasses.replace(self.scene, x_liquid=xl, x_stick=xs, traj=trajectory)
        self.engine.scenes[0] = self.scene
        self.engine.upload_scene(0, self.scene)
        if target is not None:
            target = np.asarray(target, dtype=np.float64)
            if target.shape[0] != self.max_steps:
                raise ValueError(f"the next clip's target must have {self.max_steps} rows, got {target.shape}")
            self.operator_forces_seq = target
        self.initialize_ti_field()

    # ------------------------------------------------------------------ substep
    def forward(self, s: int):
        """One substep s -> s+1 (sf:341-348).  Steps must be taken in order from 0."""
        if s != self.engine.step:
            raise ValueError(f"forward({s}) called at step {self.engine.step}: substeps run in order from 0")
        if s == 0:
            self._sync_params()
        self.engine.forward(1)
        if self._recording is not None:
            self._recording.append(s)

    def forward_steps(self, nsteps: Optional[int] = None):
        """Several substeps in one call (the same work as nsteps forward(s) calls)."""
        s0 = self.engine.step
        if s0 == 0:
            self._sync_params()
        self.engine.forward(nsteps)
        if self._recording is not None:
            self._recording.extend(range(s0, self.engine.step))

    def backward(self, s: int):
        """The adjoint of forward(s).  The adjoint of a rollout is checkpointed, so it runs as a
        whole when the tape that recorded the forward calls exits; calling it directly is an
        error (as calling backward(s) outside a Tape is meaningless in the reference)."""
        raise RuntimeError("backward(s) runs inside Simulator.tape(): the adjoint of the recorded rollout is "
                           "computed at tape exit")

    def compute_loss(self) -> float:
        """sf:330-339: stick force per step (m_stick/dt * sum dv) and the loss against
        operator_forces_seq (reference: L1 of the x force; MSE of x and z for the synthetic
        workloads)."""
        F = self.engine.forces()[0]
        self.total_operator_force = F
        tgt = self.operator_forces_seq
        self.loss, self._dl_dforces = _loss(F, tgt, self.scene.cfg.loss_kind)
        return self.loss

    @contextlib.contextmanager
    def tape(self):
        """ti.ad.Tape(loss=loss) (gdm:103-107): clears the parameter gradient, records the
        forward substeps of the body and, on exit, runs the adjoint of the recorded rollout
        for the loss computed by compute_loss() inside the body."""
        self.constitutive_parameters_grad = np.zeros(4)
        self._recording = []
        self._dl_dforces = None
        try:
            yield self
        finally:
            recorded = self._recording
            self._recording = None
        if not recorded:
            return
        if recorded != list(range(len(recorded))):
            raise RuntimeError("the tape recorded substeps out of order")
        if self._dl_dforces is None:
            self.compute_loss()
        self.constitutive_parameters_grad = self.engine.backward(self._dl_dforces[None])[0]


# ---------------------------------------------------------------------- module-level interface

SIM: Optional[Simulator] = None


def setup(scene: Scene, **kw) -> Simulator:
    """Install the simulator the module-level functions act on."""
    global SIM
    if SIM is not None:
        SIM.close()
    SIM = Simulator(scene, **kw)
    return SIM


def _sim() -> Simulator:
    if SIM is None:
        raise RuntimeError("call diffstir_amd.simulator.setup(scene) first")
    return SIM


def initialize_objects():
    _sim().initialize_objects()


def initialize_ti_field():
    _sim().initialize_ti_field()


def forward(s: int):
    _sim().forward(s)


def backward(s: int):
    _sim().backward(s)


def compute_loss() -> float:
    return _sim().compute_loss()


def Tape():
    return _sim().tape()
