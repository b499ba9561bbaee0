"""The reference's module interface (diffstir_amd.reference_api) on the GPU, driven through the
calls the reference's driver makes on it: the model choice of main.py:9-14, the taped epoch loop
with the step rule (gradient_descent_module.py:99-122), the clip hand-off (:131-138) and the force
replay (force_calculation_and_animation_module.py:5-19), on the reference scene (30 substeps of
data/traj.npy), checked against the FP64 oracle.

Tolerances as in test_gpu_parity: forces within 1e-4 of the largest force (rows >= 2), gradients
within 1e-3 of the largest component, the loss within 1e-4 relative.
"""
import dataclasses

import numpy as np
import pytest

from diffstir_amd import reference_api as R
from diffstir_amd import scene as S
from diffstir_amd.gradient_descent import single_parameter_fitting1
from oracle.oracle import Oracle, loss_and_grad
from tests.helpers import force_err, grad_err, ref_traj, reference_layout

pytestmark = pytest.mark.gpu

FRAMES = 31


@pytest.fixture(scope="module")
def layout(tmp_path_factory):
    root = reference_layout(tmp_path_factory.mktemp("reference"), FRAMES, frames_per_clip=10,
                            force=np.linspace(0.0, 0.02, FRAMES))
    assert R.load(root)
    return root


@pytest.fixture(scope="module")
def o64():
    return Oracle("f64")


def oracle_scene(model, params, x_liquid=None, x_stick=None, traj=None):
    sc = S.reference_scene(ref_traj()[:FRAMES], seed=0, params=params, model=model)
    kw = {}
    if x_liquid is not None:
        kw.update(x_liquid=np.asarray(x_liquid, np.float32), x_stick=np.asarray(x_stick, np.float32))
    if traj is not None:
        kw.update(traj=np.asarray(traj, np.float32))
    return dataclasses.replace(sc, **kw)


def oracle_loss_grad(o64, sc, target):
    """Forces, loss and dL/dparams of the reference's L1 x-force loss against `target` (per step)."""
    F = o64.rollout(sc)
    t3 = np.zeros((F.shape[0], 3))
    t3[:, 0] = target
    L, g = loss_and_grad(F, t3, 0)
    _, dth = o64.rollout_grad(sc, g, ckpt=10)
    return F, L, dth


def taped_epoch():
    """One epoch of gradient_descent_module.py:102-107: the taped rollout and the loss."""
    R.loss[None] = 0
    with R.ti.ad.Tape(loss=R.loss):
        R.loss[None] = 0
        for f in range(R.max_steps - 1):
            R.forward(f)
        R.compute_loss()
    return R.loss[None], [R.constitutive_parameters.grad[j] for j in range(R.n_constitutive_parameters)]


def test_model_choice_epoch_loop_and_clip_hand_off(layout, o64):
    print("[gpu] reference_api: models, taped epochs, clip hand-off", flush=True)
    R.initialize_objects()
    start = np.array([[2.0, 0.5, 0.7, 3.0], [2.0, 0.5, 0.6, 2.0], [1.5, 2.0, 0.8, 1.0]])
    losses = []
    for i in range(3):
        # main.py:10-14 as the loop evidently means it (entry j of the parameters)
        R.model_index[None] = i
        for j in range(4):
            R.constitutive_parameters[j] = start[i, j]
        R.initialize_ti_field()
        L, g = taped_epoch()
        _, Lo, go = oracle_loss_grad(o64, oracle_scene(i, start[i]), np.zeros(FRAMES))
        print(f"  model {i}: loss {L:.6g} vs {Lo:.6g}, grad {g} vs {go}", flush=True)
        assert abs(L - Lo) <= 1e-4 * abs(Lo)
        assert grad_err(g, go) <= 1e-3
        losses.append(L)
    assert len(set(losses)) == 3
    # gradient_descent_module.py:95-122 for Herschel-Bulkley: two epochs with the step rule
    R.model_index[None] = 2
    lo, hi = [0.1, 0.0, 0.5, 0.0], [5.0, 20.0, 1.5, 1.0]
    for j in range(4):
        R.learning_rate[j] = 1
        R.constitutive_parameters[j] = [1.5, 2.0, 0.8, 0.0][j]
    R.initialize_ti_field()
    epochs, hist = 2, []
    for i in range(epochs):
        L, g = taped_epoch()
        p = R.constitutive_parameters.to_numpy().astype(np.float64)
        hist.append((L, p))
        _, Lo, go = oracle_loss_grad(o64, oracle_scene(2, p), np.zeros(FRAMES))
        assert abs(L - Lo) <= 1e-4 * abs(Lo) and grad_err(g[:3], go[:3]) <= 1e-3
        for j in range(3):
            pj, lr = single_parameter_fitting1(R.constitutive_parameters[j], g[j], R.learning_rate[j], lo[j], hi[j],
                                               epochs, i)
            R.constitutive_parameters[j] = pj
            R.learning_rate[j] = lr
        R.initialize_ti_field()
    assert not np.array_equal(hist[0][1], hist[1][1])
    # the hand-off (gradient_descent_module.py:131-138): the last rollout's final frame is still
    # readable after the reset, a mid-rollout frame is not
    T = R.max_steps
    xs = R.operator_x.to_numpy()[T - 1]
    xl = R.x.to_numpy()[T - 1]
    with pytest.raises(IndexError):
        R.x.to_numpy()[5]
    _, st = o64.rollout(oracle_scene(2, hist[1][1]), return_state=True)
    np.testing.assert_allclose(xl, st["lx"], atol=2e-5)
    np.testing.assert_allclose(xs, st["sx"], atol=2e-5)
    R.initial_operator_x.from_numpy(xs)
    R.initial_x.from_numpy(xl)
    fpc = R.frames_per_clip
    R.operator_kinematic_seq.from_numpy(R.operator_trajectory[0:fpc])
    R.operator_forces_seq.from_numpy(R.force_magnitudes[0:fpc])
    R.initialize_ti_field()
    np.testing.assert_array_equal(R.x.to_numpy()[0], xl)
    L, g = taped_epoch()
    traj = R.operator_kinematic_seq.to_numpy()
    target = R.operator_forces_seq.to_numpy().astype(np.float64)
    assert np.array_equal(target[:fpc], R.force_magnitudes[:fpc].astype(np.float32)) and np.all(target[fpc:] == 0)
    _, Lo, go = oracle_loss_grad(o64, oracle_scene(2, R.constitutive_parameters.to_numpy().astype(np.float64),
                                                  xl, xs, traj), target)
    print(f"  after the hand-off: loss {L:.6g} vs {Lo:.6g}", flush=True)
    assert abs(L - Lo) <= 1e-4 * abs(Lo) and grad_err(g[:3], go[:3]) <= 1e-3


def test_force_replay_fields(layout, o64):
    """force_calculation_and_animation_module.py:5-19 with main.py:30's presentation time: the
    substeps run outside a tape, so every frame is kept."""
    print("[gpu] reference_api: force replay", flush=True)
    assert R.load(layout)          # the fields as a fresh import leaves them (zero target)
    R.model_index[None] = 1
    for j in range(4):
        R.constitutive_parameters[j] = [2.0, 0.5, 0.8, 1.5][j]
    R.initialize_objects()
    R.loss[None] = 0
    n = int((R.max_steps - 1) * R.dt / R.dt - 1)
    for f in range(n):
        R.forward(f)
    R.compute_loss()
    pos = R.x.to_numpy()
    op = R.operator_x.to_numpy()
    cont = R.container_x.to_numpy()
    force = R.total_operator_dv.to_numpy() * R.operator_p_vol * R.operator_density / R.dt
    assert pos.shape == (R.max_steps, R.n_particles, 3) and op.shape == (R.max_steps, R.n_operator_particles, 3)
    assert cont.shape == (R.max_steps, R.n_container_particles, 3)
    assert np.mean(op, axis=1)[:-1].shape == (R.max_steps - 1, 3)
    sc = oracle_scene(1, np.array([2.0, 0.5, 0.8, 1.5]))
    Fo, st = o64.rollout(sc, nsteps=n, return_state=True)
    F = np.zeros_like(Fo)
    F[:, 0], F[:, 2] = force[:n + 1, 0], force[:n + 1, 2]
    assert force_err(F[:, [0, 2]], Fo[:, [0, 2]]) <= 1e-4
    np.testing.assert_allclose(pos[n], st["lx"], atol=2e-5)
    np.testing.assert_array_equal(pos[0], R.liquid_particles)
    assert np.all(force[R.max_steps - 1] == 0)
    L = R.loss[None]
    Lo, _ = loss_and_grad(Fo if n + 1 == R.max_steps else np.vstack([Fo, np.zeros((R.max_steps - n - 1, 3))]),
                          np.zeros((R.max_steps, 3)), 0)
    assert abs(L - Lo) <= 1e-4 * abs(Lo)
