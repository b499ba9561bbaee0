"""The reference's driver API on the GPU: forward(s)/Tape/compute_loss, parameters_fitting, the
force replay and export, the torch autograd wrapper and the batched identification loop."""
import dataclasses
import os

import numpy as np
import pytest

from diffstir_amd import scene as S
from diffstir_amd import simulator as simmod
from diffstir_amd.engine import Engine
from diffstir_amd.force_calculation import force_calculation_and_animation
from diffstir_amd.gradient_descent import parameters_fitting
from diffstir_amd.identify import identify
from diffstir_amd.loss import compute_loss
from diffstir_amd.simulator import Simulator
from oracle.oracle import Oracle, loss_and_grad
from tests.helpers import force_err, grad_err, ref_traj

pytestmark = pytest.mark.gpu


def test_tape_over_forward_calls_matches_oracle():
    """gdm:103-107 restated: forward(f) for every substep inside the tape, compute_loss, and the
    parameter gradient on exit."""
    sc = S.tiny_scene(16, 20, seed=1)
    print("[gpu] tape over forward(s)", flush=True)
    sim = simmod.setup(sc, ckpt_every=6)
    o = Oracle("f64")
    Fo = o.rollout(sc)
    sim.operator_forces_seq = 0.6 * Fo
    with simmod.Tape():
        simmod.initialize_ti_field()
        for f in range(sc.max_steps - 1):
            simmod.forward(f)
        L = simmod.compute_loss()
    Lo, go = loss_and_grad(Fo, 0.6 * Fo, sc.cfg.loss_kind)
    _, dth_o = o.rollout_grad(sc, go, ckpt=6)
    assert L == pytest.approx(Lo, rel=1e-4)
    assert grad_err(sim.constitutive_parameters_grad, dth_o) <= 1e-3
    with pytest.raises(ValueError):
        simmod.forward(3)           # out of order
    with pytest.raises(RuntimeError):
        simmod.backward(0)          # only through the tape
    sim.close()
    simmod.SIM = None


def test_parameters_fitting_runs_the_reference_epoch_loop():
    sc = S.tiny_scene(16, 12, seed=2)
    print("[gpu] parameters_fitting", flush=True)
    sim = Simulator(sc, ckpt_every=6)
    truth = sc.params.copy()
    with Engine([sc], ckpt_every=6) as e:
        e.reset()
        e.forward()
        sim.operator_forces_seq = e.forces()[0]
    sim.constitutive_parameters[:3] = truth[:3] * np.array([1.5, 0.7, 1.2])
    lo = [0.1, 0.0, 0.5, 0.0]
    hi = [5.0, 20.0, 1.5, 1e4]
    best, hist = parameters_fitting(sim, 2, lo, hi, n_fit=3)
    assert len(hist["losses"]) == 6
    assert np.all(np.isfinite(hist["losses"]))
    assert min(hist["losses"]) <= hist["losses"][0]
    for p, a, b in zip(best[:3], lo, hi):
        assert a <= p <= b
    sim.close()


def test_force_replay_exports_reference_files(tmp_path):
    sc = S.reference_scene(ref_traj(), seed=0, frames=41)
    print("[gpu] force replay", flush=True)
    sim = Simulator(sc, ckpt_every=10)
    out = force_calculation_and_animation(sim, 40 * sc.cfg.dt, out_dir=str(tmp_path), save_point_clouds=True)
    fx = np.load(os.path.join(tmp_path, "force_profile_x.npy"))
    fz = np.load(os.path.join(tmp_path, "force_profile_z.npy"))
    sp = np.load(os.path.join(tmp_path, "stick_positions.npy"))
    assert fx.dtype == np.float32 and fx.shape == (41,) and fz.shape == (41,)
    assert fx[0] == 0 and fz[0] == 0 and fx[-1] == 0 and fz[-1] == 0
    # 40 * dt -> 39 substeps fill frames 0..39 of the 41-frame field; fcam:56 drops frame 40
    assert out["steps"] == 39
    assert sp.shape == (40, 3)
    # the stick moves rigidly with the trajectory (sf:266-269)
    np.testing.assert_allclose(sp - sp[0], sc.traj[:40] - sc.traj[0], atol=2e-6)
    pc = out["liquid_point_clouds"]
    assert pc.shape == (41, sc.n_liquid, 3)
    assert np.all(pc[40] == 0) and np.all(pc[39] != 0)
    o = Oracle("f64")
    Fo = o.rollout(sc, nsteps=39)
    assert force_err(np.stack([fx[:40], np.zeros(40), fz[:40]], 1), np.stack([Fo[:, 0], 0 * Fo[:, 0], Fo[:, 2]], 1)) <= 1e-4
    sim.close()


def test_autograd_function_matches_engine_backward():
    torch = pytest.importorskip("torch")
    from diffstir_amd.autograd import StirRollout
    scenes = [S.tiny_scene(16, 15, seed=s) for s in (0, 1)]
    print("[gpu] autograd", flush=True)
    with Engine(scenes, ckpt_every=5) as e:
        p = torch.tensor(np.stack([s.params for s in scenes]), dtype=torch.float64, requires_grad=True)
        F = StirRollout.apply(p, e)
        tgt = torch.zeros_like(F)
        loss = ((F[:, :-1, 0] - tgt[:, :-1, 0]) ** 2).sum() + ((F[:, :-1, 2] - tgt[:, :-1, 2]) ** 2).sum()
        loss.backward()
        Fn = F.detach().numpy()
        g = np.zeros_like(Fn)
        g[:, :-1, 0] = 2 * Fn[:, :-1, 0]
        g[:, :-1, 2] = 2 * Fn[:, :-1, 2]
        e.set_params(p.detach().numpy())
        e.reset()
        e.forward()
        d = e.backward(g)
    np.testing.assert_allclose(p.grad.numpy(), d, rtol=1e-6)


def test_model_selection_writes_the_reference_outputs(tmp_path):
    """main.py:9-30 on a 30-substep reference scene: each model fitted (3 epochs of the first
    clip), parameters_and_losses_result.json with the reference's keys and layout
    (main.py:19-24), the best model replayed with its force profile and stick positions."""
    import json
    from diffstir_amd.main import model_selection
    traj = ref_traj()[:31]
    cfg = np.array([[[2.0, 0.5, 0.7, 3.0], [2.0, 0.5, 0.6, 2.0], [1.5, 2.0, 0.8, 1.0]],
                    [[0.5, 0.1, 0.3, 0.5], [0.5, 0.1, 0.3, 0.5], [0.5, 0.5, 0.5, 0.0]],
                    [[5.0, 1.0, 1.0, 5.0], [5.0, 1.0, 1.0, 5.0], [5.0, 5.0, 1.2, 2.0]]])
    print("[gpu] model selection", flush=True)
    r = model_selection(traj, cfg, epochs=1, out_dir=str(tmp_path), ckpt_every=10)
    with open(os.path.join(tmp_path, "parameters_and_losses_result.json")) as f:
        text = f.read()
    d = json.loads(text)
    assert list(d) == ["parameters", "losses"] and text.startswith('{\n    "parameters": [')
    assert len(d["parameters"]) == 3 and all(len(p) == 4 for p in d["parameters"]) and len(d["losses"]) == 3
    assert r["best_model"] == int(np.argmin(d["losses"]))
    for i in range(3):
        assert all(lo <= p <= hi for p, lo, hi in zip(d["parameters"][i][:3], cfg[1, i, :3], cfg[2, i, :3]))
    fx = np.load(os.path.join(tmp_path, "force_profile_x.npy"))
    sp = np.load(os.path.join(tmp_path, "stick_positions.npy"))
    assert fx.shape == (31,) and sp.shape == (30, 3)
    # the replay of the best model against the oracle (29 substeps, main.py:30)
    best = r["best_model"]
    sc = S.reference_scene(traj, seed=0, params=d["parameters"][best], model=best)
    Fo = Oracle("f64").rollout(sc, nsteps=29)
    assert force_err(np.stack([fx[:30], np.zeros(30), fx[:30]], 1)[:, [0]], Fo[:, [0]]) <= 1e-4


def test_stale_rollout_adjoint_is_an_error():
    """An engine records one rollout.  Its adjoint after a second rollout (StirRollout applied
    twice before loss.backward()) or after new parameters raises instead of returning the
    gradient of another rollout; ds_backward itself refuses with DS_ESTATE after ds_set_params."""
    torch = pytest.importorskip("torch")
    from diffstir_amd._lib import DiffStirError
    from diffstir_amd.autograd import StirRollout
    sc = S.tiny_scene(16, 8, seed=2)
    print("[gpu] stale rollout adjoint", flush=True)
    with Engine([sc], ckpt_every=4) as e:
        p1 = torch.tensor(sc.params[None], dtype=torch.float64, requires_grad=True)
        F1 = StirRollout.apply(p1, e)
        p2 = torch.tensor(sc.params[None] * 1.1, dtype=torch.float64, requires_grad=True)
        F2 = StirRollout.apply(p2, e)
        with pytest.raises(RuntimeError, match="another rollout"):
            F1.sum().backward()
        F2.sum().backward()                       # the latest rollout's adjoint still works
        assert np.all(np.isfinite(p2.grad.numpy()))
        g = np.zeros((1, e.T, 3))
        g[0, 1:-1, 0] = 1.0
        e.set_params(sc.params[None] * 1.2)
        with pytest.raises(DiffStirError) as ei:
            e.backward(g)
        assert ei.value.code == 4                 # DS_ESTATE
        e.reset()
        e.forward()
        assert np.all(np.isfinite(e.backward(g)))
        e.set_params(sc.params[None] * 1.2)       # unchanged values keep the tape valid
        assert np.all(np.isfinite(e.backward(g)))


def test_engine_on_a_caller_stream_gives_the_same_result():
    """ds_set_stream: the engine's work queued on a torch stream (and back on its own) gives the
    same forces and gradient as the engine's own stream (same kernels, same order)."""
    torch = pytest.importorskip("torch")
    sc = S.tiny_scene(16, 12, seed=5)
    print("[gpu] caller stream", flush=True)
    out = []
    side = torch.cuda.Stream()
    for stream in (None, side.cuda_stream, None):
        with Engine([sc], ckpt_every=4) as e:
            e.set_stream(stream)
            e.reset()
            e.forward()
            F = e.forces()
            g = np.zeros_like(F)
            g[0, 1:-1, 0] = F[0, 1:-1, 0]
            out.append((F, e.backward(g)))
    for F, d in out[1:]:
        np.testing.assert_allclose(F, out[0][0], rtol=1e-12, atol=1e-12 * np.abs(out[0][0]).max())
        np.testing.assert_allclose(d, out[0][1], rtol=1e-12)


def test_batched_identification_reduces_the_loss():
    truth = S.hb_params_draw(np.random.default_rng(5))
    base = S.tiny_scene(16, 20, seed=0, params=truth)
    print("[gpu] identification", flush=True)
    from diffstir_amd.identify import synthetic_target
    tgt = synthetic_target(base, truth, noise=0.01, ckpt_every=5)
    rng = np.random.default_rng(9)
    inits = np.stack([np.r_[truth[:3] * np.exp(rng.normal(0, 0.3, 3)), truth[3]] for _ in range(4)])
    res = identify([base] * 4, tgt, inits, iterations=12, lr=5e-2, ckpt_every=5)
    L = res["losses"]
    assert L.shape == (12, 4)
    assert np.all(L[-1] < L[0])


def test_levenberg_marquardt_recovers_noiseless_parameters():
    """identify_lm from two starts 20-30 % off the ground truth, against the noiseless profile:
    the best loss never rises and falls by at least 1e4 within 8 iterations, to parameters within
    2 % of the truth (measured on MI355X: 0.6-0.8 % in the worst-determined parameter, where the
    FP32 forces' rounding leaves the loss flat)."""
    from diffstir_amd.identify import identify_lm, synthetic_target
    truth = S.hb_params_draw(np.random.default_rng(5))
    base = S.tiny_scene(16, 20, seed=0, params=truth)
    print("[gpu] Levenberg-Marquardt identification", flush=True)
    tgt = synthetic_target(base, truth, noise=0.0, ckpt_every=5)
    inits = np.stack([np.r_[truth[:3] * np.array([1.3, 0.8, 1.2]), truth[3]],
                      np.r_[truth[:3] * np.array([0.75, 1.25, 0.9]), truth[3]]])
    res = identify_lm(base, tgt, inits, iterations=8, ckpt_every=5)
    L = res["losses"]
    print(f"  best losses per iteration {L.tolist()}\n  final {res['final'][:, :3]} truth {truth[:3]}", flush=True)
    assert L.shape == (9, 2)
    assert np.all(np.diff(L, axis=0) <= 0)
    assert np.all(L[-1] <= 1e-4 * L[0])
    assert np.all(np.abs(res["final"][:, :3] - truth[:3]) <= 0.02 * truth[:3])


def test_sensitivity_jacobian_agrees_with_the_adjoint():
    """identify.sensitivity's central-difference Jacobian of the x/z force entries, projected on
    a weight vector, against the adjoint's gradient of the same weighted sum (dL/dlog theta =
    theta * dL/dtheta), for the weights w = the forces themselves over x and z, over x only and
    over z only (a swapped or shifted entry layout fails the last two)."""
    from diffstir_amd.identify import sensitivity
    truth = S.hb_params_draw(np.random.default_rng(5))
    base = S.tiny_scene(16, 20, seed=0, params=truth)
    print("[gpu] sensitivity Jacobian vs adjoint", flush=True)
    res = sensitivity(base, truth, ckpt_every=5)
    J = res["jacobian"]
    T = base.max_steps
    assert J.shape == (2 * (T - 1), 3)
    with Engine([base], ckpt_every=5) as e:
        e.set_params(truth[None])
        e.reset()
        e.forward()
        F = e.forces()[0]
        w = np.concatenate([F[:T - 1, 0], F[:T - 1, 2]])
        for mx, mz in ((1, 1), (1, 0), (0, 1)):
            wm = w * np.repeat([mx, mz], T - 1)
            g = np.zeros((1, T, 3))
            g[0, :T - 1, 0] = wm[:T - 1]
            g[0, :T - 1, 2] = wm[T - 1:]
            adj = truth[:3] * e.backward(g)[0][:3]
            fd = J.T @ wm
            print(f"  weights x{mx} z{mz}: adjoint {adj}, central {fd}", flush=True)
            # the FP32 oracle's own central differences at this step differ from its adjoint by
            # at most 1.6e-3 of the largest entry on this scene
            assert np.max(np.abs(adj - fd)) <= 1e-2 * np.max(np.abs(fd))
    G = np.asarray(res["gauss_newton"])
    np.testing.assert_allclose(G, J.T @ J / (T - 1), rtol=1e-12)
    assert np.all(np.asarray(res["eigvals"]) > 0)


def test_clip_hand_off_continues_from_the_last_state():
    """gdm:131-138 as intended: the second clip starts where the first clip's last rollout ended
    (liquid and stick positions), with the next trajectory frames and target."""
    from diffstir_amd.gradient_descent import fit_clips
    sc = S.tiny_scene(16, 10, seed=3)
    # a 2-clip trajectory: the first clip's frames, then the path continued (same circle)
    full = S.tiny_scene(16, 19, seed=3).traj
    traj = np.concatenate([full[:10], full[9:19]])
    target = np.zeros((20, 3))
    print("[gpu] clip hand-off", flush=True)
    sim = Simulator(dataclasses.replace(sc, traj=traj[:10]), ckpt_every=5)
    sim.initialize_objects()
    sim.forward_steps()
    end = sim.engine.read_liquid(0)["x"].copy()
    end_stick = sim.engine.read_stick(0)["x"].copy()
    sim.hand_off_clip(traj[10:20], target[10:20])
    assert sim.step == 0
    np.testing.assert_array_equal(sim.engine.read_liquid(0)["x"], end)
    np.testing.assert_array_equal(sim.engine.read_stick(0)["x"], end_stick)
    sim.forward_steps()
    F2 = sim.engine.forces()[0]
    assert np.all(np.isfinite(F2))
    # the second clip against the oracle rolled from the first clip's end state (gdm:131-138)
    sc2 = dataclasses.replace(sc, x_liquid=end, x_stick=end_stick, traj=traj[10:20])
    Fo = Oracle("f64").rollout(sc2)
    assert force_err(F2, Fo) <= 1e-4
    with pytest.raises(ValueError):
        sim.hand_off_clip(traj[10:15])
    best, hists = fit_clips(sim, traj, target, 2, 1, [0.1, 0.0, 0.5, 0.0], [5.0, 20.0, 1.5, 1e4], n_fit=3)
    assert len(hists) == 2 and len(hists[0]["losses"]) == 3 and len(hists[1]["losses"]) == 1
    sim.close()


def test_host_checks_and_the_device_failure_flag():
    """Inputs the kernels cannot take are refused on the host before any launch (a NaN
    parameter, a particle whose stencil leaves the grid); a rollout that blows up (a bulk modulus
    far past the explicit time step's stability limit) ends with DS_ENONFINITE from the device
    flag instead of NaN forces (the reference propagates NaN and its driver skips the update,
    gradient_descent_module.py:18)."""
    from diffstir_amd._lib import DiffStirError
    sc = S.tiny_scene(16, 40, seed=1)
    print("[gpu] host checks and the device failure flag", flush=True)
    with Engine([sc], ckpt_every=5) as e:
        with pytest.raises(DiffStirError) as ei:
            e.set_params(np.array([[np.nan, 1.0, 1.0, 1e3]]))
        assert ei.value.code == 1 and "non-finite" in str(ei.value)
        for bad in (-0.05, 1.0 - 1.0 / 32, np.inf):
            x = sc.x_liquid.copy()
            x[7, 1] = bad
            with pytest.raises(DiffStirError) as ei:
                e.upload_scene(0, dataclasses.replace(sc, x_liquid=x))
            assert ei.value.code == 1 and "liquid particle 7" in str(ei.value)
        e.upload_scene(0, sc)
        e.set_params(np.array([[1.0, 1.0, 1.0, 1e9]]))
        e.reset()
        with pytest.raises(DiffStirError) as ei:
            e.forward()
        assert ei.value.code == 5
        # the context stays usable: a stable parameter set runs afterwards
        e.set_params(sc.params[None])
        e.reset()
        e.forward()
        assert np.all(np.isfinite(e.forces()))
