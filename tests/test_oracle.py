"""Pins of the FP64 CPU oracle (oracle/mpm_oracle.c) by facts that do not restate its code:
closed forms, conservation laws, exact reconstruction, branch enumeration, the reference's
own output files, an independent autograd restatement, and central finite differences."""
import dataclasses
import itertools

import numpy as np
import pytest

from diffstir_amd import scene as S
from oracle.oracle import Oracle, loss_and_grad
from tests.helpers import golden, gold_scene, grad_err, ref_traj


@pytest.fixture(scope="module")
def o64():
    return Oracle("f64")


def _random_state(rng, sc, scale_v=0.3, scale_C=2.0, scale_F=0.1):
    nl, ns, nc = sc.n_liquid, sc.n_stick, sc.n_container
    st = Oracle.initial_state(sc)
    st["lv"] = scale_v * rng.normal(size=(nl, 3))
    st["lC"] = scale_C * rng.normal(size=(nl, 9))
    st["lF"] = (np.eye(3) + scale_F * rng.normal(size=(nl, 3, 3))).reshape(nl, 9)
    st["sv"] = scale_v * rng.normal(size=(ns, 3))
    st["sC"] = scale_C * rng.normal(size=(ns, 9))
    st["cC"] = scale_C * rng.normal(size=(nc, 9))
    return st


# ----------------------------------------------------------------------------- B-spline / G2P

def test_gather_reconstructs_affine_field_exactly(o64):
    """A grid velocity v(node) = A x_node + b is returned exactly by G2P at any particle, with
    C = A: quadratic B-splines reproduce linear fields and 4/dx^2 * sum w (x_i - x_p)(x_i - x_p)^T = I
    (simulation_functions.py:237-244)."""
    sc = S.tiny_scene()
    cfg = sc.cfg
    n = cfg.n_grid
    rng = np.random.default_rng(0)
    A = rng.normal(size=(3, 3))
    b = rng.normal(size=3)
    idx = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    V = (idx / n) @ A.T + b
    x = rng.uniform(2.0 / n, 1 - 3.0 / n, size=(500, 3))
    nv, nC = o64.gather(cfg, V, x)
    np.testing.assert_allclose(nv, x @ A.T + b, atol=1e-12)
    np.testing.assert_allclose(nC.reshape(-1, 3, 3), np.broadcast_to(A, (500, 3, 3)), atol=1e-10)


def test_gather_of_constant_field_has_zero_affine(o64):
    sc = S.tiny_scene()
    n = sc.cfg.n_grid
    V = np.tile([0.3, -1.0, 2.0], (n ** 3, 1))
    x = np.random.default_rng(1).uniform(2.0 / n, 1 - 3.0 / n, size=(200, 3))
    nv, nC = o64.gather(sc.cfg, V, x)
    np.testing.assert_allclose(nv, np.tile([0.3, -1.0, 2.0], (200, 1)), atol=1e-13)
    np.testing.assert_allclose(nC, 0.0, atol=1e-11)


# ----------------------------------------------------------------------------- P2G

def test_p2g_conserves_mass_and_momentum(o64):
    """Sum over nodes of the scatter: mass = sum m_p, momentum = sum m_p (v_p + dt g) for the
    liquid and sum m_p v_p for the other species; the stress and APIC terms sum to zero because
    sum_i w_i (x_i - x_p) = 0 (simulation_functions.py:148-196)."""
    traj = ref_traj()
    sc = S.reference_scene(traj, seed=2, frames=3)
    rng = np.random.default_rng(2)
    st = _random_state(rng, sc)
    P = o64.p2g(sc.cfg, sc.params, st)
    c = sc.cfg
    mass = c.m_liq * sc.n_liquid + c.m_stick * sc.n_stick + c.m_cont * sc.n_container
    mom = c.m_liq * (st["lv"].sum(0) + np.array([0, c.dt * c.gravity, 0]) * sc.n_liquid) \
        + c.m_stick * st["sv"].sum(0) + c.m_cont * st["cv"].sum(0)
    np.testing.assert_allclose(P[..., 3].sum(), mass, rtol=1e-12)
    np.testing.assert_allclose(P[..., :3].reshape(-1, 3).sum(0), mom, rtol=1e-9, atol=1e-12 * mass)


def test_p2g_stress_is_internal_force(o64):
    """Isotropic stress alone moves no net momentum but pushes a free surface outward: with v = 0
    and C = 0 the rest pressure term (k/2)(J - 1/J - 1)J = -k/2 at J = 1 gives net zero momentum
    and an outward first moment."""
    sc = S.tiny_scene()
    sc = dataclasses.replace(sc, x_stick=np.zeros((0, 3), np.float32), cfg=dataclasses.replace(sc.cfg, gravity=0.0))
    st = Oracle.initial_state(sc)
    P = o64.p2g(sc.cfg, sc.params, st)
    mom = P[..., :3].reshape(-1, 3)
    assert np.abs(mom.sum(0)).max() < 1e-12 * np.abs(mom).max()
    n = sc.cfg.n_grid
    idx = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3) / n
    cen = sc.x_liquid.mean(0)
    radial = ((idx - cen) * mom).sum()
    assert radial > 0


def _one_particle_affine(o64, cfg, params, F, Cm, xp=(0.503, 0.4981, 0.5117)):
    """P2G of one liquid particle (v = 0, no gravity, no other species) and the affine matrix A
    read back from the grid: sum_i w_i (x_i - x_p)(x_i - x_p)^T = dx^2/4 I for quadratic
    B-splines (per axis sum w (i - fx)^2 = 1/4, sum w (i - fx) = 0), and the node values are
    w_i (m v + A (x_i - x_p)), so 4/dx^2 * sum_i P_i (x_i - x_p)^T = A exactly, and the mass
    moment sum_i m_i (x_i - x_p) is 0.  The base is int(x/dx - 0.5) (sf:108)."""
    cfg = dataclasses.replace(cfg, gravity=0.0)
    n, dx = cfg.n_grid, 1.0 / cfg.n_grid
    st = {"lx": np.array([xp], np.float64), "lv": np.zeros((1, 3)), "lC": np.asarray(Cm, np.float64).reshape(1, 9),
          "lF": np.asarray(F, np.float64).reshape(1, 9), "sx": np.zeros((0, 3)), "sv": np.zeros((0, 3)),
          "sC": np.zeros((0, 9)), "cx": np.zeros((0, 3)), "cv": np.zeros((0, 3)), "cC": np.zeros((0, 9))}
    P = o64.p2g(cfg, params, st).reshape(-1, 4)
    idx = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    d = idx * dx - np.asarray(xp)
    hit = P[:, 3] != 0
    assert hit.sum() == 27
    np.testing.assert_allclose(P[:, 3].sum(), cfg.m_liq, rtol=1e-13)
    assert np.abs((P[hit, 3:4] * d[hit]).sum(0)).max() < 1e-12 * cfg.m_liq
    return 4.0 / dx ** 2 * P[hit, :3].T @ d[hit]


def _stress_cases():
    """(name, scene config, params, k, eta(gamma) closed form) for the reference scene (k = 20000
    hard-coded, sf:112, Carreau with n = 1: eta = eta0 for every gamma, sf:124) and a synthetic
    workload scene (k = params[3], Herschel-Bulkley with n = 1: eta = tau_y / gamma + K,
    sf:139-146)."""
    ref = S.reference_scene(ref_traj(), seed=0, frames=3, params=(3.0, 0.2, 1.0, 5.0), model=0)
    syn = S.workload("plain_n32")[0]
    return [("reference_carreau", ref.cfg, np.array([3.0, 0.2, 1.0, 5.0]), 20000.0, lambda g: 3.0),
            ("synthetic_hb", syn.cfg, np.array([2.0, 7.0, 1.0, 1000.0]), 1000.0, lambda g: 7.0 / g + 2.0)]


@pytest.mark.parametrize("case", range(2))
def test_stress_assembly_closed_forms(o64, case):
    """sf:148-151, read back from the grid through the second-moment identity above, on inputs
    whose stress has a closed form (alpha = -dt V 4/dx^2, p(J) = (k/2)(J - 1/J - 1) J):

      F = I, C = 0:                J = 1, p = -k/2, A = alpha p I = dt V 2k/dx^2 I
      F = diag(2, 1, 1), C = 0:    J = 2, p = +k/2
      F = I, C = s e0 e1^T:        J = det(I + dt C) = 1, gamma = |s| + 1e-6,
                                   A = alpha (eta(gamma) (C + C^T)/2 - k/2 I) + m C

    Each fails if the '- 1' of p is dropped (p(1) = 0), if the trailing J is dropped (p(2) = k/4),
    if the 1/2 of sym(C) is dropped (A[1,0] doubles), or if 4/dx is used for 4/dx^2 (A scales by
    dx)."""
    name, cfg, th, k, eta = _stress_cases()[case]
    dx = 1.0 / cfg.n_grid
    alpha = -cfg.dt * cfg.vol_liq * 4.0 / dx ** 2
    m = cfg.m_liq
    eye = np.eye(3)
    A = _one_particle_affine(o64, cfg, th, eye, np.zeros((3, 3)))
    tol = 1e-12 * abs(alpha * k)
    np.testing.assert_allclose(A, cfg.dt * cfg.vol_liq * 2.0 * k / dx ** 2 * eye, rtol=1e-10, atol=tol)
    np.testing.assert_allclose(A, alpha * (-k / 2) * eye, rtol=1e-10, atol=tol)
    A = _one_particle_affine(o64, cfg, th, np.diag([2.0, 1.0, 1.0]), np.zeros((3, 3)))
    np.testing.assert_allclose(A, alpha * (k / 2) * eye, rtol=1e-10, atol=tol)
    for s in (0.75, -3.0):
        Cm = np.zeros((3, 3))
        Cm[0, 1] = s
        assert np.linalg.det(eye + cfg.dt * Cm) == 1.0
        g = abs(s) + 1e-6
        A = _one_particle_affine(o64, cfg, th, eye, Cm)
        want = alpha * (eta(g) * 0.5 * (Cm + Cm.T) - k / 2 * eye) + m * Cm
        np.testing.assert_allclose(A, want, rtol=1e-10, atol=1e-12 * np.abs(want).max())
        np.testing.assert_allclose(A[1, 0], alpha * eta(g) * s / 2, rtol=1e-10)


_STRESS_MUTANTS = {
    "pressure without the -1": ("(J - ((REAL)1 / J) - (REAL)1) * J", "(J - ((REAL)1 / J)) * J"),
    "pressure without the trailing J": ("(J - ((REAL)1 / J) - (REAL)1) * J", "(J - ((REAL)1 / J) - (REAL)1)"),
    "sym(C) without the 1/2": ("q->eta * (REAL)0.5 * (C[i * 3 + j] + C[j * 3 + i])",
                               "q->eta * (C[i * 3 + j] + C[j * 3 + i])"),
    "4/dx for 4/dx^2": ("4.0 * sc->inv_dx * sc->inv_dx", "4.0 * sc->inv_dx"),
}


@pytest.mark.parametrize("mutant", sorted(_STRESS_MUTANTS))
def test_stress_pins_catch_planted_mistakes(tmp_path, mutant):
    """Plant each plausible misreading of sf:148-151 in a copy of the oracle, build it, and check
    that the closed forms above reject it for both scenes."""
    import ctypes as C
    import os
    import subprocess
    src = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                            "mpm_oracle.c")).read()
    old, new = _STRESS_MUTANTS[mutant]
    # the forward's stress assembly is liquid_pre(); plant the mistake there only
    a, b = src.index("static int liquid_pre("), src.index("static void scatter(")
    body = src[a:b]
    assert body.count(old) == 1, f"oracle source changed: update the {mutant!r} mutant"
    c_file = tmp_path / "mutant.c"
    c_file.write_text(src[:a] + body.replace(old, new) + src[b:])
    so = tmp_path / "libmutant.so"
    subprocess.check_call(["gcc", "-O2", "-fPIC", "-ffp-contract=off", "-shared", "-o", str(so), str(c_file), "-lm"])
    o = Oracle("f64")
    o._p2g = C.CDLL(str(so)).or_p2g
    for case in range(2):
        with pytest.raises(AssertionError):
            test_stress_assembly_closed_forms(o, case)


# ----------------------------------------------------------------------------- rheology

@pytest.mark.parametrize("g", [1e-3, 0.5, 3.0, 40.0])
def test_viscosity_closed_forms(o64, g):
    # Carreau, n = 1: eta = eta0 for every shear rate
    e, _, _ = o64.eta(0, [2.5, 0.3, 1.0, 7.0], g)
    assert e == pytest.approx(2.5, rel=1e-14)
    # Carreau at lambda*gamma = sqrt(3): (1+3)^((n-1)/2) = 2^(n-1)
    lam = np.sqrt(3.0) / g
    e, _, _ = o64.eta(0, [2.0, 0.5, 0.6, lam], g)
    assert e == pytest.approx(0.5 + 1.5 * 2 ** (0.6 - 1), rel=1e-12)
    # Cross at lambda*gamma = 1: the (n-1) power is 1 for any n -> mean of eta0 and eta_inf
    e, _, _ = o64.eta(1, [3.0, 1.0, 0.37, 1.0 / g], g)
    assert e == pytest.approx(2.0, rel=1e-12)
    # Cross exponent is n - 1 (simulation_functions.py:134): n = 2 -> 1/(1 + lambda*gamma)
    e, _, _ = o64.eta(1, [3.0, 1.0, 2.0, 0.7], g)
    assert e == pytest.approx(1.0 + 2.0 / (1.0 + 0.7 * g), rel=1e-12)
    # Herschel-Bulkley, n = 1: tau_y/gamma + K; K = 0 leaves the yield term only
    e, _, _ = o64.eta(2, [1.7, 4.0, 1.0, 0.0], g)
    assert e == pytest.approx(4.0 / g + 1.7, rel=1e-12)
    e, _, _ = o64.eta(2, [0.0, 4.0, 0.5, 0.0], g)
    assert e == pytest.approx(4.0 / g, rel=1e-12)


@pytest.mark.parametrize("model,th", [(0, [2.0, 0.4, 0.7, 1.3]), (1, [2.0, 0.4, 0.6, 1.7]),
                                      (2, [1.5, 3.0, 0.8, 0.0])])
def test_viscosity_partials_match_finite_differences(o64, model, th):
    th = np.array(th)
    for g in (0.3, 2.0, 11.0):
        e, dg, dth = o64.eta(model, th, g)
        h = 1e-6 * g
        fd = (o64.eta(model, th, g + h)[0] - o64.eta(model, th, g - h)[0]) / (2 * h)
        assert dg == pytest.approx(fd, rel=1e-6, abs=1e-9)
        for i in range(4):
            if th[i] == 0:
                continue
            hp = 1e-6 * abs(th[i])
            tp, tm = th.copy(), th.copy()
            tp[i] += hp
            tm[i] -= hp
            fd = (o64.eta(model, tp, g)[0] - o64.eta(model, tm, g)[0]) / (2 * hp)
            assert dth[i] == pytest.approx(fd, rel=1e-6, abs=1e-9)


# ----------------------------------------------------------------------------- grid update

def test_grid_update_box_bounds_enumerated(o64):
    """Every box-bound branch of simulation_functions.py:211-224 on a grid with the cylinder wall
    disabled: outward components are zeroed on the low/high faces, inward ones kept, and the
    top face (j > n - bound) with upward velocity zeroes all three components."""
    cfg = dataclasses.replace(S.tiny_scene().cfg, wall_r=100.0)
    n, bd = cfg.n_grid, cfg.bound
    lo, mid, hi = 1, n // 2, n - 1
    for (i, j, k), sv in itertools.product(itertools.product((lo, mid, hi), repeat=3),
                                           itertools.product((-1.0, 1.0), repeat=3)):
        v = np.array(sv) * np.array([0.3, 0.7, 1.1])
        m = 2.0
        out = o64.gridop_one(cfg, np.r_[v * (m + 1e-10), m], i, j, k)
        exp = v.copy()
        if (i < bd and v[0] < 0) or (i > n - bd and v[0] > 0):
            exp[0] = 0
        if (k < bd and v[2] < 0) or (k > n - bd and v[2] > 0):
            exp[2] = 0
        if j < bd and v[1] < 0:
            exp[1] = 0
        if j > n - bd and v[1] > 0:
            exp[:] = 0
        np.testing.assert_allclose(out, exp, atol=1e-12)


def test_grid_update_cylinder_wall_removes_radial_component(o64):
    cfg = S.tiny_scene().cfg
    n = cfg.n_grid
    rng = np.random.default_rng(3)
    for _ in range(200):
        i, j, k = rng.integers(cfg.bound + 1, n - cfg.bound - 1, size=3)
        v = rng.normal(size=3)
        out = o64.gridop_one(cfg, np.r_[v, 1.0], i, j, k)
        d = np.array([i / n - cfg.wall_cx, 0.0, k / n - cfg.wall_cz])
        if (d ** 2).sum() > cfg.wall_r ** 2:
            nh = d / np.linalg.norm(d)
            assert abs(out @ nh) < 1e-9
            np.testing.assert_allclose(out[1], v[1], rtol=1e-9)
            np.testing.assert_allclose(out - nh * (out @ nh), v - nh * (v @ nh), atol=1e-9)
        else:
            np.testing.assert_allclose(out, v / (1 + 1e-10), rtol=1e-12)


def test_grid_update_divides_by_mass_plus_epsilon(o64):
    cfg = dataclasses.replace(S.tiny_scene().cfg, wall_r=100.0)
    out = o64.gridop_one(cfg, [1e-10, 0.0, 0.0, 0.0], 8, 8, 8)
    np.testing.assert_allclose(out, [1.0, 0.0, 0.0], rtol=1e-12)


# ----------------------------------------------------------------------------- stick kinematics

def test_stick_alone_gives_centripetal_force(o64):
    """A stick with no liquid moves rigidly, so the grid returns its own velocity: F(1) =
    M v_kin(0)/dt (start-up, v = 0 at step 0) and F(s) = M (v_kin(s-1) - v_kin(s-2))/dt -> M w^2 R
    (simulation_functions.py:266-267, 330-336)."""
    sc = S.stirring_scene(32, 60, fill_radius=0.2, fill_height=0.2, stick_radius=0.05, path_rx=0.15,
                          path_rz=0.15, omega=2 * np.pi * 5)
    sc = dataclasses.replace(sc, cfg=dataclasses.replace(sc.cfg, wall_r=10.0), x_liquid=np.zeros((0, 3), np.float32))
    F = o64.rollout(sc)
    M = sc.cfg.m_stick * sc.n_stick
    dt = sc.cfg.dt
    tr = sc.traj.astype(np.float64)
    vk = (tr[1:] - tr[:-1]) / dt
    np.testing.assert_allclose(F[1], M * vk[0] / dt, rtol=1e-6)
    acc = (vk[1:] - vk[:-1]) / dt
    np.testing.assert_allclose(F[2:], M * acc[:F.shape[0] - 2], rtol=1e-4, atol=1e-4 * np.abs(M * acc).max())
    w = 2 * np.pi * 5
    np.testing.assert_allclose(np.linalg.norm(F[5:, [0, 2]], axis=1), M * w * w * 0.15, rtol=2e-2)


def test_trajectory_transform_matches_reference_output():
    """The reference's output/stick_positions.npy (mean stick position per step) equals the
    trajectory transform of data/traj.npy (initialization.py:74-78) up to a constant offset:
    the stick is moved rigidly by the transformed trajectory's differences
    (simulation_functions.py:266-269)."""
    traj = ref_traj()
    t, _ = S.transform_trajectory(traj)
    z = golden("ref_stick_positions.npz")
    rows, pos = z["rows"], z["pos"].astype(np.float64)
    off = pos - t[rows]
    np.testing.assert_allclose(off, np.broadcast_to(off[0], off.shape), atol=5e-6)
    np.testing.assert_allclose(off[0], [0.50001, 0.171079, 0.49974], atol=2e-5)


def test_reference_scene_stick_follows_trajectory(o64):
    sc = S.reference_scene(ref_traj(), seed=1, frames=21)
    _, st = o64.rollout(sc, return_state=True)
    shift = st["sx"].mean(0) - sc.x_stick.astype(np.float64).mean(0)
    np.testing.assert_allclose(shift, sc.traj[20].astype(np.float64) - sc.traj[0], atol=1e-9)


def test_reference_force_profile_indexing():
    """The reference's exported force profile has row 0 = 0 (dv[0] is never written) and its
    last row 0 (compute_loss stops at max_steps - 2): simulation_functions.py:330-336."""
    z = golden("ref_force_profile_ends.npz")
    assert z["x_head"][0] == 0 and z["z_head"][0] == 0
    assert z["x_tail"][-1] == 0 and z["z_tail"][-1] == 0
    assert np.all(z["x_head"][1:] != 0)


# ----------------------------------------------------------------------------- adjoint

def _torch_state(st, sc):
    import torch
    def t(a, shape):
        return torch.tensor(np.asarray(a, dtype=np.float64).reshape(shape), requires_grad=True)
    nl, ns, nc = sc.n_liquid, sc.n_stick, sc.n_container
    return {"lx": t(st["lx"], (nl, 3)), "lv": t(st["lv"], (nl, 3)), "lC": t(st["lC"], (nl, 3, 3)),
            "lF": t(st["lF"], (nl, 3, 3)), "sx": t(st["sx"], (ns, 3)), "sv": t(st["sv"], (ns, 3)),
            "sC": t(st["sC"], (ns, 3, 3)), "cx": t(st["cx"], (nc, 3)), "cv": t(st["cv"], (nc, 3)),
            "cC": t(st["cC"], (nc, 3, 3))}


@pytest.mark.parametrize("model", [0, 1, 2])
def test_step_adjoint_matches_torch_autograd(o64, model):
    """The oracle's hand-derived adjoint of one substep equals reverse-mode autodiff of an
    independent PyTorch restatement (tests/torch_step.py), for every species and model."""
    torch = pytest.importorskip("torch")
    from tests import torch_step
    sc = S.reference_scene(ref_traj(), seed=4, frames=3, model=model,
                           params={0: (2.0, 0.5, 0.7, 3.0), 1: (2.0, 0.5, 0.6, 2.0), 2: (1.5, 2.0, 0.8, 1.0)}[model])
    sc = dataclasses.replace(sc, x_liquid=sc.x_liquid[:400], x_stick=sc.x_stick[:60], x_container=sc.x_container[:80])
    rng = np.random.default_rng(10 + model)
    st = _random_state(rng, sc, scale_v=0.05, scale_C=3.0, scale_F=0.05)
    th = torch.tensor(sc.params, dtype=torch.float64, requires_grad=True)
    ts = _torch_state(st, sc)
    out, dv = torch_step.step(sc.cfg, th, ts, sc.traj[0].astype(np.float64), sc.traj[1].astype(np.float64))
    # forward agreement first
    o_out, o_dv = o64.step(sc.cfg, sc.params, st, sc.traj[0], sc.traj[1])
    for k in ("lx", "lv", "lC", "lF", "sC", "cC"):
        np.testing.assert_allclose(out[k].detach().numpy().reshape(o_out[k].shape), o_out[k], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(dv.detach().numpy(), o_dv, rtol=1e-9, atol=1e-12)
    # random output adjoints
    ob = {k: rng.normal(size=o_out[k].shape) for k in ("lx", "lv", "lC", "lF", "sC", "cC")}
    dvb = rng.normal(size=3)
    L = sum((out[k].reshape(o_out[k].shape) * torch.tensor(ob[k])).sum() for k in ob) + (dv * torch.tensor(dvb)).sum()
    wrt = [th, ts["lx"], ts["lv"], ts["lC"], ts["lF"], ts["sC"], ts["cC"]]
    g = torch.autograd.grad(L, wrt)
    ib, thb = o64.step_adjoint(sc.cfg, sc.params, st, sc.traj[0], sc.traj[1], ob, dvb)
    np.testing.assert_allclose(thb, g[0].numpy(), rtol=1e-8, atol=1e-10 * np.abs(g[0].numpy()).max())
    for name, gt in zip(("lx", "lv", "lC", "lF", "sC", "cC"), g[1:]):
        gt = gt.numpy().reshape(ib[name].shape)
        np.testing.assert_allclose(ib[name], gt, rtol=1e-7, atol=1e-9 * max(np.abs(gt).max(), 1e-30),
                                   err_msg=name)


@pytest.mark.parametrize("case", ["tiny_hb", "ref_carreau", "ref_cross", "ref_hb"])
def test_rollout_gradient_matches_central_differences(o64, case):
    """dL/dparams of whole rollouts against central finite differences of the FP64 forward."""
    sc, z = gold_scene(case)
    n = int(z["nsteps"])
    n = min(n, 20)
    sc = sc.truncated(n + 1)
    F = o64.rollout(sc)
    tgt = 0.8 * F + 0.05 * np.abs(F).mean() * np.random.default_rng(0).normal(size=F.shape)
    _, g = loss_and_grad(F, tgt, sc.cfg.loss_kind)
    _, thb = o64.rollout_grad(sc, g, ckpt=6)

    def L(th):
        return loss_and_grad(o64.rollout(sc, params=th), tgt, sc.cfg.loss_kind)[0]

    fd = np.zeros(4)
    for i in range(4):
        if sc.cfg.model == 2 and i == 3 and not sc.cfg.k_from_param:
            continue
        h = 1e-6 * max(abs(sc.params[i]), 1e-3)
        tp, tm = sc.params.copy(), sc.params.copy()
        tp[i] += h
        tm[i] -= h
        fd[i] = (L(tp) - L(tm)) / (2 * h)
    assert grad_err(thb, fd) < 1e-5, (thb, fd)


def test_checkpoint_interval_is_exact_in_the_oracle(o64):
    sc = S.tiny_scene(16, 18, seed=2)
    F = o64.rollout(sc)
    _, g = loss_and_grad(F, 0.5 * F, 1)
    ref = o64.rollout_grad(sc, g, ckpt=1)[1]
    for k in (4, 7, 18, 50):
        assert np.array_equal(o64.rollout_grad(sc, g, ckpt=k)[1], ref)


def test_particle_order_does_not_matter(o64):
    sc = S.tiny_scene(16, 15, seed=6)
    perm = np.random.default_rng(0).permutation(sc.n_liquid)
    sc2 = dataclasses.replace(sc, x_liquid=sc.x_liquid[perm])
    F1, F2 = o64.rollout(sc), o64.rollout(sc2)
    np.testing.assert_allclose(F2, F1, rtol=1e-10, atol=1e-10 * np.abs(F1).max())


# ----------------------------------------------------------------------------- fixtures

@pytest.mark.parametrize("name", ["tiny_hb", "ref_carreau", "ref_cross", "ref_hb", "plain_n32_20"])
def test_golden_fixtures_reproduce(o64, name):
    sc, z = gold_scene(name)
    n = int(z["nsteps"])
    F = o64.rollout(sc, nsteps=n)
    np.testing.assert_allclose(F, z["forces"], rtol=1e-12, atol=1e-12 * np.abs(z["forces"]).max())
    L, g = loss_and_grad(F, z["target"], int(z["loss_kind"]))
    assert L == pytest.approx(float(z["loss"]), rel=1e-12)
    if name != "plain_n32_20":
        _, thb = o64.rollout_grad(sc, g, nsteps=n, ckpt=5)
        np.testing.assert_allclose(thb, z["dl_dparams"], rtol=1e-10, atol=1e-12 * np.abs(z["dl_dparams"]).max())


# ----------------------------------------------------------------------------- scenes

def test_workload_particle_counts():
    """8 particles per cell: about 8 pi r^2 h n^3 liquid particles."""
    for name, r, h, n in [("plain_n32", 0.3, 0.25, 32), ("single_n64", 0.3, 0.25, 64)]:
        sc = S.workload(name)[0]
        expect = 8 * np.pi * r * r * h * n ** 3
        assert abs(sc.n_liquid - expect) < 0.01 * expect


def test_unit_density_is_unstable(o64):
    """At liquid density 1 (with k = 1e3 and Herschel-Bulkley parameters in the stated ranges)
    the explicit viscous update is far past its stability limit (eta dt / (rho dx^2) ~ 4) and
    the rollout leaves the grid within a few substeps; density 1000 is stable."""
    bad = S.stirring_scene(32, 10, density=1.0, stick_density=1.0)
    with pytest.raises(RuntimeError):
        o64.rollout(bad)
    good = S.stirring_scene(32, 10)
    assert np.all(np.isfinite(o64.rollout(good)))
