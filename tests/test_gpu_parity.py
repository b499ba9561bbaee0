"""GPU parity of the HIP substep and its adjoint against the FP64 CPU oracle.

Tolerances (FP32 bulk arithmetic on the GPU against the FP64 oracle):
  forces     max_s |F - F_oracle| <= 1e-4 * max_s |F_oracle|  (rows s >= 2; row 1 is the
             stick's start-up spike, which would make any relative figure look small)
  gradients  max_i |g_i - g_oracle_i| <= 1e-3 * max_i |g_oracle_i|
These are the force-profile and parameter-gradient tolerances the design targets.
"""
import dataclasses

import numpy as np
import pytest

from diffstir_amd import scene as S
from diffstir_amd.engine import Engine, svd3, svd3_backward, svd3_values
from diffstir_amd.loss import compute_loss
from oracle.oracle import Oracle, loss_and_grad
from tests.helpers import force_err, gold_scene, grad_err

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-4
GRAD_TOL = 1e-3


@pytest.fixture(scope="module")
def o64():
    return Oracle("f64")


# kernel families; "tiled_fused": the tiled kernels with the forward's G2P fused into the next
# step's P2G (DS_OPT_FUSED_STEP)
KERNELS = ["tiled", "per_particle", "tiled_fused"]


def make_engine(scenes, kernels="tiled", **kw):
    e = Engine(scenes, kernels="tiled" if kernels == "tiled_fused" else kernels, **kw)
    if kernels == "tiled_fused":
        e.configure(fused_step=True)
    return e


def gpu_rollout_grad(sc, nsteps, target, ckpt=10, kernels="tiled"):
    with make_engine([sc], kernels, ckpt_every=ckpt, max_steps=nsteps + 1) as e:
        e.reset()
        e.forward(nsteps)
        F = e.forces()[0]
        L, g = compute_loss(F, target, sc.cfg.loss_kind)
        dth = e.backward(g[None])[0]
    return F, L, dth


@pytest.mark.parametrize("kernels", KERNELS)
@pytest.mark.parametrize("name", ["tiny_hb", "ref_carreau", "ref_cross", "ref_hb", "plain_n32_20"])
def test_golden_forces_and_gradient(name, kernels):
    sc, z = gold_scene(name)
    n = int(z["nsteps"])
    print(f"[gpu] golden {name} ({kernels}): {sc.n_liquid} liquid, {sc.n_stick} stick, {sc.n_container} container, "
          f"{n} steps", flush=True)
    F, L, g = gpu_rollout_grad(sc, n, z["target"], ckpt=7, kernels=kernels)
    assert F[0].tolist() == [0.0, 0.0, 0.0]
    fe = force_err(F, z["forces"])
    ge = grad_err(g, z["dl_dparams"])
    print(f"  force err {fe:.3e}  loss {L:.8g} vs {float(z['loss']):.8g}  grad err {ge:.3e}")
    assert fe <= FORCE_TOL
    assert abs(L - float(z["loss"])) <= 1e-4 * abs(float(z["loss"]))
    assert ge <= GRAD_TOL


@pytest.mark.parametrize("kernels", KERNELS)
@pytest.mark.parametrize("seed,model", [(0, 2), (1, 0), (2, 1)])
def test_live_oracle_random_scenes(o64, seed, model, kernels):
    """Seeded variations checked against the oracle run live (not only the fixtures)."""
    rng = np.random.default_rng(100 + seed)
    if model == 2:
        params = S.hb_params_draw(rng)
        sc = S.tiny_scene(16, 30, seed=seed, params=params)
    else:
        sc = S.tiny_scene(16, 30, seed=seed)
        cfg = dataclasses.replace(sc.cfg, model=model, k_from_param=False, k_bulk=1e3)
        params = np.array([rng.uniform(0.5, 3.0), rng.uniform(0.0, 0.4), rng.uniform(0.5, 1.5),
                           rng.uniform(0.5, 5.0)])
        sc = dataclasses.replace(sc, cfg=cfg, params=params)
    print(f"[gpu] live oracle seed {seed} model {model} params {params}", flush=True)
    Fo = o64.rollout(sc, nsteps=30)
    tgt = 0.7 * Fo
    L0, g0 = loss_and_grad(Fo, tgt, sc.cfg.loss_kind)
    _, dth_o = o64.rollout_grad(sc, g0, nsteps=30, ckpt=6)
    F, L, dth = gpu_rollout_grad(sc, 30, tgt, ckpt=6, kernels=kernels)
    print(f"  force err {force_err(F, Fo):.3e} grad err {grad_err(dth, dth_o):.3e}")
    assert force_err(F, Fo) <= FORCE_TOL
    assert grad_err(dth, dth_o) <= GRAD_TOL


@pytest.mark.parametrize("kernels", KERNELS)
def test_batch_matches_single_instances(kernels):
    """Uneven instances advanced together give each instance's own result."""
    scenes = [S.tiny_scene(16, 25, seed=s, fill_height=h, params=S.hb_params_draw(np.random.default_rng(s)))
              for s, h in [(0, 0.15), (1, 0.25), (2, 0.2)]]
    print(f"[gpu] batch of 3 vs singles ({kernels})", [s.n_liquid for s in scenes], flush=True)
    with make_engine(scenes, kernels, ckpt_every=5) as e:
        e.reset()
        e.forward()
        Fb = e.forces()
        tg = [0.5 * Fb[b] for b in range(3)]
        gs = np.stack([compute_loss(Fb[b], tg[b], 1)[1] for b in range(3)])
        db = e.backward(gs)
    for b, sc in enumerate(scenes):
        F, L, d = gpu_rollout_grad(sc, 25, tg[b], ckpt=5, kernels=kernels)
        assert force_err(Fb[b], F) <= 1e-5
        assert grad_err(db[b], d) <= 1e-5


@pytest.mark.parametrize("kernels", KERNELS)
def test_checkpoint_interval_does_not_change_gradient(kernels):
    sc = S.tiny_scene(16, 24, seed=5)
    print(f"[gpu] checkpoint interval invariance ({kernels})", flush=True)
    F0, _, _ = gpu_rollout_grad(sc, 24, np.zeros((25, 3)), ckpt=24, kernels=kernels)
    tgt = 0.5 * F0
    grads = [gpu_rollout_grad(sc, 24, tgt, ckpt=k, kernels=kernels)[2] for k in (1, 5, 24)]
    for g in grads[1:]:
        assert grad_err(g, grads[0]) <= 1e-5


@pytest.mark.parametrize("kernels", KERNELS)
def test_stick_alone_gives_centripetal_force(kernels):
    """A stick with no liquid or container moves rigidly: its grid velocity is its own velocity,
    so F(s) = M (v_kin(s-1) - v_kin(s-2)) / dt -> M w^2 R towards the circle's centre, and
    F(1) = M v_kin(0) / dt (the start-up step, v = 0 at step 0)."""
    sc = S.stirring_scene(32, 60, fill_radius=0.2, fill_height=0.2, stick_radius=0.05, path_rx=0.15,
                          path_rz=0.15, omega=2 * np.pi * 5)
    cfg = dataclasses.replace(sc.cfg, wall_r=10.0)
    sc = dataclasses.replace(sc, cfg=cfg, x_liquid=np.zeros((0, 3), np.float32))
    print(f"[gpu] stick alone: {sc.n_stick} particles", flush=True)
    with make_engine([sc], kernels, ckpt_every=10) as e:
        e.reset()
        e.forward()
        F = e.forces()[0]
    M = sc.cfg.m_stick * sc.n_stick
    dt = sc.cfg.dt
    tr = sc.traj.astype(np.float64)
    vk = (tr[1:] - tr[:-1]) / np.float32(dt)
    np.testing.assert_allclose(F[1], M * vk[0] / dt, rtol=1e-4)
    acc = (vk[1:] - vk[:-1]) / dt      # F(s) for s >= 2 equals M * acc[s-2]
    np.testing.assert_allclose(F[2:], M * acc[:F.shape[0] - 2], rtol=2e-2, atol=2e-3 * np.abs(M * acc).max())
    w = 2 * np.pi * 5
    R = 0.15
    mag = np.linalg.norm(F[5:, [0, 2]], axis=1)
    np.testing.assert_allclose(mag, M * w * w * R, rtol=2e-2)


def test_every_execution_option_gives_the_oracle_result():
    """Every candidate of the autotuner's search space (rows per wavefront 1/2/4 of the scatter
    kernels, 1/2 particles per thread of the gather kernels, fused or split adjoint P2G) gives the
    oracle's forces and gradient, on one context reconfigured between runs."""
    from diffstir_amd.autotune import candidates
    sc, z = gold_scene("plain_n32_20")
    with Engine([sc], ckpt_every=7, max_steps=21) as e:
        for cand in candidates():
            print(f"[gpu] tiled, options {cand}", flush=True)
            e.configure(**cand)
            e.reset()
            e.forward()
            F = e.forces()[0]
            _, g = compute_loss(F, z["target"], sc.cfg.loss_kind)
            d = e.backward(g[None])[0]
            assert force_err(F, z["forces"]) <= FORCE_TOL, cand
            assert grad_err(d, z["dl_dparams"]) <= GRAD_TOL, cand


def test_graph_replay_matches_launches():
    """A whole rollout and its adjoint replayed as captured hipGraphs give the launched
    sequence's results (bitwise where no stencil takes the global fall-back scatter, see
    test_repeat_runs_are_bitwise_identical), on the capturing call and on later replays with other
    parameters; changing an option drops the graphs and the next call recaptures."""
    sc = S.tiny_scene(16, 25, seed=6)
    print("[gpu] graph replay", flush=True)
    out = {}
    for graph in (False, True):
        with Engine([sc], ckpt_every=6, graph=graph) as e:
            res = []
            for scale in (1.0, 1.3, 1.0):
                e.set_params((sc.params * np.array([scale, 1.0, 1.0, 1.0]))[None])
                e.reset()
                e.forward()
                F = e.forces()
                d = e.backward(np.ones_like(F))
                res.append((F.copy(), d.copy()))
            e.configure(particles_per_thread=1)
            e.reset()
            e.forward()
            F = e.forces()
            res.append((F.copy(), e.backward(np.ones_like(F)).copy()))
            out[graph] = res
    for (Fa, da), (Fb, db) in zip(out[False], out[True]):
        np.testing.assert_allclose(Fb, Fa, rtol=1e-10, atol=1e-10 * np.abs(Fa).max())
        np.testing.assert_allclose(db, da, rtol=1e-10, atol=1e-10 * np.abs(da).max())
    # and the launched sequence is reproducible to the same figure (deterministic sort ranks)
    for (Fa, da), (Fb, db) in ((out[False][0], out[False][2]), (out[True][0], out[True][2])):
        np.testing.assert_allclose(Fb, Fa, rtol=1e-10, atol=1e-10 * np.abs(Fa).max())
        np.testing.assert_allclose(db, da, rtol=1e-10, atol=1e-10 * np.abs(da).max())
    assert not np.array_equal(out[True][0][0], out[True][1][0])   # the parameters did change


def test_overlapped_adjoint_matches_one_stream():
    """The adjoint with segment k-1 recomputed on a second stream during segment k's adjoint (two
    segment buffer sets, separate staging) gives the one-stream results, launched and as a replayed
    hipGraph, over an odd and an even number of checkpoint segments and several calls in a row."""
    scs = [S.tiny_scene(16, 40, seed=8), S.tiny_scene(16, 40, seed=9)]
    for K in (6, 13):   # 7 and 3 segments
        out = {}
        for overlap, graph in ((False, False), (True, False), (True, True)):
            print(f"[gpu] overlap {overlap}, graph {graph}, K {K}", flush=True)
            with Engine(scs, ckpt_every=K, overlap=overlap, graph=graph) as e:
                res = []
                for scale in (1.0, 0.8, 1.0):
                    e.set_params(np.stack([s.params * np.array([scale, 1.0, 1.0, 1.0]) for s in scs]))
                    e.reset()
                    e.forward()
                    F = e.forces()
                    res.append((F.copy(), e.backward(np.ones_like(F)).copy()))
                out[(overlap, graph)] = res
        ref = out[(False, False)]
        for key in ((True, False), (True, True)):
            for (Fa, da), (Fb, db) in zip(ref, out[key]):
                np.testing.assert_allclose(Fb, Fa, rtol=1e-10, atol=1e-10 * np.abs(Fa).max())
                np.testing.assert_allclose(db, da, rtol=1e-10, atol=1e-10 * np.abs(da).max(), err_msg=str((K, key)))


def test_repeat_runs_are_bitwise_identical():
    """Forces and parameter gradient have the same bits in every run: launched twice in fresh
    contexts, replayed as hipGraphs, and with the adjoint's recompute on a second stream or on the
    main one.  The sort ranks are fixed by particle index, every grid node sums its staging slots
    in a fixed order, the stick-force partials are added as exact multiples of 2^-32 (order-free)
    and the parameter-gradient partials are reduced in a fixed order.  This scene never takes the
    fall-back scatter (checked; its own repeat test is below)."""
    sc = S.tiny_scene(16, 25, seed=6)
    out = []
    for graph, overlap in ((False, True), (False, True), (True, True), (False, False), (True, False)):
        print(f"[gpu] bitwise repeat, graph {graph}, overlap {overlap}", flush=True)
        with Engine([sc], ckpt_every=6, graph=graph, overlap=overlap) as e:
            e.set_diagnostics(count_fallback=True)
            e.reset()
            e.forward()
            F = e.forces()
            d = e.backward(np.ones_like(F))
            assert e.fallback_count() == 0
            out.append((F.copy(), d.copy()))
    F0, d0 = out[0]
    assert np.abs(F0).max() > 0 and np.abs(d0).max() > 0
    for F, d in out[1:]:
        np.testing.assert_array_equal(F, F0)
        np.testing.assert_array_equal(d, d0)


@pytest.mark.parametrize("fused", [False, True])
def test_fallback_scatter_is_bitwise_repeatable(fused):
    """With the tile window narrowed to 0 almost every stencil takes the fall-back scatter, and
    each fall-back node sums contributions of many particles with global atomics in whatever
    order the hardware runs them.  Those contributions are rounded to a per-particle grain and
    summed in FP64, where the sums are exact: forces and parameter gradient have the same bits in
    fresh contexts, with graph replay and with the adjoint's recompute on one stream."""
    sc, z = gold_scene("plain_n32_20")
    n = int(z["nsteps"])
    out = []
    for graph, overlap in ((False, True), (False, True), (True, True), (False, False)):
        print(f"[gpu] fall-back bitwise repeat, fused {fused}, graph {graph}, overlap {overlap}", flush=True)
        with Engine([sc], ckpt_every=7, max_steps=n + 1, graph=graph, overlap=overlap) as e:
            e.configure(fused_step=fused)
            e.set_diagnostics(tile_window=0, count_fallback=True)
            e.reset()
            e.forward()
            F = e.forces()
            d = e.backward(np.ones_like(F))
            assert e.fallback_count() > 0.5 * n * (sc.n_liquid + sc.n_stick + sc.n_container)
            out.append((F.copy(), d.copy()))
    F0, d0 = out[0]
    assert np.abs(F0).max() > 0 and np.abs(d0).max() > 0
    assert force_err(F0[0], z["forces"]) <= FORCE_TOL
    for F, d in out[1:]:
        np.testing.assert_array_equal(F, F0)
        np.testing.assert_array_equal(d, d0)


def test_chunk_grid_does_not_change_results():
    """The chunk kernels give the same forces and gradient whatever their launch grid: one
    workgroup per chunk, or fewer workgroups each looping over several chunks."""
    sc = S.tiny_scene(16, 20, seed=3)
    out = []
    with Engine([sc], ckpt_every=6) as e:
        for grid in (0, 64, 1 << 20):
            print(f"[gpu] chunk grid {grid}", flush=True)
            e.configure(chunk_grid=grid)
            e.reset()
            e.forward()
            F = e.forces()
            out.append((F.copy(), e.backward(np.ones_like(F)).copy()))
    for F, d in out[1:]:
        np.testing.assert_allclose(F, out[0][0], rtol=1e-10, atol=1e-10 * np.abs(out[0][0]).max())
        np.testing.assert_allclose(d, out[0][1], rtol=1e-10, atol=1e-10 * np.abs(out[0][1]).max())


def test_autotune_picks_a_candidate_on_the_device():
    from diffstir_amd import autotune as A
    A.clear_cache()
    sc = S.tiny_scene(16, 30, seed=4)
    print("[gpu] autotune", flush=True)
    res = A.tune([sc], ckpt_every=6, substeps=12, repeats=1)
    opts = dict(res["options"])
    assert opts.pop("sort_every") in A.sort_candidates(6)
    assert opts.pop("reduce_parts") in (1, 2, 4)
    assert opts.pop("xgrad_stream") in (False, True)
    assert opts.pop("g2p_split") <= opts["gather_particles_per_thread"]
    assert opts.pop("chunk_grid") in (0, 2048)
    assert opts.pop("fused_step") in ((False, True) if opts["scatter_chunks"] == "rows" else (False,))
    assert opts in A.candidates()
    assert res["warm_steps"] == min((sc.max_steps - 1) // 2, sc.max_steps - 1 - 12)
    n_split = len(A.stage_candidates("g2p_split", 6, opts)) - 1
    n_fused = len(A.stage_candidates("fused_step", 6, opts)) - 1
    # the last pass: the other two rows-per-wavefront values (row chunks only), the other two
    # reduction splits
    n_last = (2 if opts["scatter_chunks"] == "rows" else 0) + 2
    assert len(res["timings_ms"]) == (len(A.candidates()) + len(A.sort_candidates(6)) - 1 + 2 + 1 + n_split + n_fused
                                      + 1 + n_last)
    assert all(np.isfinite(t["ms"]) and t["ms"] > 0 for t in res["timings_ms"])
    A.clear_cache()


def test_tiled_matches_per_particle_kernels_over_a_longer_rollout():
    """The two kernel families (the tiled one with and without the fused forward step) on the
    18.5k-particle n_grid 32 scene for 150 substeps with two re-sorts per checkpoint interval of 60:
    state, forces and gradient agree to FP32 accumulation-order noise."""
    sc = S.workload("plain_n32")[0].truncated(151)
    print(f"[gpu] tiled vs per-particle: {sc.n_liquid} particles, 150 steps", flush=True)
    out = {}
    for kernels in KERNELS:
        with make_engine([sc], kernels, ckpt_every=60) as e:
            e.reset()
            e.forward()
            F = e.forces()[0]
            st = e.read_liquid(0)
            stick = e.read_stick(0)
            _, g = compute_loss(F, 0.5 * F, 1)
            d = e.backward(g[None])[0]
        out[kernels] = (F, st, stick, d)
    Fp, st_p, sk_p, dp = out["per_particle"]
    for kernels in ("tiled", "tiled_fused"):
        Ft, st_t, sk_t, dt_ = out[kernels]
        assert force_err(Ft, Fp) <= 1e-5, kernels
        for k in ("x", "v", "C", "F"):
            np.testing.assert_allclose(st_t[k], st_p[k], rtol=1e-4, atol=1e-5 * np.abs(st_p[k]).max(),
                                       err_msg=kernels)
        np.testing.assert_allclose(sk_t["x"], sk_p["x"], atol=1e-6, err_msg=kernels)
        assert grad_err(dt_, dp) <= 1e-4, kernels


# ------------------------------------------------------------------------- FP64 SVD

def _random_F(rng, n):
    F = np.eye(3) + 0.3 * rng.normal(size=(n, 3, 3))
    F[: n // 8] = np.eye(3)                                     # exactly degenerate
    F[n // 8: n // 4] = np.eye(3) + 1e-7 * rng.normal(size=(n // 4 - n // 8, 3, 3))
    F[n // 4: n // 4 + 8, :, 0] *= -1.0                          # inverted
    F[n // 4 + 8: n // 4 + 16] *= 1e-3                           # tiny
    return F


def test_svd3_matches_lapack():
    rng = np.random.default_rng(0)
    F = _random_F(rng, 512)
    U, Sg, V = svd3(F)
    rec = U @ (Sg[:, :, None] * np.swapaxes(V, 1, 2))
    assert np.abs(rec - F).max() <= 1e-12 * max(1.0, np.abs(F).max())
    eye = np.eye(3)
    assert np.abs(np.swapaxes(U, 1, 2) @ U - eye).max() < 1e-12
    assert np.abs(np.swapaxes(V, 1, 2) @ V - eye).max() < 1e-12
    np.testing.assert_allclose(np.linalg.det(U), 1.0, atol=1e-12)
    np.testing.assert_allclose(np.linalg.det(V), 1.0, atol=1e-12)
    ref = np.linalg.svd(F, compute_uv=False)
    np.testing.assert_allclose(np.abs(Sg), ref, atol=1e-13, rtol=1e-12)
    np.testing.assert_allclose(Sg.prod(1), np.linalg.det(F), rtol=1e-11, atol=1e-15)


def test_svd3_values_match_lapack_and_the_full_svd():
    """The values-only SVD of the forward substep: sorted, signed like svd3, S0 S1 S2 = det F."""
    rng = np.random.default_rng(2)
    F = _random_F(rng, 512)
    S = svd3_values(F)
    _, Sf, _ = svd3(F)
    np.testing.assert_allclose(S, Sf, atol=1e-13, rtol=1e-12)
    np.testing.assert_allclose(np.abs(S), np.linalg.svd(F, compute_uv=False), atol=1e-13, rtol=1e-12)
    np.testing.assert_allclose(S.prod(1), np.linalg.det(F), rtol=1e-11, atol=1e-15)
    assert np.all(S[:, 0] >= S[:, 1]) and np.all(S[:, 1] >= np.abs(S[:, 2]))
    assert np.all(np.sign(S[:, 2]) * np.sign(np.linalg.det(F)) >= 0)


def _rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


def test_svd3_values_on_hard_spectra():
    """Spectra where a closed form is most fragile, with known singular values: an equal leading
    pair (the triple-angle cubic's double root at cos = 1/2), an equal trailing pair (its simple
    root at 1), three equal values, rank one and two, six decades of range, and plain shears
    (whose smallest value comes from the determinant). Each F = Q1 diag(s) Q2^T with random
    rotations, one column sign flipped for half of them. The forward substep uses only their
    product J = S0 S1 S2, which comes from det F and must hold to rounding: 1e-15 s0^3. Each value
    is held to 3e-8 s0. The closed form is less accurate than LAPACK for single values here
    (measured on gfx950: about 2e-9 s0 at an equal leading pair, where the cubic has a double root,
    and 1.1e-8 s0 at a condition number of 1e6), but a dropped term, a wrong root of the cubic or a
    swapped pair misses either bound by orders of magnitude. """
    rng = np.random.default_rng(5)
    spectra = [(2.0, 2.0, 0.5), (1.0, 1.0, 1e-6), (3.0, 1e-4, 1e-4), (1.5, 0.7, 0.7),
               (1.0, 1.0, 1.0), (1.0, 0.0, 0.0), (2.0, 1.0, 0.0), (1e3, 1.0, 1e-3),
               (1.0, 1.0 - 1e-9, 1.0 - 2e-9), (1.0 + 1e-12, 1.0, 1e-12)]
    F, S_true = [], []
    for s in spectra:
        for k in range(8):
            d = np.diag(s)
            if k % 2:
                d = d @ np.diag([1.0, 1.0, -1.0])
            F.append(_rotation(rng) @ d @ _rotation(rng).T)
            S_true.append(s)
    for k in (0.0, 0.3, 1.0, 5.0, 40.0):        # shears, unit determinant
        F.append(np.array([[1.0, k, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]))
        S_true.append(np.linalg.svd(F[-1], compute_uv=False))
    F = np.ascontiguousarray(np.array(F))
    S_true = np.array(S_true, dtype=np.float64)
    S = svd3_values(F)
    s0 = S_true[:, 0]
    det = np.linalg.det(F)
    assert np.all(np.abs(S.prod(1) - det) <= 1e-15 * s0 ** 3)
    tol = 3e-8 * s0[:, None]
    assert np.all(np.abs(np.abs(S) - S_true) <= tol)
    assert np.all(S[:, 0] >= S[:, 1]) and np.all(S[:, 1] >= np.abs(S[:, 2]))
    clear = np.abs(det) > 1e-10 * s0 ** 3
    assert np.all(np.sign(S[clear, 2]) == np.sign(det[clear]))
    _, Sf, _ = svd3(F)
    assert np.all(np.abs(S - Sf) <= 2 * tol)


def test_svd3_backward_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    n = 64
    F = np.eye(3) + 0.5 * rng.normal(size=(n, 3, 3))
    Ft = torch.tensor(F, dtype=torch.float64, requires_grad=True)
    U, Sg, Vh = torch.linalg.svd(Ft)
    V = Vh.transpose(1, 2)
    gU = rng.normal(size=(n, 3, 3))
    gS = rng.normal(size=(n, 3))
    gV = rng.normal(size=(n, 3, 3))
    L = (U * torch.tensor(gU)).sum() + (Sg * torch.tensor(gS)).sum() + (V * torch.tensor(gV)).sum()
    (gF_t,) = torch.autograd.grad(L, Ft)
    gF = svd3_backward(U.detach().numpy(), Sg.detach().numpy(), V.detach().numpy(), gU, gS, gV)
    np.testing.assert_allclose(gF, gF_t.numpy(), rtol=1e-8, atol=1e-10)


def test_particle_order_does_not_matter():
    """Shuffling the liquid particles of a scene changes only the FP32 summation order inside the
    tiles (the sort ranks each cell's particles by their given index): forces, gradient and the
    final state (read back in the given order) agree to that rounding."""
    sc = S.tiny_scene(16, 30, seed=7)
    perm = np.random.default_rng(3).permutation(sc.n_liquid)
    shuffled = dataclasses.replace(sc, x_liquid=np.ascontiguousarray(sc.x_liquid[perm]))
    out = []
    for scene in (sc, shuffled):
        print(f"[gpu] particle order ({'shuffled' if scene is shuffled else 'given'})", flush=True)
        with Engine([scene], ckpt_every=8) as e:
            e.reset()
            e.forward()
            F = e.forces()[0]
            st = e.read_liquid(0)
            _, g = compute_loss(F, 0.5 * F, 1)
            out.append((F, st, e.backward(g[None])[0]))
    (Fa, sa, da), (Fb, sb, db) = out
    assert force_err(Fb, Fa) <= 1e-5
    inv = np.argsort(perm)
    for k in ("x", "v", "C", "F"):
        np.testing.assert_allclose(sb[k][inv], sa[k], rtol=1e-4, atol=1e-5 * np.abs(sa[k]).max())
    assert grad_err(db, da) <= 1e-4


def test_sorting_more_often_gives_the_same_result():
    """Bucket sorts every 30 or 12 substeps inside a 60-substep checkpoint interval (the adjoint
    replays the forward's sorts: it permutes the recomputed state with the forward's maps and
    un-permutes the particle adjoint at each sort) give the once-per-interval forces, gradient and
    state to FP32 summation-order noise: launched in several forward calls, replayed as a hipGraph,
    with and without the overlapped recompute."""
    sc = S.stirring_scene(32, 150, fill_radius=0.25, fill_height=0.2, stick_radius=0.05, path_rx=0.15, path_rz=0.15,
                          omega=2 * np.pi * 3, seed=2)
    res = {}
    for R, graph, overlap in ((0, False, True), (30, False, True), (12, False, False), (12, True, True)):
        print(f"[gpu] sort every {R or 60}, graph {graph}, overlap {overlap}", flush=True)
        with Engine([sc], ckpt_every=60, graph=graph, overlap=overlap) as e:
            e.configure(sort_every=R)
            for rep in range(2):
                e.reset()
                if graph:
                    e.forward()
                else:
                    e.forward(7)
                    e.forward(50)
                    e.forward()
                F = e.forces()[0]
                st = e.read_liquid(0)
                _, g = compute_loss(F, 0.5 * F, 1)
                d = e.backward(g[None])[0]
            res[(R, graph, overlap)] = (F, st, d)
    Fa, sa, da = res[(0, False, True)]
    for key, (Fb, sb, db) in res.items():
        assert force_err(Fb, Fa) <= 1e-5, key
        for k in ("x", "v", "C", "F"):
            np.testing.assert_allclose(sb[k], sa[k], rtol=1e-4, atol=1e-5 * np.abs(sa[k]).max(), err_msg=str(key))
        assert grad_err(db, da) <= 1e-4, key


def test_split_grid_reductions_give_the_same_result():
    """Node-block reductions split over 2 or 4 wavefronts (each sums every n-th round of the
    covering chunk slots; the partial sums are added in a fixed order) give the one-wavefront
    forces, gradient and state to FP32 summation-order noise.  Small chunks (4 rows) on a
    fast-stirred scene give buckets several chunks, so the reductions have several rounds."""
    sc = S.stirring_scene(32, 150, fill_radius=0.25, fill_height=0.2, stick_radius=0.05, path_rx=0.15, path_rz=0.15,
                          omega=2 * np.pi * 3, seed=2)
    res = {}
    for parts in (1, 2, 4):
        print(f"[gpu] reduce parts {parts}", flush=True)
        with Engine([sc], ckpt_every=60, particles_per_thread=1) as e:
            e.configure(reduce_parts=parts)
            assert e.options()["reduce_parts"] == parts
            e.reset()
            e.forward()
            F = e.forces()[0]
            st = e.read_liquid(0)
            _, g = compute_loss(F, 0.5 * F, 1)
            res[parts] = (F, st, e.backward(g[None])[0])
    Fa, sa, da = res[1]
    for parts, (Fb, sb, db) in res.items():
        assert force_err(Fb, Fa) <= 1e-5, parts
        for k in ("x", "v", "C", "F"):
            np.testing.assert_allclose(sb[k], sa[k], rtol=1e-4, atol=1e-5 * np.abs(sa[k]).max(), err_msg=str(parts))
        assert grad_err(db, da) <= 1e-4, parts


def test_position_gradient_stream_gives_the_same_gradient():
    """The adjoint P2G's position gradient computed by its own kernel on a third stream, beside
    the adjoint G2P and grid reduction of the same step, gives the gradient computed inside the
    adjoint P2G: launched and replayed as a hipGraph, with and without the overlapped recompute,
    with one sort per checkpoint interval and with sorts inside it."""
    sc = S.stirring_scene(32, 150, fill_radius=0.25, fill_height=0.2, stick_radius=0.05, path_rx=0.15, path_rz=0.15,
                          omega=2 * np.pi * 3, seed=2)
    res = {}
    for xs, graph, overlap, R in ((False, False, True, 0), (True, False, True, 0), (True, True, True, 0),
                                  (False, False, False, 30), (True, False, False, 30), (True, True, True, 30)):
        print(f"[gpu] position-gradient stream {xs}, graph {graph}, overlap {overlap}, sort every {R or 60}",
              flush=True)
        with Engine([sc], ckpt_every=60, graph=graph, overlap=overlap) as e:
            e.configure(sort_every=R, xgrad_stream=xs)
            assert e.options()["xgrad_stream"] == xs
            for rep in range(2):
                e.reset()
                e.forward()
                F = e.forces()[0]
                _, g = compute_loss(F, 0.5 * F, 1)
                d = e.backward(g[None])[0]
            res[(xs, graph, overlap, R)] = (F, d)
    for R in (0, 30):
        Fa, da = res[(False, False, R == 0, R)]
        for key, (Fb, db) in res.items():
            if key[3] != R:
                continue
            assert force_err(Fb, Fa) <= 1e-6, key
            assert grad_err(db, da) <= 1e-6, key


def test_g2p_split_gives_the_same_result():
    """The forward's G2P running each gather chunk as 2 or 4 workgroups (chunks of 2 or 4
    particles per thread, as the adjoint P2G runs them) gives the one-workgroup state exactly and
    the forces and gradient to the rounding of the FP64 stick-force atomics."""
    sc = S.tiny_scene(16, 20, seed=8)
    out = {}
    with Engine([sc], ckpt_every=6) as e:
        for gppt, split in ((2, 1), (2, 2), (4, 1), (4, 2), (4, 4)):
            print(f"[gpu] gather particles per thread {gppt}, G2P split {split}", flush=True)
            e.configure(gather_particles_per_thread=gppt, g2p_split=split)
            e.reset()
            e.forward()
            F = e.forces()
            st = e.read_liquid(0)
            out[(gppt, split)] = (F.copy(), st, e.backward(np.ones_like(F)).copy())
    Fa, sa, da = out[(2, 1)]
    for key, (F, st, d) in out.items():
        for k in ("x", "v", "C", "F"):
            np.testing.assert_array_equal(st[k], sa[k], err_msg=str(key))
        np.testing.assert_allclose(F, Fa, rtol=1e-10, atol=1e-10 * np.abs(Fa).max(), err_msg=str(key))
        np.testing.assert_allclose(d, da, rtol=1e-10, atol=1e-10 * np.abs(da).max(), err_msg=str(key))


def test_fused_g2p_p2g_gives_the_same_result():
    """The forward's G2P fused with the next step's P2G (one kernel on the row chunks wherever no
    sort falls between the two steps) gives the separate kernels' particle state bit for bit at the
    end of a rollout and, with a loss gradient that does not depend on the forces, the same
    gradient bit for bit (the recompute of the adjoint replays the same states): the per-particle
    arithmetic, the rows and the per-node sum order are the same.  The forces
    agree to the rounding of the stick-force partials, which the two forms group into different
    workgroups.  Cases: several forward calls (each ends on a separate G2P), sorts inside the
    checkpoint interval, 1 and 4 rows per wavefront, the overlapped recompute, graph replay; and two
    fused runs have the same bits."""
    sc = S.stirring_scene(32, 150, fill_radius=0.25, fill_height=0.2, stick_radius=0.05, path_rx=0.15, path_rz=0.15,
                          omega=2 * np.pi * 3, seed=2)
    res = {}
    for fused, R, ppt, graph, overlap in ((False, 0, 2, False, True), (True, 0, 2, False, True),
                                          (True, 0, 2, True, True), (False, 30, 1, False, False),
                                          (True, 30, 1, False, False), (True, 30, 1, True, True),
                                          (False, 12, 4, False, True), (True, 12, 4, False, True)):
        print(f"[gpu] fused G2P + P2G {fused}, sort every {R or 60}, rows per wavefront {ppt}, graph {graph}, "
              f"overlap {overlap}", flush=True)
        with Engine([sc], ckpt_every=60, graph=graph, overlap=overlap) as e:
            e.configure(sort_every=R, particles_per_thread=ppt, fused_step=fused)
            assert e.options()["fused_step"] == fused
            e.reset()
            if graph:
                e.forward()
            else:
                e.forward(7)
                e.forward(50)
                e.forward()
            F = e.forces()[0]
            st = e.read_liquid(0)
            d = e.backward(np.ones_like(F)[None])[0]
            res[(fused, R, ppt, graph, overlap)] = (F, st, d)
    for (fused, R, ppt, graph, overlap), (F, st, d) in res.items():
        if not fused:
            continue
        key = (fused, R, ppt, graph, overlap)
        Fa, sa, da = next(v for k, v in res.items() if not k[0] and k[1] == R and k[2] == ppt)
        for k in ("x", "v", "C", "F"):
            np.testing.assert_array_equal(st[k], sa[k], err_msg=str(key))
        np.testing.assert_array_equal(d, da, err_msg=str(key))
        np.testing.assert_allclose(F, Fa, rtol=1e-10, atol=1e-10 * np.abs(Fa).max(), err_msg=str(key))
    # repeat runs of the fused form: the same bits, launched and replayed
    Fa, _, da = res[(True, 0, 2, False, True)]
    Fb, _, db = res[(True, 0, 2, True, True)]
    np.testing.assert_array_equal(Fb, Fa)
    np.testing.assert_array_equal(db, da)


@pytest.mark.parametrize("opts", [
    dict(sort_every=5, reduce_parts=4, xgrad_stream=True, g2p_split=2, gather_particles_per_thread=2),
    dict(sort_every=5, reduce_parts=2, fused_step=True, particles_per_thread=1),
    dict(fused_step=True, particles_per_thread=4, xgrad_stream=True, adjoint_p2g="split"),
    dict(sort_every=7, reduce_parts=2, xgrad_stream=True, adjoint_p2g="split", scatter_chunks="packed"),
    dict(reduce_parts=2, xgrad_stream=False, g2p_split=4, gather_particles_per_thread=4, particles_per_thread=1),
])
def test_later_stage_options_give_the_oracle_result(opts):
    """Combinations of the autotuner's later-stage options (sort interval, reduction parts, the
    position-gradient stream, G2P split, G2P fused with the next step's P2G) with the kernel options
    give the oracle's forces and gradient on the golden n_grid 32 scene, launched and replayed as a
    hipGraph."""
    sc, z = gold_scene("plain_n32_20")
    for graph in (False, True):
        print(f"[gpu] options {opts}, graph {graph}", flush=True)
        with Engine([sc], ckpt_every=35 if opts.get("sort_every") == 7 else 10, max_steps=21, graph=graph) as e:
            e.configure(**opts)
            for rep in range(2 if graph else 1):
                e.reset()
                e.forward()
                F = e.forces()[0]
                _, g = compute_loss(F, z["target"], sc.cfg.loss_kind)
                d = e.backward(g[None])[0]
                assert force_err(F, z["forces"]) <= FORCE_TOL, (opts, graph)
                assert grad_err(d, z["dl_dparams"]) <= GRAD_TOL, (opts, graph)
