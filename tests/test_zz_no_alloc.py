"""ds_substep_noalloc against ds_forward of the per-particle kernel family (the entry point it is a
variant of), through the C ABI on device memory the caller owns.

Three calls in a row share the library's scratch array: a substep of the 18.5k-particle n_grid-32
scene (0 -> 1), the next substep of that scene from the first call's output (1 -> 2), then a
substep of a smaller n_grid-16 scene in place (state_out = state_in).  Each is checked against a
context running the same scene with DS_KERNELS_PER_PARTICLE: the same kernels, whose global float
atomics add in a run-dependent order, so the comparison is to FP32 rounding (particle state within
1e-5 of its largest component, forces within 1e-4 of the largest force), not bitwise.
"""
import ctypes as C

import numpy as np
import pytest

from diffstir_amd import _lib
from diffstir_amd import scene as S
from diffstir_amd.engine import Engine, scene_desc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _runs(n):
    return (n + 63) // 64 * 64


def initial_soa(sc) -> np.ndarray:
    """The step-0 state (v = 0, C = 0, F = I) in the [24][P] layout of ds_substep_noalloc."""
    counts = (sc.n_liquid, sc.n_stick, sc.n_container)
    P = sum(_runs(n) for n in counts)
    st = np.zeros((24, P), np.float32)
    st[0:3] = 0.5
    st[[15, 19, 23]] = 1.0
    o = 0
    for n, x in zip(counts, (sc.x_liquid, sc.x_stick, sc.x_container)):
        st[0:3, o:o + n] = np.asarray(x, np.float32).T
        o += _runs(n)
    return st


def context_steps(sc, nsteps):
    """The per-particle context after nsteps substeps: liquid and stick state, force rows."""
    with Engine([sc], ckpt_every=4, max_steps=nsteps + 1, kernels="per_particle") as e:
        e.reset()
        e.forward(nsteps)
        return e.read_liquid(0), e.read_stick(0), e.forces()[0]


def noalloc_step(lib, sc, st_in, st_out, f, params):
    dev = "cuda"
    traj2 = torch.tensor(np.asarray(sc.traj[f:f + 2], np.float32), device=dev)
    force = torch.zeros(3, dtype=torch.float64, device=dev)
    status = torch.full((1,), 7, dtype=torch.int32, device=dev)
    desc = scene_desc(sc.cfg)
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.ds_substep_noalloc(C.byref(desc), sc.n_liquid, sc.n_stick, sc.n_container, params.data_ptr(),
                                st_in.data_ptr(), st_out.data_ptr(), traj2.data_ptr(), force.data_ptr(),
                                status.data_ptr(), C.c_void_p(stream))
    _lib.check(rc, None)
    torch.cuda.synchronize()
    return force.cpu().numpy(), int(status.cpu().item())


def unpack(st, sc):
    st = st.cpu().numpy()
    nl, ns = sc.n_liquid, sc.n_stick
    o = _runs(nl)
    liq = {"x": st[0:3, :nl].T, "v": st[3:6, :nl].T, "C": st[6:15, :nl].T, "F": st[15:24, :nl].T}
    stick = {"x": st[0:3, o:o + ns].T, "v": st[3:6, o:o + ns].T, "C": st[6:15, o:o + ns].T}
    return liq, stick


def close(a, b, rel=1e-5):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    return float(np.abs(a - b).max() / scale) <= rel


def check_against_context(sc, st, force, nsteps):
    liq_c, stick_c, F_c = context_steps(sc, nsteps)
    liq, stick = unpack(st, sc)
    for k in ("x", "v", "C", "F"):
        assert close(liq[k], liq_c[k]), k
    for k in ("x", "v", "C"):
        assert close(stick[k], stick_c[k]), "stick " + k
    fe = np.abs(force - F_c[nsteps]).max() / np.abs(F_c[1:nsteps + 1]).max()
    print(f"  substep {nsteps}: force {force} vs {F_c[nsteps]} (rel {fe:.2e})", flush=True)
    assert fe <= 1e-4


def test_noalloc_substep_matches_the_per_particle_forward():
    lib = _lib.load()
    big = S.workload("plain_n32")[0].truncated(3)
    small = S.tiny_scene(16, 2, seed=4)
    assert small.n_liquid < big.n_liquid and small.cfg.n_grid < big.cfg.n_grid
    print(f"[gpu] noalloc substep: {big.n_liquid} liquid (n_grid 32), then {small.n_liquid} (n_grid 16)", flush=True)
    dev = "cuda"
    pb = torch.tensor(np.asarray(big.params, np.float64), device=dev)
    s0 = torch.tensor(initial_soa(big), device=dev)
    s1 = torch.full_like(s0, float("nan"))
    s2 = torch.full_like(s0, float("nan"))
    f1, st1 = noalloc_step(lib, big, s0, s1, 0, pb)
    assert st1 == 0
    check_against_context(big, s1, f1, 1)
    f2, st2 = noalloc_step(lib, big, s1, s2, 1, pb)
    assert st2 == 0
    check_against_context(big, s2, f2, 2)
    # the second input is smaller, on a smaller grid, and updated in place
    ps = torch.tensor(np.asarray(small.params, np.float64), device=dev)
    t0 = torch.tensor(initial_soa(small), device=dev)
    f3, st3 = noalloc_step(lib, small, t0, t0, 0, ps)
    assert st3 == 0
    check_against_context(small, t0, f3, 1)


def test_noalloc_substep_rejects_sizes_beyond_its_scratch():
    lib = _lib.load()
    sc = S.tiny_scene(16, 2, seed=4)
    dev = "cuda"
    buf = torch.zeros(64, dtype=torch.float64, device=dev)
    p = buf.data_ptr()
    desc = scene_desc(sc.cfg)
    desc.n_grid = 128
    assert lib.ds_substep_noalloc(C.byref(desc), 10, 10, 0, p, p, p, p, p, p, None) == 1
    assert b"no-allocation" in lib.ds_last_error(None)
    desc = scene_desc(sc.cfg)
    assert lib.ds_substep_noalloc(C.byref(desc), 300000, 0, 0, p, p, p, p, p, p, None) == 1
    assert b"no-allocation limit" in lib.ds_last_error(None)
