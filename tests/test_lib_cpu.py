"""The C-ABI library without a GPU: it loads, exports every entry point include/diffstir.h
declares, and rejects bad arguments before touching the device.  Plus the host-side loss."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from diffstir_amd import _lib
from diffstir_amd import scene as S
from diffstir_amd.engine import scene_desc
from diffstir_amd.loss import compute_loss
from oracle.oracle import loss_and_grad

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "diffstir.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ds_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_function():
    lib = _lib.load()
    names = header_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTS)


def test_library_is_in_tree():
    assert os.path.dirname(_lib.LIB_PATH) == os.path.join(ROOT, "diffstir_amd")


def _create(desc, counts=((10,), (5,), (0,)), T=11, K=5):
    lib = _lib.load()
    B = len(counts[0])
    arr = [(C.c_int32 * B)(*c) for c in counts]
    h = C.c_void_p()
    rc = lib.ds_create(C.byref(desc), B, arr[0], arr[1], arr[2], T, K, C.byref(h))
    return rc, lib.ds_last_error(None).decode()


@pytest.mark.parametrize("field,value,msg", [
    ("n_grid", 4, "n_grid"), ("model", 7, "model"), ("dt", 0.0, "dt"),
])
def test_create_rejects_bad_scene(field, value, msg):
    d = scene_desc(S.tiny_scene().cfg)
    setattr(d, field, value)
    rc, err = _create(d)
    assert rc == 1 and msg in err


def test_create_rejects_k_from_param_outside_herschel_bulkley():
    d = scene_desc(S.tiny_scene().cfg)
    d.model = 0
    d.k_from_param = 1
    rc, err = _create(d)
    assert rc == 1 and "Herschel-Bulkley" in err


def test_create_rejects_bad_sizes():
    d = scene_desc(S.tiny_scene().cfg)
    assert _create(d, T=1)[0] == 1
    assert _create(d, K=0)[0] == 1
    assert _create(d, counts=((-1,), (5,), (0,)))[0] == 1


def test_null_context_is_an_error():
    lib = _lib.load()
    assert lib.ds_forward(None, 1) == 1
    assert lib.ds_reset(None) == 1


@pytest.mark.parametrize("kind", [0, 1])
def test_host_loss_matches_oracle_loss(kind):
    rng = np.random.default_rng(kind)
    F = rng.normal(size=(50, 3))
    F[0] = 0
    t = rng.normal(size=(50, 3))
    L1, g1 = compute_loss(F, t, kind)
    L2, g2 = loss_and_grad(F, t, kind)
    assert L1 == pytest.approx(L2, rel=1e-14)
    np.testing.assert_allclose(g1, g2, rtol=1e-14)
    # gradient by central differences (away from the |.| kink)
    h = 1e-7
    for s, a in [(3, 0), (10, 2), (49, 0)]:
        Fp, Fm = F.copy(), F.copy()
        Fp[s, a] += h
        Fm[s, a] -= h
        fd = (compute_loss(Fp, t, kind)[0] - compute_loss(Fm, t, kind)[0]) / (2 * h)
        assert g1[s, a] == pytest.approx(fd, rel=1e-5, abs=1e-9)


def test_reference_l1_loss_counts_steps_0_to_T_minus_2():
    """sf:332-339: the loop runs over steps 0..T-2 and divides by T-1."""
    T = 7
    F = np.zeros((T, 3))
    F[:, 0] = np.arange(T)
    L, g = compute_loss(F, np.zeros(T), 0)
    assert L == pytest.approx(sum(range(T - 1)) / (T - 1))
    assert g[T - 1, 0] == 0 and g[1, 0] == pytest.approx(1 / (T - 1))


def test_engine_option_numbers_match_the_header():
    """The Python engine passes the option and value numbers that include/diffstir.h declares."""
    from diffstir_amd.engine import Engine
    txt = open(os.path.join(ROOT, "include", "diffstir.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    enum = {k: int(v) for k, v in re.findall(r"\b(DS_[A-Z0-9_]+)\s*=\s*(-?\d+)", txt)}
    want = {"OPT_KERNELS": "DS_OPT_KERNELS", "OPT_PPT": "DS_OPT_PARTICLES_PER_THREAD",
            "OPT_TIMING": "DS_OPT_KERNEL_TIMING", "OPT_GATHER_PPT": "DS_OPT_GATHER_PARTICLES_PER_THREAD",
            "OPT_ADJ_P2G": "DS_OPT_ADJOINT_P2G", "OPT_SCATTER": "DS_OPT_SCATTER_CHUNKS", "OPT_GRAPH": "DS_OPT_GRAPH",
            "OPT_OVERLAP": "DS_OPT_OVERLAP", "OPT_CHUNK_GRID": "DS_OPT_CHUNK_GRID",
            "OPT_SORT_EVERY": "DS_OPT_SORT_EVERY", "OPT_REDUCE_PARTS": "DS_OPT_REDUCE_PARTS",
            "OPT_XGRAD_STREAM": "DS_OPT_XGRAD_STREAM", "OPT_G2P_SPLIT": "DS_OPT_G2P_SPLIT",
            "OPT_TILE_WINDOW": "DS_OPT_TILE_WINDOW", "OPT_COUNT_FALLBACK": "DS_OPT_COUNT_FALLBACK",
            "OPT_FUSED_STEP": "DS_OPT_FUSED_STEP"}
    for attr, name in want.items():
        assert getattr(Engine, attr) == enum[name], name
    assert len([k for k in enum if k.startswith("DS_OPT_")]) == len(want)
    assert Engine.KERNELS == {"tiled": enum["DS_KERNELS_TILED"], "per_particle": enum["DS_KERNELS_PER_PARTICLE"]}
    assert Engine.ADJ_P2G == {"fused": enum["DS_ADJ_P2G_FUSED"], "split": enum["DS_ADJ_P2G_SPLIT"]}
    assert Engine.SCATTER == {"rows": enum["DS_SCATTER_ROWS"], "packed": enum["DS_SCATTER_PACKED"]}
    assert len(_lib.KERNEL_NAMES) == enum["DS_NKERNELS"]


def test_noalloc_substep_rejects_bad_arguments_on_the_host():
    """ds_substep_noalloc validates its sizes before touching the device: beyond its scratch
    (n_grid above 64, more than 262144 padded slots) or with NULL buffers it returns DS_EINVAL."""
    lib = _lib.load()
    desc = scene_desc(S.tiny_scene(16, 2).cfg)
    p = C.c_void_p(4096)   # never dereferenced: every case fails validation first
    desc.n_grid = 128
    assert lib.ds_substep_noalloc(C.byref(desc), 10, 10, 0, p, p, p, p, p, p, None) == 1
    assert b"no-allocation" in lib.ds_last_error(None)
    desc.n_grid = 16
    assert lib.ds_substep_noalloc(C.byref(desc), 262144, 1, 0, p, p, p, p, p, p, None) == 1
    assert b"no-allocation limit" in lib.ds_last_error(None)
    assert lib.ds_substep_noalloc(C.byref(desc), 10, 10, 0, p, None, p, p, p, p, None) == 1
    assert lib.ds_substep_noalloc(C.byref(desc), -1, 10, 0, p, p, p, p, p, p, None) == 1
