"""`Engine`: a batch of stirring instances on one GPU, over the C ABI of libdiffstir_hip.so.

One Engine owns one device context.  All instances share the grid, time step, rheology
model and species masses of their SceneConfig; each has its own particles, stick trajectory
and constitutive parameters (the reference runs exactly one instance,
src/simulator/initialization.py:123-165).
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .scene import Scene, SceneConfig


def scene_desc(cfg: SceneConfig) -> _lib.SceneDesc:
    return _lib.SceneDesc(int(cfg.n_grid), int(cfg.bound), int(cfg.model), int(bool(cfg.k_from_param)),
                          float(cfg.dt), float(cfg.gravity), float(cfg.k_bulk), float(cfg.wall_r),
                          float(cfg.wall_cx), float(cfg.wall_cz), float(cfg.m_liq), float(cfg.vol_liq),
                          float(cfg.m_stick), float(cfg.m_cont))


def _f32(a, shape):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(shape))


def _p32(a):
    return a.ctypes.data_as(_lib.F32P) if a is not None and a.size else None


def _p64(a):
    return a.ctypes.data_as(_lib.F64P)


_SHARED = ("n_grid", "dt", "bound", "gravity", "model", "k_bulk", "k_from_param", "wall_r", "wall_cx", "wall_cz",
           "m_liq", "vol_liq", "m_stick", "m_cont")


class Engine:
    """B instances advanced together; forward/backward run every instance."""

    KERNELS = {"tiled": 0, "per_particle": 1}

    # execution options of the tiled kernels (ds_set_option ids)
    OPT_KERNELS, OPT_PPT, OPT_TIMING, OPT_GATHER_PPT, OPT_ADJ_P2G, OPT_SCATTER, OPT_GRAPH = 1, 2, 3, 4, 5, 6, 7
    OPT_OVERLAP, OPT_CHUNK_GRID, OPT_SORT_EVERY, OPT_REDUCE_PARTS, OPT_XGRAD_STREAM = 8, 9, 10, 11, 12
    OPT_G2P_SPLIT, OPT_TILE_WINDOW, OPT_COUNT_FALLBACK, OPT_FUSED_STEP = 13, 14, 15, 16
    ADJ_P2G = {"fused": 0, "split": 1}
    SCATTER = {"rows": 0, "packed": 1}

    def __init__(self, scenes: Sequence[Scene], ckpt_every: int = 50, max_steps: Optional[int] = None,
                 kernels: str = "tiled", particles_per_thread: int = 2, kernel_timing_every: int = 0,
                 gather_particles_per_thread: int = 1, adjoint_p2g: str = "fused", scatter_chunks: str = "rows",
                 graph: bool = False, overlap: bool = True):
        self.lib = _lib.load()
        # bumped by every call that invalidates a recorded rollout (reset, new parameters, new
        # particles or trajectory): a consumer holding a rollout checks it before the adjoint
        self.generation = 0
        scenes = list(scenes)
        if not scenes:
            raise ValueError("need at least one scene")
        self.scenes = scenes
        cfg = scenes[0].cfg
        for s in scenes[1:]:
            for k in _SHARED:
                if getattr(s.cfg, k) != getattr(cfg, k):
                    raise ValueError(f"instances must share SceneConfig.{k}")
        self.cfg = cfg
        self.B = len(scenes)
        self.T = int(max_steps if max_steps is not None else min(s.max_steps for s in scenes))
        self.K = int(ckpt_every)
        nl = (C.c_int32 * self.B)(*[s.n_liquid for s in scenes])
        ns = (C.c_int32 * self.B)(*[s.n_stick for s in scenes])
        nc = (C.c_int32 * self.B)(*[s.n_container for s in scenes])
        self._desc = scene_desc(cfg)
        h = C.c_void_p()
        _lib.check(self.lib.ds_create(C.byref(self._desc), self.B, nl, ns, nc, self.T, self.K, C.byref(h)), None)
        self.ctx = h
        if kernels not in self.KERNELS:
            raise ValueError(f"kernels must be one of {list(self.KERNELS)}")
        self.kernels = kernels
        self._check(self.lib.ds_set_option(self.ctx, self.OPT_KERNELS, self.KERNELS[kernels]))
        self._check(self.lib.ds_set_option(self.ctx, self.OPT_TIMING, int(kernel_timing_every)))
        self.configure(particles_per_thread=particles_per_thread,
                       gather_particles_per_thread=gather_particles_per_thread, adjoint_p2g=adjoint_p2g,
                       scatter_chunks=scatter_chunks, overlap=overlap)
        self.sort_every = 0
        self.reduce_parts = 1
        self.xgrad_stream = False
        self.g2p_split = 1
        self.fused_step = False
        self.chunk_grid = 0
        self.set_graph(graph)
        for b, s in enumerate(scenes):
            self.upload_scene(b, s)
        self.set_params(np.stack([np.asarray(s.params, dtype=np.float64) for s in scenes]))

    def upload_scene(self, b: int, s: Scene):
        """Initial particle positions and stick trajectory of instance b (same particle counts as
        the scene the engine was built with); a trajectory shorter than the engine's T frames is
        held at its last frame.  Takes effect at the next reset."""
        xl = _f32(s.x_liquid, (-1, 3))
        xs = _f32(s.x_stick, (-1, 3))
        xc = _f32(s.x_container, (-1, 3))
        self._check(self.lib.ds_set_particles(self.ctx, b, _p32(xl), _p32(xs), _p32(xc)))
        self.generation += 1
        tr = np.asarray(s.traj[:self.T], dtype=np.float32)
        if tr.shape[0] < self.T:
            tr = np.concatenate([tr, np.repeat(tr[-1:], self.T - tr.shape[0], axis=0)])
        tr = _f32(tr, (self.T, 3))
        self._check(self.lib.ds_set_trajectory(self.ctx, b, _p32(tr)))

    def configure(self, particles_per_thread: Optional[int] = None, gather_particles_per_thread: Optional[int] = None,
                  adjoint_p2g: Optional[str] = None, scatter_chunks: Optional[str] = None,
                  overlap: Optional[bool] = None, chunk_grid: Optional[int] = None,
                  sort_every: Optional[int] = None, reduce_parts: Optional[int] = None,
                  xgrad_stream: Optional[bool] = None, g2p_split: Optional[int] = None,
                  fused_step: Optional[bool] = None):
        """Execution options of the tiled kernels (they change how the substep runs, never what it
        computes); particle-per-thread changes take effect at the next reset.  `overlap`: the
        adjoint recomputes the next checkpoint segment on a second stream while the current one's
        adjoint runs (a second set of segment buffers).  `chunk_grid`: workgroups per chunk-kernel
        launch (0: the library default).  `sort_every`: substeps between bucket sorts (0: the
        checkpoint interval; else a divisor of it), from the next reset.  `reduce_parts`:
        wavefronts per node block of the grid reductions (1, 2 or 4).  `xgrad_stream`: the adjoint
        P2G's position gradient in its own kernel on a third stream, beside the adjoint G2P and
        grid reduction of the same step.  `g2p_split`: workgroups per gather chunk in the forward's G2P
        (1, 2 or 4, at most the gather particles per thread).  `fused_step`: each forward G2P runs
        together with the next step's P2G in one kernel on the scatter's row chunks, where no sort
        falls between them and both steps are in one forward() call."""
        if particles_per_thread is not None:
            self._check(self.lib.ds_set_option(self.ctx, self.OPT_PPT, int(particles_per_thread)))
            self.particles_per_thread = int(particles_per_thread)
        if gather_particles_per_thread is not None:
            self._check(self.lib.ds_set_option(self.ctx, self.OPT_GATHER_PPT, int(gather_particles_per_thread)))
            self.gather_particles_per_thread = int(gather_particles_per_thread)
        if adjoint_p2g is not None:
            if adjoint_p2g not in self.ADJ_P2G:
                raise ValueError(f"adjoint_p2g must be one of {list(self.ADJ_P2G)}")
            self._check(self.lib.ds_set_option(self.ctx, self.OPT_ADJ_P2G, self.ADJ_P2G[adjoint_p2g]))
            self.adjoint_p2g = adjoint_p2g
        if scatter_chunks is not None:
            if scatter_chunks not in self.SCATTER:
                raise ValueError(f"scatter_chunks must be one of {list(self.SCATTER)}")
            self._check(self.lib.ds_set_option(self.ctx, self.OPT_SCATTER, self.SCATTER[scatter_chunks]))
            self.scatter_chunks = scatter_chunks
        if overlap is not None:
            self._check(self.lib.ds_set_option(self.ctx, self.OPT_OVERLAP, 1 if overlap else 0))
            self.overlap = bool(overlap)
        if chunk_grid is not None:
            self._check(self.lib.ds_set_option(self.ctx, self.OPT_CHUNK_GRID, int(chunk_grid)))
            self.chunk_grid = int(chunk_grid)
        if sort_every is not None:
            self._check(self.lib.ds_set_option(self.ctx, self.OPT_SORT_EVERY, int(sort_every)))
            self.sort_every = int(sort_every)
        if reduce_parts is not None:
            self._check(self.lib.ds_set_option(self.ctx, self.OPT_REDUCE_PARTS, int(reduce_parts)))
            self.reduce_parts = int(reduce_parts)
        if xgrad_stream is not None:
            self._check(self.lib.ds_set_option(self.ctx, self.OPT_XGRAD_STREAM, 1 if xgrad_stream else 0))
            self.xgrad_stream = bool(xgrad_stream)
        if g2p_split is not None:
            self._check(self.lib.ds_set_option(self.ctx, self.OPT_G2P_SPLIT, int(g2p_split)))
            self.g2p_split = int(g2p_split)
        if fused_step is not None:
            self._check(self.lib.ds_set_option(self.ctx, self.OPT_FUSED_STEP, 1 if fused_step else 0))
            self.fused_step = bool(fused_step)

    def set_diagnostics(self, tile_window: Optional[int] = None, count_fallback: Optional[bool] = None):
        """Diagnostics of the tiled kernels: `tile_window` (0..5, default 5) narrows the window of
        stencil bases that use their chunk's LDS tile, sending the others through the global
        fall-back grid; `count_fallback` counts the stencils P2G sends there (fallback_count())."""
        if tile_window is not None:
            self._check(self.lib.ds_set_option(self.ctx, self.OPT_TILE_WINDOW, int(tile_window)))
        if count_fallback is not None:
            self._check(self.lib.ds_set_option(self.ctx, self.OPT_COUNT_FALLBACK, 1 if count_fallback else 0))

    def fallback_count(self, reset: bool = False) -> int:
        n = C.c_int64()
        self._check(self.lib.ds_fallback_count(self.ctx, C.byref(n), int(bool(reset))))
        return int(n.value)

    def set_stream(self, stream: Optional[int] = None):
        """Queue the engine's work on a caller's HIP stream (a handle, e.g.
        torch.cuda.current_stream().cuda_stream; None or 0: the engine's own stream)."""
        self._check(self.lib.ds_set_stream(self.ctx, C.c_void_p(int(stream) if stream else None)))

    def set_kernel_timing(self, every: int):
        """Sample every N-th substep's kernels with HIP events (0: off)."""
        self._check(self.lib.ds_set_option(self.ctx, self.OPT_TIMING, int(every)))

    def set_graph(self, on: bool):
        """Replay a whole-rollout forward() and the backward() after it as hipGraphs captured at
        their first call (per-kernel timing does not sample inside them)."""
        self._check(self.lib.ds_set_option(self.ctx, self.OPT_GRAPH, 1 if on else 0))
        self.graph = bool(on)

    def options(self) -> dict:
        return {"particles_per_thread": self.particles_per_thread,
                "gather_particles_per_thread": self.gather_particles_per_thread, "adjoint_p2g": self.adjoint_p2g,
                "scatter_chunks": self.scatter_chunks, "overlap": self.overlap, "sort_every": self.sort_every,
                "reduce_parts": self.reduce_parts, "xgrad_stream": self.xgrad_stream, "g2p_split": self.g2p_split,
                "fused_step": self.fused_step, "chunk_grid": self.chunk_grid}

    # ------------------------------------------------------------------ lifecycle
    def close(self):
        if getattr(self, "ctx", None):
            self.lib.ds_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, code):
        _lib.check(code, self.ctx)

    # ------------------------------------------------------------------ API
    @property
    def n_liquid(self) -> List[int]:
        return [s.n_liquid for s in self.scenes]

    def set_params(self, params):
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(self.B, 4))
        self._check(self.lib.ds_set_params(self.ctx, _p64(p)))
        if not np.array_equal(p, getattr(self, "params", None)):
            self.generation += 1
        self.params = p.copy()

    def reset(self):
        """initialize_ti_field (simulation_functions.py:301-319)."""
        self._check(self.lib.ds_reset(self.ctx))
        self.generation += 1

    def forward(self, nsteps: Optional[int] = None):
        """Advance nsteps substeps (default: to the end of the trajectory)."""
        if nsteps is None:
            nsteps = self.T - 1 - self.step
        self._check(self.lib.ds_forward(self.ctx, int(nsteps)))

    @property
    def step(self) -> int:
        s = C.c_int32()
        self._check(self.lib.ds_current_step(self.ctx, C.byref(s)))
        return int(s.value)

    def forces(self) -> np.ndarray:
        """[B, T, 3] stick forces; row s is m_stick/dt * sum_p dv[s, p]."""
        out = np.zeros((self.B, self.T, 3))
        self._check(self.lib.ds_get_forces(self.ctx, _p64(out)))
        return out

    def backward(self, dl_dforces) -> np.ndarray:
        """dL/dparams [B, 4] for the recorded rollout, given dL/dforces [B, T, 3]."""
        g = np.ascontiguousarray(np.asarray(dl_dforces, dtype=np.float64).reshape(self.B, self.T, 3))
        out = np.zeros((self.B, 4))
        self._check(self.lib.ds_backward(self.ctx, _p64(g), _p64(out)))
        return out

    def timing(self):
        f = C.c_double()
        b = C.c_double()
        self._check(self.lib.ds_last_timing(self.ctx, C.byref(f), C.byref(b)))
        return f.value, b.value

    def kernel_timing(self, reset: bool = False) -> dict:
        """{kernel name: (total ms, launches)} from the sampled HIP-event timing."""
        n = len(_lib.KERNEL_NAMES)
        ms = (C.c_double * n)()
        cnt = (C.c_int64 * n)()
        self._check(self.lib.ds_kernel_timing(self.ctx, ms, cnt, int(bool(reset))))
        return {name: (ms[i], cnt[i]) for i, name in enumerate(_lib.KERNEL_NAMES)}

    def read_liquid(self, inst: int = 0) -> dict:
        n = self.scenes[inst].n_liquid
        x = np.zeros((n, 3), np.float32)
        v = np.zeros((n, 3), np.float32)
        Cm = np.zeros((n, 9), np.float32)
        F = np.zeros((n, 9), np.float32)
        self._check(self.lib.ds_read_liquid(self.ctx, inst, _p32(x), _p32(v), _p32(Cm), _p32(F)))
        return {"x": x, "v": v, "C": Cm, "F": F}

    def read_stick(self, inst: int = 0) -> dict:
        n = self.scenes[inst].n_stick
        x = np.zeros((n, 3), np.float32)
        v = np.zeros((n, 3), np.float32)
        Cm = np.zeros((n, 9), np.float32)
        self._check(self.lib.ds_read_stick(self.ctx, inst, _p32(x), _p32(v), _p32(Cm)))
        return {"x": x, "v": v, "C": Cm}

    def synchronize(self):
        self._check(self.lib.ds_synchronize(self.ctx))


def svd3(F: np.ndarray):
    """FP64 signed SVD of [n, 3, 3] matrices on the device: returns U, S, V."""
    lib = _lib.load()
    F = np.ascontiguousarray(np.asarray(F, dtype=np.float64).reshape(-1, 9))
    n = F.shape[0]
    U = np.zeros((n, 9))
    S = np.zeros((n, 3))
    V = np.zeros((n, 9))
    _lib.check(lib.ds_svd3(_p64(F), n, _p64(U), _p64(S), _p64(V)), None)
    return U.reshape(n, 3, 3), S, V.reshape(n, 3, 3)


def svd3_values(F: np.ndarray) -> np.ndarray:
    """The FP64 singular values the forward substep uses (signed, as svd3), [n, 3]."""
    lib = _lib.load()
    F = np.ascontiguousarray(np.asarray(F, dtype=np.float64).reshape(-1, 9))
    n = F.shape[0]
    S = np.zeros((n, 3))
    _lib.check(lib.ds_svd3_values(_p64(F), n, _p64(S)), None)
    return S


def svd3_backward(U, S, V, gU, gS, gV):
    """FP64 SVD backward on the device (simulation_functions.py:70-87)."""
    lib = _lib.load()
    arrs = [np.ascontiguousarray(np.asarray(a, dtype=np.float64)) for a in (U, S, V, gU, gS, gV)]
    n = arrs[1].reshape(-1, 3).shape[0]
    gF = np.zeros((n, 9))
    _lib.check(lib.ds_svd3_backward(*[_p64(a) for a in arrs[:6]], n, _p64(gF)), None)
    return gF.reshape(n, 3, 3)
