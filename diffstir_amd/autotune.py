"""Autotuner for the execution options of the tiled kernels.

The options change how a substep runs on the device, never what it computes:

* ``particles_per_thread`` -- rows per wavefront of the scatter kernels (P2G and the adjoint G2P);
  a chunk is 4 * ppt rows of one bucket, a row holding at most one particle per cell;
* ``gather_particles_per_thread`` -- particles per thread of the gather kernels (G2P, J, the
  adjoint P2G); a gather chunk is 256 * gppt particles of one bucket;
* ``adjoint_p2g`` -- the adjoint P2G as one kernel ("fused") or as an LDS-gather kernel followed
  by a per-particle stress kernel ("split");
* ``scatter_chunks`` -- the scatter kernels on rows ("rows": lanes of a wavefront never share a
  stencil base, but rows of sparse buckets leave lanes idle) or on the packed gather chunks
  ("packed": full wavefronts, lanes sharing a base served in successive passes).  With packed
  chunks the rows-per-wavefront choice does not apply.

Later stages then refine the chosen set one option at a time (and a last pass re-checks the rows
per wavefront and then the reduction split, whose best values depend on the options chosen after
them):

* ``sort_every`` -- sorting the particles into buckets more often than once per checkpoint
  interval (half or a quarter/fifth of it): where particles drift far within an interval (the
  n_grid 128 workload) it keeps their stencils inside their chunk's tile;
* ``reduce_parts`` -- 1, 2 or 4 wavefronts per node block of the grid reductions: where buckets
  hold many chunks (one clumped instance) the rounds of slot loads split over the wavefronts;
  where they hold one or two (large batches) the extra wavefronts only take up slots;
* ``xgrad_stream`` -- the adjoint P2G's position gradient in its own kernel on a third stream,
  beside the adjoint G2P and grid reduction of the same step: it shortens the adjoint's kernel
  chain where the chip has room for a third stream (large batches) and lengthens it where the
  extra tile load and stencil cost more than the overlap gains;
* ``g2p_split`` -- the forward's G2P running each gather chunk as 2 or 4 workgroups: with 2 gather
  particles per thread chosen for the adjoint P2G, the forward's G2P can still run 256-particle
  workgroups (only values up to the chosen particles per thread are tried);
* ``fused_step`` -- each forward G2P fused with the next step's P2G in one kernel on the
  scatter's row chunks (row chunks only): it saves a launch and a state round trip per substep
  where the chip is latency-bound (one instance), and costs where G2P's full gather chunks keep
  the lanes busier than the rows do (large batches);
* ``chunk_grid`` -- the launch grid of the chunk kernels: the library's default (16384
  workgroups, of which those past the device-resident chunk count exit at once) or 2048 (a chunk
  count above it loops).  With few chunks (one instance) the smaller grid leaves the two streams
  of the overlapped adjoint fewer empty workgroups to dispatch; with many (large batches) the
  larger grid lets the dispatcher spread chunks of uneven cost.

Which choice is fastest depends on the workload's size and on how the particles cluster, so each
candidate is timed on the workload itself: a short checkpointed forward + adjoint rollout
(``substeps`` substeps of every instance) on one Engine, best of ``repeats``.  The particles of a
freshly filled container sit on a regular lattice, which is not how they are spread over most of
a rollout, so the timed rollout starts from the positions the workload reaches after
``warm_steps`` substeps (half of the rollout by default; velocities restart from rest).  Results
are cached by a workload signature, in memory and optionally in a JSON file.
"""
from __future__ import annotations

import itertools
import json
import os
from typing import Callable, Dict, Iterable, List, Optional, Sequence

import numpy as np

SPACE = {
    "particles_per_thread": (1, 2, 4),
    "gather_particles_per_thread": (1, 2),
    "adjoint_p2g": ("fused", "split"),
    "scatter_chunks": ("rows", "packed"),
}
DEFAULT = {"particles_per_thread": 2, "gather_particles_per_thread": 1, "adjoint_p2g": "fused",
           "scatter_chunks": "rows"}

# options refined after the first stage, in order, with their default (first) values
LATER_STAGES = ("sort_every", "reduce_parts", "xgrad_stream", "g2p_split", "fused_step", "chunk_grid")
# a last pass over the options whose best value depends most on the later ones: the rows per
# wavefront (with the chosen sort interval and reduction split) and then the reduction split
FINAL_STAGES = ("particles_per_thread", "reduce_parts")

_CACHE: Dict[str, dict] = {}


def candidates(space: Optional[dict] = None) -> List[dict]:
    """Every distinct combination of the option values in `space` (default SPACE): with packed
    scatter chunks, particles_per_thread (rows per wavefront) has no effect and keeps its default."""
    space = space or SPACE
    keys = list(space)
    out, seen = [], set()
    for vals in itertools.product(*(space[k] for k in keys)):
        d = dict(zip(keys, vals))
        if d.get("scatter_chunks") == "packed" and "particles_per_thread" in d:
            d["particles_per_thread"] = DEFAULT["particles_per_thread"]
        key = tuple(sorted(d.items()))
        if key not in seen:
            seen.add(key)
            out.append(d)
    return out


def sort_candidates(ckpt_every: int) -> List[int]:
    """Sort intervals of the second stage: the checkpoint interval itself (0) and up to two of its
    divisors K/2, K/4, K/5."""
    K = int(ckpt_every)
    return [0] + [K // d for d in (2, 4, 5) if K % d == 0 and K // d >= 2][:2]


def stage_candidates(key: str, ckpt_every: int, chosen: Optional[dict] = None) -> List:
    """Values of a later-stage option, its default first (`chosen`: the options chosen so far,
    which rule out values that would change nothing)."""
    if key == "particles_per_thread":
        return list(SPACE["particles_per_thread"])
    if key == "sort_every":
        return sort_candidates(ckpt_every)
    if key == "reduce_parts":
        return [1, 2, 4]
    if key == "xgrad_stream":
        return [False, True]
    if key == "g2p_split":
        g = (chosen or {}).get("gather_particles_per_thread", 4)
        return [1] + [v for v in (2, 4) if v <= g]
    if key == "fused_step":
        # the fused kernel runs on row chunks only
        return [False, True] if (chosen or {}).get("scatter_chunks", "rows") == "rows" else [False]
    if key == "chunk_grid":
        return [0, 2048]
    raise KeyError(key)


def later_defaults(ckpt_every: int) -> dict:
    return {k: stage_candidates(k, ckpt_every)[0] for k in LATER_STAGES}


def signature(scenes, ckpt_every: int) -> str:
    """What the choice depends on: grid size, instance count, particle counts (to 2 digits), K."""
    n = scenes[0].cfg.n_grid
    total = sum(s.n_liquid + s.n_stick + s.n_container for s in scenes)
    return f"n{n}_b{len(scenes)}_p{float(f'{total:.2g}'):g}_k{ckpt_every}"


def pick(timings: Iterable[tuple]) -> dict:
    """The candidate with the smallest time; ties go to the earlier candidate."""
    best, best_t = None, float("inf")
    for cand, t in timings:
        if np.isfinite(t) and t < best_t:
            best, best_t = cand, t
    if best is None:
        raise RuntimeError("no candidate ran")
    return dict(best)


def time_candidate(engine, cand: dict, substeps: int, repeats: int = 2) -> float:
    """Milliseconds of a forward + adjoint rollout of `substeps` substeps with options `cand`
    (the library's own HIP-event timing), best of `repeats`."""
    engine.configure(**cand)
    best = float("inf")
    for _ in range(repeats):
        engine.reset()
        engine.forward(substeps)
        F = engine.forces()
        engine.backward(np.ones_like(F))
        fwd_ms, adj_ms = engine.timing()
        best = min(best, fwd_ms + adj_ms)
    return best


def warmed_scenes(scenes, engine, warm_steps: int):
    """`scenes` restarted from the particle positions `engine` (an Engine over `scenes`, default
    options) reaches after `warm_steps` substeps: liquid and stick positions of that step, the
    stick trajectory from that frame on."""
    import dataclasses
    engine.configure(**DEFAULT, sort_every=0, reduce_parts=1, xgrad_stream=False, g2p_split=1, fused_step=False)
    engine.reset()
    engine.forward(warm_steps)
    out = []
    for b, s in enumerate(scenes):
        xl = engine.read_liquid(b)["x"]
        xs = engine.read_stick(b)["x"] if s.n_stick else s.x_stick
        out.append(dataclasses.replace(s, x_liquid=xl, x_stick=xs,
                                       traj=np.ascontiguousarray(s.traj[warm_steps:])))
    return out


def tune(scenes, ckpt_every: int = 50, substeps: Optional[int] = None, space: Optional[dict] = None,
         repeats: int = 2, cache_file: Optional[str] = None, engine=None, warm_steps: Optional[int] = None,
         timer: Optional[Callable[[object, dict, int, int], float]] = None, verbose: bool = False) -> dict:
    """Choose the execution options for `scenes`; returns {"options": ..., "timings_ms": [...],
    "signature": ..., "substeps": ..., "warm_steps": ...}.  `engine` (an Engine over `scenes`) is
    used when given (its scenes are restored afterwards and the chosen options applied), else a
    temporary one."""
    scenes = list(scenes)
    sig = signature(scenes, ckpt_every)
    if sig in _CACHE:
        return _CACHE[sig]
    if cache_file and os.path.exists(cache_file):
        with open(cache_file) as f:
            disk = json.load(f)
        if sig in disk:
            _CACHE[sig] = disk[sig]
            return disk[sig]
    from .engine import Engine
    T = min(s.max_steps for s in scenes)
    if substeps is None:
        substeps = min(T - 1, 2 * ckpt_every)
    if warm_steps is None:
        warm_steps = (T - 1) // 2
    warm_steps = max(0, min(warm_steps, T - 1 - substeps))
    timer = timer or time_candidate
    own = engine is None
    eng = Engine(scenes, ckpt_every=ckpt_every) if own else engine
    try:
        if warm_steps > 0:
            # time on the warmed positions in the same context, then put the scenes back
            for b, s in enumerate(warmed_scenes(scenes, eng, warm_steps)):
                eng.upload_scene(b, s)
        timings = []

        def run(cand):
            t = timer(eng, cand, substeps, repeats)
            timings.append((cand, t))
            if verbose:
                print(f"[autotune] {cand}: {t:.2f} ms", flush=True)
            return t

        # first stage: the kernel options, with the later stages' options at their defaults
        first = [dict(c, **later_defaults(ckpt_every)) for c in candidates(space)]
        best = pick((c, run(c)) for c in first)
        best_t = min(t for c, t in timings if c == best)
        # later stages: one option at a time, with the options chosen so far
        for key in LATER_STAGES + FINAL_STAGES:
            if key == "particles_per_thread" and best.get("scatter_chunks") == "packed":
                continue   # rows per wavefront do not apply to packed chunks
            vals = stage_candidates(key, ckpt_every, best)
            if space and key in space:
                vals = [v for v in vals if v in space[key]]
            stage = [(best, best_t)] + [(dict(best, **{key: v}), None) for v in vals if v != best[key]]
            stage = [(c, t if t is not None else run(c)) for c, t in stage]
            best = pick(stage)
            best_t = min(t for c, t in stage if c == best)
        if warm_steps > 0:
            for b, s in enumerate(scenes):
                eng.upload_scene(b, s)
        eng.configure(**best)
    finally:
        if own:
            eng.close()
    res = {"signature": sig, "substeps": substeps, "warm_steps": warm_steps, "options": best,
           "timings_ms": [{"options": c, "ms": t} for c, t in timings]}
    _CACHE[sig] = res
    if cache_file:
        disk = {}
        if os.path.exists(cache_file):
            with open(cache_file) as f:
                disk = json.load(f)
        disk[sig] = res
        os.makedirs(os.path.dirname(os.path.abspath(cache_file)), exist_ok=True)
        with open(cache_file, "w") as f:
            json.dump(disk, f, indent=1)
    return res


def clear_cache():
    _CACHE.clear()
