"""The autotuner's search logic without a GPU: candidate enumeration, the choice, the cache."""
import json

import numpy as np
import pytest

from diffstir_amd import autotune as A
from diffstir_amd import scene as S


class _FakeEngine:
    def __init__(self):
        self.configured = []

    def configure(self, **kw):
        self.configured.append(kw)

    def close(self):
        pass


def test_candidates_cover_the_space_once():
    c = A.candidates()
    # rows: 3 rows-per-wavefront x 2 gather sizes x 2 adjoint forms; packed: 2 x 2
    assert len(c) == 3 * 2 * 2 + 2 * 2
    assert len({tuple(sorted(d.items())) for d in c}) == len(c)
    assert A.DEFAULT in c
    assert all(d["particles_per_thread"] == A.DEFAULT["particles_per_thread"]
               for d in c if d["scatter_chunks"] == "packed")


def test_pick_takes_the_fastest_and_skips_failed_runs():
    c = A.candidates()
    times = [5.0, float("nan"), 3.0, 3.0, float("inf")] + [9.0] * (len(c) - 5)
    best = A.pick(zip(c, times))
    assert best == c[2]
    with pytest.raises(RuntimeError):
        A.pick([(c[0], float("nan"))])


def test_tune_uses_the_timer_configures_the_winner_and_caches(tmp_path):
    A.clear_cache()
    scenes = [S.tiny_scene(16, 20, seed=0)]
    want = {"particles_per_thread": 4, "gather_particles_per_thread": 2, "adjoint_p2g": "split",
            "scatter_chunks": "rows"}
    calls = []

    def timer(engine, cand, substeps, repeats):
        calls.append((dict(cand), substeps))
        kernel_opts = {k: cand[k] for k in want}
        if kernel_opts != want:
            return 2.0 + len(calls)
        # among the later stages' values: sorting once per interval wins, 2 reduction parts win,
        # the position-gradient stream wins, the fused G2P + P2G wins
        return (1.0 + (0.5 if cand["sort_every"] else 0.0) + (0.0 if cand["reduce_parts"] == 2 else 0.25)
                + (0.0 if cand["xgrad_stream"] else 0.125) + 0.0625 * (cand["g2p_split"] - 1)
                + (0.0 if cand["fused_step"] else 0.03125))

    eng = _FakeEngine()
    f = tmp_path / "tune.json"
    res = A.tune(scenes, ckpt_every=10, engine=eng, timer=timer, cache_file=str(f), warm_steps=0)
    want2 = dict(want, sort_every=0, reduce_parts=2, xgrad_stream=True, g2p_split=1, fused_step=True, chunk_grid=0)
    assert res["options"] == want2
    assert eng.configured[-1] == want2
    nc = len(A.candidates())
    n2 = len(A.sort_candidates(10)) - 1
    assert n2 == 2
    T = scenes[0].max_steps
    # later stages: 2 sort intervals, 2 reduction splits, the stream, G2P split 2 (gppt 2 chosen),
    # the fused G2P + P2G, the smaller launch grid; then the last pass: rows per wavefront 1 and 2,
    # reduction splits 1 and 4
    assert len(calls) == nc + n2 + 6 + 4 and all(s == min(T - 1, 20) for _, s in calls)   # 2 checkpoint segments at most
    # first stage: every kernel candidate with the later options at their defaults
    assert [c for c, _ in calls[:nc]] == [dict(c, sort_every=0, reduce_parts=1, xgrad_stream=False, g2p_split=1,
                                               fused_step=False, chunk_grid=0) for c in A.candidates()]
    # then the sort intervals, the reduction parts and the position-gradient stream, each with the
    # options chosen so far
    assert [c for c, _ in calls[nc:nc + n2]] == [dict(want, sort_every=R, reduce_parts=1, xgrad_stream=False,
                                                      g2p_split=1, fused_step=False, chunk_grid=0)
                                                 for R in A.sort_candidates(10)[1:]]
    assert [c for c, _ in calls[nc + n2:nc + n2 + 2]] == [dict(want, sort_every=0, reduce_parts=p, xgrad_stream=False,
                                                               g2p_split=1, fused_step=False, chunk_grid=0)
                                                          for p in (2, 4)]
    assert calls[nc + n2 + 2][0] == dict(want2, fused_step=False)
    assert calls[nc + n2 + 3][0] == dict(want2, g2p_split=2, fused_step=False)
    assert calls[nc + n2 + 4][0] == want2
    assert calls[nc + n2 + 5][0] == dict(want2, chunk_grid=2048)
    assert [c for c, _ in calls[nc + n2 + 6:]] == [dict(want2, particles_per_thread=1), dict(want2, particles_per_thread=2),
                                                  dict(want2, reduce_parts=1), dict(want2, reduce_parts=4)]
    # cached in memory and on disk by signature
    again = A.tune(scenes, ckpt_every=10, engine=eng, timer=lambda *a: pytest.fail("re-timed"))
    assert again["options"] == want2
    disk = json.loads(f.read_text())
    assert disk[res["signature"]]["options"] == want2
    A.clear_cache()
    from_disk = A.tune(scenes, ckpt_every=10, engine=eng, timer=lambda *a: pytest.fail("re-timed"),
                       cache_file=str(f))
    assert from_disk["options"] == want2
    A.clear_cache()


def test_last_pass_rechecks_rows_per_wavefront():
    """A choice of rows per wavefront made with the later options at their defaults is revisited:
    here 4 rows win at the default sort interval, 2 rows once the interval is shorter."""
    A.clear_cache()
    scenes = [S.tiny_scene(16, 20, seed=1)]

    def timer(engine, cand, substeps, repeats):
        if cand["gather_particles_per_thread"] != 1 or cand["adjoint_p2g"] != "fused" or cand["scatter_chunks"] != "rows":
            return 9.0
        t = {4: 1.0, 2: 1.1, 1: 1.2}[cand["particles_per_thread"]]
        if cand["sort_every"]:
            t -= 0.2 if cand["particles_per_thread"] == 4 else 0.4
        return t + (0.0 if cand["reduce_parts"] == 1 else 0.05) + (0.05 if cand["xgrad_stream"] else 0.0)

    res = A.tune(scenes, ckpt_every=10, engine=_FakeEngine(), timer=timer, warm_steps=0)
    assert res["options"]["sort_every"] != 0
    assert res["options"]["particles_per_thread"] == 2
    assert res["options"]["reduce_parts"] == 1
    A.clear_cache()


def test_sort_candidates_divide_the_interval():
    assert A.sort_candidates(100) == [0, 50, 25]
    assert A.sort_candidates(50) == [0, 25, 10]
    assert A.sort_candidates(5) == [0]
    for K in (5, 6, 7, 12, 50, 100):
        assert all(K % R == 0 for R in A.sort_candidates(K)[1:])


def test_later_stage_defaults_are_the_library_defaults():
    assert A.later_defaults(50) == {"sort_every": 0, "reduce_parts": 1, "xgrad_stream": False, "g2p_split": 1,
                                    "fused_step": False, "chunk_grid": 0}
    assert A.stage_candidates("fused_step", 50, {"scatter_chunks": "rows"}) == [False, True]
    assert A.stage_candidates("fused_step", 50, {"scatter_chunks": "packed"}) == [False]
    assert A.stage_candidates("g2p_split", 50, {"gather_particles_per_thread": 1}) == [1]
    assert A.stage_candidates("g2p_split", 50, {"gather_particles_per_thread": 2}) == [1, 2]
    assert A.stage_candidates("reduce_parts", 50) == [1, 2, 4]
    assert A.stage_candidates("chunk_grid", 50) == [0, 2048]
    with pytest.raises(KeyError):
        A.stage_candidates("overlap", 50)


def test_signature_separates_workloads():
    a = A.signature([S.tiny_scene(16, 20, seed=0)], 5)
    b = A.signature([S.tiny_scene(16, 20, seed=0)] * 2, 5)
    c = A.signature([S.tiny_scene(16, 20, seed=0)], 10)
    assert len({a, b, c}) == 3
