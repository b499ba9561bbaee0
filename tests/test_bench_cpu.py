"""bench.py's host-side contract without a GPU: the one-GPU refusal and the kernel-trace summary
that ties the bench's event-sampled kernel times to a rocprofv3 trace."""
import json
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def test_trace_summary_samples_the_first_dispatch_after_each_hold(tmp_path, capsys):
    """The sampled dispatches are the first of each kernel after each k_hold on the same queue;
    every dispatch counts in the overall average; kernels are ordered by total time."""
    rows = [  # (queue, name, start ns, end ns)
        (1, "void ds::k_hold(long)", 0, 500_000),
        (1, "void ds::k_a<2>(float const*)", 500_000, 510_000),      # sampled: 10 us
        (1, "void ds::k_b(int)", 510_000, 530_000),                   # sampled: 20 us
        (1, "void ds::k_a<2>(float const*)", 530_000, 542_000),      # second after the hold: not sampled
        (2, "void ds::k_b(int)", 535_000, 565_000),                   # other queue, no hold there
        (1, "void ds::k_hold(long)", 600_000, 1_100_000),
        (1, "void ds::k_a<2>(float const*)", 1_100_000, 1_114_000),  # sampled: 14 us
    ]
    p = tmp_path / "trace.csv"
    with open(p, "w") as f:
        f.write("Kernel_Name,Queue_Id,Start_Timestamp,End_Timestamp\n")
        for q, n, s, e in rows:
            f.write(f"\"{n}\",{q},{s},{e}\n")
    bench.trace_summary(str(p))
    out = json.loads(capsys.readouterr().out)
    assert list(out) == ["k_b", "k_a<2>"]
    a, b = out["k_a<2>"], out["k_b"]
    assert a["dispatches"] == 3 and abs(a["avg_us"] - 12.0) < 1e-9
    assert a["sampled"] == 2 and abs(a["sampled_avg_us"] - 12.0) < 1e-9
    assert b["dispatches"] == 2 and abs(b["avg_us"] - 25.0) < 1e-9
    assert b["sampled"] == 1 and abs(b["sampled_avg_us"] - 20.0) < 1e-9
    assert "k_hold" not in out


def test_roofline_of_the_dominant_sampled_kernel():
    """The dominant kernel is the one with the most sampled time; its launch time is the event
    sample minus the dispatch gap; achieved = algorithmic bytes per launch / launch time; the
    traffic comes from the committed counter summary only when kernel and options match."""
    with open(os.path.join(ROOT, "profiles", "traffic_single_n64.json")) as f:
        t = json.load(f)
    opts = dict(t["options"])
    n_all = 149932
    dom = t["kernel"]
    other = "k_g2p_p2g_tiled" if dom != "k_g2p_p2g_tiled" else "k_p2g_tiled"
    r = {"workload": "single_n64", "options": opts, "n_all": n_all,
         "kernels": {dom: (240 * 34.7e-3, 240), other: (200 * 31.0e-3, 200),
                     "k_g2p_tiled": (0.0, 0), "bucket_sort": (10 * 90e-3, 10)}}
    out = bench.roofline(r)
    assert out["kernel"] == dom
    assert abs(out["event_sample_us"] - 34.7) < 1e-9
    assert abs(out["avg_launch_us"] - (34.7 - bench.DISPATCH_GAP_US)) < 1e-9
    per_launch = bench.KERNEL_BYTES_PER_PARTICLE[dom] * n_all
    assert out["algorithmic_bytes_per_launch"] == per_launch
    gbs = per_launch / ((34.7 - bench.DISPATCH_GAP_US) * 1e-6) / 1e9
    assert abs(out["achieved"] - gbs) < 1e-6 and abs(out["frac"] - gbs / 8000.0) < 1e-12
    assert out["peak"] == 8000.0 and out["unit"] == "GB/s" and out["bound"] == "hbm"
    assert out["traffic"] == t["bytes_per_launch"]
    assert out["kernel_avg_us"]["k_g2p_tiled"] is None
    # other options: the committed counters do not describe this run
    r2 = dict(r, options=dict(opts, particles_per_thread=4))
    assert bench.roofline(r2)["traffic"] is None
