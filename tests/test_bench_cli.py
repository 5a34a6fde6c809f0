"""bench.py's command line, checked without a GPU: more than one GPU is refused with a
message (one process per GPU is not built), and the help lists the workloads."""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)


def test_help_lists_workloads():
    r = _run("--help")
    assert r.returncode == 0
    for name in ("fdk1024", "precision2048", "gradient512", "refine768", "rowfilter1024", "--steps", "--warmup"):
        assert name in r.stdout


def test_every_workload_choice_has_a_runner():
    """The --workload choices are the default FDK line plus exactly the runners of
    bench_workloads.py (importing it touches no GPU)."""
    import re
    sys.path.insert(0, ROOT)
    import bench_workloads
    text = open(os.path.join(ROOT, "bench.py")).read()
    choices = re.search(r'choices=\[("fdk1024"[^\]]*)\]', text).group(1)
    names = set(re.findall(r'"([a-z0-9]+)"', choices))
    assert names == {"fdk1024"} | set(bench_workloads.WORKLOADS)


def test_cpu_baseline_timing_build_matches_the_pinned_oracle():
    """The CPU baseline times oracle.c built at -O3 -march=native on the measuring host;
    that build reconstructs a small scan as the pinned (-O2, no contraction) build does,
    to float rounding."""
    import numpy as np
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import bench
    import oracle as orc
    from helpers import rel_rmse, small_scan
    scan = small_scan(W=48, H=32, P=20, seed=3)
    _, g, _, _ = orc.volume_geometry(48, 32, scan["fcd"], 0.0, None)
    ref = orc.fdk_f32(g, scan["raw"], "u16", 1.0, "hann", scan["angles"])
    path = bench._timing_build()
    assert path is not None
    try:
        orc.use(path)
        got = orc.fdk_f32(g, scan["raw"], "u16", 1.0, "hann", scan["angles"])
    finally:
        orc.use(None)
    assert np.isfinite(got).all()
    assert rel_rmse(got, ref) < 1e-5
