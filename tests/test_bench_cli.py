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
