"""The measurement modes of bench.py (bench_workloads.py) at small sizes: they run end
to end on the GPU and their outputs satisfy the properties the full-size runs report."""
from __future__ import annotations

import argparse
import json
import os
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def _run(capsys, name, **kw):
    import bench_workloads
    args = argparse.Namespace(det_w=None, det_h=None, nproj=None, grid=None, slab=None, steps=1)
    for k, v in kw.items():
        setattr(args, k, v)
    bench_workloads.WORKLOADS[name](args)
    out = capsys.readouterr().out.strip().splitlines()[-1]
    return json.loads(out)


def test_precision_workload_small(gpu, capsys):
    r = _run(capsys, "precision2048", det_w=192, det_h=128, nproj=90, grid=160, slab=48)
    st = r["stores"]
    print({k: (v["rel_rmse_vs_f64"], v["maxabs_vs_f64"]) for k, v in st.items()})
    assert st["f32"]["rel_rmse_vs_f64"] < 1e-5
    assert st["f16"]["rel_rmse_vs_f64"] < 1e-3
    assert st["bf16"]["rel_rmse_vs_f64"] < 4e-3
    assert st["f16"]["rel_rmse_vs_f64"] < st["bf16"]["rel_rmse_vs_f64"]


def test_gradient_workload_small(gpu, capsys):
    r = _run(capsys, "gradient512", det_w=96, nproj=40, grid=96, steps=1)
    assert r["whole"]["dot_rel"] < 1e-5 and r["slabs"]["dot_rel"] < 1e-5


def test_refinement_workload_small(gpu, capsys):
    r = _run(capsys, "refine768", det_w=192, det_h=144, nproj=30, grid=128, slab=48)
    losses = [s["loss_before"] for s in r["steps"][1:]]
    # steepest descent with exact line search never increases the loss
    assert all(b <= a * (1 + 1e-9) for a, b in zip(losses, losses[1:]))
    assert r["rmse_fp16_fdk_vs_truth"] == pytest.approx(r["rmse_fp32_fdk_vs_truth"], rel=1e-3)
    assert 0.5 < r["fdk_seed_least_squares_scale"] < 2.0
