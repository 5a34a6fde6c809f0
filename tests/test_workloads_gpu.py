"""The measurement modes of bench.py (bench_workloads.py) at their full configured sizes,
with the properties their runs report:

* precision: 2048^3 from 1440 projections of 2048 x 1536 (hann), filtered projections
  stored in FP32, FP16 and BF16 with per-projection scales, each against the FP64 GPU
  reconstruction over all voxels, in z-slabs;
* refinement: 768^3 from 120 projections of 1024 x 768 (Poisson I0 = 5e3), FP16 FDK
  seed, 10 steepest-descent steps with per-slab checkpoints: the loss never rises, the
  RMSE against the phantom falls, the FP16 seed matches the FP32 FDK;
* the adjoint-pair mode at a small size (the 512^3 case is tests/test_autograd_gpu.py);
* the per-slab checkpoint: a refinement stopped after a slab and resumed from its
  checkpoint equals an uninterrupted one bit for bit.
"""
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
    args = argparse.Namespace(det_w=None, det_h=None, nproj=None, grid=None, slab=None, steps=1, checkpoint_dir="")
    for k, v in kw.items():
        setattr(args, k, v)
    bench_workloads.WORKLOADS[name](args)
    out = capsys.readouterr().out.strip().splitlines()[-1]
    return json.loads(out)


def test_precision_workload_2048cubed(gpu, capsys):
    import torch
    r = _run(capsys, "precision2048")
    torch.cuda.empty_cache()
    st = r["stores"]
    with capsys.disabled():
        print("\n2048^3 from 1440x1536x2048 vs FP64:",
              {k: (v["rel_rmse_vs_f64"], v["maxabs_vs_f64"], round(v["bp_gups"])) for k, v in st.items()})
    assert r["updates"] == 2048.0 ** 3 * 1440
    assert st["f32"]["rel_rmse_vs_f64"] < 1e-5
    assert st["f16"]["rel_rmse_vs_f64"] < 1e-3
    assert st["bf16"]["rel_rmse_vs_f64"] < 4e-3
    assert st["f16"]["rel_rmse_vs_f64"] < st["bf16"]["rel_rmse_vs_f64"]


def test_gradient_workload_small(gpu, capsys):
    r = _run(capsys, "gradient512", det_w=96, nproj=40, grid=96, steps=1)
    assert r["whole"]["dot_rel"] < 1e-5 and r["slabs"]["dot_rel"] < 1e-5


def test_refinement_workload_768cubed_with_checkpoints(gpu, capsys, tmp_path):
    import torch
    r = _run(capsys, "refine768", checkpoint_dir=str(tmp_path / "ck"))
    torch.cuda.empty_cache()
    steps = r["steps"]
    with capsys.disabled():
        print(f"\n768^3 refinement: {r['ms_per_step']:.0f} ms per step (+{r['ms_checkpoint_per_step']:.0f} ms "
              f"checkpoint writes), RMSE vs phantom {steps[0]['rmse_truth']:.3e} -> {steps[-1]['rmse_truth']:.3e}, "
              f"vs FP32 FDK {steps[-1]['rmse_vs_fp32_fdk']:.3e}")
    assert len(steps) == 11
    losses = [s["loss_before"] for s in steps[1:]]
    # steepest descent with exact line search never increases the loss
    assert all(b <= a * (1 + 1e-9) for a, b in zip(losses, losses[1:]))
    assert steps[-1]["rmse_truth"] < steps[0]["rmse_truth"]
    assert r["rmse_fp16_fdk_vs_truth"] == pytest.approx(r["rmse_fp32_fdk_vs_truth"], rel=1e-3)
    assert 0.5 < r["fdk_seed_least_squares_scale"] < 2.0
    state = json.load(open(str(tmp_path / "ck" / "state.json")))
    assert state["step"] == 10 and state["slab"] == -1


def test_refinement_resumes_from_a_slab_checkpoint(gpu, tmp_path):
    """Stop after slab 1 of step 2 (of 3 slabs), resume in a fresh call: the result and
    the loss history equal an uninterrupted run's exactly."""
    import torch
    import ct_amd
    from ct_amd.autograd import ConeBeamOperator
    from ct_amd.phantom import circular_angles, phantom_unit_radius, random_ellipsoids
    from ct_amd.refine import Interrupted, refine
    W, H, P, n = 160, 120, 30, 96
    g = ct_amd.geometry_grid(W, H, 800.0, 0.0, n, n, n)
    angles = circular_angles(P)
    unit = phantom_unit_radius(g.radius, H, 800.0)
    ell = [e[:9] + [e[9] * 0.5 / unit] for e in random_ellipsoids(12, 9, unit)]
    with ct_amd.Context(0) as ctx:
        ctx.set_geometry(g)
        ctx.set_projections(angles)
        lineint = torch.empty((P, H, W), dtype=torch.float32, device="cuda")
        raw = torch.empty((P, H, W), dtype=torch.int16, device="cuda")
        ctx.phantom_images(ell, 0, P, 0.0, 65535.0, 9, raw, lineint)
    op = ConeBeamOperator(g, angles, slab_nz=32)
    x0 = op.backproject(lineint) * 1e-4
    ref, hist_ref = refine(op, x0, lineint, 3)
    ck = str(tmp_path / "ck")
    with pytest.raises(Interrupted):
        refine(op, x0, lineint, 3, checkpoint_dir=ck, stop_after=(2, 1))
    state = json.load(open(os.path.join(ck, "state.json")))
    assert (state["step"], state["slab"]) == (1, 1)
    got, hist = refine(op, torch.zeros_like(x0), lineint, 3, checkpoint_dir=ck)
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
    assert [h["loss_before"] for h in hist] == [h["loss_before"] for h in hist_ref]
    assert [h["alpha"] for h in hist] == [h["alpha"] for h in hist_ref]
    # a run over a finished checkpoint returns its iterate without further steps
    again, _ = refine(op, x0, lineint, 3, checkpoint_dir=ck)
    assert torch.equal(again, ref)
    # checkpoints every second step: a stop in a step that writes nothing (3) and one
    # part-way through a checkpointed step (4) both resume from step 2's iterate
    ref4, hist4 = refine(op, x0, lineint, 4)
    for stop in ((3, 0), (4, 1)):
        ck2 = str(tmp_path / f"ck_{stop[0]}")
        with pytest.raises(Interrupted):
            refine(op, x0, lineint, 4, checkpoint_dir=ck2, stop_after=stop, every=2)
        state = json.load(open(os.path.join(ck2, "state.json")))
        assert state["step"] == 2
        got, hist = refine(op, x0, lineint, 4, checkpoint_dir=ck2, every=2)
        torch.cuda.synchronize()
        assert torch.equal(got, ref4), stop
        assert [h["loss_before"] for h in hist] == [h["loss_before"] for h in hist4]
