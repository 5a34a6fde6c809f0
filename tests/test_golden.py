"""Golden fixtures (tests/golden/, written by make_golden.py from the oracle):
the oracle and the C ABI's host geometry must reproduce them on the CPU, and the HIP
path must reproduce the reconstructions on the GPU to the parity bar."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from helpers import rel_rmse

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def small():
    with np.load(os.path.join(GOLDEN, "fdk_small.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def geometry_cases():
    with open(os.path.join(GOLDEN, "geometry.json")) as f:
        return json.load(f)


def _cases(small):
    return [k for k in small if k.split("_")[0] in ("ramlak", "shepplogan", "hann") and k.count("_") == 1]


def test_oracle_reproduces_golden(oracle, small):
    W, H = small["raw"].shape[2], small["raw"].shape[1]
    _, g, _, _ = oracle.volume_geometry(W, H, float(small["fcd"]), float(small["u_offset"]))
    assert len(_cases(small)) == 4
    for key in _cases(small):
        filt, form = key.split("_")
        vol = oracle.fdk_f32(g, small["raw"], "u16", float(small["base"]), filt, small["angles"], small["heights"],
                             log_form=form)
        assert vol.shape == small[key].shape
        assert rel_rmse(vol, small[key]) < 1e-6, key
    rows = oracle.preprocess_f32(g, small["raw"], "u16", float(small["base"]), "ramlak", "cpu")
    assert rel_rmse(rows, small["filtered_ramlak_cpu"]) < 1e-6
    _, gr, roi, _ = oracle.volume_geometry(W, H, float(small["fcd"]), float(small["u_offset"]),
                                           list(small["roi_bounds"]))
    assert [gr.nx, gr.ny, gr.nz] == list(small["roi_grid"]) and list(roi) == list(small["roi_ranges"])


def test_oracle_geometry_golden(oracle, geometry_cases):
    for case in geometry_cases:
        rc, g, roi, _ = oracle.volume_geometry(case["det_width"], case["det_height"], case["fcd_px"],
                                               case["u_offset_px"], case["bounds"])
        assert rc == case["status"]
        assert list(roi) == case["roi"]
        for k, v in case["geometry"].items():
            assert float(getattr(g, k)) == v, (case, k)


def test_ctg_host_geometry_matches_golden(ctg_lib, geometry_cases):
    """The library's own CtVolume geometry (host code, no GPU needed) gives the same grids."""
    import ct_amd
    for case in geometry_cases:
        g, roi = ct_amd.geometry_ctvolume(case["det_width"], case["det_height"], case["fcd_px"],
                                          case["u_offset_px"], case["bounds"])
        assert list(roi) == case["roi"], case
        for k, v in case["geometry"].items():
            assert float(getattr(g, k)) == v, (case, k)


@pytest.mark.gpu
def test_hip_fdk_matches_golden(gpu, small):
    import torch
    import ct_amd
    raw = small["raw"]
    P, H, W = raw.shape
    ctx = ct_amd.Context(0)
    stream = torch.cuda.current_stream().cuda_stream
    d_raw = torch.from_numpy(raw.astype(np.int16)).to(gpu)  # same bytes as uint16
    for bounds, keys in ((None, _cases(small)), (list(small["roi_bounds"]), ["roi_shepplogan_cpu"])):
        g, _ = ct_amd.geometry_ctvolume(W, H, float(small["fcd"]), float(small["u_offset"]), bounds)
        ctx.set_geometry(g)
        ctx.set_projections(small["angles"], small["heights"])
        for key in keys:
            filt, form = key.split("_")[-2:]
            filtered = torch.empty(ctx.filtered_bytes("f32", P), dtype=torch.uint8, device=gpu)
            scales = torch.empty(P, dtype=torch.float32, device=gpu)
            vol = torch.empty((g.nx, g.ny, g.nz), dtype=torch.float32, device=gpu)
            ctx.preprocess(d_raw, "u16", P, float(small["base"]), form, filt, "f32", filtered, scales, stream)
            ctx.backproject(filtered, "f32", scales, 0, P, vol, 0, g.nz, False, stream)
            torch.cuda.synchronize()
            err = rel_rmse(vol.cpu().numpy(), small[key])
            print(f"golden {key}: rel RMSE {err:.3e}")
            assert err < 1e-4, key
