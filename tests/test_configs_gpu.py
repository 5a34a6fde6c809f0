"""Parity at the configured workload sizes.

* 256 x 192 detector (0.4 mm at the isocentre, FCD 1250 px), 360 projections, 10 seeded
  ellipsoids, intensities round(65535 e^-p), ramlak: the whole scan through the CLI
  (PGM images + CSV, ``-j zfastest``) on CtVolume's own 254 x 254 x 192 grid, and
  through the C ABI on the explicit 256^3 grid, each against the CPU oracle's FDK
  (Ct/CtVolume.cpp:923-1054) at <= 1e-4 relative RMSE.
* 1024^2 detector (FCD 2500 px), 720 projections, 3-D Shepp-Logan + 64 ellipsoids,
  Poisson counts at I0 = 1e4, shepplogan: the 1024^3 reconstruction's central 64
  slices against the oracle's FDK of the same slab (oracle preprocessing of the same
  raw counts), the region-of-interest run on CtVolume's grid bounded to x, y in
  [0.25, 0.75], z in [0.1, 0.9] (501 x 501 x 820) against the oracle on 64 of its
  slices, and the FP32 and FP16-store volumes against the FP64 GPU reconstruction over
  all voxels (<= 1e-4 and <= 1e-3).
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import rel_rmse
from scan_io import write_scan

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "ct_amd", "ct_cli")


def _phantom_raw(ctx, torch, g, angles, ell, photons, seed):
    """uint16 images of the ellipsoid phantom on the GPU (ctg_phantom_images), as numpy."""
    P, H, W = len(angles), g.det_height, g.det_width
    ctx.set_geometry(g)
    ctx.set_projections(angles)
    raw = torch.empty((P, H, W), dtype=torch.int16, device="cuda")  # uint16 bytes
    ctx.phantom_images(ell, 0, P, photons, 65535.0, seed, raw, None)
    torch.cuda.synchronize()
    return raw.cpu().numpy().view(np.uint16)


@pytest.fixture(scope="module")
def scan256(gpu):
    import torch
    import ct_amd
    from ct_amd.phantom import circular_angles, phantom_unit_radius, random_ellipsoids
    W, H, P, fcd = 256, 192, 360, 1250.0
    g, _ = ct_amd.geometry_ctvolume(W, H, fcd, 0.0, None)
    angles = circular_angles(P)
    unit = phantom_unit_radius(g.radius, H, fcd)
    # attenuation U[0.1, 1.0] per phantom unit (the unit sphere's radius)
    ell = [e[:9] + [e[9] / unit] for e in random_ellipsoids(10, 1, unit)]
    with ct_amd.Context(0) as ctx:
        raw = _phantom_raw(ctx, torch, g, angles, ell, 0.0, 1)
    return dict(W=W, H=H, P=P, fcd=fcd, angles=np.asarray(angles), raw=raw)


def test_scan256_cli_zfastest_matches_oracle(scan256, oracle, tmp_path):
    s = scan256
    pixel = 0.4
    csv = write_scan(str(tmp_path), s["raw"], s["angles"], pixel_size=pixel, fcd_px=s["fcd"], base=1.0)
    r = subprocess.run([CLI, "-i", csv, "-o", "vol.raw", "-j", "zfastest", "-f", "ramlak"], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:])
    assert r.returncode == 0, r.stdout + r.stderr
    # the facade's FCD in pixels, as the loader computes it (mm / pixel size in float)
    fcd = float(np.float32(np.float32(s["fcd"] * pixel) / np.float32(pixel)))
    _, g, _, _ = oracle.volume_geometry(s["W"], s["H"], fcd, 0.0, None)
    assert (g.nx, g.ny, g.nz) == (254, 254, 192)
    ref = oracle.fdk_f32(g, s["raw"], "u16", 1.0, "ramlak", s["angles"], log_form="gpu")
    got = np.fromfile(str(tmp_path / "vol.raw"), "<f4").reshape(g.nx, g.ny, g.nz)
    err = rel_rmse(got, ref)
    print(f"256x192x360 CLI (254x254x192, zfastest): rel RMSE vs oracle {err:.3e}")
    assert err < 1e-4


def test_scan256_abi_256_cube_matches_oracle(scan256, oracle):
    import torch
    import ct_amd
    s = scan256
    P = s["P"]
    g = ct_amd.geometry_grid(s["W"], s["H"], s["fcd"], 0.0, 256, 256, 256)
    raw = torch.from_numpy(s["raw"].view(np.int16)).cuda()
    out = {}
    with ct_amd.Context(0) as ctx:
        ctx.set_geometry(g)
        ctx.set_projections(list(s["angles"]))
        vol = torch.empty((256, 256, 256), dtype=torch.float32, device="cuda")
        for store in ("f32", "f16"):
            filtered = torch.empty(ctx.filtered_bytes(store, P), dtype=torch.uint8, device="cuda")
            scales = torch.empty(P, dtype=torch.float32, device="cuda")
            ctx.preprocess(raw, "u16", P, 1.0, "cpu", "ramlak", store, filtered, scales)
            ctx.backproject(filtered, store, scales, 0, P, vol, 0, 256)
            torch.cuda.synchronize()
            out[store] = vol.cpu().numpy()
    og = oracle.geometry_from(g)
    ref = oracle.fdk_f32(og, s["raw"], "u16", 1.0, "ramlak", s["angles"])
    ref64 = oracle.fdk_f64(og, s["raw"], "u16", 1.0, "ramlak", s["angles"])
    e32, e16 = rel_rmse(out["f32"], ref), rel_rmse(out["f16"], ref64)
    print(f"256^3 from 360x192x256: FP32 rel RMSE vs oracle {e32:.3e}; FP16 store vs FP64 oracle {e16:.3e}")
    assert e32 < 1e-4
    assert e16 < 1e-3


# ------------------------------------------------------------------ 1024^3 from 720 x 1024^2

@pytest.fixture(scope="module")
def scan1024(gpu, oracle):
    import torch
    import ct_amd
    sys.path.insert(0, ROOT)
    import bench
    ctx = ct_amd.Context(0)
    args = argparse.Namespace(det=1024, nproj=720, grid=1024)
    g, raw, angles, base = bench.build_workload(ctx, torch, args)
    host_raw = raw.cpu().numpy().view(np.uint16)
    # the oracle's preprocessing of the same raw counts (CPU log form, shepplogan)
    og = oracle.geometry_from(g)
    filtered = oracle.preprocess_f32(og, host_raw, "u16", base, "shepplogan", "cpu")
    yield dict(ctx=ctx, g=g, raw=raw, angles=np.asarray(angles), base=base, filtered=filtered, host_raw=host_raw)
    ctx.close()
    del raw
    torch.cuda.empty_cache()


def _gpu_fdk(ctx, torch, g, raw, P, base, store):
    ctx.set_geometry(g)
    filtered = torch.empty(ctx.filtered_bytes(store, P), dtype=torch.uint8, device="cuda")
    scales = torch.empty(P, dtype=torch.float32, device="cuda")
    vol = torch.empty((g.nx, g.ny, g.nz), dtype=torch.float32, device="cuda")
    ctx.preprocess(raw, "u16", P, base, "cpu", "shepplogan", store, filtered, scales)
    ctx.backproject(filtered, store, scales, 0, P, vol, 0, g.nz)
    torch.cuda.synchronize()
    del filtered, scales
    return vol


def _oracle_slab(oracle, g, filtered, angles, z0, nz):
    """The oracle's backprojection of slices [z0, z0+nz) of grid g: the same grid with
    nz slices whose world z starts at z0 - z_pre."""
    sub = oracle.geometry_from(g)
    sub.nz = nz
    sub.z_pre = g.z_pre - z0
    return oracle.backproject_f32(sub, filtered, angles)


def test_scan1024_central_slab_matches_oracle(scan1024, oracle):
    import torch
    c = scan1024
    g, P = c["g"], len(c["angles"])
    vol = _gpu_fdk(c["ctx"], torch, g, c["raw"], P, c["base"], "f32")
    z0, nz = g.nz // 2 - 32, 64
    got = vol[:, :, z0:z0 + nz].cpu().numpy()
    del vol
    ref = _oracle_slab(oracle, g, c["filtered"], c["angles"], z0, nz)
    err = rel_rmse(got, ref)
    print(f"1024^3 from 720x1024^2, slices {z0}..{z0 + nz}: rel RMSE vs oracle {err:.3e}")
    assert err < 1e-4


def test_scan1024_roi_matches_oracle(scan1024, oracle):
    import torch
    import ct_amd
    c = scan1024
    g0, P = c["g"], len(c["angles"])
    g, roi = ct_amd.geometry_ctvolume(g0.det_width, g0.det_height, g0.fcd, 0.0, [0.25, 0.75, 0.25, 0.75, 0.1, 0.9])
    assert (g.nx, g.ny, g.nz) == (501, 501, 820) and tuple(roi) == (251, 752, 251, 752, 102, 922)
    c["ctx"].set_projections(list(c["angles"]))
    vol = _gpu_fdk(c["ctx"], torch, g, c["raw"], P, c["base"], "f32")
    z0, nz = g.nz // 2 - 32, 64
    got = vol[:, :, z0:z0 + nz].cpu().numpy()
    del vol
    ref = _oracle_slab(oracle, g, c["filtered"], c["angles"], z0, nz)
    err = rel_rmse(got, ref)
    print(f"ROI 501x501x820, slices {z0}..{z0 + nz}: rel RMSE vs oracle {err:.3e}")
    assert err < 1e-4
    c["ctx"].set_geometry(g0)


def test_scan1024_fp32_and_fp16_against_fp64_all_voxels(scan1024):
    import torch
    c = scan1024
    g, P = c["g"], len(c["angles"])
    ctx = c["ctx"]
    ctx.set_geometry(g)
    vol64 = torch.empty((g.nx, g.ny, g.nz), dtype=torch.float64, device="cuda")
    ctx.fdk_f64(c["raw"], "u16", P, c["base"], "shepplogan", vol64, 0, g.nz)
    torch.cuda.synchronize()
    ref2 = float(torch.sum(vol64 * vol64))
    res = {}
    for store in ("f32", "f16"):
        vol = _gpu_fdk(ctx, torch, g, c["raw"], P, c["base"], store)
        d = vol.double() - vol64
        res[store] = (float(torch.sqrt(torch.sum(d * d) / ref2)), float(torch.max(torch.abs(d))))
        del vol, d
    del vol64
    torch.cuda.empty_cache()
    print("1024^3 against FP64 (rel RMSE, max-abs):", res)
    assert res["f32"][0] < 1e-4
    assert res["f16"][0] < 1e-3
