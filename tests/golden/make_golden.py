#!/usr/bin/env python3
"""Writes the golden fixtures of this directory from the CPU oracle (oracle/).

    python tests/golden/make_golden.py

fdk_small.npz -- one small scan and its reconstructions:
  raw (12 x 16 x 24 uint16, I = round(65535 * B * exp(-p)) of seeded ellipsoids),
  angles (deg), heights (px), scalars fcd, u_offset, base; for each case
  "<filter>_<logform>" the FP32 oracle volume on CtVolume's grid (Z_FASTEST
  [x][y][z]), the filtered line integrals of ramlak/cpu, and one region-of-interest
  case (bounds roi_bounds) with its grid.
geometry.json -- CtVolume geometry (radius, grid, precomputes, detector window, ROI
  voxel ranges) for the detector and FCD of every configured workload, with and
  without bounds, including bounds that hit the reference's yTo/zTo clamp.

The fixtures freeze the oracle's output: the oracle is pinned independently
(closed forms, numpy FFT, brute force; tests/test_oracle.py), and these vectors
let the HIP path and later oracle builds be checked against the same numbers
without recomputing them.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

W, H, P = 24, 16, 12
FCD = W * 1250.0 / 256.0
U_OFFSET = 1.25
BASE = 0.8
ROI = [0.1, 0.9, 0.2, 0.7, 0.25, 0.8]
CASES = [("ramlak", "cpu"), ("shepplogan", "gpu"), ("hann", "cpu"), ("hann", "gpu")]

GEOMETRY_CASES = [
    # (W, H, FCD px, uOffset px, bounds or None)
    (256, 192, 1250.0, 0.0, None),
    (1024, 1024, 2500.0, 0.0, None),
    (1024, 1024, 2500.0, 0.0, [0.25, 0.75, 0.25, 0.75, 0.1, 0.9]),
    (2048, 1536, 5000.0, 0.0, None),
    (512, 512, 2500.0, 0.0, None),
    (1024, 768, 5000.0, 0.0, None),
    (256, 192, 1250.0, 3.5, [0.5, 0.9, 0.1, 0.3, 0.2, 0.8]),  # yTo clamped against xFrom
    (300, 200, 900.0, -2.0, [0.3, 0.3, 0.0, 1.0, 0.0, 1.0]),  # equal bounds are ignored
]


def scan():
    import oracle as orc
    from ct_amd.phantom import circular_angles, phantom_unit_radius, random_ellipsoids
    _, g, _, _ = orc.volume_geometry(W, H, FCD, U_OFFSET)
    unit = phantom_unit_radius(g.radius, H, FCD)
    ell = random_ellipsoids(5, 17, unit, mu_range=(0.02, 0.06))
    angles = np.asarray(circular_angles(P), np.float64)
    heights = np.linspace(-0.75, 0.75, P)
    p = orc.phantom_project(g, ell, angles, heights)
    raw = np.clip(np.floor(65535.0 * BASE * np.exp(-p) + 0.5), 0, 65535).astype(np.uint16)
    return raw, angles, heights


def main():
    import oracle as orc
    orc.load()
    raw, angles, heights = scan()
    out = dict(raw=raw, angles=angles, heights=heights, fcd=np.float64(FCD), u_offset=np.float64(U_OFFSET),
               base=np.float64(BASE), roi_bounds=np.asarray(ROI, np.float64))
    _, g, _, _ = orc.volume_geometry(W, H, FCD, U_OFFSET)
    for filt, form in CASES:
        out[f"{filt}_{form}"] = orc.fdk_f32(g, raw, "u16", BASE, filt, angles, heights, log_form=form)
    out["filtered_ramlak_cpu"] = orc.preprocess_f32(g, raw, "u16", BASE, "ramlak", "cpu")
    _, gr, roi, _ = orc.volume_geometry(W, H, FCD, U_OFFSET, ROI)
    out["roi_grid"] = np.asarray([gr.nx, gr.ny, gr.nz], np.int64)
    out["roi_ranges"] = np.asarray(roi, np.int64)
    out["roi_shepplogan_cpu"] = orc.fdk_f32(gr, raw, "u16", BASE, "shepplogan", angles, heights, log_form="cpu")
    np.savez_compressed(os.path.join(HERE, "fdk_small.npz"), **out)

    geo = []
    for (w, h, fcd, uoff, bounds) in GEOMETRY_CASES:
        rc, gg, r, eff = orc.volume_geometry(w, h, fcd, uoff, bounds)
        geo.append({"det_width": w, "det_height": h, "fcd_px": fcd, "u_offset_px": uoff, "bounds": bounds,
                    "status": rc, "geometry": {k: float(v) for k, v in gg.as_dict().items()},
                    "roi": [int(v) for v in r]})
    with open(os.path.join(HERE, "geometry.json"), "w") as f:
        json.dump(geo, f, indent=1)
    print("wrote fdk_small.npz and geometry.json")


if __name__ == "__main__":
    main()
