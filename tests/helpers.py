"""Shared helpers for the tests: small scans, device buffers, error metrics."""
from __future__ import annotations

import numpy as np


def rel_rmse(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.sqrt(np.sum(b * b))
    return float(np.sqrt(np.sum((a - b) ** 2)) / den) if den > 0 else float(np.sqrt(np.sum(a * a)))


def small_scan(W=64, H=48, P=24, fcd=None, seed=11, n_ell=6, photons=0.0, base=1.0, u_offset=0.0,
               heights=None, angles=None):
    """Analytic ellipsoid scan computed by the oracle's CPU projector, as uint16 images."""
    import oracle as orc
    from ct_amd.phantom import circular_angles, phantom_unit_radius, random_ellipsoids
    if fcd is None:
        fcd = W * 1250.0 / 256.0
    _, g, _, _ = orc.volume_geometry(W, H, fcd, u_offset)
    unit = phantom_unit_radius(g.radius, H, fcd)
    ell = random_ellipsoids(n_ell, seed, unit, mu_range=(0.005, 0.02))
    ang = circular_angles(P) if angles is None else list(angles)
    hts = np.zeros(P) if heights is None else np.asarray(heights, dtype=np.float64)
    p = orc.phantom_project(g, ell, ang, hts)
    I = np.clip(np.floor(65535.0 * base * np.exp(-p) + 0.5), 0, 65535).astype(np.uint16)
    return dict(W=W, H=H, P=P, fcd=fcd, u_offset=u_offset, angles=np.asarray(ang, np.float64), heights=hts,
                raw=I, lineint=p, ell=ell, base=base)


def entries_to_rows(entries: np.ndarray) -> np.ndarray:
    """FP32 backprojection layout (P, W, H, 4) -> row-major images (P, H, W) (the A term)."""
    return np.ascontiguousarray(np.transpose(entries[..., 0], (0, 2, 1)))
