"""Synthetic-scan tooling (ct_amd.phantom) on the CPU: the voxelised phantom that
scores reconstructions must use the same placement and rotation convention as the
analytic line integrals the scans are made from."""
from __future__ import annotations

import math

import numpy as np


def test_voxelize_moments_pin_centre_and_rotation():
    """Mass, centroid and second-moment tensor of a voxelised ellipsoid against the
    closed forms V = 4/3 pi abc, centroid c, M2 = R diag(a^2, b^2, c^2)/5 R^T V (a
    transposed rotation or a sign slip in the centre fails these)."""
    import ct_amd
    from ct_amd.phantom import _rotation, voxelize
    g = ct_amd.geometry_grid(160, 160, 800.0, 0.0, 96, 96, 96)
    c, ax, ang = [5.0, -7.0, 3.0], [30.0, 18.0, 10.0], [0.4, 0.9, 1.3]
    v = voxelize([c + ax + ang + [1.0]], g, "cpu", dtype=None).double().numpy()
    i, j, k = np.meshgrid(np.arange(g.nx) - g.x_pre, np.arange(g.ny) - g.y_pre, np.arange(g.nz) - g.z_pre,
                          indexing="ij")
    m = v.sum()
    vol = 4.0 / 3.0 * math.pi * ax[0] * ax[1] * ax[2]
    assert abs(m - vol) / vol < 0.01
    cen = np.array([(v * i).sum(), (v * j).sum(), (v * k).sum()]) / m
    assert np.allclose(cen, c, atol=0.05)
    d = [i - cen[0], j - cen[1], k - cen[2]]
    M2 = np.array([[(v * d[a] * d[b]).sum() / m for b in range(3)] for a in range(3)])
    R = np.array(_rotation(*ang))
    ref = R @ np.diag(np.square(ax) / 5.0) @ R.T
    assert np.abs(M2 - ref).max() < 0.02 * np.abs(ref).max()
    assert np.abs(M2 - R.T @ np.diag(np.square(ax) / 5.0) @ R).max() > 0.1 * np.abs(ref).max()


def test_voxelized_phantom_projects_to_the_analytic_line_integrals(oracle):
    """The oracle's FP64 projector applied to the voxelised phantom reproduces the
    analytic cone-beam line integrals of the same ellipsoids (discretisation error only)."""
    import ct_amd
    from ct_amd.phantom import voxelize
    _, g, _, _ = oracle.volume_geometry(96, 72, 480.0, 0.0)
    ell = [[4.0, -3.0, 2.0, 22.0, 12.0, 9.0, 0.5, 1.1, 0.3, 0.05],
           [-6.0, 5.0, -4.0, 8.0, 14.0, 6.0, 2.0, 0.4, 1.7, 0.08]]
    angles = [0.0, 37.0, 90.0, 151.0]
    ref = oracle.phantom_project(g, ell, angles)
    vol = voxelize(ell, ct_amd.geometry_ctvolume(96, 72, 480.0, 0.0)[0], "cpu").double().numpy()
    got = oracle.forward_project_f64(g, vol, angles)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 0.05, err
    # the same ellipsoids with the rotation transposed project visibly differently
    swapped = [e[:6] + [-e[8], -e[7], -e[6]] + [e[9]] for e in ell]
    vol2 = voxelize(swapped, ct_amd.geometry_ctvolume(96, 72, 480.0, 0.0)[0], "cpu").double().numpy()
    err2 = np.linalg.norm(oracle.forward_project_f64(g, vol2, angles) - ref) / np.linalg.norm(ref)
    assert err2 > 3 * err, (err, err2)
