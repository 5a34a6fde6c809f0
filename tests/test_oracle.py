"""Pins of the CPU oracle (oracle/oracle.c) against facts fixed independently of it:
literal sizes of the configured scans, closed forms, numpy's FFT, an independent
vectorised brute force of the geometry model, and invariants.

The reference ships no tests or golden vectors and cannot be built here (OpenCV,
CUDA/cuFFT and Qt are absent), so these checks are what holds the oracle.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from helpers import rel_rmse, small_scan


# ------------------------------------------------------------------ geometry

@pytest.mark.parametrize("W,H,fcd,grid", [
    (256, 192, 1250.0, (254, 254, 192)),     # 256x192 detector, 0.4 mm iso pitch, SOD 500 mm
    (1024, 1024, 2500.0, (1002, 1002, 1024)),
    (2048, 1536, 5000.0, (2006, 2006, 1536)),
    (512, 512, 2500.0, (508, 508, 512)),
    (1024, 768, 5000.0, (1018, 1018, 768)),
])
def test_ctvolume_grid_sizes(oracle, W, H, fcd, grid):
    rc, g, roi, _ = oracle.volume_geometry(W, H, fcd)
    assert rc == 0
    assert (g.nx, g.ny, g.nz) == grid
    assert roi == (0, grid[0], 0, grid[1], 0, grid[2])
    assert g.x_pre == grid[0] / 2 and g.z_pre == H / 2


def test_roi_rounding_literal(oracle):
    # x,y in [0.25,0.75], z in [0.1,0.9] on the 1002x1002x1024 grid
    rc, g, roi, _ = oracle.volume_geometry(1024, 1024, 2500.0, 0.0, [0.25, 0.75, 0.25, 0.75, 0.1, 0.9])
    assert rc == 0
    assert roi == (251, 752, 251, 752, 102, 922)
    assert (g.nx, g.ny, g.nz) == (501, 501, 820)
    assert g.x_pre == pytest.approx(501 - 251) and g.z_pre == pytest.approx(512 - 102)


def test_set_volume_bounds_clamp_quirk(oracle):
    # yTo and zTo are clamped against xFrom (not yFrom/zFrom) in the reference
    rc, g, roi, eff = oracle.volume_geometry(256, 192, 1250.0, 0.0, [0.6, 0.9, 0.1, 0.3, 0.0, 0.5])
    assert eff[3] == pytest.approx(0.6)      # yTo raised to xFrom
    assert roi[2:4] == (25, 152)
    # equal bounds are ignored (full grid kept)
    rc, g, roi, eff = oracle.volume_geometry(256, 192, 1250.0, 0.0, [0.3, 0.3, 0.0, 1.0, 0.0, 1.0])
    assert roi == (0, 254, 0, 254, 0, 192)
    # a yTo below yFrom through the clamp wraps in the reference; the oracle reports it
    rc, g, roi, eff = oracle.volume_geometry(256, 192, 1250.0, 0.0, [0.1, 0.9, 0.8, 0.2, 0.0, 1.0])
    assert rc == -1


def test_detector_window(oracle):
    _, g, _, _ = oracle.volume_geometry(256, 192, 1250.0)
    assert (g.u_pre, g.v_pre) == (128.0, 96.0)
    assert g.u_lo == -128.0 and g.u_hi == pytest.approx(126.9, abs=1e-5)
    assert g.v_lo == pytest.approx(-94.9, abs=1e-5) and g.v_hi == 96.0


# ------------------------------------------------------------- preprocessing

def test_log_closed_form(oracle):
    p = np.linspace(0.0, 6.0, 97)
    for B in (1.0, 0.7, 1e4 / 65535.0):
        I = np.round(65535.0 * B * np.exp(-p)).astype(np.uint16)
        got_cpu = oracle.log_cpu(I, B, "u16")
        got_gpu = oracle.log_gpu(I, B, "u16")
        # quantisation of I bounds |dp| by 0.5/I, the 1e-6 epsilon adds at most ~1e-6/(I/(65535B))
        Ip = np.maximum(I.astype(np.float64), 1.0)
        tol = 0.5 / Ip + 2e-6 * 65535.0 * B / Ip + 1e-6
        assert np.all(np.abs(got_cpu - p) <= tol)
        assert np.all(np.abs(got_gpu - got_cpu) <= 2e-6 * np.maximum(1.0, np.abs(got_cpu)))
    # zero intensity stays finite thanks to the epsilon
    assert np.isfinite(oracle.log_cpu(np.zeros(4, np.uint16), 1.0, "u16")).all()
    # 8-bit images use 255 as full scale
    I8 = np.array([255, 128, 1], np.uint8)
    assert np.allclose(oracle.log_cpu(I8, 1.0, "u8"), -np.log(I8 / 255.0 + 1e-6) + math.log(1 + 1e-6), atol=1e-6)


def test_feldkamp_is_cosine_of_ray_angle(oracle):
    W, H, fcd = 16, 12, 40.0
    img = np.ones((1, H, W), np.float32)
    got = oracle.feldkamp(img, fcd, W / 2, H / 2)[0]
    # cosine between the central ray (fcd,0,0) and the ray to pixel (c, r)
    c = np.arange(W)[None, :] - W / 2
    r = H / 2 - np.arange(H)[:, None]
    rays = np.stack(np.broadcast_arrays(np.full((H, W), fcd), c, r), -1)
    cosang = rays[..., 0] / np.linalg.norm(rays, axis=-1)
    assert np.allclose(got, cosang, rtol=1e-6)
    assert got[H // 2, W // 2] == 1.0


def test_window_values_literal(oracle):
    W = 8  # N = W/2 + 1 = 5, r = k/5
    assert np.allclose(oracle.filter_window("ramlak", W), [0.0, 0.2, 0.4, 0.6, 0.8])
    assert np.allclose(oracle.filter_window("shepplogan", W),
                       [0.0, 0.1967263, 0.3741957, 0.5150362, 0.6054614], atol=1e-6)
    assert np.allclose(oracle.filter_window("hann", W),
                       [0.0, 0.1809017, 0.2618034, 0.2072949, 0.0763932], atol=1e-6)


def _numpy_filter(rows, filt, oracle):
    W = rows.shape[-1]
    h = oracle.filter_window(filt, W).astype(np.float64)
    return np.fft.irfft(np.fft.rfft(rows, axis=-1) * h, n=W, axis=-1) * W


@pytest.mark.parametrize("W", [8, 64, 90, 255, 256])
@pytest.mark.parametrize("filt", ["ramlak", "shepplogan", "hann"])
def test_filter_rows_matches_numpy_fft(oracle, W, filt):
    rows = np.random.default_rng(W).standard_normal((7, W))
    ref = _numpy_filter(rows, filt, oracle)
    for direct in (False, True):
        got = oracle.filter_rows_f64(rows, filt, direct=direct)
        assert np.max(np.abs(got - ref)) < 1e-9 * max(1.0, np.max(np.abs(ref)))
    got32 = oracle.filter_rows(rows.astype(np.float32), filt)
    assert rel_rmse(got32, ref) < 1e-6


def test_filter_invariants(oracle):
    W = 64
    # DC row -> 0 (H(0) = 0)
    assert np.max(np.abs(oracle.filter_rows_f64(np.full((1, W), 3.0), "ramlak"))) < 1e-10
    # impulse response = unnormalised inverse DFT of the even-extended window
    delta = np.zeros((1, W))
    delta[0, 0] = 1.0
    h = oracle.filter_window("hann", W).astype(np.float64)
    full = np.concatenate([h, h[1:W // 2][::-1]])
    g = np.real(np.fft.ifft(full)) * W
    assert np.allclose(oracle.filter_rows_f64(delta, "hann")[0], g, atol=1e-12)
    # shift equivariance (circular) and linearity
    x = np.random.default_rng(1).standard_normal((2, W))
    fx = oracle.filter_rows_f64(x, "shepplogan")
    assert np.allclose(oracle.filter_rows_f64(np.roll(x, 5, axis=1), "shepplogan"), np.roll(fx, 5, axis=1))
    assert np.allclose(oracle.filter_rows_f64(2 * x[:1] - x[1:], "shepplogan"), 2 * fx[:1] - fx[1:])


# ------------------------------------------------------------ backprojection

def _brute_bp(g, q, angles, heights):
    """Vectorised restatement of the geometry model (independent of the oracle's loops)."""
    x = np.arange(g.nx)[:, None, None] - np.float64(g.x_pre)
    y = np.arange(g.ny)[None, :, None] - np.float64(g.y_pre)
    z = np.arange(g.nz)[None, None, :] - np.float64(g.z_pre)
    vol = np.zeros((g.nx, g.ny, g.nz))
    for p, ang in enumerate(np.deg2rad(angles)):
        sn, cs = np.sin(ang), np.cos(ang)
        t = -x * sn + y * cs + g.u_offset
        s = x * cs + y * sn
        U = g.fcd / (g.fcd - s)
        u = t * U
        v = (z + heights[p]) * U
        ok = (x ** 2 + y ** 2 < np.float64(g.radius) ** 2) & (u >= g.u_lo) & (u <= g.u_hi) & \
             (v >= g.v_lo) & (v <= g.v_hi)
        uI = np.broadcast_to(u + g.u_pre, ok.shape)[ok]
        vI = np.broadcast_to(g.v_pre - v, ok.shape)[ok]
        u0 = np.floor(uI).astype(int)
        v0 = np.floor(vI).astype(int)
        fu, fv = uI - u0, vI - v0
        img = q[p]
        val = (1 - fv) * ((1 - fu) * img[v0, u0] + fu * img[v0, u0 + 1]) + \
            fv * ((1 - fu) * img[v0 + 1, u0] + fu * img[v0 + 1, u0 + 1])
        vol[ok] += np.broadcast_to(U * U, ok.shape)[ok] * val
    return vol


def test_backprojection_brute_force(oracle):
    W, H, fcd = 16, 12, 40.0
    rc, g, _, _ = oracle.volume_geometry(W, H, fcd, 0.75)
    assert (g.nx, g.nz) == (14, 12)
    rng = np.random.default_rng(5)
    angles = np.array([0.0, 37.5, 211.0])
    heights = np.array([0.0, 0.4, -0.7])
    q = rng.standard_normal((3, H, W))
    ref = _brute_bp(g, q, angles, heights)
    got64 = oracle.backproject_f64(g, q, angles, heights)
    assert np.max(np.abs(got64 - ref)) < 1e-12
    got32 = oracle.backproject_f32(g, q.astype(np.float32), angles, heights)
    assert rel_rmse(got32, ref) < 1e-5


def test_backprojection_closed_form_single_view(oracle):
    """theta = 0 and a constant detector c: every voxel inside the masks gets U^2 c with
    U = FCD/(FCD - x)."""
    W, H, fcd, c = 32, 24, 80.0, 1.5
    _, g, _, _ = oracle.volume_geometry(W, H, fcd)
    q = np.full((1, H, W), c)
    vol = oracle.backproject_f64(g, q, [0.0])
    for xi, yi, zi in [(g.nx // 2, g.ny // 2, g.nz // 2), (3, g.ny // 2, 5), (g.nx - 4, g.ny // 2 + 2, g.nz - 3)]:
        x = xi - g.x_pre
        U = fcd / (fcd - x)
        assert vol[xi, yi, zi] == pytest.approx(U * U * c, rel=1e-12)
    # outside the cylinder the volume is exactly zero
    assert vol[0, 0, 0] == 0.0


def test_forward_projector_adjoint_f64(oracle):
    W, H, fcd = 24, 20, 60.0
    _, g, _, _ = oracle.volume_geometry(W, H, fcd, -0.5)
    rng = np.random.default_rng(9)
    angles = np.array([0.0, 45.0, 100.0, 300.0])
    heights = np.array([0.3, 0.0, -0.2, 0.1])
    x = rng.standard_normal((g.nx, g.ny, g.nz))
    y = rng.standard_normal((4, H, W))
    lhs = np.sum(oracle.forward_project_f64(g, x, angles, heights) * y)
    rhs = np.sum(x * oracle.backproject_f64(g, y, angles, heights))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def test_linearity_and_f32_vs_f64(oracle):
    scan = small_scan(W=48, H=32, P=12, seed=4)
    _, g, _, _ = oracle.volume_geometry(48, 32, scan["fcd"])
    q = oracle.preprocess_f32(g, scan["raw"], "u16", 1.0, "ramlak")
    a = oracle.backproject_f64(g, q.astype(np.float64), scan["angles"])
    b = oracle.backproject_f64(g, 2.0 * q.astype(np.float64), scan["angles"])
    assert np.allclose(b, 2.0 * a)
    c = oracle.backproject_f32(g, q, scan["angles"])
    assert rel_rmse(c, a) < 1e-5


def test_backprojection_in_z_slabs_equals_the_whole_grid(oracle):
    """The full-grid CPU baseline backprojects in z-slabs (each slab a grid of nz' slices
    with zPre shifted by its first slice): every voxel's arithmetic is the same, so the
    slabs put together equal the whole grid bit for bit."""
    scan = small_scan(W=64, H=48, P=20, seed=9)
    _, g, _, _ = oracle.volume_geometry(64, 48, scan["fcd"])
    q = oracle.preprocess_f32(g, scan["raw"], "u16", 1.0, "shepplogan")
    whole = oracle.backproject_f32(g, q, scan["angles"])
    parts = []
    for z0 in range(0, g.nz, 20):
        sub = oracle.geometry_from(g)
        sub.nz = min(20, g.nz - z0)
        sub.z_pre = g.z_pre - z0
        parts.append(oracle.backproject_f32(sub, q, scan["angles"]))
    assert np.array_equal(np.concatenate(parts, axis=2), whole)
    assert float(np.abs(whole).max()) > 0.0


# --------------------------------------------------------------------- phantom

def test_phantom_sphere_chords(oracle):
    W, H, fcd, r0 = 32, 32, 100.0, 6.0
    _, g, _, _ = oracle.volume_geometry(W, H, fcd)
    p = oracle.phantom_project(g, [[0, 0, 0, r0, r0, r0, 0, 0, 0, 1.0]], [0.0])[0]
    for c, r in [(16, 16), (18, 15), (21, 19), (10, 12)]:
        u, v = c - g.u_pre, g.v_pre - r
        d = fcd * math.hypot(u, v) / math.sqrt(fcd * fcd + u * u + v * v)
        expect = 2 * math.sqrt(r0 * r0 - d * d) if d < r0 else 0.0
        assert p[r, c] == pytest.approx(expect, abs=1e-9)


def test_fdk_scale_constant_uniform_sphere(oracle):
    """mu ~= out * (pi/P) * (W/2+1)/W^2: the reference omits the angular factor, the
    1/W of the inverse DFT and uses k/(W/2+1) for the ramp (sanity bound, not parity).
    The frequency-sampled ramp without zero padding (circular convolution, H(0) = 0)
    lowers the DC level by a few percent, hence the 8 % band."""
    W, H, P, mu, r0 = 64, 64, 90, 0.02, 12.0
    fcd = 400.0
    _, g, _, _ = oracle.volume_geometry(W, H, fcd)
    ang = np.arange(P) * 360.0 / P
    p = oracle.phantom_project(g, [[0, 0, 0, r0, r0, r0, 0, 0, 0, mu]], ang)
    raw = np.clip(np.round(65535.0 * np.exp(-p)), 0, 65535).astype(np.uint16)
    vol = oracle.fdk_f32(g, raw, "u16", 1.0, "ramlak", ang)
    c = g.nx // 2
    centre = vol[c - 3:c + 3, c - 3:c + 3, g.nz // 2 - 3:g.nz // 2 + 3].mean()
    est = centre * (math.pi / P) * (W / 2 + 1) / W ** 2
    assert est == pytest.approx(mu, rel=0.08)
