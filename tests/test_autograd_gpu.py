"""Differentiation form on the GPU: the forward projector A (exact adjoint of the
backprojection), the FDK operator's backward, and the torch.autograd harness.

Full-size checks use the 512^3 grid with 360 projections of 512x512 pixels
(FCD 2500 px, i.e. SOD 500 mm at 0.2 mm isocentre pitch), with the volume processed in
z-slabs and the backprojection layout recomputed per slab:
  * dot-product test <A x, y> = <x, A^T y>, x ~ N(0,1) (seed 4), y ~ N(0,1) (seed 5),
    inner products in FP64, bar 1e-5;
  * gradient of 0.5 ||A x - y||^2 by central finite differences (h = 1e-3) on 64 random
    in-cylinder voxels (seed 6);
  * the same gradient entries from the CPU FP64 oracle projector.
A small case compares the whole gradient against the FP64 oracle."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FCD = 2500.0


def _op(g, angles, **kw):
    from ct_amd.autograd import ConeBeamOperator
    return ConeBeamOperator(g, angles, **kw)


def _randn(shape, seed, device):
    import torch
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    return torch.randn(shape, generator=gen, device=device, dtype=torch.float32)


def _dot(a, b):
    import torch
    return float(torch.sum(a.double() * b.double()))


def _sub_geometry(g, xi, yi, z0, nz):
    """Geometry of the 1 x 1 x nz column of grid g starting at voxel (xi, yi, z0): the
    same world coordinates, so the same per-voxel arithmetic when z0 is a multiple of 32
    (the projectors evaluate detector rows from 32-slice anchors of the grid)."""
    import ct_amd
    h = ct_amd.Geometry()
    for name, _ in ct_amd.Geometry._fields_:
        setattr(h, name, getattr(g, name))
    h.nx = h.ny = 1
    h.nz = nz
    h.x_pre = g.x_pre - xi
    h.y_pre = g.y_pre - yi
    h.z_pre = g.z_pre - z0
    return h


def _in_cylinder_voxels(g, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        xi, yi, zi = (int(rng.integers(0, g.nx)), int(rng.integers(0, g.ny)), int(rng.integers(0, g.nz)))
        x, y = xi - g.x_pre, yi - g.y_pre
        if x * x + y * y < g.radius * g.radius:
            out.append((xi, yi, zi))
    return out


@pytest.fixture(scope="module")
def full():
    import torch
    import ct_amd
    from ct_amd.phantom import circular_angles
    dev = torch.device("cuda:0")
    g = ct_amd.geometry_grid(512, 512, FCD, 0.0, 512, 512, 512)
    angles = circular_angles(360)
    op = _op(g, angles, slab_nz=128, max_entry_bytes=256 << 20)
    print(f"512^3 operator: {op.nproj} projections, slab {op.slab_nz}, layout batch {op.batch}", flush=True)
    x = _randn(op.volume_shape, 4, dev)
    y = _randn(op.proj_shape, 5, dev)
    return dict(g=g, angles=angles, op=op, x=x, y=y, dev=dev)


def test_dot_product_adjoint_512(full):
    import torch
    op, x, y = full["op"], full["x"], full["y"]
    ax = op.project(x)
    aty = op.backproject(y)
    torch.cuda.synchronize()
    lhs, rhs = _dot(ax, y), _dot(x, aty)
    rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"<Ax,y> = {lhs:.9e}  <x,A^T y> = {rhs:.9e}  rel {rel:.3e}")
    assert rel <= 1e-5


def test_slabs_and_batches_do_not_change_results(full):
    import torch
    g, angles, x, y = full["g"], full["angles"], full["x"], full["y"]
    whole = _op(g, angles)
    a1, a2 = whole.project(x), full["op"].project(x)
    b1, b2 = whole.backproject(y), full["op"].backproject(y)
    torch.cuda.synchronize()
    ea = float(torch.linalg.vector_norm((a1 - a2).double()) / torch.linalg.vector_norm(a1.double()))
    eb = float(torch.linalg.vector_norm((b1 - b2).double()) / torch.linalg.vector_norm(b1.double()))
    print(f"slabbed vs whole: A {ea:.2e}, A^T {eb:.2e}")
    assert ea < 1e-6 and eb < 1e-6


def test_gradient_finite_differences_and_cpu_fp64(full, oracle):
    """Autograd gradient of 0.5||Ax - y||^2 at 64 voxels against (1) central finite
    differences of the loss with h = 1e-3 and (2) the CPU FP64 projector gradient.

    The loss at x +- h e_i is evaluated through the forward projector using A's
    linearity: A(x +- h e_i) = A x +- h A e_i, where A e_i is the same GPU projector
    run on the 32-slice column block of the grid that holds voxel i (identical world
    coordinates and anchors, so identical arithmetic).  Differences are accumulated in
    FP64 over every detector pixel.

    The CPU FP64 gradient differs from the FP32 one by FP32 detector-position rounding
    (one ulp of a coordinate near 500 pixels is 3e-5 pixel, times the residual's
    pixel-to-pixel variation, over 360 projections), about 1e-4 of the gradient's rms;
    the bound is 1e-3 of the rms for the largest of the 64 entries."""
    import torch
    import ct_amd
    from ct_amd.autograd import ConeBeamOperator, least_squares_loss
    g, angles, op, x, y = full["g"], full["angles"], full["op"], full["x"], full["y"]
    xv = x.clone().requires_grad_(True)
    loss = least_squares_loss(xv, y, op)
    loss.backward()
    grad = xv.grad
    with torch.no_grad():
        r = (op.project(x) - y).double()
    voxels = _in_cylinder_voxels(g, 64, 6)
    h = 1e-3
    fd, ag = [], []
    for (xi, yi, zi) in voxels:
        za = zi - zi % 32
        one = ConeBeamOperator(_sub_geometry(g, xi, yi, za, 32), angles)
        e = torch.zeros((1, 1, 32), dtype=torch.float32, device=full["dev"])
        e[0, 0, zi - za] = 1.0
        a = one.project(e).double()
        lp = 0.5 * torch.sum((r + h * a) ** 2)
        lm = 0.5 * torch.sum((r - h * a) ** 2)
        fd.append(float((lp - lm) / (2 * h)))
        ag.append(float(grad[xi, yi, zi]))
        del one
    fd, ag = np.array(fd), np.array(ag)
    scale = np.sqrt(np.mean(fd * fd))
    err_fd = np.max(np.abs(fd - ag)) / scale
    print(f"FD vs autograd at 64 voxels: max err {err_fd:.3e} of rms {scale:.3e}")
    assert err_fd < 1e-4

    # CPU FP64 projector gradient: (A^T r)_i in double for each voxel, same residual
    r_np = r.cpu().numpy()
    cpu = []
    for (xi, yi, zi) in voxels:
        og = oracle.geometry_from(_sub_geometry(g, xi, yi, zi, 1))
        cpu.append(float(oracle.backproject_f64(og, r_np, angles)[0, 0, 0]))
    cpu = np.array(cpu)
    err_cpu = np.max(np.abs(cpu - ag)) / np.sqrt(np.mean(cpu * cpu))
    print(f"CPU FP64 gradient vs autograd: max err {err_cpu:.3e}")
    assert err_cpu < 1e-3


def test_full_gradient_matches_fp64_oracle_small(gpu, oracle):
    """Whole gradient A^T(Ax - y) on a small grid against the CPU FP64 projector pair."""
    import torch
    import ct_amd
    from ct_amd.autograd import ConeBeamOperator, least_squares_loss
    from ct_amd.phantom import circular_angles
    g = ct_amd.geometry_grid(48, 40, 240.0, 1.25, 44, 40, 36)
    angles = circular_angles(30)
    heights = np.linspace(-2.0, 2.0, 30)
    op = ConeBeamOperator(g, angles, heights, slab_nz=16, max_entry_bytes=7 * 48 * 40 * 16)
    x = _randn(op.volume_shape, 14, gpu)
    y = _randn(op.proj_shape, 15, gpu)
    xv = x.clone().requires_grad_(True)
    least_squares_loss(xv, y, op).backward()
    og = oracle.geometry_from(g)
    r64 = oracle.forward_project_f64(og, x.cpu().numpy(), angles, heights) - y.cpu().numpy().astype(np.float64)
    ref = oracle.backproject_f64(og, r64, angles, heights)
    got = xv.grad.cpu().numpy()
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"small gradient vs FP64 oracle: rel RMSE {rel:.3e}")
    assert rel < 5e-5  # FP32 positions and sums against FP64; a wrong term costs O(1)


def test_fdk_operator_and_its_adjoint(gpu, oracle):
    """fdk() on line integrals equals the preprocessing path on the same images, and
    ctg_fdk_backward is its adjoint (dot test), with and without slabs."""
    import torch
    import ct_amd
    from ct_amd.autograd import ConeBeamOperator, fdk
    from helpers import small_scan
    scan = small_scan(W=64, H=48, P=40, seed=21)
    _, g, _, _ = oracle.volume_geometry(scan["W"], scan["H"], scan["fcd"], 0.0)
    cg = ct_amd.geometry_ctvolume(scan["W"], scan["H"], scan["fcd"], 0.0)[0]
    lineint = oracle.log_cpu(scan["raw"], scan["base"], "u16").reshape(scan["raw"].shape)
    ref = oracle.fdk_f32(g, scan["raw"], "u16", scan["base"], "hann", scan["angles"])
    for kw in ({}, {"slab_nz": 16, "max_entry_bytes": 9 * 64 * 48 * 16}):
        op = ConeBeamOperator(cg, scan["angles"], filt="hann", **kw)
        p = torch.from_numpy(lineint).to(gpu)
        vol = op.fdk(p).cpu().numpy()
        err = np.linalg.norm(vol - ref) / np.linalg.norm(ref)
        print(f"fdk() vs oracle FDK ({kw}): {err:.3e}")
        assert err < 1e-4
        gv = _randn(op.volume_shape, 22, gpu)
        pp = _randn(op.proj_shape, 23, gpu)
        lhs = _dot(op.fdk(pp), gv)
        rhs = _dot(pp, op.fdk_adjoint(gv))
        rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
        print(f"FDK dot test: {rel:.3e}")
        assert rel < 1e-5
        # autograd wiring: d/dp <F p, gv> = F^T gv
        pv = pp.clone().requires_grad_(True)
        torch.sum(fdk(pv, op) * gv).backward()
        # the projector sums in a fixed order (no atomics): a repeated run is bit-identical
        ref_g = op.fdk_adjoint(gv)
        assert torch.equal(pv.grad, ref_g)


def test_autograd_functions_are_adjoint_pairs(gpu):
    import torch
    import ct_amd
    from ct_amd.autograd import ConeBeamOperator, backproject, project
    from ct_amd.phantom import circular_angles
    g = ct_amd.geometry_grid(40, 32, 200.0, 0.0, 36, 36, 30)
    op = ConeBeamOperator(g, circular_angles(20))
    x = _randn(op.volume_shape, 31, gpu).requires_grad_(True)
    w = _randn(op.proj_shape, 32, gpu)
    torch.sum(project(x, op) * w).backward()
    assert torch.equal(x.grad, op.backproject(w))
    y = _randn(op.proj_shape, 33, gpu).requires_grad_(True)
    v = _randn(op.volume_shape, 34, gpu)
    torch.sum(backproject(y, op) * v).backward()
    ref = op.project(v)  # deterministic: the same call gives the same bits
    assert torch.equal(y.grad, ref)
