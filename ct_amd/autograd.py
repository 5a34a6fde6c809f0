"""torch.autograd harness over the cone-beam operators of libctg.so.

Three linear operators on one circular cone-beam geometry (CtVolume's model,
Ct/CtVolume.cpp:979-1039), each with its exact adjoint as the backward:

* ``A``  (:meth:`ConeBeamOperator.project`): forward projector, volume -> detector
  images.  It transposes the voxel-driven backprojection term by term
  (``ctg_forward_project``), so ``<A x, y> = <x, A^T y>`` up to FP32 rounding.
* ``A^T`` (:meth:`ConeBeamOperator.backproject`): the reference's backprojection of
  row-major detector images (``ctg_layout_projections`` + ``ctg_backproject``).
* ``F = B . H . Wf`` (:meth:`ConeBeamOperator.fdk`): FDK on line integrals -- Feldkamp
  weight (Ct/cuda_ct.cu:102-131), ramp-family row filter (Ct/CtVolume.cpp:832-858),
  backprojection.  Its adjoint ``Wf . H . A`` is ``ctg_fdk_backward``.

Volumes are ``(nx, ny, nz)`` FP32 tensors, z fastest (Common/Volume.h:260); detector
stacks are ``(P, H, W)`` FP32 tensors.  All operators are linear, so the autograd
Functions save no activations: the backward recomputes what it needs from the
incoming gradient.  Memory for the backprojection layout (16 B per detector texel) is
bounded by ``max_entry_bytes``: projections are then laid out and backprojected in
batches, and re-laid out for every z-slab when the volume is processed in slabs of
``slab_nz`` (per-slab recomputation instead of keeping the filtered projections).

There is no CPU path: every call runs the HIP kernels and raises if the library or a
GPU is missing.
"""
from __future__ import annotations

import torch

from ._ctg import Context, Geometry


def _stream():
    return torch.cuda.current_stream().cuda_stream


class ConeBeamOperator:
    """A, A^T and the FDK operator for one geometry and set of projection angles.

    geometry: :class:`ct_amd.Geometry` (``ct_amd.geometry_grid`` / ``geometry_ctvolume``).
    angles_deg, heights_px: per-projection angle (degrees) and height offset (pixels).
    slab_nz: process the volume in z-slabs of this many slices (None = whole volume).
    max_entry_bytes: cap on the backprojection layout buffer (None = all projections).
    filt: ramp-family window of the FDK operator ("ramlak", "shepplogan", "hann").
    """

    def __init__(self, geometry: Geometry, angles_deg, heights_px=None, device: int = 0,
                 slab_nz: int | None = None, max_entry_bytes: int | None = None, filt: str = "ramlak"):
        if not torch.cuda.is_available():
            raise RuntimeError("ct_amd.autograd needs a HIP device; there is no CPU path")
        self.device = torch.device("cuda", device)
        self.ctx = Context(device)
        self.ctx.set_geometry(geometry)
        self.ctx.set_projections(list(angles_deg), None if heights_px is None else list(heights_px))
        self.g = geometry
        self.nproj = len(angles_deg)
        self.slab_nz = int(slab_nz) if slab_nz else geometry.nz
        if self.slab_nz <= 0:
            raise ValueError("slab_nz must be positive")
        per_proj = self.ctx.filtered_bytes("f32", 1)
        self.batch = self.nproj if max_entry_bytes is None else max(1, min(self.nproj, int(max_entry_bytes) // per_proj))
        self.filt = filt

    # ------------------------------------------------------------------ shapes
    @property
    def volume_shape(self):
        return (self.g.nx, self.g.ny, self.g.nz)

    @property
    def proj_shape(self):
        return (self.nproj, self.g.det_height, self.g.det_width)

    def _check(self, t, shape, name):
        if t.device != self.device or t.dtype != torch.float32 or tuple(t.shape) != tuple(shape):
            raise ValueError(f"{name}: expected float32 {tuple(shape)} on {self.device}, got {t.dtype} "
                             f"{tuple(t.shape)} on {t.device}")
        return t.contiguous()

    def _slabs(self):
        for z0 in range(0, self.g.nz, self.slab_nz):
            yield z0, min(self.slab_nz, self.g.nz - z0)

    # ---------------------------------------------------------------- operators
    def project(self, volume: torch.Tensor) -> torch.Tensor:
        """A x: (nx, ny, nz) -> (P, H, W)."""
        volume = self._check(volume, self.volume_shape, "volume")
        out = torch.empty(self.proj_shape, dtype=torch.float32, device=self.device)
        s = _stream()
        for i, (z0, nz) in enumerate(self._slabs()):
            slab = volume if nz == self.g.nz else volume[:, :, z0:z0 + nz].contiguous()
            self.ctx.forward_project(slab, z0, nz, 0, self.nproj, out, accumulate=i > 0, stream=s)
        return out

    def _backproject_entries(self, make_entries, out: torch.Tensor):
        """Backprojects the layout produced by make_entries(p0, n, entries, scales) into out,
        slab by slab; the layout is recomputed per slab unless all of it is resident."""
        s = _stream()
        per = self.ctx.filtered_bytes("f32", 1)
        entries = torch.empty(per * self.batch, dtype=torch.uint8, device=self.device)
        scales = torch.empty(self.batch, dtype=torch.float32, device=self.device)
        resident = self.batch >= self.nproj
        if resident:
            make_entries(0, self.nproj, entries, scales)
        for z0, nz in self._slabs():
            whole = nz == self.g.nz
            slab = out if whole else torch.empty((self.g.nx, self.g.ny, nz), dtype=torch.float32,
                                                 device=self.device)
            for p0 in range(0, self.nproj, self.batch):
                n = min(self.batch, self.nproj - p0)
                if not resident:
                    make_entries(p0, n, entries, scales)
                self.ctx.backproject(entries, "f32", scales, p0, n, slab, z0, nz, accumulate=p0 > 0, stream=s)
            if not whole:
                out[:, :, z0:z0 + nz] = slab
        return out

    def backproject(self, images: torch.Tensor) -> torch.Tensor:
        """A^T y: (P, H, W) -> (nx, ny, nz)."""
        images = self._check(images, self.proj_shape, "images")
        out = torch.empty(self.volume_shape, dtype=torch.float32, device=self.device)
        s = _stream()

        def make(p0, n, entries, scales):
            self.ctx.layout_projections(images[p0:p0 + n], n, "f32", entries, scales, stream=s)

        return self._backproject_entries(make, out)

    def fdk(self, lineint: torch.Tensor) -> torch.Tensor:
        """F p = B H Wf p: line integrals (P, H, W) -> (nx, ny, nz)."""
        lineint = self._check(lineint, self.proj_shape, "lineint")
        out = torch.empty(self.volume_shape, dtype=torch.float32, device=self.device)
        s = _stream()
        H, W = self.g.det_height, self.g.det_width
        work = torch.empty((self.batch, H, W), dtype=torch.float32, device=self.device)

        def make(p0, n, entries, scales):
            w = work[:n]
            w.copy_(lineint[p0:p0 + n])
            self.ctx.feldkamp_weight(w, n, stream=s)
            self.ctx.filter_rows(w, w, n * H, self.filt, stream=s)
            self.ctx.layout_projections(w, n, "f32", entries, scales, stream=s)

        return self._backproject_entries(make, out)

    def fdk_adjoint(self, grad_volume: torch.Tensor) -> torch.Tensor:
        """F^T g = Wf H A g: (nx, ny, nz) -> (P, H, W)."""
        grad_volume = self._check(grad_volume, self.volume_shape, "grad_volume")
        s = _stream()
        out = torch.empty(self.proj_shape, dtype=torch.float32, device=self.device)
        if self.slab_nz >= self.g.nz:
            self.ctx.fdk_backward(grad_volume, 0, self.g.nz, 0, self.nproj, self.filt, out, stream=s)
            return out
        # slabs: A is accumulated over slabs first, then filtered and weighted once
        for i, (z0, nz) in enumerate(self._slabs()):
            slab = grad_volume[:, :, z0:z0 + nz].contiguous()
            self.ctx.forward_project(slab, z0, nz, 0, self.nproj, out, accumulate=i > 0, stream=s)
        H = self.g.det_height
        self.ctx.filter_rows(out, out, self.nproj * H, self.filt, stream=s)
        self.ctx.feldkamp_weight(out, self.nproj, stream=s)
        return out


class ConeProject(torch.autograd.Function):
    """y = A x; backward: grad_x = A^T grad_y."""

    @staticmethod
    def forward(ctx, volume, op: ConeBeamOperator):
        ctx.op = op
        return op.project(volume)

    @staticmethod
    def backward(ctx, grad):
        return ctx.op.backproject(grad.contiguous()), None


class ConeBackproject(torch.autograd.Function):
    """x = A^T y; backward: grad_y = A grad_x."""

    @staticmethod
    def forward(ctx, images, op: ConeBeamOperator):
        ctx.op = op
        return op.backproject(images)

    @staticmethod
    def backward(ctx, grad):
        return ctx.op.project(grad.contiguous()), None


class FDK(torch.autograd.Function):
    """x = B H Wf p on line integrals; backward: grad_p = Wf H A grad_x (ctg_fdk_backward)."""

    @staticmethod
    def forward(ctx, lineint, op: ConeBeamOperator):
        ctx.op = op
        return op.fdk(lineint)

    @staticmethod
    def backward(ctx, grad):
        return ctx.op.fdk_adjoint(grad.contiguous()), None


def project(volume, op):
    return ConeProject.apply(volume, op)


def backproject(images, op):
    return ConeBackproject.apply(images, op)


def fdk(lineint, op):
    return FDK.apply(lineint, op)


def least_squares_loss(volume, measured, op):
    """0.5 * ||A x - p||^2 accumulated in FP64, differentiable in ``volume``."""
    r = project(volume, op) - measured
    return 0.5 * torch.sum(r.double() * r.double())
