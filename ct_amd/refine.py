"""FDK-seeded gradient refinement with per-slab checkpoints.

Steepest descent with exact line search on ``0.5 ||A x - p||^2`` (A the cone-beam
forward projector, A^T its exact adjoint, both from :class:`ConeBeamOperator`):

    r = A x_k - p,   g = A^T r,   alpha = ||g||^2 / ||A g||^2,   x_{k+1} = x_k - alpha g.

The update is applied z-slab by z-slab (the operator's ``slab_nz``), the granularity
of the reference's chunk loop (Ct/CtVolume.cpp:1056-1320, one chunk downloaded at a
time at :1284).  With a checkpoint directory every finished slab of x_{k+1} is
written to disk together with the step state, so a run that stops (a crash, a kill,
``stop_after``) resumes where it stopped:

* ``iterate.bin``: the committed iterate x_k, slab after slab (slab s is its
  ``(nx, ny, nz_s)`` block, contiguous), float32;
* ``next.bin``: the slabs written so far of the step being checkpointed, same layout;
* ``state.json``: ``{"step": k, "slab": s, "target": t, "alpha": ..., "loss": ...,
  "history": [...]}`` -- k the committed step, and while step t is being written its
  last slab on disk and its line search (``slab`` = -1 between checkpoints) --
  replaced atomically (write + rename) after each slab.

When the last slab of a step is on disk, ``next.bin`` replaces ``iterate.bin`` (a
rename) and the state moves to that step.  Steps can be checkpointed every few
(``every``); the others write nothing.  A resumed run reloads the committed iterate,
recomputes the steps after it (the projector and the backprojection are
deterministic: no atomics, fixed summation order), and in a step stopped part-way
takes alpha from the state, reloads the slabs already written and continues with the
next slab; its result equals an uninterrupted run's bit for bit.
"""
from __future__ import annotations

import json
import os
import time

import torch

from .autograd import ConeBeamOperator


class Interrupted(RuntimeError):
    """Raised by :func:`refine` when ``stop_after`` is reached (a simulated stop)."""


def _sum_sq(t: torch.Tensor) -> float:
    return float(torch.sum(t.double() * t.double()))


class SlabCheckpoint:
    """The checkpoint files of one refinement in ``directory`` (see the module text)."""

    def __init__(self, directory: str, shape, slabs):
        self.dir = directory
        self.shape = tuple(shape)
        self.slabs = list(slabs)  # (z0, nz) per slab
        nx, ny = self.shape[0], self.shape[1]
        self.offsets = []
        off = 0
        for _, nz in self.slabs:
            self.offsets.append(off)
            off += nx * ny * nz * 4
        self.nbytes = off
        os.makedirs(directory, exist_ok=True)

    def _path(self, name):
        return os.path.join(self.dir, name)

    def state(self):
        try:
            with open(self._path("state.json")) as f:
                return json.load(f)
        except FileNotFoundError:
            return None

    def write_state(self, st: dict):
        tmp = self._path("state.json.tmp")
        with open(tmp, "w") as f:
            json.dump(st, f)
            f.flush()
            os.fsync(f.fileno())
        os.replace(tmp, self._path("state.json"))

    def _write_slab(self, name, s, block: torch.Tensor):
        data = block.contiguous().cpu().numpy().tobytes()
        mode = "r+b" if os.path.exists(self._path(name)) else "w+b"
        with open(self._path(name), mode) as f:
            f.seek(self.offsets[s])
            f.write(data)
            f.flush()
            os.fsync(f.fileno())

    def _read_slab(self, name, s, out: torch.Tensor):
        import numpy as np
        nx, ny = self.shape[0], self.shape[1]
        z0, nz = self.slabs[s]
        with open(self._path(name), "rb") as f:
            f.seek(self.offsets[s])
            buf = f.read(nx * ny * nz * 4)
        out[:, :, z0:z0 + nz] = torch.from_numpy(np.frombuffer(buf, np.float32).reshape(nx, ny, nz).copy()).to(
            out.device)

    def save_iterate(self, x: torch.Tensor, name="iterate.bin"):
        for s, (z0, nz) in enumerate(self.slabs):
            self._write_slab(name, s, x[:, :, z0:z0 + nz])

    def load_iterate(self, out: torch.Tensor, name="iterate.bin", upto=None):
        n = len(self.slabs) if upto is None else upto
        for s in range(n):
            self._read_slab(name, s, out)

    def write_next_slab(self, s, block):
        self._write_slab("next.bin", s, block)

    def commit(self):
        os.replace(self._path("next.bin"), self._path("iterate.bin"))


def refine(op: ConeBeamOperator, x0: torch.Tensor, measured: torch.Tensor, steps: int,
           checkpoint_dir: str | None = None, stop_after=None, on_step=None, every: int = 1):
    """Runs (or resumes) ``steps`` steepest-descent steps from ``x0``.

    op: the operator (its ``slab_nz`` sets the checkpoint granularity).
    x0: the seed (used only when the checkpoint directory holds no state).
    measured: line integrals p, (P, H, W) float32 on the operator's device.
    checkpoint_dir: where the per-slab checkpoints go (None: no checkpoints).
    every: checkpoint the steps whose number is a multiple of this (and the last); a
        resumed run recomputes the steps after the last committed one.
    stop_after: (step, slab) after whose write the run raises :class:`Interrupted`
        (steps count from 1, slabs from 0), to exercise the resume path.
    on_step: called as on_step(record, x) after every step.
    Returns (x, history); history holds per step the loss before it, alpha and the ms
    the step took (without the checkpoint writes: ``ms_checkpoint`` separately).
    """
    slabs = list(op._slabs())
    ck = SlabCheckpoint(checkpoint_dir, op.volume_shape, slabs) if checkpoint_dir else None
    x = torch.empty(op.volume_shape, dtype=torch.float32, device=op.device)
    # timing waits for the device (an operator on host tensors, as in the CPU tests of
    # the checkpoint logic, has nothing to wait for)
    sync = torch.cuda.synchronize if x.is_cuda else (lambda: None)
    if every < 1:
        raise ValueError("every must be at least 1")
    # state: "step" = the committed step (iterate.bin holds its iterate); while a
    # checkpointed step "target" is being written, "slab" = its last slab on disk and
    # "alpha"/"loss" its line search (otherwise slab = -1)
    st = ck.state() if ck else None
    mid = None  # (target, slab, alpha, loss) of a step stopped part-way
    if st is None:
        x.copy_(x0)
        committed, history = 0, []
        if ck:
            ck.save_iterate(x)
            ck.write_state({"step": 0, "slab": -1, "target": None, "alpha": None, "loss": None, "history": []})
    else:
        committed = st["step"]
        history = list(st["history"][:committed])
        if st["slab"] >= 0:
            if st["slab"] == len(slabs) - 1 and not os.path.exists(os.path.join(checkpoint_dir, "next.bin")):
                # stopped between the commit's rename and its state write: the step is done
                history = list(st["history"][:st["target"] - 1])
                history.append({"step": st["target"], "loss_before": st["loss"], "alpha": st["alpha"], "ms": None,
                                "ms_checkpoint": None})
                committed = st["target"]
            else:
                mid = (st["target"], st["slab"], st["alpha"], st["loss"])
        ck.load_iterate(x)
    k = committed
    while k < steps:
        k += 1
        ckpt = ck is not None and (k % every == 0 or k == steps)
        resume = mid is not None and mid[0] == k
        sync()
        t0 = time.perf_counter()
        r = op.project(x) - measured
        g = op.backproject(r)
        if resume:
            alpha, loss = mid[2], mid[3]
        else:
            ag = op.project(g)
            alpha = _sum_sq(g) / _sum_sq(ag)
            loss = 0.5 * _sum_sq(r)
            del ag
        del r
        sync()
        ms_grad = (time.perf_counter() - t0) * 1e3
        ms_ck = 0.0
        first = mid[1] + 1 if resume else 0
        if first > 0:  # slabs of this step written before the stop
            ck.load_iterate(x, name="next.bin", upto=first)
        t1 = time.perf_counter()
        for s in range(first, len(slabs)):
            z0, nz = slabs[s]
            x[:, :, z0:z0 + nz] -= alpha * g[:, :, z0:z0 + nz]
            if ckpt:
                tc = time.perf_counter()
                ck.write_next_slab(s, x[:, :, z0:z0 + nz])
                ck.write_state({"step": committed, "slab": s, "target": k, "alpha": alpha, "loss": loss,
                                "history": history})
                ms_ck += (time.perf_counter() - tc) * 1e3
            if stop_after is not None and (k, s) == tuple(stop_after):
                raise Interrupted(f"stopped after slab {s} of step {k}")
        sync()
        ms_update = (time.perf_counter() - t1) * 1e3 - ms_ck
        del g
        rec = {"step": k, "loss_before": loss, "alpha": alpha, "ms": ms_grad + ms_update, "ms_checkpoint": ms_ck}
        history.append(rec)
        if ckpt:
            tc = time.perf_counter()
            ck.commit()
            committed = k
            ck.write_state({"step": k, "slab": -1, "target": None, "alpha": None, "loss": None, "history": history})
            rec["ms_checkpoint"] += (time.perf_counter() - tc) * 1e3
        mid = None
        if on_step:
            on_step(rec, x)
    return x, history
