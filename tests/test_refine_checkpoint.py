"""The per-slab checkpoint logic of ct_amd.refine on the CPU, with a small dense linear
operator standing in for the cone-beam pair: stops at every slab of every step (with
checkpoints every step and every second step), a stop between the final rename and
the state write, and a finished checkpoint, each resumed and compared bit for bit
with an uninterrupted run."""
from __future__ import annotations

import json
import os

import pytest
import torch

from ct_amd.refine import Interrupted, refine


class DenseOperator:
    """A x = M vec(x) with a fixed random M; three z-slabs of 4 slices."""

    def __init__(self, shape=(5, 4, 12), rows=300, slab_nz=4, seed=0):
        gen = torch.Generator().manual_seed(seed)
        self.volume_shape = shape
        self.device = torch.device("cpu")
        self.slab_nz = slab_nz
        n = shape[0] * shape[1] * shape[2]
        self.M = torch.randn((rows, n), generator=gen, dtype=torch.float32)

    def _slabs(self):
        nz = self.volume_shape[2]
        for z0 in range(0, nz, self.slab_nz):
            yield z0, min(self.slab_nz, nz - z0)

    def project(self, x):
        return self.M @ x.reshape(-1)

    def backproject(self, r):
        return (self.M.t() @ r).reshape(self.volume_shape)


@pytest.fixture(scope="module")
def problem():
    op = DenseOperator()
    gen = torch.Generator().manual_seed(1)
    truth = torch.randn(op.volume_shape, generator=gen)
    measured = op.project(truth) + 0.01 * torch.randn(op.M.shape[0], generator=gen)
    x0 = torch.zeros(op.volume_shape)
    ref, hist = refine(op, x0, measured, 5)
    return op, x0, measured, ref, hist


def test_uninterrupted_runs_agree_and_descend(problem, tmp_path):
    op, x0, measured, ref, hist = problem
    losses = [h["loss_before"] for h in hist]
    assert all(b <= a for a, b in zip(losses, losses[1:]))
    got, h2 = refine(op, x0, measured, 5, checkpoint_dir=str(tmp_path / "ck"))
    assert torch.equal(got, ref)
    assert [h["alpha"] for h in h2] == [h["alpha"] for h in hist]
    st = json.load(open(str(tmp_path / "ck" / "state.json")))
    assert st["step"] == 5 and st["slab"] == -1


@pytest.mark.parametrize("every", [1, 2])
def test_resume_after_a_stop_at_any_slab(problem, tmp_path, every):
    op, x0, measured, ref, hist = problem
    for step in range(1, 6):
        for slab in range(3):
            d = str(tmp_path / f"ck_{every}_{step}_{slab}")
            with pytest.raises(Interrupted):
                refine(op, x0, measured, 5, checkpoint_dir=d, stop_after=(step, slab), every=every)
            st = json.load(open(os.path.join(d, "state.json")))
            checkpointed = step % every == 0 or step == 5
            # the committed step is the last checkpointed one before the stop
            assert st["step"] == ((step - 1) // every) * every or (every == 1 and st["step"] == step - 1)
            assert st["slab"] == (slab if checkpointed else -1)
            got, h = refine(op, torch.full_like(x0, 7.0), measured, 5, checkpoint_dir=d, every=every)
            assert torch.equal(got, ref), (every, step, slab)
            assert [r["loss_before"] for r in h] == [r["loss_before"] for r in hist]


def test_stop_between_rename_and_state_write(problem, tmp_path):
    """The last slab of step 3 is on disk and renamed to the iterate, but the state still
    says step 2 part-way: the resumed run takes step 3 as done."""
    op, x0, measured, ref, hist = problem
    d = str(tmp_path / "ck")
    with pytest.raises(Interrupted):
        refine(op, x0, measured, 5, checkpoint_dir=d, stop_after=(3, 2))
    os.replace(os.path.join(d, "next.bin"), os.path.join(d, "iterate.bin"))
    got, h = refine(op, x0, measured, 5, checkpoint_dir=d)
    assert torch.equal(got, ref)
    assert [r["loss_before"] for r in h] == [r["loss_before"] for r in hist]


def test_finished_checkpoint_returns_its_iterate(problem, tmp_path):
    op, x0, measured, ref, hist = problem
    d = str(tmp_path / "ck")
    refine(op, x0, measured, 5, checkpoint_dir=d, every=2)
    got, h = refine(op, torch.full_like(x0, 3.0), measured, 5, checkpoint_dir=d, every=2)
    assert torch.equal(got, ref)
    assert len(h) == 5
    with pytest.raises(ValueError):
        refine(op, x0, measured, 5, every=0)
