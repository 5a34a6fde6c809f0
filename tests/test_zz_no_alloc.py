"""ctg_layout_projections_noalloc, the allocation-free variant of ctg_layout_projections
(its max-abs counters live in a static device buffer instead of context memory):
through the C ABI, its entries and scales equal ctg_layout_projections' bit for bit in
every store, on two calls in a row with the second one smaller (the counters keep the
first call's values unless cleared), and it rejects more projections than its buffer
holds before any launch."""
from __future__ import annotations

import ctypes

import pytest

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def test_noalloc_layout_matches_layout(gpu):
    import torch
    import ct_amd
    from ct_amd import _ctg
    W, H = 97, 61
    g = ct_amd.geometry_grid(W, H, 400.0, 0.0, 16, 16, 8)
    gen = torch.Generator(device=gpu)
    with ct_amd.Context(0) as ctx:
        ctx.set_geometry(g)
        lib, h = ctx.lib, ctx.handle
        for P, seed, amp in ((37, 1, 50.0), (5, 2, 0.25)):
            gen.manual_seed(seed)
            img = torch.randn((P, H, W), generator=gen, device=gpu) * amp
            for store in ("f32", "f16", "bf16"):
                sid = _ctg.STORE[store]
                nbytes = ctx.filtered_bytes(store, P)
                ref = torch.zeros(nbytes, dtype=torch.uint8, device=gpu)
                got = torch.full((nbytes,), 0xAB, dtype=torch.uint8, device=gpu)
                sref = torch.zeros(P, dtype=torch.float32, device=gpu)
                sgot = torch.full((P,), -1.0, dtype=torch.float32, device=gpu)
                assert lib.ctg_layout_projections(h, _p(img), P, sid, _p(ref), _p(sref), None) == 0
                st = lib.ctg_layout_projections_noalloc(h, _p(img), P, sid, _p(got), _p(sgot), None)
                assert st == 0, _ctg.last_error()
                torch.cuda.synchronize()
                assert torch.equal(got, ref), (P, store)
                assert torch.equal(sgot, sref), (P, store, sgot, sref)
                if store != "f32":
                    # the scale is the power of two at or above each image's max |value|
                    m = img.abs().amax(dim=(1, 2))
                    assert bool(((sgot >= m) & (sgot <= 2 * m)).all())


def test_noalloc_layout_rejects_too_many_projections(gpu):
    import torch
    import ct_amd
    from ct_amd import _ctg
    g = ct_amd.geometry_grid(8, 4, 100.0, 0.0, 4, 4, 4)
    with ct_amd.Context(0) as ctx:
        ctx.set_geometry(g)
        small = torch.zeros(64, dtype=torch.float32, device=gpu)
        st = ctx.lib.ctg_layout_projections_noalloc(ctx.handle, _p(small), 65537, 0, _p(small), _p(small), None)
        assert st == _ctg.ERROR_INVALID_ARGUMENT
        assert "65536" in _ctg.last_error()
