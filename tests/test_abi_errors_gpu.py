"""Argument errors and empty inputs at the C ABI, on a live context.

The reference reports failures through a `bool&` out-parameter and stdout
(Ct/cuda_ct.cu:95-99, 138-144); here each rejected call returns
CTG_ERROR_INVALID_ARGUMENT with a message in ctg_last_error(), before any kernel
is launched.  Empty inputs (zero projections, zero rows, zero slices) are valid
no-ops, except that a zero-projection overwrite leaves a zero slab.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = 4


@pytest.fixture(scope="module")
def setup(gpu):
    import torch
    import ct_amd
    g = ct_amd.geometry_grid(32, 24, 120.0, 0.0, 16, 16, 8)
    ctx = ct_amd.Context(0)
    ctx.set_geometry(g)
    ctx.set_projections([0.0, 45.0, 90.0, 135.0])
    dev = gpu
    bufs = {
        "raw": torch.full((P, 24, 32), 30000, dtype=torch.uint16, device=dev),
        "img": torch.ones((P, 24, 32), dtype=torch.float32, device=dev),
        "filt16": torch.empty(ctx.filtered_bytes("f16", P), dtype=torch.uint8, device=dev),
        "filt32": torch.empty(ctx.filtered_bytes("f32", P), dtype=torch.uint8, device=dev),
        "scales": torch.ones(P, dtype=torch.float32, device=dev),
        "vol": torch.zeros((16, 16, 8), dtype=torch.float32, device=dev),
        "vol64": torch.zeros((16, 16, 8), dtype=torch.float64, device=dev),
        "proj": torch.zeros((P, 24, 32), dtype=torch.float32, device=dev),
        "out": torch.zeros(16 * 16 * 8, dtype=torch.float32, device=dev),
    }
    yield ctx, g, bufs
    ctx.close()


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _rejected(ctx, status, words):
    from ct_amd import _ctg
    msg = _ctg.last_error()
    assert status == _ctg.ERROR_INVALID_ARGUMENT, (status, msg)
    assert words in msg, msg


def test_reduced_precision_stores_need_scales(setup):
    ctx, g, b = setup
    lib, h = ctx.lib, ctx.handle
    st = lib.ctg_preprocess(h, _p(b["raw"]), 1, P, 0.9, 0, 0, 1, _p(b["filt16"]), None, None)
    _rejected(ctx, st, "scale")
    st = lib.ctg_layout_projections(h, _p(b["img"]), P, 2, _p(b["filt16"]), None, None)
    _rejected(ctx, st, "scale")
    st = lib.ctg_backproject(h, _p(b["filt16"]), 1, None, 0, P, _p(b["vol"]), 0, 8, 0, None)
    _rejected(ctx, st, "scale")
    # FP32 entries take a null scale array as unit scales
    assert lib.ctg_preprocess(h, _p(b["raw"]), 1, P, 0.9, 0, 0, 0, _p(b["filt32"]), None, None) == 0


def test_enum_values_out_of_range(setup):
    ctx, g, b = setup
    lib, h = ctx.lib, ctx.handle
    _rejected(ctx, lib.ctg_preprocess(h, _p(b["raw"]), 1, P, 0.9, 2, 0, 0, _p(b["filt32"]), _p(b["scales"]),
                                      None), "enum")
    _rejected(ctx, lib.ctg_preprocess(h, _p(b["raw"]), 3, P, 0.9, 0, 0, 0, _p(b["filt32"]), _p(b["scales"]),
                                      None), "enum")
    _rejected(ctx, lib.ctg_filter_rows(h, _p(b["img"]), _p(b["img"]), P * 24, 3, None), "filter")
    _rejected(ctx, lib.ctg_layout_projections(h, _p(b["img"]), P, 3, _p(b["filt32"]), _p(b["scales"]), None),
              "store")
    _rejected(ctx, lib.ctg_backproject(h, _p(b["filt32"]), -1, _p(b["scales"]), 0, P, _p(b["vol"]), 0, 8, 0,
                                       None), "store")
    _rejected(ctx, lib.ctg_fdk_backward(h, _p(b["vol"]), 0, 8, 0, P, 7, _p(b["proj"]), None), "filter")
    _rejected(ctx, lib.ctg_fdk_f64(h, _p(b["raw"]), 9, P, 0.9, 0, _p(b["vol64"]), 0, 8, None), "enum")
    _rejected(ctx, lib.ctg_convert_volume(h, _p(b["vol"]), 8, 2, 1, _p(b["out"]), None), "enum")
    host_raw = np.zeros((P, 24, 32), np.uint16)
    host_vol = np.zeros((16, 16, 8), np.float32)
    from ct_amd import _ctg
    st = lib.ctg_reconstruct(h, host_raw.ctypes.data_as(ctypes.c_void_p), 1, P, 0.9, 0, 5, 0,
                             host_vol.ctypes.data_as(ctypes.c_void_p), 1, _ctg.PROGRESS_FN(), None)
    _rejected(ctx, st, "filter")
    # an unknown store or pixel type is rejected before the memory plan (a store of 3
    # would size no filtered buffer; a pixel type of 3 would be read as 4 bytes a pixel)
    st = lib.ctg_reconstruct(h, host_raw.ctypes.data_as(ctypes.c_void_p), 1, P, 0.9, 0, 0, 3,
                             host_vol.ctypes.data_as(ctypes.c_void_p), 1, _ctg.PROGRESS_FN(), None)
    _rejected(ctx, st, "store")
    st = lib.ctg_reconstruct(h, host_raw.ctypes.data_as(ctypes.c_void_p), 3, P, 0.9, 0, 0, 0,
                             host_vol.ctypes.data_as(ctypes.c_void_p), 1, _ctg.PROGRESS_FN(), None)
    _rejected(ctx, st, "pixel")


def test_null_buffers_and_ranges(setup):
    ctx, g, b = setup
    lib, h = ctx.lib, ctx.handle
    _rejected(ctx, lib.ctg_feldkamp_weight(h, None, P, None), "buffers")
    _rejected(ctx, lib.ctg_filter_rows(h, None, _p(b["img"]), 24, 0, None), "buffers")
    _rejected(ctx, lib.ctg_layout_projections(h, None, P, 0, _p(b["filt32"]), _p(b["scales"]), None), "buffers")
    _rejected(ctx, lib.ctg_fdk_backward(h, None, 0, 8, 0, P, 0, _p(b["proj"]), None), "buffers")
    _rejected(ctx, lib.ctg_fdk_f64(h, None, 1, P, 0.9, 0, _p(b["vol64"]), 0, 8, None), "buffers")
    _rejected(ctx, lib.ctg_backproject(h, _p(b["filt32"]), 0, _p(b["scales"]), 2, P, _p(b["vol"]), 0, 8, 0,
                                       None), "projection range")
    _rejected(ctx, lib.ctg_forward_project(h, _p(b["vol"]), 4, 8, 0, P, _p(b["proj"]), 0, None), "slab")
    _rejected(ctx, lib.ctg_convert_volume(h, _p(b["vol"]), 9, 1, 1, _p(b["out"]), None), "exceeds")
    # an error leaves the device usable: a valid call afterwards succeeds
    assert lib.ctg_backproject(h, _p(b["filt32"]), 0, _p(b["scales"]), 0, P, _p(b["vol"]), 0, 8, 0, None) == 0


def test_empty_inputs_are_no_ops(setup):
    import torch
    ctx, g, b = setup
    lib, h = ctx.lib, ctx.handle
    # zero projections: an overwrite leaves a zero slab, an accumulation leaves it as it was
    vol = torch.full((16, 16, 8), 3.0, dtype=torch.float32, device="cuda")
    assert lib.ctg_backproject(h, None, 1, None, 0, 0, _p(vol), 0, 8, 1, None) == 0
    torch.cuda.synchronize()
    assert bool((vol == 3.0).all())
    assert lib.ctg_backproject(h, None, 0, None, 0, 0, _p(vol), 0, 8, 0, None) == 0
    torch.cuda.synchronize()
    assert bool((vol == 0.0).all())
    vol64 = torch.full((16, 16, 8), 5.0, dtype=torch.float64, device="cuda")
    assert lib.ctg_fdk_f64(h, None, 1, 0, 0.9, 0, _p(vol64), 0, 8, None) == 0
    torch.cuda.synchronize()
    assert bool((vol64 == 0.0).all())
    # zero projections, rows or slices elsewhere: nothing is written
    proj = torch.full((P, 24, 32), 7.0, dtype=torch.float32, device="cuda")
    assert lib.ctg_forward_project(h, _p(b["vol"]), 0, 8, 0, 0, None, 0, None) == 0
    assert lib.ctg_forward_project(h, _p(b["vol"]), 0, 0, 0, P, _p(proj), 1, None) == 0
    assert lib.ctg_fdk_backward(h, None, 0, 8, 0, 0, 0, None, None) == 0
    assert lib.ctg_preprocess(h, None, 1, 0, 0.9, 0, 0, 1, None, None, None) == 0
    assert lib.ctg_layout_projections(h, None, 0, 1, None, None, None) == 0
    assert lib.ctg_feldkamp_weight(h, None, 0, None) == 0
    assert lib.ctg_filter_rows(h, None, None, 0, 0, None) == 0
    assert lib.ctg_convert_volume(h, _p(b["vol"]), 0, 0, 1, _p(b["out"]), None) == 0
    torch.cuda.synchronize()
    assert bool((proj == 7.0).all())


def test_row_filter_wider_than_the_lds_buffer_is_unsupported(gpu):
    """W = 8193 would need a padded transform of 32768 points (256 KB of LDS): the call
    returns CTG_ERROR_UNSUPPORTED before any launch.  8191 and 8192, the widest rows that
    fit, are checked against the oracle with the other row-filter widths."""
    import torch
    import ct_amd
    from ct_amd import _ctg
    W, H = 8193, 2
    g, _ = ct_amd.geometry_ctvolume(W, H, 3.0 * W, 0.0, None)
    with ct_amd.Context(0) as ctx:
        ctx.set_geometry(g)
        ctx.set_projections([0.0])
        raw = torch.full((1, H, W), 30000, dtype=torch.uint16, device=gpu)
        filtered = torch.empty(ctx.filtered_bytes("f32", 1), dtype=torch.uint8, device=gpu)
        st = ctx.lib.ctg_preprocess(ctx.handle, _p(raw), 1, 1, 0.9, 0, 0, 0, _p(filtered), None, None)
        assert st == _ctg.ERROR_UNSUPPORTED, (st, _ctg.last_error())
        assert "LDS" in _ctg.last_error()


def test_detector_beyond_the_staging_limit_is_unsupported(gpu):
    """The backprojection stages a projection through a buffer resource with a 32-bit
    size: 8192 x 16384 FP32 entries (16 B each, 2^31 bytes in one projection) return
    CTG_ERROR_UNSUPPORTED before any launch and leave the slab untouched."""
    import torch
    import ct_amd
    from ct_amd import _ctg
    W, H = 8192, 16384
    g = ct_amd.geometry_grid(W, H, 3.0 * W, 0.0, 16, 16, 8)
    with ct_amd.Context(0) as ctx:
        ctx.set_geometry(g)
        ctx.set_projections([0.0])
        # the size check comes before any read of the entries: a small buffer stands in
        filtered = torch.zeros(4096, dtype=torch.uint8, device=gpu)
        scales = torch.ones(1, dtype=torch.float32, device=gpu)
        vol = torch.zeros((16, 16, 8), dtype=torch.float32, device=gpu)
        st = ctx.lib.ctg_backproject(ctx.handle, _p(filtered), 0, _p(scales), 0, 1, _p(vol), 0, 8, 0, None)
        assert st == _ctg.ERROR_UNSUPPORTED, (st, _ctg.last_error())
        assert "2^31" in _ctg.last_error()
        torch.cuda.synchronize()
        assert bool((vol == 0.0).all())
