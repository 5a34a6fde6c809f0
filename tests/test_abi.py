"""C-ABI checks that need no GPU: the library loads, exports every entry point that
include/ctg.h declares, and its host-side logic (geometry restatement, argument
validation, error reporting) behaves like the reference's."""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "ctg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ctg_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(ctg_lib):
    from ct_amd import _ctg
    names = _declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(ctg_lib, name), name
        assert name in _ctg.SIGNATURES, f"{name} missing from the ctypes binding"


def test_library_is_hip_code_object_for_gfx950():
    data = open(os.path.join(ROOT, "ct_amd", "libctg.so"), "rb").read()
    assert b"gfx950" in data


@pytest.mark.parametrize("W,H,fcd,uoff,bounds", [
    (256, 192, 1250.0, 0.0, None),
    (1024, 1024, 2500.0, 0.0, [0.25, 0.75, 0.25, 0.75, 0.1, 0.9]),
    (2048, 1536, 5000.0, 3.5, None),
    (90, 45, 439.5, -1.25, [0.1, 0.95, 0.0, 0.8, 0.3, 0.6]),
    (64, 64, 66.0, 0.0, [0.6, 0.9, 0.1, 0.3, 0.0, 0.5]),
])
def test_geometry_matches_oracle(oracle, W, H, fcd, uoff, bounds):
    import ct_amd
    g, roi = ct_amd.geometry_ctvolume(W, H, fcd, uoff, bounds)
    rc, og, oroi, _ = oracle.volume_geometry(W, H, fcd, uoff, bounds)
    assert rc == 0
    assert roi == oroi
    assert g.as_dict() == og.as_dict()


def test_grid_geometry_matches_oracle(oracle):
    import ct_amd
    g = ct_amd.geometry_grid(1024, 1024, 2500.0, 0.0, 1024, 1024, 1024)
    og = oracle.grid_geometry(1024, 1024, 2500.0, 0.0, 1024, 1024, 1024)
    assert g.as_dict() == og.as_dict()
    assert g.radius == 501.0 and g.x_pre == 512.0


def test_geometry_rejections():
    import ct_amd
    from ct_amd import _ctg
    # imageWidth >= FCD is rejected by the loader (Ct/CtVolume.cpp:294-301)
    with pytest.raises(_ctg.CtgError) as e:
        ct_amd.geometry_ctvolume(256, 192, 256.0)
    assert e.value.status == _ctg.ERROR_INVALID_ARGUMENT
    assert "FCD" in str(e.value)
    # bounds whose clamp makes yTo < yFrom (the reference would wrap the unsigned size)
    with pytest.raises(_ctg.CtgError):
        ct_amd.geometry_ctvolume(256, 192, 1250.0, 0.0, [0.1, 0.9, 0.8, 0.2, 0.0, 1.0])


def test_filtered_bytes(ctg_lib):
    import ct_amd
    from ct_amd import _ctg
    g = ct_amd.geometry_grid(1024, 1024, 2500.0, 0.0, 8, 8, 8)
    f = ctg_lib.ctg_filtered_bytes
    assert f(ctypes.byref(g), _ctg.STORE["f32"], 720) == 1024 * 1024 * 16 * 720
    assert f(ctypes.byref(g), _ctg.STORE["f16"], 3) == 1024 * 1024 * 8 * 3
    assert f(ctypes.byref(g), _ctg.STORE["bf16"], 0) == 0
    assert f(ctypes.byref(g), 3, 4) == 0 and f(ctypes.byref(g), -1, 4) == 0
    assert f(ctypes.byref(g), _ctg.STORE["f32"], -1) == 0
    assert f(None, _ctg.STORE["f32"], 4) == 0


def test_no_device_is_reported_not_faked():
    """Without a GPU the library must refuse to run, never fall back to the CPU."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import ct_amd
    from ct_amd import _ctg
    assert _ctg.device_count() == 0
    with pytest.raises(_ctg.CtgError) as e:
        ct_amd.Context(0)
    assert e.value.status == _ctg.ERROR_NO_DEVICE
    assert _ctg.last_error()


def test_version(ctg_lib):
    assert b"gfx950" in ctg_lib.ctg_version()


def test_autograd_harness_has_no_cpu_path():
    """The torch.autograd operators refuse to run without a HIP device."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import ct_amd
    from ct_amd.autograd import ConeBeamOperator
    g = ct_amd.geometry_grid(32, 32, 160.0, 0.0, 16, 16, 16)
    with pytest.raises(RuntimeError, match="no CPU path"):
        ConeBeamOperator(g, [0.0, 90.0])
