"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker or as the timed CPU baseline.  The product never
imports it.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_double, c_float, c_int, c_int32, c_int64, c_size_t, c_uint32, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

DEPTH = {"u8": 0, "u16": 1, "f32": 2}
FILTER = {"ramlak": 0, "shepplogan": 1, "hann": 2}


class Geometry(Structure):
    _fields_ = [
        ("det_width", c_int32), ("det_height", c_int32),
        ("nx", c_int32), ("ny", c_int32), ("nz", c_int32),
        ("fcd", c_float), ("u_offset", c_float), ("radius", c_float),
        ("x_pre", c_float), ("y_pre", c_float), ("z_pre", c_float),
        ("u_pre", c_float), ("v_pre", c_float),
        ("u_lo", c_float), ("u_hi", c_float), ("v_lo", c_float), ("v_hi", c_float),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None
_pinned = None
_D = POINTER(c_double)
_F = POINTER(c_float)


def use(path=None):
    """Routes the functions below to another build of oracle.c at `path` (bench.py's
    timing build, -O3 -march=native); None returns to the pinned build."""
    global _lib
    if path is None:
        _lib = _pinned
        load()
    else:
        _lib = _configure(ctypes.CDLL(path))
    return _lib


def load():
    global _lib, _pinned
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} missing: run `make -C oracle`")
        _lib = _pinned = _configure(ctypes.CDLL(LIB_PATH))
    return _lib


def _configure(lib):
    sig = {
        "orc_cylinder_radius": (c_uint32, [c_uint32, c_float]),
        "orc_volume_geometry": (c_int, [c_uint32, c_uint32, c_float, c_float, _F, POINTER(Geometry),
                                        POINTER(c_uint32), _F]),
        "orc_grid_geometry": (None, [c_uint32, c_uint32, c_float, c_float, c_int32, c_int32, c_int32,
                                     POINTER(Geometry)]),
        "orc_log_cpu": (None, [c_void_p, c_int, c_size_t, c_double, _F]),
        "orc_log_gpu": (None, [c_void_p, c_int, c_size_t, c_double, _F]),
        "orc_log_cpu_f64": (None, [c_void_p, c_int, c_size_t, c_double, _D]),
        "orc_feldkamp": (None, [_F, c_int32, c_int32, c_int32, c_float, c_float, c_float]),
        "orc_feldkamp_f64": (None, [_D, c_int32, c_int32, c_int32, c_double, c_double, c_double]),
        "orc_filter_window": (None, [c_int, c_int32, _F]),
        "orc_filter_rows": (None, [_F, c_int32, c_int64, c_int]),
        "orc_filter_rows_f64": (None, [_D, c_int32, c_int64, c_int]),
        "orc_filter_rows_direct_f64": (None, [_D, c_int32, c_int64, c_int]),
        "orc_backproject_f32": (None, [POINTER(Geometry), _F, _D, _D, c_int32, c_int32, _F]),
        "orc_backproject_f64": (None, [POINTER(Geometry), _D, _D, _D, c_int32, _D]),
        "orc_forward_project_f64": (None, [POINTER(Geometry), _D, _D, _D, c_int32, _D]),
        "orc_fdk_f32": (None, [POINTER(Geometry), c_void_p, c_int, c_double, c_int, c_int, _D, _D, c_int32, _F]),
        "orc_fdk_f64": (None, [POINTER(Geometry), c_void_p, c_int, c_double, c_int, _D, _D, c_int32, _D]),
        "orc_preprocess_f32": (None, [POINTER(Geometry), c_void_p, c_int, c_double, c_int, c_int, c_int32, _F]),
        "orc_phantom_project": (None, [POINTER(Geometry), _D, c_int32, _D, _D, c_int32, _D]),
        "orc_set_threads": (None, [c_int]),
        "orc_get_threads": (c_int, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def _p(a, ctype):
    return a.ctypes.data_as(POINTER(ctype))


def geometry_from(g) -> Geometry:
    """Copy any structure with the ctg_geometry fields into an oracle Geometry."""
    out = Geometry()
    for name, _ in Geometry._fields_:
        setattr(out, name, getattr(g, name))
    return out


def volume_geometry(width, height, fcd, u_offset=0.0, bounds=None):
    g = Geometry()
    roi = (c_uint32 * 6)()
    eff = (c_float * 6)()
    b = (c_float * 6)(*bounds) if bounds is not None else None
    rc = load().orc_volume_geometry(width, height, fcd, u_offset, b, ctypes.byref(g), roi, eff)
    return rc, g, tuple(roi), tuple(eff)


def grid_geometry(width, height, fcd, u_offset, nx, ny, nz):
    g = Geometry()
    load().orc_grid_geometry(width, height, fcd, u_offset, nx, ny, nz, ctypes.byref(g))
    return g


def cylinder_radius(width, fcd):
    return load().orc_cylinder_radius(width, fcd)


def log_cpu(img: np.ndarray, base: float, depth: str) -> np.ndarray:
    img = np.ascontiguousarray(img)
    out = np.empty(img.size, np.float32)
    load().orc_log_cpu(img.ctypes.data, DEPTH[depth], img.size, base, _p(out, c_float))
    return out.reshape(img.shape)


def log_gpu(img: np.ndarray, base: float, depth: str) -> np.ndarray:
    img = np.ascontiguousarray(img)
    out = np.empty(img.size, np.float32)
    load().orc_log_gpu(img.ctypes.data, DEPTH[depth], img.size, base, _p(out, c_float))
    return out.reshape(img.shape)


def feldkamp(img: np.ndarray, fcd, u_pre, v_pre) -> np.ndarray:
    out = np.ascontiguousarray(img, dtype=np.float32).copy()
    n, H, W = out.shape
    load().orc_feldkamp(_p(out, c_float), W, H, n, fcd, u_pre, v_pre)
    return out


def filter_window(filt: str, W: int) -> np.ndarray:
    out = np.empty(W // 2 + 1, np.float32)
    load().orc_filter_window(FILTER[filt], W, _p(out, c_float))
    return out


def filter_rows(rows: np.ndarray, filt: str) -> np.ndarray:
    out = np.ascontiguousarray(rows, dtype=np.float32).copy()
    W = out.shape[-1]
    load().orc_filter_rows(_p(out, c_float), W, out.size // W, FILTER[filt])
    return out


def filter_rows_f64(rows: np.ndarray, filt: str, direct=False) -> np.ndarray:
    out = np.ascontiguousarray(rows, dtype=np.float64).copy()
    W = out.shape[-1]
    fn = load().orc_filter_rows_direct_f64 if direct else load().orc_filter_rows_f64
    fn(_p(out, c_double), W, out.size // W, FILTER[filt])
    return out


def backproject_f32(g: Geometry, filtered: np.ndarray, angles, heights=None, step=1, vol=None):
    filtered = np.ascontiguousarray(filtered, dtype=np.float32)
    P = filtered.shape[0]
    a = np.ascontiguousarray(angles, dtype=np.float64)
    h = np.zeros(P) if heights is None else np.ascontiguousarray(heights, dtype=np.float64)
    if vol is None:
        vol = np.zeros((g.nx, g.ny, g.nz), np.float32)
    load().orc_backproject_f32(ctypes.byref(g), _p(filtered, c_float), _p(a, c_double), _p(h, c_double), P, step,
                               _p(vol, c_float))
    return vol


def backproject_f64(g: Geometry, filtered: np.ndarray, angles, heights=None):
    filtered = np.ascontiguousarray(filtered, dtype=np.float64)
    P = filtered.shape[0]
    a = np.ascontiguousarray(angles, dtype=np.float64)
    h = np.zeros(P) if heights is None else np.ascontiguousarray(heights, dtype=np.float64)
    vol = np.zeros((g.nx, g.ny, g.nz), np.float64)
    load().orc_backproject_f64(ctypes.byref(g), _p(filtered, c_double), _p(a, c_double), _p(h, c_double), P,
                               _p(vol, c_double))
    return vol


def forward_project_f64(g: Geometry, vol: np.ndarray, angles, heights=None):
    vol = np.ascontiguousarray(vol, dtype=np.float64)
    P = len(angles)
    a = np.ascontiguousarray(angles, dtype=np.float64)
    h = np.zeros(P) if heights is None else np.ascontiguousarray(heights, dtype=np.float64)
    proj = np.zeros((P, g.det_height, g.det_width), np.float64)
    load().orc_forward_project_f64(ctypes.byref(g), _p(vol, c_double), _p(a, c_double), _p(h, c_double), P,
                                   _p(proj, c_double))
    return proj


def preprocess_f32(g: Geometry, raw: np.ndarray, depth: str, base: float, filt: str, log_form="cpu"):
    raw = np.ascontiguousarray(raw)
    P = raw.shape[0]
    out = np.empty((P, g.det_height, g.det_width), np.float32)
    load().orc_preprocess_f32(ctypes.byref(g), raw.ctypes.data, DEPTH[depth], base, 1 if log_form == "gpu" else 0,
                              FILTER[filt], P, _p(out, c_float))
    return out


def fdk_f32(g: Geometry, raw: np.ndarray, depth: str, base: float, filt: str, angles, heights=None, log_form="cpu"):
    raw = np.ascontiguousarray(raw)
    P = raw.shape[0]
    a = np.ascontiguousarray(angles, dtype=np.float64)
    h = np.zeros(P) if heights is None else np.ascontiguousarray(heights, dtype=np.float64)
    vol = np.zeros((g.nx, g.ny, g.nz), np.float32)
    load().orc_fdk_f32(ctypes.byref(g), raw.ctypes.data, DEPTH[depth], base, 1 if log_form == "gpu" else 0,
                       FILTER[filt], _p(a, c_double), _p(h, c_double), P, _p(vol, c_float))
    return vol


def fdk_f64(g: Geometry, raw: np.ndarray, depth: str, base: float, filt: str, angles, heights=None):
    raw = np.ascontiguousarray(raw)
    P = raw.shape[0]
    a = np.ascontiguousarray(angles, dtype=np.float64)
    h = np.zeros(P) if heights is None else np.ascontiguousarray(heights, dtype=np.float64)
    vol = np.zeros((g.nx, g.ny, g.nz), np.float64)
    load().orc_fdk_f64(ctypes.byref(g), raw.ctypes.data, DEPTH[depth], base, FILTER[filt], _p(a, c_double),
                       _p(h, c_double), P, _p(vol, c_double))
    return vol


def phantom_project(g: Geometry, ellipsoids, angles, heights=None):
    ell = np.ascontiguousarray(np.asarray(ellipsoids, dtype=np.float64).reshape(-1, 10))
    P = len(angles)
    a = np.ascontiguousarray(angles, dtype=np.float64)
    h = np.zeros(P) if heights is None else np.ascontiguousarray(heights, dtype=np.float64)
    out = np.zeros((P, g.det_height, g.det_width), np.float64)
    load().orc_phantom_project(ctypes.byref(g), _p(ell, c_double), ell.shape[0], _p(a, c_double), _p(h, c_double),
                               P, _p(out, c_double))
    return out


def set_threads(n: int):
    load().orc_set_threads(n)


def get_threads() -> int:
    return load().orc_get_threads()
