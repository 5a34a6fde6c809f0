"""ctypes binding of libctg.so (the C ABI declared in include/ctg.h).

Device buffers are passed as integer device addresses (``torch.Tensor.data_ptr()``
of a tensor on the context's device) and streams as integer ``hipStream_t`` handles
(``torch.cuda.current_stream().cuda_stream``).  Every non-OK status raises
:class:`CtgError` carrying ``ctg_last_error()``.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, c_char, c_char_p, c_double, c_float, c_int32, c_int64,
                    c_size_t, c_uint16, c_uint32, c_uint64, c_void_p)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libctg.so")

# ctg_status
OK = 0
ERROR_INVALID_ARGUMENT = 1
ERROR_DEVICE = 2
ERROR_OUT_OF_MEMORY = 3
ERROR_INSUFFICIENT_VRAM = 4
ERROR_INTERRUPTED = 5
ERROR_UNSUPPORTED = 6
ERROR_IO = 7
ERROR_NO_DEVICE = 8

# enums (same values as include/ctg.h)
FILTER = {"ramlak": 0, "shepplogan": 1, "hann": 2}
PIXEL = {"u8": 0, "u16": 1, "f32": 2}
LOG_FORM = {"cpu": 0, "gpu": 1}
STORE = {"f32": 0, "f16": 1, "bf16": 2}
INDEX_ORDER = {"xfastest": 0, "zfastest": 1}
BYTE_ORDER = {"bigendian": 0, "littleendian": 1}


class CtgError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"ctg status {status}: {message}")
        self.status = status


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


class DeviceInfo(Structure):
    _fields_ = [
        ("name", c_char * 256),
        ("total_bytes", c_uint64), ("free_bytes", c_uint64),
        ("compute_units", c_int32), ("memory_bus_width_bits", c_int32),
        ("memory_clock_khz", c_int32), ("clock_khz", c_int32),
    ]


PROGRESS_FN = ctypes.CFUNCTYPE(None, c_double, c_void_p)

# name -> (restype, argtypes); the list of exported entry points of include/ctg.h
SIGNATURES = {
    "ctg_device_count": (c_int32, [POINTER(c_int32)]),
    "ctg_get_device_info": (c_int32, [c_int32, POINTER(DeviceInfo)]),
    "ctg_last_error": (c_char_p, []),
    "ctg_version": (c_char_p, []),
    "ctg_geometry_ctvolume": (c_int32, [c_uint32, c_uint32, c_float, c_float, POINTER(c_float),
                                        POINTER(Geometry), POINTER(c_uint32)]),
    "ctg_geometry_grid": (c_int32, [c_uint32, c_uint32, c_float, c_float, c_int32, c_int32, c_int32,
                                    POINTER(Geometry)]),
    "ctg_create": (c_int32, [c_int32, POINTER(c_void_p)]),
    "ctg_destroy": (c_int32, [c_void_p]),
    "ctg_set_geometry": (c_int32, [c_void_p, POINTER(Geometry)]),
    "ctg_set_projections": (c_int32, [c_void_p, POINTER(c_double), POINTER(c_double), c_int32]),
    "ctg_cancel": (None, [c_void_p]),
    "ctg_reset_cancel": (None, [c_void_p]),
    "ctg_set_spare_memory": (c_int32, [c_void_p, c_uint64]),
    "ctg_set_memory_budget": (c_int32, [c_void_p, c_uint64]),
    "ctg_plan_reconstruct": (c_int32, [c_void_p, c_int32, c_int32, c_int32, POINTER(c_int32), POINTER(c_int32)]),
    "ctg_filtered_bytes": (c_size_t, [POINTER(Geometry), c_int32, c_int32]),
    "ctg_preprocess": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_float, c_int32, c_int32, c_int32,
                                 c_void_p, c_void_p, c_void_p]),
    "ctg_feldkamp_weight": (c_int32, [c_void_p, c_void_p, c_int32, c_void_p]),
    "ctg_filter_rows": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_int32, c_void_p]),
    "ctg_layout_projections": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_void_p]),
    "ctg_layout_projections_noalloc": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_void_p, c_void_p,
                                                 c_void_p]),
    "ctg_backproject": (c_int32, [c_void_p, c_void_p, c_int32, c_void_p, c_int32, c_int32, c_void_p, c_int32,
                                  c_int32, c_int32, c_void_p]),
    "ctg_forward_project": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_int32, c_int32, c_void_p, c_int32,
                                      c_void_p]),
    "ctg_fdk_backward": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_int32, c_int32, c_int32, c_void_p,
                                   c_void_p]),
    "ctg_fdk_f64": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_float, c_int32, c_void_p, c_int32, c_int32,
                              c_void_p]),
    "ctg_convert_volume": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p]),
    "ctg_reconstruct": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_float, c_int32, c_int32, c_int32,
                                  c_void_p, c_int32, PROGRESS_FN, c_void_p]),
    "ctg_phantom_images": (c_int32, [c_void_p, POINTER(c_double), c_int32, c_int32, c_int32, c_double, c_double,
                                     c_uint64, c_void_p, c_void_p, c_void_p]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load libctg.so; raises OSError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} not found: build it with `make -C ct_amd/csrc` or __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int):
    if status != OK:
        msg = load().ctg_last_error()
        raise CtgError(status, msg.decode() if msg else "")


def last_error() -> str:
    return load().ctg_last_error().decode()


def _floats(seq, n):
    arr = (c_float * n)(*seq)
    return arr


def geometry_ctvolume(width, height, fcd_px, u_offset_px=0.0, bounds=None):
    """CtVolume semantics (radius, 2R x 2R x H grid, ROI).  Returns (Geometry, roi tuple)."""
    lib = load()
    g = Geometry()
    roi = (c_uint32 * 6)()
    b = _floats(bounds, 6) if bounds is not None else None
    check(lib.ctg_geometry_ctvolume(width, height, fcd_px, u_offset_px, b, ctypes.byref(g), roi))
    return g, tuple(roi)


def geometry_grid(width, height, fcd_px, u_offset_px, nx, ny, nz):
    lib = load()
    g = Geometry()
    check(lib.ctg_geometry_grid(width, height, fcd_px, u_offset_px, nx, ny, nz, ctypes.byref(g)))
    return g


def device_count() -> int:
    n = c_int32(0)
    check(load().ctg_device_count(ctypes.byref(n)))
    return n.value


def device_info(device: int = 0) -> DeviceInfo:
    info = DeviceInfo()
    check(load().ctg_get_device_info(device, ctypes.byref(info)))
    return info


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return c_void_p(x)
    return c_void_p(x.data_ptr())


class Context:
    """Owns a ctg_context on one device."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = c_void_p()
        check(self.lib.ctg_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device
        self.geometry = None

    def close(self):
        if self.handle:
            self.lib.ctg_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_geometry(self, g: Geometry):
        check(self.lib.ctg_set_geometry(self.handle, ctypes.byref(g)))
        self.geometry = g

    def set_projections(self, angles_deg, heights_px=None):
        n = len(angles_deg)
        a = (c_double * n)(*[float(v) for v in angles_deg])
        h = (c_double * n)(*[float(v) for v in heights_px]) if heights_px is not None else None
        check(self.lib.ctg_set_projections(self.handle, a, h, n))

    def cancel(self):
        self.lib.ctg_cancel(self.handle)

    def reset_cancel(self):
        self.lib.ctg_reset_cancel(self.handle)

    def set_spare_memory(self, mib: int):
        check(self.lib.ctg_set_spare_memory(self.handle, int(mib)))

    def set_memory_budget(self, nbytes: int):
        check(self.lib.ctg_set_memory_budget(self.handle, int(nbytes)))

    def plan_reconstruct(self, pixel, nproj, store):
        """(resident, slab_nz) that reconstruct() would use under the current memory limits."""
        r, z = c_int32(), c_int32()
        check(self.lib.ctg_plan_reconstruct(self.handle, PIXEL[pixel], nproj, STORE[store], ctypes.byref(r),
                                            ctypes.byref(z)))
        return bool(r.value), z.value

    def filtered_bytes(self, store: str, nproj: int) -> int:
        return self.lib.ctg_filtered_bytes(ctypes.byref(self.geometry), STORE[store], nproj)

    def preprocess(self, raw, pixel, nproj, base, log_form, filt, store, filtered, scales, stream=0):
        check(self.lib.ctg_preprocess(self.handle, _ptr(raw), PIXEL[pixel], nproj, float(base),
                                      LOG_FORM[log_form], FILTER[filt], STORE[store], _ptr(filtered),
                                      _ptr(scales), c_void_p(stream)))

    def feldkamp_weight(self, images, nproj, stream=0):
        check(self.lib.ctg_feldkamp_weight(self.handle, _ptr(images), nproj, c_void_p(stream)))

    def filter_rows(self, src, dst, rows, filt, stream=0):
        check(self.lib.ctg_filter_rows(self.handle, _ptr(src), _ptr(dst), int(rows), FILTER[filt],
                                       c_void_p(stream)))

    def layout_projections(self, images, nproj, store, filtered, scales, stream=0):
        check(self.lib.ctg_layout_projections(self.handle, _ptr(images), nproj, STORE[store], _ptr(filtered),
                                              _ptr(scales), c_void_p(stream)))

    def backproject(self, filtered, store, scales, p0, nproj, volume, z0, nz, accumulate=False, stream=0):
        check(self.lib.ctg_backproject(self.handle, _ptr(filtered), STORE[store], _ptr(scales), p0, nproj,
                                       _ptr(volume), z0, nz, 1 if accumulate else 0, c_void_p(stream)))

    def forward_project(self, volume, z0, nz, p0, nproj, proj, accumulate=False, stream=0):
        check(self.lib.ctg_forward_project(self.handle, _ptr(volume), z0, nz, p0, nproj, _ptr(proj),
                                           1 if accumulate else 0, c_void_p(stream)))

    def fdk_backward(self, grad_volume, z0, nz, p0, nproj, filt, grad_lineint, stream=0):
        check(self.lib.ctg_fdk_backward(self.handle, _ptr(grad_volume), z0, nz, p0, nproj, FILTER[filt],
                                        _ptr(grad_lineint), c_void_p(stream)))

    def fdk_f64(self, raw, pixel, nproj, base, filt, volume, z0, nz, stream=0):
        check(self.lib.ctg_fdk_f64(self.handle, _ptr(raw), PIXEL[pixel], nproj, float(base), FILTER[filt],
                                   _ptr(volume), z0, nz, c_void_p(stream)))

    def convert_volume(self, volume, nz, order, byte_order, out, stream=0):
        check(self.lib.ctg_convert_volume(self.handle, _ptr(volume), nz, INDEX_ORDER[order],
                                          BYTE_ORDER[byte_order], _ptr(out), c_void_p(stream)))

    def reconstruct(self, h_raw_ptr: int, pixel, nproj, base, log_form, filt, store, h_volume_ptr: int,
                    order="zfastest", progress=None):
        cb = PROGRESS_FN(progress) if progress is not None else PROGRESS_FN()
        check(self.lib.ctg_reconstruct(self.handle, c_void_p(h_raw_ptr), PIXEL[pixel], nproj, float(base),
                                       LOG_FORM[log_form], FILTER[filt], STORE[store], c_void_p(h_volume_ptr),
                                       INDEX_ORDER[order], cb, None))

    def phantom_images(self, ellipsoids, p0, nproj, photons, full_scale, seed, images=None, lineint=None,
                       stream=0):
        flat = [float(v) for e in ellipsoids for v in e]
        n = len(ellipsoids)
        arr = (c_double * max(len(flat), 1))(*flat)
        check(self.lib.ctg_phantom_images(self.handle, arr, n, p0, nproj, float(photons), float(full_scale),
                                          int(seed) & 0xFFFFFFFFFFFFFFFF, _ptr(images), _ptr(lineint),
                                          c_void_p(stream)))
