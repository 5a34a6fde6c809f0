"""CtVolume facade (libctvolume.so) on the CPU: CSV loading semantics of
CtVolume::sinogramFromImages (Ct/CtVolume.cpp:136-330), setVolumeBounds, the writer's
info files, and the image decoders (pinned against PIL as an independent encoder)."""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

from scan_io import write_image, write_pgm16, write_scan

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ct_amd", "libctvolume.so")


def load_ctv():
    """ctypes handle of libctvolume.so with the C wrapper's signatures."""
    import ct_amd
    ct_amd.load()
    lib = ctypes.CDLL(LIB)
    lib.ctv_create.restype = ctypes.c_void_p
    for name in ("ctv_destroy", "ctv_reconstruct", "ctv_stop"):
        getattr(lib, name).argtypes = [ctypes.c_void_p]
    lib.ctv_load.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    lib.ctv_set_bounds.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]
    lib.ctv_set_filter.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.ctv_set_use_cuda.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.ctv_set_use_cpu_preprocessing.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.ctv_set_skip_projections.argtypes = [ctypes.c_void_p, ctypes.c_uint]
    lib.ctv_save.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    lib.ctv_sizes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_ulonglong)]
    lib.ctv_params.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]
    lib.ctv_status.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
    lib.ctv_volume.restype = ctypes.POINTER(ctypes.c_float)
    lib.ctv_volume.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_ulonglong)]
    lib.ctv_cross_section.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, ctypes.c_void_p,
                                      ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    lib.ctv_decode_image.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_int), ctypes.c_void_p, ctypes.c_ulonglong,
                                     ctypes.c_char_p, ctypes.c_int]
    return lib


@pytest.fixture(scope="module")
def ctv():
    return load_ctv()


class Facade:
    def __init__(self, lib):
        self.lib = lib
        self.h = ctypes.c_void_p(lib.ctv_create())

    def __del__(self):
        self.lib.ctv_destroy(self.h)

    def load(self, csv):
        return bool(self.lib.ctv_load(self.h, csv.encode()))

    def sizes(self):
        out = (ctypes.c_ulonglong * 7)()
        self.lib.ctv_sizes(self.h, out)
        return tuple(out)

    def params(self):
        out = (ctypes.c_float * 4)()
        self.lib.ctv_params(self.h, out)
        return tuple(out)

    def status(self):
        buf = ctypes.create_string_buffer(1024)
        code = self.lib.ctv_status(self.h, buf, 1024)
        return code, buf.value.decode()

    def bounds(self, b):
        self.lib.ctv_set_bounds(self.h, (ctypes.c_float * 6)(*b))

    def volume(self):
        n = ctypes.c_ulonglong()
        p = self.lib.ctv_volume(self.h, ctypes.byref(n))
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0, np.float32)

    def cross_section(self, axis, index):
        """getVolumeCrossSection(axis 0/1/2 = X/Y/Z, index): None when there is no
        volume; raises IndexError for an index outside the axis."""
        rows, cols = ctypes.c_int(), ctypes.c_int()
        code = self.lib.ctv_cross_section(self.h, axis, index, None, 0, ctypes.byref(rows), ctypes.byref(cols))
        if code < 0:
            raise IndexError("Index out of bounds.")
        if code == 0:
            return None
        out = np.empty((rows.value, cols.value), np.float32)
        self.lib.ctv_cross_section(self.h, axis, index, out.ctypes.data, out.size, ctypes.byref(rows),
                                   ctypes.byref(cols))
        return out


def decode(lib, path):
    w, h, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    err = ctypes.create_string_buffer(256)
    if not lib.ctv_decode_image(path.encode(), ctypes.byref(w), ctypes.byref(h), ctypes.byref(d), None, 0, err, 256):
        raise ValueError(err.value.decode())
    dt = {0: np.uint8, 1: np.uint16, 2: np.float32}[d.value]
    out = np.empty((h.value, w.value), dt)
    lib.ctv_decode_image(path.encode(), ctypes.byref(w), ctypes.byref(h), ctypes.byref(d), out.ctypes.data,
                         out.nbytes, err, 256)
    return out


@pytest.mark.parametrize("fmt,compression,dtype", [
    ("pgm", None, np.uint16),
    ("tif", None, np.uint16), ("tif", "tiff_lzw", np.uint16), ("tif", "tiff_adobe_deflate", np.uint16),
    ("tif", "packbits", np.uint16), ("tif", None, np.float32), ("tif", "tiff_lzw", np.uint8),
    ("png", None, np.uint16), ("png", None, np.uint8),
])
def test_decoders_against_pil(ctv, tmp_path, fmt, compression, dtype):
    rng = np.random.default_rng(7)
    h, w = 37, 53
    if dtype == np.float32:
        img = rng.standard_normal((h, w)).astype(np.float32)
    else:
        hi = 65535 if dtype == np.uint16 else 255
        img = rng.integers(0, hi + 1, (h, w)).astype(dtype)
        img[5:15, 10:30] = hi // 3  # runs for the run-length and dictionary coders
    path = str(tmp_path / f"im.{fmt}")
    write_image(path, img, fmt, compression)
    got = decode(ctv, path)
    assert got.dtype == img.dtype and got.shape == img.shape
    assert np.array_equal(got, img)


def test_decoder_rejects_unsupported(ctv, tmp_path):
    p = tmp_path / "x.bin"
    p.write_bytes(b"not an image at all")
    with pytest.raises(ValueError):
        decode(ctv, str(p))


def _scan(tmp_path, **kw):
    from helpers import small_scan
    s = small_scan(W=48, H=32, P=12, seed=3)
    args = dict(pixel_size=0.5, fcd_px=s["fcd"])
    args.update(kw)
    csv = write_scan(str(tmp_path), s["raw"], s["angles"], **args)
    return s, csv


def test_load_matches_ctvolume_geometry(ctv, oracle, tmp_path):
    s, csv = _scan(tmp_path, u_offset_px=1.5, heights_px=np.linspace(-1, 1, 12), height_shift_mm=7.0)
    f = Facade(ctv)
    assert f.load(csv)
    assert f.status()[0] == 0
    nx, ny, nz, W, H, R, n = f.sizes()
    fcd, pix, uoff, base = f.params()
    assert (W, H, n) == (48, 32, 12)
    assert pix == pytest.approx(0.5) and fcd == pytest.approx(s["fcd"]) and uoff == pytest.approx(1.5)
    rc, g, _, _ = oracle.volume_geometry(W, H, fcd, uoff)
    assert (nx, ny, nz, R) == (g.nx, g.ny, g.nz, int(g.radius))
    # ROI after loading, with the reference's clamp of yTo against xFrom
    f.bounds([0.5, 0.9, 0.1, 0.3, 0.2, 0.8])
    rc, g, roi, _ = oracle.volume_geometry(W, H, fcd, uoff, [0.5, 0.9, 0.1, 0.3, 0.2, 0.8])
    assert f.sizes()[:3] == (g.nx, g.ny, g.nz)
    # identical lower and upper bounds are ignored
    f.bounds([0.3, 0.3, 0.0, 1.0, 0.0, 1.0])
    assert f.sizes()[:3] == (g.nx, g.ny, g.nz)


@pytest.mark.parametrize("edit,message", [
    (lambda lines: lines.__setitem__(1, "abc\tpixel size"), "Could not read the parameters"),
    (lambda lines: lines.__setitem__(5, "0\tbase"), "base intensity given in the configuration file must not be"),
    (lambda lines: lines.__setitem__(slice(7, None), []), "does not contain any images"),
    (lambda lines: lines.__setitem__(8, "proj_0001.pgm\tnot-a-number"), "Could not read the image parameters"),
    (lambda lines: lines.__setitem__(9, "missing.pgm\t10.0"), "Error loading the image"),
    (lambda lines: lines.__setitem__(4, "10\tFCD"), "the FCD must be greater than the image width"),
])
def test_loader_rejections(ctv, tmp_path, edit, message):
    _, csv = _scan(tmp_path)
    lines = open(csv).read().splitlines()
    edit(lines)
    open(csv, "w").write("\n".join(lines) + "\n")
    f = Facade(ctv)
    assert not f.load(csv)
    code, msg = f.status()
    assert code == 2 and message in msg, msg
    assert f.sizes()[6] == 0


def test_loader_rejects_mixed_sizes(ctv, tmp_path):
    _, csv = _scan(tmp_path)
    write_pgm16(os.path.join(os.path.dirname(csv), "images", "proj_0003.pgm"), np.zeros((10, 10), np.uint16))
    f = Facade(ctv)
    assert not f.load(csv)
    assert "dimensions differ" in f.status()[1]


def test_missing_csv(ctv, tmp_path):
    f = Facade(ctv)
    assert not f.load(str(tmp_path / "nope.csv"))
    assert f.status() == (2, "Could not open the config file.")


def test_blank_projection_line_repeats_previous(ctv, tmp_path):
    _, csv = _scan(tmp_path)
    lines = open(csv).read().splitlines()
    lines.insert(9, "")
    open(csv, "w").write("\n".join(lines) + "\n")
    f = Facade(ctv)
    assert f.load(csv)
    assert f.sizes()[6] == 13


def test_image_directory_relative_to_csv(ctv, tmp_path):
    _, csv = _scan(tmp_path)
    cwd = os.getcwd()
    try:
        os.chdir("/")
        f = Facade(ctv)
        assert f.load(os.path.relpath(csv, "/"))
    finally:
        os.chdir(cwd)


def test_reconstruct_and_save_without_volume(ctv, tmp_path):
    import torch
    _, csv = _scan(tmp_path)
    f = Facade(ctv)
    assert f.load(csv)
    if not torch.cuda.is_available():
        ctv.ctv_reconstruct(f.h)
        code, msg = f.status()
        assert code == 2 and "HIP device" in msg
    out = str(tmp_path / "out" / "vol.raw")
    os.makedirs(os.path.dirname(out))
    g = Facade(ctv)
    assert g.load(csv)
    ctv.ctv_save(g.h, out.encode(), 1, 1)
    code, msg = g.status()
    assert code == 2 and "appears to be empty" in msg
    txt = open(str(tmp_path / "out" / "vol.txt")).read()
    nx, ny, nz = g.sizes()[:3]
    assert f"X size:\t\t\t\t{nx}\n" in txt and f"Z size:\t\t\t\t{nz}\n" in txt
    assert txt.endswith("Index order:\t\tZ fastest")
    vgi = open(str(tmp_path / "out" / "vol.vgi")).read()
    assert f"size = {nx} {ny} {nz}\n" in vgi and "Name = ./vol.raw\n" in vgi and vgi.endswith("volume = volume1\n")
    assert not os.path.exists(out)


def test_cpu_path_is_not_provided(ctv, tmp_path):
    _, csv = _scan(tmp_path)
    f = Facade(ctv)
    assert f.load(csv)
    ctv.ctv_set_use_cuda(f.h, 0)
    ctv.ctv_reconstruct(f.h)
    code, msg = f.status()
    assert code == 2 and "CPU path" in msg
    assert f.volume().size == 0
    # no volume: an empty cross section for any axis (Common/Volume.h:578)
    assert f.cross_section(2, 0) is None and f.cross_section(0, 10**6) is None


@pytest.mark.parametrize("shape", [(37, 23, 41), (512, 512, 80)])
def test_raw_writer_orders_and_byte_orders(ctv, tmp_path, shape):
    """The raw stream of Volume::saveToBinaryFile (Common/Volume.h:474-574): Z fastest
    keeps the in-memory [x][y][z] order, X fastest is the transpose [z][y][x]; big endian
    swaps every float.  The 512 x 512 x 80 volume spans several 32 MiB chunks in both
    orders (the chunks are filled by worker threads and written by another thread)."""
    ctv.ctv_write_raw.argtypes = [ctypes.c_void_p, ctypes.c_ulonglong, ctypes.c_ulonglong, ctypes.c_ulonglong,
                                  ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    nx, ny, nz = shape
    vol = np.random.default_rng(5).standard_normal(shape).astype(np.float32)
    for order, expect in ((1, vol), (0, np.transpose(vol, (2, 1, 0)))):
        for bo, dt in ((1, "<f4"), (0, ">f4")):
            path = str(tmp_path / f"v{order}{bo}.raw")
            assert ctv.ctv_write_raw(vol.ctypes.data, nx, ny, nz, path.encode(), order, bo) == 0
            got = np.fromfile(path, dt)
            assert got.size == vol.size
            np.testing.assert_array_equal(got, np.ascontiguousarray(expect).ravel())
            os.remove(path)
    assert ctv.ctv_write_raw(vol.ctypes.data, nx, ny, nz, str(tmp_path / "no" / "dir.raw").encode(), 1, 1) == 1
