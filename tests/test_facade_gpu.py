"""CtVolume facade on the GPU: stacks of mixed image depths under the CPU log form, and
the cross-section that comes with every progress report of a reconstruction
(reconstructionProgress(double, cv::Mat), Ct/CtVolume.cpp:1470-1485)."""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

from helpers import rel_rmse, small_scan
from scan_io import write_image
from test_facade import Facade, load_ctv

pytestmark = pytest.mark.gpu


def _write_mixed_scan(directory, scan, pixel_size, fmts):
    """One image per projection in the format/depth fmts[p % len(fmts)]:
    'u16' (16-bit PGM), 'u8' (8-bit PNG, intensities/257) or 'f32' (32-bit float TIFF
    holding intensities as fractions of full scale).  Returns the CSV path and the
    images as written."""
    os.makedirs(os.path.join(directory, "images"), exist_ok=True)
    lines = ["images\tdir", f"{pixel_size!r}\tpixel size", "ccw\tdirection", "0.0\tu offset",
             f"{float(scan['fcd'] * pixel_size)!r}\tFCD", f"{float(scan['base'])!r}\tbase", "ignored"]
    images = []
    for p in range(scan["P"]):
        kind = fmts[p % len(fmts)]
        raw = scan["raw"][p]
        if kind == "u16":
            img, name, fmt = raw, f"p{p:04d}.pgm", "pgm"
        elif kind == "u8":
            img, name, fmt = np.floor(raw / 257.0 + 0.5).astype(np.uint8), f"p{p:04d}.png", "png"
        else:
            img, name, fmt = (raw.astype(np.float64) / 65535.0 + 1e-3).astype(np.float32), f"p{p:04d}.tif", "tif"
        write_image(os.path.join(directory, "images", name), img, fmt)
        images.append((kind, img))
        lines.append(f"{name}\t{float(scan['angles'][p])!r}")
    csv = os.path.join(directory, "scan.csv")
    with open(csv, "w") as f:
        f.write("\n".join(lines) + "\n")
    return csv, images


def test_mixed_depth_stack_with_zero_pixels_cpu_log_form(gpu, oracle, tmp_path):
    """Integer images keep the 1e-6*B that convertTo32bit adds (Ct/CtVolume.cpp:802-812)
    when the stack mixes depths: zero pixels in 8- and 16-bit images give finite line
    integrals, and the volume matches the oracle run image by image at its own depth."""
    scan = small_scan(W=64, H=48, P=24, seed=41, base=0.9)
    scan["raw"][0, 20:24, 30:34] = 0   # 16-bit image
    scan["raw"][1, 10:12, 8:12] = 0    # 8-bit image
    csv, images = _write_mixed_scan(str(tmp_path), scan, 0.5, ["u16", "u8", "f32"])
    lib = load_ctv()
    f = Facade(lib)
    assert f.load(csv)
    lib.ctv_set_use_cpu_preprocessing(f.h, 1)
    lib.ctv_set_filter(f.h, 0)
    lib.ctv_reconstruct(f.h)
    code, msg = f.status()
    assert code == 0, msg
    nx, ny, nz = f.sizes()[:3]
    got = f.volume().reshape(nx, ny, nz)
    assert np.isfinite(got).all()
    fcd = float(f.params()[0])
    _, g, _, _ = oracle.volume_geometry(scan["W"], scan["H"], fcd, 0.0, None)
    filtered = np.concatenate([oracle.preprocess_f32(g, img[None], kind, scan["base"], "ramlak", "cpu")
                               for kind, img in images])
    assert np.isfinite(filtered).all()
    ref = oracle.backproject_f32(g, filtered, scan["angles"])
    err = rel_rmse(got, ref)
    print(f"mixed-depth stack: rel RMSE vs oracle {err:.3e}")
    assert err < 1e-4


CS_FN = ctypes.CFUNCTYPE(None, ctypes.c_double, ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_int,
                         ctypes.c_void_p)


def test_progress_reports_carry_the_cross_section(gpu, tmp_path):
    """A 68 x 68 x 256 grid is reconstructed in x ranges of 16 columns (the pipelined plan
    for 256 or more slices): the callback runs after each range with the Z cross-section
    (rows = x, columns = y) of the volume finished so far.
    Each report holds the earlier report's rows unchanged plus new rows, and the last
    one equals getVolumeCrossSection after the run."""
    from scan_io import write_scan
    scan = small_scan(W=70, H=256, P=24, seed=42)
    csv = write_scan(str(tmp_path), scan["raw"], scan["angles"], pixel_size=0.5, fcd_px=scan["fcd"])
    lib = load_ctv()
    lib.ctv_on_cross_section_progress.argtypes = [ctypes.c_void_p, CS_FN, ctypes.c_void_p]
    lib.ctv_set_cross_section.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint]
    f = Facade(lib)
    assert f.load(csv)
    nx, ny, nz = f.sizes()[:3]
    assert nz == 256
    # the central slice: the phantom lies within about 30 slices of it, and detector rows
    # outside its shadow filter to exactly zero
    lib.ctv_set_cross_section(f.h, 2, 128)
    reports = []

    def on_progress(pct, data, rows, cols, user):
        img = np.ctypeslib.as_array(data, shape=(rows, cols)).copy() if data else None
        reports.append((pct, img))

    cb = CS_FN(on_progress)
    lib.ctv_on_cross_section_progress(f.h, cb, None)
    lib.ctv_reconstruct(f.h)
    code, msg = f.status()
    assert code == 0, msg
    assert len(reports) >= 2, [r[0] for r in reports]
    assert reports[-1][0] == 100.0
    final = f.cross_section(2, 128)
    assert final.shape == (nx, ny)
    np.testing.assert_array_equal(reports[-1][1], final)
    filled_prev = 0
    for (pct, img), (_, nxt) in zip(reports, reports[1:] + [(None, final)]):
        assert img.shape == (nx, ny)
        filled = int(np.count_nonzero(np.any(img != 0, axis=1)))
        assert filled >= filled_prev
        rows = np.any(img != 0, axis=1)
        np.testing.assert_array_equal(img[rows], nxt[rows])
        filled_prev = filled
    assert np.count_nonzero(np.any(reports[0][1] != 0, axis=1)) < np.count_nonzero(np.any(final != 0, axis=1))
    lib.ctv_on_cross_section_progress(f.h, CS_FN(), None)


def test_images_kept_from_the_load_give_the_same_volume(gpu, tmp_path, monkeypatch):
    """The load keeps its decoded images and the reconstruction takes them (gathered when
    setSkipProjections(1) uses every second projection); with CTG_FACADE_IMAGE_CACHE=0
    every file is decoded again at the reconstruction.  Both give the same volume bit
    for bit, and a second load of another scan replaces the kept images."""
    from scan_io import write_scan
    lib = load_ctv()

    def run(csv, skip):
        f = Facade(lib)
        assert f.load(csv)
        lib.ctv_set_filter(f.h, 1)
        lib.ctv_set_skip_projections(f.h, skip)
        lib.ctv_reconstruct(f.h)
        code, msg = f.status()
        assert code == 0, msg
        return f.volume().copy()

    scans = []
    for i, seed in enumerate((51, 52)):
        scan = small_scan(W=64, H=48, P=30, seed=seed)
        d = str(tmp_path / f"s{i}")
        scans.append(write_scan(d, scan["raw"], scan["angles"], pixel_size=0.5, fcd_px=scan["fcd"]))
    for skip in (0, 1):  # every projection, every second one
        kept = run(scans[0], skip)
        monkeypatch.setenv("CTG_FACADE_IMAGE_CACHE", "0")
        decoded = run(scans[0], skip)
        monkeypatch.delenv("CTG_FACADE_IMAGE_CACHE")
        assert np.isfinite(kept).all() and np.abs(kept).max() > 0
        assert np.array_equal(kept, decoded), skip
    # one facade, two loads: the second scan's images replace the first's
    f = Facade(lib)
    assert f.load(scans[0]) and f.load(scans[1])
    lib.ctv_set_filter(f.h, 1)
    lib.ctv_reconstruct(f.h)
    assert f.status()[0] == 0
    assert np.array_equal(f.volume(), run(scans[1], 0))
