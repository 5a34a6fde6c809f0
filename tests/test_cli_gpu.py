"""End to end on the GPU: Ct's console flags (Ct/main.cpp:106-189) through ct_cli and the
CtVolume facade, checked against the CPU oracle's FDK on the same images, plus the
output files (Ct/CtVolume.cpp:634-721, Common/Volume.h:474-574)."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from helpers import rel_rmse, small_scan
from scan_io import write_scan

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "ct_amd", "ct_cli")


def _run(args, cwd):
    r = subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:])
    return r


def _oracle_volume(oracle, scan, pixel_size, filt, log_form, bounds=None, step=1):
    fcd = np.float32(np.float32(scan["fcd"] * pixel_size) / np.float32(pixel_size))
    _, g, _, _ = oracle.volume_geometry(scan["W"], scan["H"], float(fcd), 0.0, bounds)
    raw = scan["raw"][::step]
    ang = scan["angles"][::step]
    return oracle.fdk_f32(g, raw, "u16", scan["base"], filt, ang, log_form=log_form), g


def test_cli_zfastest_shepplogan_cpu_formulas(gpu, oracle, tmp_path):
    scan = small_scan(W=64, H=48, P=36, seed=31)
    csv = write_scan(str(tmp_path), scan["raw"], scan["angles"], pixel_size=0.5, fcd_px=scan["fcd"])
    # like the reference writer (Ct/CtVolume.cpp:641-645), the output directory is not created
    r = _run(["-i", csv, "-o", "out/vol.raw", "-j", "zfastest"], str(tmp_path))
    assert r.returncode != 0 and "Could not write the info file." in r.stdout
    os.makedirs(str(tmp_path / "out"))
    r = _run(["-i", csv, "-o", "out/vol.raw", "-f", "shepplogan", "-j", "zfastest", "-c"], str(tmp_path))
    assert r.returncode == 0, r.stdout
    ref, g = _oracle_volume(oracle, scan, 0.5, "shepplogan", "cpu")
    got = np.fromfile(str(tmp_path / "out" / "vol.raw"), "<f4").reshape(g.nx, g.ny, g.nz)
    err = rel_rmse(got, ref)
    print(f"CLI FDK rel RMSE {err:.3e}")
    assert err < 1e-4
    txt = open(str(tmp_path / "out" / "vol.txt")).read()
    assert f"X size:\t\t\t\t{g.nx}\nY size:\t\t\t\t{g.ny}\nZ size:\t\t\t\t{g.nz}\n" in txt
    assert "Byte order:\t\t\tLittle endian\n" in txt and txt.endswith("Index order:\t\tZ fastest")
    vgi = open(str(tmp_path / "out" / "vol.vgi")).read()
    lo, hi = [float(v) for v in vgi.split("datarange = ")[1].split("\n")[0].split()]
    assert lo == pytest.approx(float(got.min()), rel=1e-5) and hi == pytest.approx(float(got.max()), rel=1e-5)


def test_cli_default_xfastest_bigendian_and_roi(gpu, oracle, tmp_path):
    scan = small_scan(W=64, H=48, P=30, seed=32)
    csv = write_scan(str(tmp_path), scan["raw"], scan["angles"], pixel_size=0.5, fcd_px=scan["fcd"])
    bounds = [0.2, 0.8, 0.1, 0.9, 0.25, 0.75]
    r = _run(["-i", csv, "-o", "vol.raw", "-e", "bigendian", "--xmin", "0.2", "--xmax", "0.8", "--ymin", "0.1",
              "--ymax", "0.9", "--zmin", "0.25", "--zmax", "0.75"], str(tmp_path))
    assert r.returncode == 0, r.stdout
    ref, g = _oracle_volume(oracle, scan, 0.5, "ramlak", "gpu", bounds)
    got = np.fromfile(str(tmp_path / "vol.raw"), ">f4").astype(np.float32).reshape(g.nz, g.ny, g.nx)
    got = np.transpose(got, (2, 1, 0))
    err = rel_rmse(got, ref)
    print(f"CLI ROI FDK rel RMSE {err:.3e}")
    assert err < 1e-4
    txt = open(str(tmp_path / "vol.txt")).read()
    assert "Byte order:\t\t\tBig endian\n" in txt and txt.endswith("Index order:\t\tX fastest")
    assert "X range relative:\t[0.200 - 0.800]\n" in txt


@pytest.mark.parametrize("fmt,compression", [("tif", "tiff_lzw"), ("png", None)])
def test_cli_tiff_png_inputs_match_pgm(gpu, tmp_path, fmt, compression):
    scan = small_scan(W=48, H=40, P=20, seed=33)
    a = tmp_path / "pgm"
    b = tmp_path / fmt
    csv_a = write_scan(str(a), scan["raw"], scan["angles"], pixel_size=0.5, fcd_px=scan["fcd"])
    csv_b = write_scan(str(b), scan["raw"], scan["angles"], pixel_size=0.5, fcd_px=scan["fcd"], fmt=fmt,
                       compression=compression)
    assert _run(["-i", csv_a, "-o", "v.raw", "-j", "zfastest"], str(a)).returncode == 0
    assert _run(["-i", csv_b, "-o", "v.raw", "-j", "zfastest"], str(b)).returncode == 0
    assert open(str(a / "v.raw"), "rb").read() == open(str(b / "v.raw"), "rb").read()


def test_cli_reports_failures(gpu, tmp_path):
    r = _run(["-i", str(tmp_path / "missing.csv"), "-o", "v.raw"], str(tmp_path))
    assert r.returncode != 0
    assert "Volume was not reconstructed" in r.stdout
    r = _run(["-i", "x"], str(tmp_path))
    assert r.returncode != 0 and "You must provide" in r.stdout
    r = _run(["--bogus"], str(tmp_path))
    assert r.returncode != 0 and "Unknown or misplaced parameter" in r.stdout


def test_facade_skip_projections_and_clockwise(gpu, oracle, tmp_path):
    """setSkipProjections(1) uses every second projection (Ct/CtVolume.cpp:945); 'cw' with
    increasing angles negates them (Ct/CtVolume.cpp:755-767)."""
    from test_facade import Facade, load_ctv
    lib = load_ctv()
    scan = small_scan(W=64, H=48, P=24, seed=34)
    csv = write_scan(str(tmp_path), scan["raw"], scan["angles"], pixel_size=0.5, fcd_px=scan["fcd"], direction="cw")
    f = Facade(lib)
    assert f.load(csv)
    lib.ctv_set_skip_projections(f.h, 1)
    lib.ctv_set_filter(f.h, 2)
    lib.ctv_reconstruct(f.h)
    assert f.status()[0] == 0, f.status()
    nx, ny, nz = f.sizes()[:3]
    got = f.volume().reshape(nx, ny, nz)
    cw = dict(scan)
    cw["angles"] = -scan["angles"]
    ref, g = _oracle_volume(oracle, cw, 0.5, "hann", "gpu", step=2)
    err = rel_rmse(got, ref)
    print(f"facade skip/cw FDK rel RMSE {err:.3e}")
    assert err < 1e-4

    # getVolumeCrossSection, left-handed (Common/Volume.h:600-625): X -> (z descending, y),
    # Y -> (z descending, x descending), Z -> (x, y)
    np.testing.assert_array_equal(f.cross_section(0, 5), got[5].T[::-1, :])
    np.testing.assert_array_equal(f.cross_section(1, ny - 1), got[:, ny - 1, :].T[::-1, ::-1])
    np.testing.assert_array_equal(f.cross_section(2, nz // 2), got[:, :, nz // 2])
    for axis, n in enumerate((nx, ny, nz)):
        assert f.cross_section(axis, n - 1).size == got.size // n
        with pytest.raises(IndexError):
            f.cross_section(axis, n)
