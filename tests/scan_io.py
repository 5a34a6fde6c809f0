"""Writes synthetic scans in Ct's input format: a CSV geometry file (6 parameter lines,
one ignored line, then `filename<TAB>angle[<TAB>heightOffset]`) and one image per
projection.  PGM is written here; TIFF and PNG through PIL (an independent encoder)."""
from __future__ import annotations

import os

import numpy as np


def write_pgm16(path, img):
    h, w = img.shape
    with open(path, "wb") as f:
        f.write(f"P5\n{w} {h}\n65535\n".encode())
        f.write(np.ascontiguousarray(img, dtype=">u2").tobytes())


def write_image(path, img, fmt, compression=None):
    if fmt == "pgm":
        write_pgm16(path, img)
        return
    from PIL import Image
    if img.dtype == np.uint16:
        im = Image.frombytes("I;16", (img.shape[1], img.shape[0]), np.ascontiguousarray(img, "<u2").tobytes())
    elif img.dtype == np.uint8:
        im = Image.fromarray(img, mode="L")
    else:
        im = Image.fromarray(img.astype(np.float32), mode="F")
    if fmt == "tif":
        im.save(path, format="TIFF", compression=compression)
    else:
        im.save(path, format="PNG")


def write_scan(directory, raw, angles, pixel_size, fcd_px, u_offset_px=0.0, base=1.0, heights_px=None,
               direction="ccw", fmt="pgm", compression=None, image_subdir="images", height_shift_mm=0.0):
    """raw: (P, H, W) images.  Returns the CSV path."""
    os.makedirs(os.path.join(directory, image_subdir), exist_ok=True)
    ext = {"pgm": "pgm", "tif": "tif", "png": "png"}[fmt]
    lines = [
        f"{image_subdir}\tdirectory of the projections",
        f"{float(pixel_size)!r}\tpixel size [mm]",
        f"{direction}\trotation direction",
        f"{float(u_offset_px * pixel_size)!r}\tu offset [mm]",
        f"{float(fcd_px * pixel_size)!r}\tFCD [mm]",
        f"{float(base)!r}\tbase intensity",
        "ignored line",
    ]
    for p in range(raw.shape[0]):
        name = f"proj_{p:04d}.{ext}"
        write_image(os.path.join(directory, image_subdir, name), raw[p], fmt, compression)
        if heights_px is None:
            lines.append(f"{name}\t{float(angles[p])!r}")
        else:
            lines.append(f"{name}\t{float(angles[p])!r}\t{float(heights_px[p] * pixel_size + height_shift_mm)!r}")
    csv = os.path.join(directory, "scan.csv")
    with open(csv, "w") as f:
        f.write("\n".join(lines) + "\n")
    return csv
