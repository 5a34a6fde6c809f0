"""Measurement modes beside bench.py's default FDK line (python bench.py --workload ...).

precision2048  2048^3 FP32 volume from 1440 projections of 2048x1536 pixels (0.2 mm
               pitch, FCD 5000 px), Hann filter; filtered projections stored in FP32, FP16
               and BF16 (per-projection power-of-two scales), interpolation and
               accumulation in FP32; every result against the FP64 GPU reconstruction
               (relative RMSE and max-abs over all voxels).  The volume is processed in
               z-slabs.  Phantom: 256 seeded ellipsoids (attenuation 0.01..2.0 per phantom
               unit, scaled) plus 8 needle rods of radius 2 voxels, noise-free intensities.
gradient512    512^3 volume, 360 projections of 512x512 (FCD 2500 px): forward projector A,
               its adjoint A^T, the FDK operator and its backward, with and without
               z-slab recomputation; dot-product test <Ax,y> vs <x,A^T y> on N(0,1) data.
refine768      sparse view: 120 projections of 1024x768 (FCD 5000 px), 768^3 volume,
               128 seeded ellipsoids, Poisson counts at I0 = 5e3.  FDK seed from FP16
               filtered projections, then 10 steepest-descent steps (exact line search) on
               0.5 ||A x - p||^2 with the adjoint pair, in z-slabs (ct_amd.refine; with
               --checkpoint-dir every finished slab of each step is checkpointed there);
               RMSE against the voxelised analytic phantom and against the plain FP32
               FDK, ms per step.
files1024      the CtVolume facade (the path of ct_cli) from 720 PGM files of 1024^2: load,
               reconstruct and save, each timed, on CtVolume's 1002 x 1002 x 1024 grid.
rowfilter1024  the ramp-family row filter over 720 x 1024 rows of 1024 (Shepp-Logan):
               ctg_filter_rows (the register FFT) against the same circular filter as a
               dense circulant GEMM on the matrix cores (torch.matmul with FP32 inputs,
               and with BF16 inputs for scale), with the GEMMs' error against the FFT.

Each mode prints one JSON line.  Synthetic data are generated on the GPU.
"""
from __future__ import annotations

import json
import math
import os
import sys
import time


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _sync(torch):
    torch.cuda.synchronize()


def _errors(torch, v, ref64):
    """(sum of squared differences, sum of squared reference, max |difference|) in FP64."""
    d = v.double() - ref64
    return float(torch.sum(d * d)), float(torch.sum(ref64 * ref64)), float(torch.max(torch.abs(d)))


def precision2048(args):
    import torch
    import ct_amd
    from ct_amd.phantom import circular_angles, needles, phantom_unit_radius, random_ellipsoids

    dev = torch.device("cuda:0")
    W, H, P, n = args.det_w or 2048, args.det_h or 1536, args.nproj or 1440, args.grid or 2048
    fcd = 500.0 / (0.2 * 500.0 / 1000.0)
    slab = min(args.slab or 256, n)
    g = ct_amd.geometry_grid(W, H, fcd, 0.0, n, n, n)
    angles = circular_angles(P)
    ctx = ct_amd.Context(0)
    ctx.set_geometry(g)
    ctx.set_projections(angles)
    unit = phantom_unit_radius(g.radius, H, fcd)
    k = 0.25 / unit
    ell = [e[:9] + [e[9] * k] for e in random_ellipsoids(256, 3, unit, mu_range=(0.01, 2.0))]
    ell += needles(8, 33, unit, mu=2.0 * k * 4.0)
    raw = torch.empty((P, H, W), dtype=torch.int16, device=dev)  # uint16 bytes
    t0 = time.perf_counter()
    ctx.phantom_images(ell, 0, P, 0.0, 65535.0, 3, raw, None)
    _sync(torch)
    imin = int(torch.min(raw[:, ::8, ::8].to(torch.int32) & 0xFFFF))
    log(f"phantom {P}x{H}x{W}: {time.perf_counter() - t0:.1f} s, max line integral {-math.log(max(imin, 1) / 65535.0):.2f}")
    stream = torch.cuda.current_stream().cuda_stream
    filt, base = "hann", 1.0

    stores = ["f32", "f16", "bf16"]
    entries, scales, t_pre = {}, {}, {}
    for s in stores:
        entries[s] = torch.empty(ctx.filtered_bytes(s, P), dtype=torch.uint8, device=dev)
        scales[s] = torch.empty(P, dtype=torch.float32, device=dev)
        _sync(torch)
        t0 = time.perf_counter()
        ctx.preprocess(raw, "u16", P, base, "cpu", filt, s, entries[s], scales[s], stream)
        _sync(torch)
        t_pre[s] = time.perf_counter() - t0
        log(f"{s}: preprocessing {t_pre[s] * 1e3:.0f} ms, {entries[s].numel() / 2**30:.1f} GiB of entries")

    ctx64 = ct_amd.Context(0)
    ctx64.set_geometry(g)
    ctx64.set_projections(angles)
    acc = {s: [0.0, 0.0, 0.0] for s in stores}
    t_bp = {s: 0.0 for s in stores}
    t64 = 0.0
    vol64 = torch.empty((n, n, slab), dtype=torch.float64, device=dev)
    vol = torch.empty((n, n, slab), dtype=torch.float32, device=dev)
    for z0 in range(0, n, slab):
        nz = min(slab, n - z0)
        v64 = vol64[:, :, :nz] if nz == slab else torch.empty((n, n, nz), dtype=torch.float64, device=dev)
        _sync(torch)
        t0 = time.perf_counter()
        ctx64.fdk_f64(raw, "u16", P, base, filt, v64, z0, nz, stream)
        _sync(torch)
        t64 += time.perf_counter() - t0
        for s in stores:
            v = vol[:, :, :nz] if nz == slab else torch.empty((n, n, nz), dtype=torch.float32, device=dev)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            ctx.backproject(entries[s], s, scales[s], 0, P, v, z0, nz, False, stream)
            e1.record()
            _sync(torch)
            t_bp[s] += e0.elapsed_time(e1) / 1e3
            d2, r2, mx = _errors(torch, v.contiguous() if nz != slab else v, v64.contiguous() if nz != slab else v64)
            acc[s][0] += d2
            acc[s][1] += r2
            acc[s][2] = max(acc[s][2], mx)
        log(f"slab z {z0}..{z0 + nz}: FP64 {t64:.1f} s cumulative")
    updates = float(n) * n * n * P
    out = {"workload": f"precision: {n}^3 from {P}x{H}x{W}, hann, z-slabs of {slab}", "updates": updates,
           "fp64_seconds": t64, "stores": {}}
    for s in stores:
        out["stores"][s] = {"rel_rmse_vs_f64": math.sqrt(acc[s][0] / acc[s][1]), "maxabs_vs_f64": acc[s][2],
                            "bp_seconds": t_bp[s], "bp_gups": updates / t_bp[s] / 1e9,
                            "preprocess_seconds": t_pre[s]}
    print(json.dumps(out), flush=True)


def gradient512(args):
    import torch
    import ct_amd
    from ct_amd.autograd import ConeBeamOperator
    from ct_amd.phantom import circular_angles

    dev = torch.device("cuda:0")
    n, W, P = args.grid or 512, args.det_w or 512, args.nproj or 360
    g = ct_amd.geometry_grid(W, W, 2500.0, 0.0, n, n, n)
    angles = circular_angles(P)
    gen = torch.Generator(device=dev)
    gen.manual_seed(4)
    x = torch.randn((n, n, n), generator=gen, device=dev)
    gen.manual_seed(5)
    y = torch.randn((P, W, W), generator=gen, device=dev)
    res = {"workload": f"differentiation: {n}^3, {P}x{W}x{W}", "updates": float(n) ** 3 * P}
    for label, kw in (("whole", {}), ("slabs", {"slab_nz": 128, "max_entry_bytes": 256 << 20})):
        op = ConeBeamOperator(g, angles, filt="ramlak", **kw)
        times = {}
        for name, fn, arg in (("A", op.project, x), ("At", op.backproject, y), ("fdk", op.fdk, y),
                              ("fdk_adjoint", op.fdk_adjoint, x)):
            fn(arg)
            _sync(torch)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                r = fn(arg)
            _sync(torch)
            times[name] = (time.perf_counter() - t0) / args.steps * 1e3
            del r
        ax, aty = op.project(x), op.backproject(y)
        lhs = float(torch.sum(ax.double() * y.double()))
        rhs = float(torch.sum(x.double() * aty.double()))
        res[label] = {"ms": times, "A_gups": res["updates"] / times["A"] / 1e6,
                      "At_gups": res["updates"] / times["At"] / 1e6,
                      "dot_rel": abs(lhs - rhs) / max(abs(lhs), abs(rhs))}
        log(f"{label}: {times}, dot test {res[label]['dot_rel']:.2e}")
        del op, ax, aty
    print(json.dumps(res), flush=True)


def refine768(args):
    import torch
    import ct_amd
    from ct_amd.autograd import ConeBeamOperator
    from ct_amd.phantom import circular_angles, phantom_unit_radius, random_ellipsoids, voxelize
    from ct_amd.refine import refine

    dev = torch.device("cuda:0")
    W, H, P, n = args.det_w or 1024, args.det_h or 768, args.nproj or 120, args.grid or 768
    fcd = 500.0 / (0.2 * 500.0 / 1000.0)
    g = ct_amd.geometry_grid(W, H, fcd, 0.0, n, n, n)
    angles = circular_angles(P)
    ctx = ct_amd.Context(0)
    ctx.set_geometry(g)
    ctx.set_projections(angles)
    unit = phantom_unit_radius(g.radius, H, fcd)
    # attenuation 0.1..1.0 per phantom unit, halved: line integrals of a few units, so
    # that I0 = 5e3 photons leave tens to hundreds of counts behind the densest paths
    ell = [e[:9] + [e[9] * 0.5 / unit] for e in random_ellipsoids(128, 7, unit)]
    photons = 5.0e3
    raw = torch.empty((P, H, W), dtype=torch.int16, device=dev)
    ctx.phantom_images(ell, 0, P, photons, 65535.0, 7, raw, None)
    base = photons / 65535.0
    stream = torch.cuda.current_stream().cuda_stream
    truth = voxelize(ell, g, dev)
    r2 = g.radius * g.radius
    xs = torch.arange(n, device=dev, dtype=torch.float32) - g.x_pre
    ys = torch.arange(n, device=dev, dtype=torch.float32) - g.y_pre
    mask = ((xs[:, None] ** 2 + ys[None, :] ** 2) < r2)[:, :, None].expand(n, n, n)

    def rmse(a, b):
        d = (a - b)[mask].double()
        return float(torch.sqrt(torch.mean(d * d)))

    # FDK in the reference's units, then the scale to attenuation per voxel
    scale = (math.pi / P) * (W / 2 + 1) / (W * W)
    fdk = {}
    for s in ("f32", "f16"):
        ent = torch.empty(ctx.filtered_bytes(s, P), dtype=torch.uint8, device=dev)
        sc = torch.empty(P, dtype=torch.float32, device=dev)
        v = torch.empty((n, n, n), dtype=torch.float32, device=dev)
        ctx.preprocess(raw, "u16", P, base, "cpu", "ramlak", s, ent, sc, stream)
        ctx.backproject(ent, s, sc, 0, P, v, 0, n, False, stream)
        fdk[s] = v * scale
        del ent, sc, v
    # measured line integrals, the data term of the refinement
    lineint = torch.empty((P, H, W), dtype=torch.float32, device=dev)
    raw_f = (raw.to(torch.int32) & 0xFFFF).to(torch.float32)
    lineint.copy_(-torch.log(torch.clamp(raw_f, min=1.0) / photons))
    op = ConeBeamOperator(g, angles, slab_nz=args.slab or 256, max_entry_bytes=512 << 20)
    x = fdk["f16"].clone()
    # least-squares scale of the seed against the data: how well the FDK scale constant
    # matches the projector's normalisation (1 = exactly)
    ax = op.project(x)
    ls_scale = float(torch.sum(ax.double() * lineint.double()) / torch.sum(ax.double() * ax.double()))
    del ax
    hist = [{"step": 0, "rmse_truth": rmse(x, truth), "rmse_vs_fp32_fdk": rmse(x, fdk["f32"])}]
    step_ms, ck_ms = [], []

    def on_step(rec, xk):
        step_ms.append(rec["ms"])
        ck_ms.append(rec["ms_checkpoint"])
        hist.append({"step": rec["step"], "loss_before": rec["loss_before"], "rmse_truth": rmse(xk, truth),
                     "rmse_vs_fp32_fdk": rmse(xk, fdk["f32"])})
        log(f"step {rec['step']}: loss {rec['loss_before']:.4e}, RMSE vs phantom {hist[-1]['rmse_truth']:.4e}, "
            f"{rec['ms']:.0f} ms (+{rec['ms_checkpoint']:.0f} ms checkpoint writes)")

    ckdir = getattr(args, "checkpoint_dir", "") or None
    if ckdir and os.path.exists(os.path.join(ckdir, "state.json")):
        raise SystemExit(f"{ckdir} holds a checkpoint of an earlier run; give an empty directory")
    every = max(1, getattr(args, "checkpoint_every", 1) or 1)
    x, _ = refine(op, x, lineint, 10, checkpoint_dir=ckdir, on_step=on_step, every=every)
    out = {"workload": f"sparse-view refinement: {n}^3 from {P}x{H}x{W}, Poisson I0=5e3, FP16 FDK seed, "
                       f"10 steepest-descent steps in z-slabs of {op.slab_nz}"
                       + (f", per-slab checkpoints every {every} step(s)" if ckdir else ""),
           "fdk_scale_to_mu_per_voxel": scale, "fdk_seed_least_squares_scale": ls_scale,
           "rmse_ls_scaled_fdk_vs_truth": rmse(fdk["f16"] * ls_scale, truth),
           "rmse_fp32_fdk_vs_truth": rmse(fdk["f32"], truth), "rmse_fp16_fdk_vs_truth": rmse(fdk["f16"], truth),
           "truth_rms": float(torch.sqrt(torch.mean(truth[mask].double() ** 2))),
           "ms_per_step": sum(step_ms) / len(step_ms),
           "ms_checkpoint_per_step": sum(ck_ms) / len(ck_ms) if ckdir else None, "steps": hist}
    print(json.dumps(out), flush=True)


def rowfilter1024(args):
    import torch
    import ct_amd

    dev = torch.device("cuda:0")
    W, P = args.det_w or 1024, args.nproj or 720
    H = args.det_h or W
    rows = P * H
    filt = "shepplogan"
    g = ct_amd.geometry_grid(W, H, 2500.0, 0.0, 16, 16, 16)  # the row filter reads only the width
    ctx = ct_amd.Context(0)
    ctx.set_geometry(g)
    stream = torch.cuda.current_stream()
    gen = torch.Generator(device=dev)
    gen.manual_seed(7)
    x = torch.randn((rows, W), generator=gen, device=dev)
    y_fft = torch.empty_like(x)
    # The filter is a circular convolution, y[j] = sum_i x[i] h[(j - i) mod W], with h the
    # library's own response to a unit impulse: y = x C with C[i][j] = h[(j - i) mod W].
    delta = torch.zeros((1, W), device=dev)
    delta[0, 0] = 1.0
    h = torch.empty_like(delta)
    ctx.filter_rows(delta, h, 1, filt, stream.cuda_stream)
    ar = torch.arange(W, device=dev)
    C = h[0][(ar[None, :] - ar[:, None]) % W].contiguous()
    torch.backends.cuda.matmul.allow_tf32 = False
    xb, Cb = x.bfloat16(), C.bfloat16()

    def timed(fn):
        fn()
        _sync(torch)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            r = fn()
        e1.record(stream)
        _sync(torch)
        return e0.elapsed_time(e1) / args.steps, r

    t_fft, _ = timed(lambda: ctx.filter_rows(x, y_fft, rows, filt, stream.cuda_stream))
    t_f32, y32 = timed(lambda: torch.matmul(x, C))
    t_bf16, yb = timed(lambda: torch.matmul(xb, Cb))

    def rel(a):
        d = a.double() - y_fft.double()
        return float(torch.sqrt(torch.sum(d * d) / torch.sum(y_fft.double() ** 2)))

    flops = 2.0 * rows * W * W
    out = {"workload": f"row filter: {P} x {H} rows of {W}, {filt}", "rows": rows,
           "fft_ms": t_fft, "fft_gbs": rows * W * 8 / t_fft / 1e6,
           "gemm_f32_ms": t_f32, "gemm_f32_tflops": flops / t_f32 / 1e9, "gemm_f32_peak_tflops": 157.3,
           "gemm_f32_rel_rmse_vs_fft": rel(y32),
           "gemm_bf16_ms": t_bf16, "gemm_bf16_tflops": flops / t_bf16 / 1e9, "gemm_bf16_peak_tflops": 2500.0,
           "gemm_bf16_rel_rmse_vs_fft": rel(yb)}
    log(f"FFT {t_fft:.2f} ms; FP32 GEMM {t_f32:.2f} ms ({out['gemm_f32_tflops']:.0f} TFLOP/s, "
        f"rel RMSE {out['gemm_f32_rel_rmse_vs_fft']:.1e}); BF16 GEMM {t_bf16:.2f} ms "
        f"({out['gemm_bf16_tflops']:.0f} TFLOP/s, rel RMSE {out['gemm_bf16_rel_rmse_vs_fft']:.1e})")
    print(json.dumps(out), flush=True)
    ctx.close()


def _write_scan_files(directory, host_raw, angles, pixel_mm, fcd_px, base):
    """Ct's input format (Ct/CtVolume.cpp:136-330): a CSV of six parameter lines, one
    ignored line, then `file<TAB>angle` per projection, and one 16-bit PGM per image
    (big-endian samples).  Returns the CSV path and the bytes written."""
    import numpy as np
    img_dir = os.path.join(directory, "images")
    os.makedirs(img_dir, exist_ok=True)
    P, H, W = host_raw.shape
    lines = ["images\tprojection directory", f"{pixel_mm!r}\tpixel size [mm]", "ccw\trotation direction",
             "0.0\tu offset [mm]", f"{fcd_px * pixel_mm!r}\tFCD [mm]", f"{base!r}\tbase intensity", "ignored"]
    header = f"P5\n{W} {H}\n65535\n".encode()
    nbytes = 0
    for p in range(P):
        name = f"proj_{p:04d}.pgm"
        with open(os.path.join(img_dir, name), "wb") as f:
            f.write(header)
            f.write(np.ascontiguousarray(host_raw[p], dtype=">u2").tobytes())
        nbytes += len(header) + 2 * W * H
        lines.append(f"{name}\t{float(angles[p])!r}")
    csv = os.path.join(directory, "scan.csv")
    with open(csv, "w") as f:
        f.write("\n".join(lines) + "\n")
    return csv, nbytes


def files1024(args):
    """The CtVolume facade (libctvolume.so, the path of ct_cli) from image files: the
    default workload's 720 uint16 projections of 1024^2 written as PGM files with a CSV
    (0.2 mm pixels, FCD 2500 px), then sinogramFromImages, reconstructVolume (Shepp-Logan
    filter, GPU) and saveVolumeToBinaryFile (Z fastest, little endian) on CtVolume's own
    1002 x 1002 x 1024 grid, each timed.  The files are read through the page cache
    (they were just written), so the load and decode figures are CPU-bound decode rates,
    not disk rates."""
    import argparse
    import ctypes
    import shutil
    import tempfile
    import numpy as np
    import torch
    import ct_amd
    import bench
    P = args.nproj or 720
    W = args.det_w or 1024
    ctx = ct_amd.Context(0)
    g, raw, angles, base = bench.build_workload(ctx, torch, argparse.Namespace(det=W, nproj=P, grid=1024))
    host = raw.cpu().numpy().view(np.uint16)
    del raw
    ctx.close()
    torch.cuda.empty_cache()
    scratch = tempfile.mkdtemp(prefix="ct_files_")
    try:
        t = time.perf_counter()
        csv, in_bytes = _write_scan_files(scratch, host, angles, 0.2, float(g.fcd), base)
        log(f"wrote {P} PGM files ({in_bytes / 1e9:.2f} GB) in {time.perf_counter() - t:.1f} s")
        lib = ctypes.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "ct_amd", "libctvolume.so"))
        lib.ctv_create.restype = ctypes.c_void_p
        for name in ("ctv_destroy", "ctv_reconstruct"):
            getattr(lib, name).argtypes = [ctypes.c_void_p]
        lib.ctv_load.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        lib.ctv_set_filter.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.ctv_set_use_cuda.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.ctv_save.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
        lib.ctv_sizes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_ulonglong)]
        lib.ctv_status.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
        h = lib.ctv_create()

        def checked(what):
            msg = ctypes.create_string_buffer(512)
            if lib.ctv_status(h, msg, 512) != 0:  # 0 success, 1 interrupted, 2 error
                raise RuntimeError(f"{what}: {msg.value.decode()}")

        t0 = time.perf_counter()
        if not lib.ctv_load(h, csv.encode()):
            checked("sinogramFromImages")
        t1 = time.perf_counter()
        lib.ctv_set_filter(h, 1)  # shepplogan
        lib.ctv_set_use_cuda(h, 1)
        lib.ctv_reconstruct(h)
        t2 = time.perf_counter()
        checked("reconstructVolume")
        out_path = os.path.join(scratch, "vol.raw")
        lib.ctv_save(h, out_path.encode(), 0, 0)  # Z fastest, little endian
        t3 = time.perf_counter()
        checked("saveVolumeToBinaryFile")
        sizes = (ctypes.c_ulonglong * 7)()
        lib.ctv_sizes(h, sizes)
        lib.ctv_destroy(h)
        nx, ny, nz = int(sizes[0]), int(sizes[1]), int(sizes[2])
        updates = float(nx) * ny * nz * P
        out = {"workload": f"CtVolume facade from files: {P} PGM images of {W}x{W} (16-bit), shepplogan, "
                           f"CtVolume grid {nx}x{ny}x{nz}, raw volume Z fastest LE",
               "input_bytes": in_bytes, "volume_bytes": 4 * nx * ny * nz, "updates": updates,
               "ms_load": (t1 - t0) * 1e3, "ms_reconstruct": (t2 - t1) * 1e3, "ms_save": (t3 - t2) * 1e3,
               "ms_total": (t3 - t0) * 1e3, "gups_reconstruct": updates / (t2 - t1) / 1e9,
               "gups_total": updates / (t3 - t0) / 1e9,
               "load_gbs": in_bytes / (t1 - t0) / 1e9, "save_gbs": 4.0 * nx * ny * nz / (t3 - t2) / 1e9}
        log(f"load {out['ms_load']:.0f} ms, reconstruct {out['ms_reconstruct']:.0f} ms "
            f"({out['gups_reconstruct']:.0f} GUPS), save {out['ms_save']:.0f} ms")
        print(json.dumps(out), flush=True)
    finally:
        shutil.rmtree(scratch, ignore_errors=True)


WORKLOADS = {"precision2048": precision2048, "gradient512": gradient512, "refine768": refine768,
             "rowfilter1024": rowfilter1024, "files1024": files1024}
