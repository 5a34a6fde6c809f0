#!/usr/bin/env python3
"""FDK reconstruction benchmark on one MI355X.

Workload (default): 1024^3 FP32 volume from 720 uint16 projections of 1024x1024
pixels (0.4 mm pitch, 0.2 mm at the isocentre, SOD 500 mm -> FCD 2500 px),
equiangular full circular scan, Shepp-Logan filter; phantom = 3-D Shepp-Logan plus
64 seeded random ellipsoids, Poisson counts at I0 = 1e4.

One step = the whole hot path on the GPU with the raw projections already in HBM:
log + Feldkamp weight + ramp-family row filter (ctg_preprocess) and the voxel-driven
backprojection of all 720 projections into the 1024^3 volume (ctg_backproject).
value = voxel-updates (nx*ny*nz*P) / step time, in GUPS.  Also reported: the
backprojection-only rate, the same step with FP16-stored filtered projections, the
error of both against the FP64 GPU reconstruction, the backprojection's roofline (VALU
issue, with its measured HBM traffic beside the compulsory bytes), the time to write
the volume in both index orders, and the CPU restatement of the reference's
reconstruction timed on the host's cores (a bounded region of the same scan, plus the
256x192x360 scan in full; reported only).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, 8 TB/s
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "backproject_traffic_1024cubed_720proj.json")
BYTES_PER_UPDATE = 8.0     # one FP32 voxel load + one store per (voxel, projection)
# VALU issue peak at 4 cycles per wave64 instruction per SIMD: 256 CUs x 64 lanes x
# 2.4 GHz, in 1e12 lane-instructions/s.  ct_amd/isa_rates measures 4.2-4.4 cycles for
# v_fract_f32, v_cvt_flr_i32_f32, v_lshl_add_u32 and v_fma_mix_f32 but 2.8 for v_fma_f32,
# so for an FMA-heavy mix this is a lower bound of the issue rate.
VALU_PEAK_T = 256 * 64 * 2.4e9 / 1e12
# MI355X's published FP32 vector peak (TFLOP/s) and the flops of one voxel-update
FP32_PEAK_TFLOPS = 157.3
FLOPS_PER_UPDATE = 8.0
# lane-instructions per voxel-update of the register-accumulating formulation: one FMA
# for the detector row, floor and fraction of it, three FMAs for the bilinear form and
# one FMA into the accumulator (the LDS address is not counted)
VALU_PER_UPDATE = 7.0
# LDS read bandwidth: 256 bytes per clock per CU (a ds_read_b128 serves 16 lanes per
# 4-cycle pass: ct_amd/isa_rates measures at most 6.6 cycles per conflict-free wave
# read and 16.9 for a 4-way conflict), 256 CUs at 2.4 GHz, in TB/s
LDS_PEAK_TBS = 256 * 256 * 2.4e9 / 1e12
# the same at the measured conflict-free ds_read_b128 rate: 1 KB per wave instruction
# every 6.64 cycles per CU (ct_amd/isa_rates)
DS_READ_B128_CYCLES = 6.64
LDS_READ_MEASURED_TBS = 256 * (64 * 16 / DS_READ_B128_CYCLES) * 2.4e9 / 1e12


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_workload(ctx, torch, args):
    import ct_amd
    from ct_amd.phantom import circular_angles, phantom_unit_radius, random_ellipsoids, shepp_logan_3d

    W = H = args.det
    P = args.nproj
    n = args.grid
    fcd = 500.0 / (0.4 * 500.0 / 1000.0)  # SOD / isocentre pitch = 2500 px
    g = ct_amd.geometry_grid(W, H, fcd, 0.0, n, n, n)
    ctx.set_geometry(g)
    angles = circular_angles(P)
    ctx.set_projections(angles)
    unit = phantom_unit_radius(g.radius, H, fcd)
    ell = shepp_logan_3d(unit, scale=1.0 / unit)
    ell += [e[:9] + [e[9] * 0.1 / unit] for e in random_ellipsoids(64, 2, unit)]
    raw = torch.empty((P, H, W), dtype=torch.uint16, device="cuda")
    t0 = time.perf_counter()
    ctx.phantom_images(ell, 0, P, 1.0e4, 65535.0, 2, raw, None)
    torch.cuda.synchronize()
    log(f"phantom: {P}x{H}x{W} uint16 in {time.perf_counter() - t0:.2f} s")
    return g, raw, angles, 1.0e4 / 65535.0


def _host_cpu():
    """(logical CPUs usable by this process, model name, SMT state) of the host."""
    model, smt = "", None
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    try:
        with open("/sys/devices/system/cpu/smt/active") as f:
            smt = f.read().strip() == "1"
    except OSError:
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count()
    return n, model, smt


def _timing_build():
    """The oracle's source built at -O3 -march=native on this host (oracle/Makefile's
    `timing` target) into a temporary directory; None if the build fails."""
    import subprocess
    import tempfile
    out = tempfile.mkdtemp(prefix="oracle_timing_")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "timing", f"OUT={out}"],
                       capture_output=True, text=True)
    path = os.path.join(out, "liboracle_timing.so")
    return path if r.returncode == 0 and os.path.exists(path) else None


def cpu_full_grid(orc, host2, base2, angles2, g2, slab=128):
    """The same CPU restatement on the scan's full CtVolume grid (1002x1002x1024 for
    1024^2 x 720): the CPU preprocessing once, then the backprojection in z-slabs of
    `slab` slices (each slab takes every projection, as the reference's chunks do,
    Ct/CtVolume.cpp:1206), with a line per slab.  Returns (grid, seconds, updates)."""
    import ct_amd
    gf, _ = ct_amd.geometry_ctvolume(g2.det_width, g2.det_height, g2.fcd, 0.0, None)
    og = orc.geometry_from(gf)
    t0 = time.perf_counter()
    filt = orc.preprocess_f32(og, host2, "u16", base2, "shepplogan", "cpu")
    log(f"CPU full grid {gf.nx}x{gf.ny}x{gf.nz}: preprocessing {time.perf_counter() - t0:.1f} s")
    for z0 in range(0, gf.nz, slab):
        sub = orc.geometry_from(gf)
        sub.nz = min(slab, gf.nz - z0)
        sub.z_pre = gf.z_pre - z0
        orc.backproject_f32(sub, filt, angles2)
        log(f"CPU full grid: slices {z0}..{z0 + sub.nz} done, {time.perf_counter() - t0:.1f} s")
    t = time.perf_counter() - t0
    return [gf.nx, gf.ny, gf.nz], t, float(gf.nx) * gf.ny * gf.nz * len(angles2)


def cpu_baseline(torch, raw2, base2, angles2, g2, full=False):
    """The CPU restatement of the reference's reconstruction (CPU preprocessing:
    Ct/CtVolume.cpp:787-885; reconstructionCore: :923-1054, OpenMP dynamic over x)
    timed on all host threads with the -O3 -march=native build: the 256x192 scan with
    360 projections on CtVolume's 254x254x192 grid in full, and the 1024^2 x 720 scan
    on CtVolume's grid bounded to x, y in [0.25, 0.75], z in [0.1, 0.9] (501x501x820),
    both from the raw uint16 images.  `value` is the second: the bounded sample of the
    GPU workload's scan.  Also: the rate on one thread (8 slices of the first)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as orc
    import ct_amd
    from ct_amd.phantom import circular_angles, phantom_unit_radius, random_ellipsoids

    orc.load()
    fast = _timing_build()
    build = "g++ -O3 -march=native -fopenmp (oracle.c)" if fast else "pinned oracle build (-O2)"
    if fast:
        orc.use(fast)
    try:
        # 256x192 detector, FCD 1250 px, 360 projections, 10 ellipsoids, I = round(65535 e^-p)
        W1, H1, P1, fcd1 = 256, 192, 360, 1250.0
        g1, _ = ct_amd.geometry_ctvolume(W1, H1, fcd1, 0.0, None)
        ang1 = circular_angles(P1)
        unit = phantom_unit_radius(g1.radius, H1, fcd1)
        ell = [e[:9] + [e[9] / unit] for e in random_ellipsoids(10, 1, unit)]
        with ct_amd.Context(0) as c1:
            c1.set_geometry(g1)
            c1.set_projections(ang1)
            r1 = torch.empty((P1, H1, W1), dtype=torch.int16, device="cuda")
            c1.phantom_images(ell, 0, P1, 0.0, 65535.0, 1, r1, None)
            raw1 = r1.cpu().numpy().view(np.uint16)
        og1 = orc.geometry_from(g1)
        t0 = time.perf_counter()
        orc.fdk_f32(og1, raw1, "u16", 1.0, "ramlak", ang1)
        t1 = time.perf_counter() - t0
        ups1 = float(g1.nx) * g1.ny * g1.nz * P1
        # one thread: the same scan's backprojection into 8 central slices
        filt1 = orc.preprocess_f32(og1, raw1, "u16", 1.0, "ramlak", "cpu")
        sub = orc.geometry_from(g1)
        sub.nz = 8
        sub.z_pre = g1.z_pre - (g1.nz // 2 - 4)
        threads = orc.get_threads()
        orc.set_threads(1)
        t0 = time.perf_counter()
        orc.backproject_f32(sub, filt1, ang1)
        ts = time.perf_counter() - t0
        orc.set_threads(0)
        # the GPU workload's scan on the bounded CtVolume grid
        g2r, roi = ct_amd.geometry_ctvolume(g2.det_width, g2.det_height, g2.fcd, 0.0, [0.25, 0.75, 0.25, 0.75, 0.1, 0.9])
        host2 = raw2.view(torch.int16).cpu().numpy().view(np.uint16)
        t0 = time.perf_counter()
        orc.fdk_f32(orc.geometry_from(g2r), host2, "u16", base2, "shepplogan", angles2)
        t2 = time.perf_counter() - t0
        ups2 = float(g2r.nx) * g2r.ny * g2r.nz * len(angles2)
        fullrun = cpu_full_grid(orc, host2, base2, angles2, g2) if full else None
    finally:
        orc.use(None)
    n, model, smt = _host_cpu()
    extra = {}
    if fullrun is not None:
        grid, tf, upsf = fullrun
        extra = {"full_grid": grid, "full_seconds": tf, "full_gups": upsf / tf / 1e9,
                 "full_workload": "the same scan on CtVolume's full grid, CPU preprocessing once, then the "
                                  "backprojection in z-slabs of 128 slices"}
    return {"value": ups2 / t2 / 1e9, "unit": "GUPS", "cores": threads, "kind": "port",
            "sample": f"CPU restatement of the reference's FDK (CPU preprocessing + reconstructionCore, OpenMP) of "
                      f"the same 720x1024x1024 scan on CtVolume's grid bounded to x,y in [0.25,0.75], z in "
                      f"[0.1,0.9] ({g2r.nx}x{g2r.ny}x{g2r.nz}), from the raw uint16 images: {t2:.1f} s wall",
            "build": build, "cpu_model": model, "host_logical_cpus": n, "smt_active": smt,
            "roi_seconds": t2,
            "scan256_grid": [g1.nx, g1.ny, g1.nz], "scan256_seconds": t1, "scan256_gups": ups1 / t1 / 1e9,
            "scan256_workload": f"{P1}x{H1}x{W1} uint16, ramlak, CtVolume grid {g1.nx}x{g1.ny}x{g1.nz}, "
                               f"preprocessing + backprojection",
            "single_thread_gups": float(sub.nx) * sub.ny * sub.nz * P1 / ts / 1e9, **extra}


def write_timing(vol_host, nx, ny, nz):
    """ms to write the volume's raw stream (libctvolume's writer: Z fastest, X fastest,
    little endian) next to one plain fwrite of the same bytes (the host's write rate,
    the bound for the writer), into a temporary file that is removed afterwards."""
    import ctypes
    import tempfile
    lib = ctypes.CDLL(os.path.join(ROOT, "ct_amd", "libctvolume.so"))
    lib.ctv_write_raw.argtypes = [ctypes.c_void_p, ctypes.c_ulonglong, ctypes.c_ulonglong, ctypes.c_ulonglong,
                                  ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    d = tempfile.mkdtemp(prefix="ct_write_")
    path = os.path.join(d, "vol.raw")
    out = {}
    try:
        for name, order in (("zfastest", 1), ("xfastest", 0)):
            t0 = time.perf_counter()
            rc = lib.ctv_write_raw(vol_host.ctypes.data, nx, ny, nz, path.encode(), order, 1)
            out[f"ms_write_{name}"] = (time.perf_counter() - t0) * 1e3
            os.remove(path)
            if rc != 0:
                raise RuntimeError(f"volume write failed ({rc})")
        t0 = time.perf_counter()
        with open(path, "wb") as f:
            f.write(memoryview(vol_host).cast("B"))
        out["ms_write_plain_fwrite"] = (time.perf_counter() - t0) * 1e3
        os.remove(path)
    finally:
        os.rmdir(d)
    out["write_bytes"] = int(vol_host.nbytes)
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--det", type=int, default=None, help="square detector size (default 1024)")
    ap.add_argument("--nproj", type=int, default=None, help="projections (default: the workload's)")
    ap.add_argument("--grid", type=int, default=None, help="cubic grid size (default: the workload's)")
    ap.add_argument("--no-precision", action="store_true", help="skip the FP64 reference and FP16 error")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-full", action="store_true",
                    help="also time the CPU baseline on the full CtVolume grid (about 3-4 minutes of CPU work)")
    ap.add_argument("--no-write", action="store_true", help="skip the volume write timing")
    ap.add_argument("--only", choices=["f32", "f16", "both"], default="both")
    ap.add_argument("--no-e2e", action="store_true", help="skip the end-to-end run from pinned host memory")
    ap.add_argument("--no-roi", action="store_true", help="skip the region-of-interest run")
    ap.add_argument("--workload", default="fdk1024",
                    choices=["fdk1024", "precision2048", "gradient512", "refine768", "rowfilter1024", "files1024"],
                    help="fdk1024: the headline line below; the others: bench_workloads.py")
    ap.add_argument("--det-w", type=int, default=0, help="detector width (other workloads)")
    ap.add_argument("--det-h", type=int, default=0, help="detector height (other workloads)")
    ap.add_argument("--slab", type=int, default=0, help="z-slab height (other workloads)")
    ap.add_argument("--checkpoint-dir", default="",
                    help="refine768: write per-slab checkpoints to this (empty) directory")
    ap.add_argument("--checkpoint-every", type=int, default=1,
                    help="refine768: checkpoint every this many steps (and the last)")
    ap.add_argument("--calib-copy-gib", type=float, default=0.0,
                    help="run a device-to-device copy of this size first (HBM counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: ct_amd has no multi-GPU path "
              "(FDK would shard over z-slabs with no exchange step). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.calib_copy_gib > 0:
        # a device copy of known size first, for calibrating the HBM byte counters
        import torch
        n = int(args.calib_copy_gib * 2**30) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {n * 4} bytes read and written, twice")
        del src, dst
    if args.workload != "fdk1024":
        import bench_workloads
        bench_workloads.WORKLOADS[args.workload](args)
        return
    args.det = args.det or 1024
    args.nproj = args.nproj or 720
    args.grid = args.grid or 1024

    import torch
    import ct_amd

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    info = ct_amd._ctg.device_info(0)
    log(f"device: {info.name.decode()} CUs={info.compute_units} HBM={info.total_bytes / 2**30:.0f} GiB")
    # measured HBM copy rate beside the published 8 TB/s: 1 GiB read and 1 GiB written
    # per copy, timed with events over 5 copies after one warm-up
    n = (1 << 30) // 4
    src = torch.ones(n, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = 5 * 2 * n * 4 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    log(f"device copy: {copy_gbs:.0f} GB/s (read + write)")
    del src, dst
    ctx = ct_amd.Context(0)
    g, raw, angles, base = build_workload(ctx, torch, args)
    P = args.nproj
    updates = float(g.nx) * g.ny * g.nz * P
    stream = torch.cuda.current_stream().cuda_stream
    filt = "shepplogan"

    def make_bufs(store):
        return (torch.empty(ctx.filtered_bytes(store, P), dtype=torch.uint8, device=dev),
                torch.empty(P, dtype=torch.float32, device=dev))

    vol = torch.empty((g.nx, g.ny, g.nz), dtype=torch.float32, device=dev)

    def run_store(store):
        filtered, scales = make_bufs(store)
        ev = []

        def step():
            ctx.preprocess(raw, "u16", P, base, "cpu", filt, store, filtered, scales, stream)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            ctx.backproject(filtered, store, scales, 0, P, vol, 0, g.nz, False, stream)
            e1.record()
            ev.append((e0, e1))

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        ev.clear()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / args.steps
        bp = sum(a.elapsed_time(b) for a, b in ev) / len(ev) / 1e3
        log(f"{store}: step {t * 1e3:.1f} ms ({updates / t / 1e9:.1f} GUPS), backprojection {bp * 1e3:.1f} ms "
            f"({updates / bp / 1e9:.1f} GUPS)")
        return t, bp, filtered, scales

    def e2e_from_pinned_host(store, chunks=8, groups=4):
        """H2D of the raw uint16 projections from pinned host memory in chunks on an upload
        stream; each chunk is preprocessed on the compute stream as soon as it lands, and
        each group of chunks is backprojected (accumulating) as soon as it is
        preprocessed, so later uploads run under earlier backprojections: the whole
        reconstruction from host images to the volume in HBM."""
        filtered, scales = make_bufs(store)
        raw16 = raw.view(torch.int16)  # same bytes; int16 has the full set of copy kernels
        host = torch.empty(raw.shape, dtype=torch.int16, pin_memory=True)
        host.copy_(raw16)
        dev_raw = torch.empty_like(raw16)
        per = ctx.filtered_bytes(store, 1)
        up = torch.cuda.Stream(device=dev)
        cmp = torch.cuda.current_stream()
        bounds = [P * i // chunks for i in range(chunks + 1)]

        def run():
            evs = []
            up.wait_stream(cmp)  # the previous run's preprocessing has finished reading dev_raw
            with torch.cuda.stream(up):
                for i in range(chunks):
                    a, b = bounds[i], bounds[i + 1]
                    dev_raw[a:b].copy_(host[a:b], non_blocking=True)
                    e = torch.cuda.Event()
                    e.record(up)
                    evs.append(e)
            for grp in range(groups):
                c0, c1 = chunks * grp // groups, chunks * (grp + 1) // groups
                for i in range(c0, c1):
                    a, b = bounds[i], bounds[i + 1]
                    cmp.wait_event(evs[i])
                    ctx.preprocess(dev_raw[a:b], "u16", b - a, base, "cpu", filt, store, filtered[a * per:b * per],
                                   scales[a:b], cmp.cuda_stream)
                a, b = bounds[c0], bounds[c1]
                ctx.backproject(filtered[a * per:b * per], store, scales[a:b], a, b - a, vol, 0, g.nz, grp > 0,
                                cmp.cuda_stream)

        run()
        torch.cuda.synchronize()
        reps = max(1, args.steps)
        t0 = time.perf_counter()
        for _ in range(reps):
            run()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / reps
        h2d = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(up):
            h2d[0].record(up)
            dev_raw.copy_(host, non_blocking=True)
            h2d[1].record(up)
        torch.cuda.synchronize()
        th2d = h2d[0].elapsed_time(h2d[1]) / 1e3
        log(f"{store} end to end from pinned host: {t * 1e3:.1f} ms (H2D alone {th2d * 1e3:.1f} ms, "
            f"{raw.numel() * 2 / th2d / 1e9:.1f} GB/s)")
        del host, dev_raw, filtered, scales
        return t, th2d

    def reconstruct_host_to_host():
        """ctg_reconstruct, the C-ABI pipeline behind the CtVolume facade: pageable host
        uint16 images in, Z_FASTEST host FP32 volume out (one warm-up and one timed run)."""
        import numpy as np
        host_raw = raw.view(torch.int16).cpu().numpy().view(np.uint16)
        out_vol = np.empty(g.nx * g.ny * g.nz, np.float32)
        ctx.reconstruct(host_raw.ctypes.data, "u16", P, base, "cpu", filt, "f32", out_vol.ctypes.data, "zfastest")
        t0 = time.perf_counter()
        ctx.reconstruct(host_raw.ctypes.data, "u16", P, base, "cpu", filt, "f32", out_vol.ctypes.data, "zfastest")
        t = time.perf_counter() - t0
        log(f"ctg_reconstruct host to host: {t * 1e3:.1f} ms")
        del host_raw, out_vol
        return t

    res = {}
    if args.only == "f16":
        t16, bp16, _, _ = run_store("f16")
        print(json.dumps({"bp_ms_f16": bp16 * 1e3, "bp_gups_f16": updates / bp16 / 1e9}), flush=True)
        return
    t32, bp32, filt32, sc32 = run_store("f32")
    res["f32"] = (t32, bp32)
    e2e = None if args.no_e2e else e2e_from_pinned_host("f32")
    h2h = None if args.no_e2e else reconstruct_host_to_host()
    vol32 = vol.clone() if not args.no_precision else None
    writes = None
    if not args.no_write:
        writes = write_timing(vol.cpu().numpy(), g.nx, g.ny, g.nz)
        log(f"volume write ({writes['write_bytes'] / 2**30:.1f} GiB): Z fastest {writes['ms_write_zfastest']:.0f} ms, "
            f"X fastest {writes['ms_write_xfastest']:.0f} ms, one plain fwrite {writes['ms_write_plain_fwrite']:.0f} ms")
    del filt32, sc32
    if args.only in ("f16", "both"):
        t16, bp16, f16buf, sc16 = run_store("f16")
        res["f16"] = (t16, bp16)
        del f16buf, sc16
    vol16 = vol.clone() if not args.no_precision and "f16" in res else None

    prec = {}
    if not args.no_precision:
        import numpy as np
        vol64 = torch.empty((g.nx, g.ny, g.nz), dtype=torch.float64, device=dev)
        t0 = time.perf_counter()
        ctx.fdk_f64(raw, "u16", P, base, filt, vol64, 0, g.nz, stream)
        torch.cuda.synchronize()
        log(f"FP64 GPU reconstruction: {time.perf_counter() - t0:.1f} s")

        def errs(v):
            d = v.double() - vol64
            rr = math.sqrt(float(torch.sum(d * d)) / float(torch.sum(vol64 * vol64)))
            return rr, float(torch.max(torch.abs(d)))

        prec["f32_rel_rmse_vs_f64"], prec["f32_maxabs_vs_f64"] = errs(vol32)
        if vol16 is not None:
            prec["f16_rel_rmse_vs_f64"], prec["f16_maxabs_vs_f64"] = errs(vol16)
        del vol64

    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline ...")
        cpu = cpu_baseline(torch, raw, base, angles, g, full=args.cpu_full)
        log(f"CPU: {cpu['value']:.3f} GUPS on {cpu['cores']} threads ({cpu['build']}); 254x254x192 from "
            f"360x192x256 in {cpu['scan256_seconds']:.2f} s")

    t32, bp32 = res["f32"]
    # The FP32 backprojection keeps its voxels in registers, so its binder is the VALU
    # issue rate, not HBM: achieved = VALU_PER_UPDATE lane-instructions per voxel-update
    # (the formulation's count, see the constant) / kernel time, against the non-packed
    # FP32 issue peak of 256 CUs x 64 lanes x 2.4 GHz.
    achieved = VALU_PER_UPDATE * updates / bp32 / 1e12
    traffic = None
    if os.path.exists(TRAFFIC_FILE) and (args.det, args.nproj, args.grid) == (1024, 720, 1024):
        # HBM bytes per launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated
        # against a copy kernel of known size (see the profile summary beside the file)
        traffic = json.load(open(TRAFFIC_FILE))["traffic_bytes_per_launch"]
    # compulsory HBM bytes of one launch: every FP32 entry (16 B a texel) read once and
    # the volume written once
    compulsory = 16.0 * g.det_width * g.det_height * P + 4.0 * g.nx * g.ny * g.nz
    out = {
        "metric": "FDK voxel-update rate (backprojection GUPS), 1024^3 from 720x1024^2; FP32 vs FP16-stored projections",
        "value": updates / t32 / 1e9,
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (3-D Shepp-Logan + 64 seeded ellipsoids, Poisson I0=1e4, generated on the GPU)",
        "config": {"workload": f"FDK {g.nx}x{g.ny}x{g.nz} FP32 from {P}x{g.det_height}x{g.det_width} uint16, "
                               f"{filt}, full circular scan, FCD {g.fcd:.0f} px",
                   "step": "preprocess (log, Feldkamp weight, row filter) + backprojection, inputs in HBM"},
        "ms_per_recon_f32": t32 * 1e3,
        "bp_ms_f32": bp32 * 1e3,
        "bp_gups_f32": updates / bp32 / 1e9,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_T, "unit": "T lane-instr/s",
                     "frac": achieved / VALU_PEAK_T, "traffic": traffic, "compulsory_bytes": compulsory,
                     "traffic_over_compulsory": traffic / compulsory if traffic else None,
                     "valu_per_update": VALU_PER_UPDATE},
        # against the published FP32 vector peak: the update's arithmetic is the three
        # FMAs of the bilinear form and the accumulating FMA, 8 flops
        "roofline_fp32_peak": {"bound": "fp32", "achieved": FLOPS_PER_UPDATE * updates / bp32 / 1e12,
                               "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                               "frac": FLOPS_PER_UPDATE * updates / bp32 / 1e12 / FP32_PEAK_TFLOPS,
                               "flops_per_update": FLOPS_PER_UPDATE},
        # the other on-chip ceiling: one 16-byte FP32 entry read from LDS per update,
        # against 256 B per clock per CU (256 CUs at 2.4 GHz), and against the measured
        # rate of conflict-free 16-byte LDS reads
        "roofline_lds": {"bound": "lds", "achieved": 16.0 * updates / bp32 / 1e12, "peak": LDS_PEAK_TBS,
                         "unit": "TB/s", "frac": 16.0 * updates / bp32 / 1e12 / LDS_PEAK_TBS, "bytes_per_update": 16,
                         "peak_measured_ds_read_b128": LDS_READ_MEASURED_TBS,
                         "frac_of_measured": 16.0 * updates / bp32 / 1e12 / LDS_READ_MEASURED_TBS},
        # the 8-byte read-modify-write model of the metric's definition (the reference
        # kernel's per-update voxel load and store, Ct/cuda_ct.cu:193-199): 1000 GUPS
        # at 8 TB/s, against which a 60 % HBM-roofline target reads 600 GUPS
        "hbm_8b_per_update_frac": BYTES_PER_UPDATE * updates / bp32 / 1e9 / HBM_PEAK_GBS,
        "hbm_copy_measured_gbs": copy_gbs,
        "cpu_baseline": cpu,
    }
    if "f16" in res:
        t16, bp16 = res["f16"]
        out.update({"ms_per_recon_f16": t16 * 1e3, "bp_ms_f16": bp16 * 1e3, "bp_gups_f16": updates / bp16 / 1e9,
                    "f16_over_f32_bp": bp32 / bp16, "f16_over_f32_step": t32 / t16,
                    # the FP16-entry kernel is VALU-bound with the same 7 lane-instructions per
                    # update (interpolation in FP32): at 100 % of the issue rate it would run at
                    # VALU_PEAK_T / 7 updates/s, which caps its ratio to this FP32 run
                    "roofline_f16": {"bound": "valu", "achieved": VALU_PER_UPDATE * updates / bp16 / 1e12,
                                     "peak": VALU_PEAK_T, "unit": "T lane-instr/s",
                                     "frac": VALU_PER_UPDATE * updates / bp16 / 1e12 / VALU_PEAK_T,
                                     "valu_per_update": VALU_PER_UPDATE,
                                     "f16_over_f32_bp_ceiling": VALU_PEAK_T * 1e12 / VALU_PER_UPDATE / (updates / bp32)}})
    if e2e is not None:
        out["ms_e2e_from_pinned_host_f32"] = e2e[0] * 1e3
        out["ms_reconstruct_host_to_host_f32"] = h2h * 1e3
        out["ms_h2d_raw"] = e2e[1] * 1e3
    if not args.no_roi:
        out.update(roi_run(ctx, torch, raw, P, base, filt, stream, args.steps))
    out.update(prec)
    if writes:
        out.update(writes)
    print(json.dumps(out), flush=True)


def roi_run(ctx, torch, raw, P, base, filt, stream, steps):
    """The same scan reconstructed on CtVolume's own grid (2R x 2R x H) restricted to
    x, y in [0.25, 0.75] and z in [0.1, 0.9] with setVolumeBounds semantics."""
    import ct_amd
    g0 = ctx.geometry
    g, roi = ct_amd.geometry_ctvolume(g0.det_width, g0.det_height, g0.fcd, g0.u_offset,
                                      [0.25, 0.75, 0.25, 0.75, 0.1, 0.9])
    ctx.set_geometry(g)
    dev = raw.device
    filtered = torch.empty(ctx.filtered_bytes("f32", P), dtype=torch.uint8, device=dev)
    scales = torch.empty(P, dtype=torch.float32, device=dev)
    vol = torch.empty((g.nx, g.ny, g.nz), dtype=torch.float32, device=dev)

    def step():
        ctx.preprocess(raw, "u16", P, base, "cpu", filt, "f32", filtered, scales, stream)
        ctx.backproject(filtered, "f32", scales, 0, P, vol, 0, g.nz, False, stream)

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(1, steps)):
        step()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / max(1, steps)
    ups = float(g.nx) * g.ny * g.nz * P
    log(f"ROI {g.nx}x{g.ny}x{g.nz} (x {roi[0]}..{roi[1]}, z {roi[4]}..{roi[5]}): {t * 1e3:.1f} ms, "
        f"{ups / t / 1e9:.1f} GUPS")
    ctx.set_geometry(g0)
    return {"roi_grid": [g.nx, g.ny, g.nz], "roi_ms_f32": t * 1e3, "roi_gups_f32": ups / t / 1e9}


if __name__ == "__main__":
    main()
