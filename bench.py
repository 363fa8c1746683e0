"""Benchmark: Siddon forward + exact-adjoint backprojection on one MI355X.

Workload (one "step"): forward-project a 256^3 float32 phantom of 20 random
ellipsoids along every ray of a 3-D cone-beam scan -- 360 views over 2*pi,
source-isocentre 500 mm, source-detector 1000 mm, 512 x 384 detector at
1.0 mm, 0.5 mm voxels -- and backproject the resulting sinogram with the
exact adjoint.  70,778,880 rays per step; inputs resident in HBM.

Prints ONE JSON line (rank 0): value = rays/s = rays * steps / timed seconds,
roofline of the dominant kernel from HIP-event kernel timings, and the CPU
baseline (the C oracle port, single thread, on a bounded ray sample).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--views", type=int, default=360, help="views to use (default: all 360)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target CPU time of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run a 1 GiB device copy (PMC calibration)")
    return ap.parse_args()


def measured_traffic(kernel):
    """HBM-side bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs of this workload (calibrated against a
    device copy of known size), or None when not profiled."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "cone256_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        k = t["kernels"][kernel]
        return (k["fetch_kib"] * t["fetch_scale"] + k["write_kib"] * t["write_scale"]) * 1024.0
    except (OSError, KeyError, ValueError):
        return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(src, trg, grid, target_s):
    """The reference's CPU path -- siddons() trace to CSR, CSR forward
    projection and CSR backprojection -- as the C oracle port (single
    thread), on a bounded contiguous sample of rays of the same workload."""
    root = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, root)
    from oracle import projection as op
    from oracle import siddon as osd
    nP, dP = grid["nPixels"], grid["dPixels"]
    s = src.reshape(-1, 3)
    t = trg.reshape(-1, 3)
    rng = np.random.default_rng(0)
    vol = rng.uniform(0, 1, nP).astype(np.float32)
    nvox = int(np.prod(nP))

    def run(n, start):
        t0 = time.perf_counter()
        ind, ln, off = osd.siddons(s[start:start + n], t[start:start + n], nP, dP, 0.0, flat=True)
        sino = op.fp64(vol, ind, ln, off).astype(np.float32)
        op.bp64(sino, ind, ln, off, nvox)
        return time.perf_counter() - t0

    start = s.shape[0] // 2
    n = 2048
    dt = run(n, start)
    n = int(min(s.shape[0] - start, max(n, n * target_s / max(dt, 1e-3))))
    dt = run(n, start)
    return {"value": n / dt, "unit": "rays/s", "cores": 1, "kind": "port",
            "sample": f"{n} consecutive rays of the same scan (view {start // (512 * 384)}), "
                      f"C oracle siddons() trace + float64 CSR FP + BP, {dt:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it "
              "(use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import torch
    from vir_amd import _lib
    from vir_amd.sys_mat.siddon import SiddonProjector
    from vir_amd.workloads import cone256_geometry, phantom

    if args.calib_copy:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        b = torch.empty_like(a)
        a.fill_(1.0)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes": a.numel() * 4}))
        return

    t0 = time.time()
    views = np.arange(args.views) if args.views < 360 else None
    src, trg, grid = cone256_geometry(views)
    log(f"[bench] geometry {src.shape} in {time.time() - t0:.1f}s")
    P = SiddonProjector(src, trg, grid["nPixels"], grid["dPixels"], structured=True)
    log(f"[bench] projector ready in {time.time() - t0:.1f}s")
    vol = phantom(grid["nPixels"], 20, seed=1002)
    nrays = P.nrays
    nvox = P.nvox

    # exact intersection count nnz (algorithmic bytes) from the trace kernel
    counts = torch.empty(nrays, dtype=torch.int32, device="cuda")
    flags = torch.empty(nrays, dtype=torch.uint8, device="cuda")
    _lib.check(_lib.lib().vir_siddon_count(_lib.ptr(P.src), _lib.ptr(P.trg), nrays, P.grid,
                                           _lib.ptr(counts), _lib.ptr(flags), _lib.stream_ptr()),
               "count")
    nnz = int(counts.to(torch.int64).sum().item())
    bad = int(((flags & 1) != 0).sum().item())
    del counts, flags
    log(f"[bench] nrays={nrays} nnz={nnz} out-of-grid={bad}")

    sino = torch.empty(P.ray_shape, dtype=torch.float32, device="cuda")
    back = torch.empty(P.nPixels, dtype=torch.float32, device="cuda")
    for i in range(args.warmup):
        P.fp(vol, out=sino)
        P.bp(sino, out=back)
        torch.cuda.synchronize()
        log(f"[bench] warmup {i + 1}/{args.warmup}")

    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        e0, e1, e2 = ev[i]
        e0.record()
        P.fp(vol, out=sino)
        e1.record()
        P.bp(sino, out=back)
        e2.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    fp_ms = float(np.mean([a.elapsed_time(b) for a, b, _ in ev]))
    bp_ms = float(np.mean([b.elapsed_time(c) for _, b, c in ev]))
    log(f"[bench] steps={args.steps} wall={wall:.3f}s fp={fp_ms:.2f}ms bp={bp_ms:.2f}ms")

    # algorithmic bytes: one 4-byte operand fetch per (ray, voxel) entry plus
    # one 4-byte store per output element (+ 24 B of ray endpoints per ray)
    fp_bytes = 4.0 * nnz + 4.0 * nrays + 24.0 * nrays
    bp_bytes = 4.0 * nnz + 4.0 * nvox
    dom = "bp" if bp_ms >= fp_ms else "fp"
    dom_kernel = "k_siddon_bp_wave" if dom == "bp" else "k_siddon_fp"
    dom_bytes, dom_ms = (bp_bytes, bp_ms) if dom == "bp" else (fp_bytes, fp_ms)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    rays_per_s = nrays * args.steps / wall
    out = {
        "metric": "Siddon FP+BP rays/s; SIRT jobs/s + p99 latency at RMSE<=1e-5 vs FP64, 1 GPU",
        "value": rays_per_s,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "cone3d_siddon_fp_bp_256",
                   "volume": [256, 256, 256], "voxel_mm": 0.5, "views": int(P.layout[0]),
                   "detector": [512, 384], "pitch_mm": 1.0, "sid_mm": 500, "sdd_mm": 1000,
                   "rays": nrays, "nnz": nnz, "fp_ms": fp_ms, "bp_ms": bp_ms},
        "roofline": {"bound": "hbm", "kernel": dom_kernel,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom_kernel),
                     "alg_bytes": dom_bytes},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(src, trg, grid, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
