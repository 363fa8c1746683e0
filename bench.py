"""Benchmark: Siddon forward + exact-adjoint backprojection on one MI355X.

Workload (one "step"): forward-project a 256^3 float32 phantom of 20 random
ellipsoids along every ray of a 3-D cone-beam scan -- 360 views over 2*pi,
source-isocentre 500 mm, source-detector 1000 mm, 512 x 384 detector at
1.0 mm, 0.5 mm voxels -- and backproject the resulting sinogram with the
exact adjoint.  70,778,880 rays per step; inputs resident in HBM.

Prints ONE JSON line (rank 0): value = rays/s = rays * steps / timed seconds,
roofline of the dominant kernel from HIP-event kernel timings, and the CPU
baseline (the C oracle port on the host cores, on a bounded ray sample).

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
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched-volume, fan-beam SIRT, differentiable-projector and serving sections")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving section (profiling runs)")
    ap.add_argument("--no-fp64", action="store_true",
                    help="skip the FP64 error checks (profiling runs)")
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="requests of the Poisson trace replayed by the serving section")
    ap.add_argument("--plain-requests", type=int, default=6,
                    help="requests run through the kernel-per-call path for comparison")
    ap.add_argument("--plain-small", type=int, default=12,
                    help="small requests (at most 128^2 / 128^3) compared with the per-call path")
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


def l2_hit_rate(kernel):
    """L2 (TCC) hit rate of `kernel` from the committed rocprofv3
    TCC_HIT / TCC_MISS run of this workload, or None when not profiled."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "l2_hit.json")
    try:
        with open(path) as f:
            return float(json.load(f)["kernels"][kernel]["hit_rate"])
    except (OSError, KeyError, ValueError, TypeError):
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# --------------------------------------------------------- FP64 references
#
# The metric is quoted at RMSE <= 1e-5 against FP64.  The references below
# evaluate the same operators in float64 with torch on the device: the
# Siddon matrix from the device trace (siddons(), bit-exact with the
# reference's), and the distance-driven fan and separable-footprint
# parallel-3-D weights restated in float64 (on the float32 cos/sin the
# projectors use, as the reference evaluates np.cos of float32 angles).

def _err(got, ref):
    import torch
    g = got.reshape(-1).double()
    r = ref.reshape(-1).double()
    d = g - r
    return {"rel_rmse": float(d.norm() / r.norm()), "max_abs": float(d.abs().max()),
            "max_abs_over_max": float(d.abs().max() / r.abs().max())}


def _bnd64(n, d, dev):
    import torch
    return d * (torch.arange(n + 1, dtype=torch.float64, device=dev) - n / 2.0)


def _coo_apply(rows, cols, vals, x, n_out):
    """y[rows] += vals * x[cols] in float64."""
    import torch
    y = torch.zeros(n_out, dtype=torch.float64, device=x.device)
    return y.index_add_(0, rows, vals * x[cols])


def dd_fan_arc_fp64_matrix(img_shape, cosv, sinv, nu, du, DSO, DSD, d=1.0, su=0.0, dev="cuda"):
    """COO (rows = view*nu + bin, cols = pixel, vals) of the distance-driven
    fan projector with an equiangular detector in float64: per view the
    driving axis is x when |sin| >= |cos|; each line's driving-axis pixel
    boundaries project to DSD*atan2(p, DSO - o); a pixel weighs a bin by the
    overlap of its projected interval times r/((DSO - o_line)|sin|) (sys_mat/
    dd.py:164-221, 526-585 on an arc), and the forward output is scaled by
    d/du.  cosv, sinv: the projector's float32 cos/sin per view."""
    import torch
    nx, ny = img_shape
    ub = du * (torch.arange(nu + 1, dtype=torch.float64, device=dev) - nu / 2.0 + su)
    R, C, V = [], [], []
    for a, (c, s) in enumerate(zip(np.asarray(cosv, np.float64), np.asarray(sinv, np.float64))):
        c, s = float(c), float(s)
        xd = abs(s) >= abs(c)
        if xd:
            pd_, po, od, oo, rs, nd, no = -s, c, c, s, abs(s), nx, ny
        else:
            pd_, po, od, oo, rs, nd, no = c, -s, s, c, abs(c), ny, nx
        db = _bnd64(nd, d, dev)[None, :]
        ob = _bnd64(no, d, dev)
        oc = ((ob[:-1] + ob[1:]) / 2)[:, None]
        pb = DSD * torch.atan2(pd_ * db + po * oc, DSO - (od * db + oo * oc))   # (no, nd+1)
        den = DSO - oo * oc
        sc = torch.sqrt(den * den + (po * oc) ** 2) / (den * rs)
        lo = torch.minimum(pb[:, :-1], pb[:, 1:])
        hi = torch.maximum(pb[:, :-1], pb[:, 1:])
        io = torch.arange(no, device=dev)[:, None]
        k = torch.arange(nd, device=dev)[None, :]
        pix = (k * ny + io) if xd else (io * ny + k)
        j0 = torch.floor((lo - ub[0]) / du).long()
        span = int(torch.ceil((hi - lo).max() / du).item()) + 1
        for t in range(span + 1):
            j = j0 + t
            jc = j.clamp(0, nu - 1)
            ov = torch.minimum(hi, ub[jc + 1]) - torch.maximum(lo, ub[jc])
            ok = (j >= 0) & (j < nu) & (ov > 0)
            R.append(a * nu + jc[ok])
            C.append(pix.expand_as(j)[ok])
            V.append((ov * sc.expand_as(ov))[ok])
    return torch.cat(R), torch.cat(C), torch.cat(V) * (d / du)


def sf_par3d_fp64(vol, cosv, sinv, nu, nv, du=1.0, dv=1.0, d=1.0, su=0.0, sv=0.0, y=None):
    """Separable-footprint parallel 3-D projector in float64 (one volume):
    pixel footprint [p_c - h, p_c + h], h = d(|sin|+|cos|)/2, amplitude
    1/(|sin|+|cos|) (sys_mat/pd.py:129-234), times the overlap of each slab
    with the detector rows; forward scaled by d.  cosv, sinv: the
    projector's float32 cos/sin per view.  Returns A vol, and A^T y when y
    is given."""
    import torch
    dev = vol.device
    nx, ny, nz = vol.shape
    ub = du * (torch.arange(nu + 1, dtype=torch.float64, device=dev) - nu / 2.0 + su)
    vb = dv * (torch.arange(nv + 1, dtype=torch.float64, device=dev) - nv / 2.0 + sv)
    zb = _bnd64(nz, d, dev)
    Wv = (torch.minimum(zb[1:, None], vb[None, 1:]) - torch.maximum(zb[:-1, None], vb[None, :-1]))
    Wv = Wv.clamp(min=0) / dv                                              # (nz, nv)
    X = vol.double().reshape(nx * ny, nz) @ Wv                              # (npix, nv)
    xb, yb = _bnd64(nx, d, dev), _bnd64(ny, d, dev)
    xc, yc = ((xb[:-1] + xb[1:]) / 2)[:, None], ((yb[:-1] + yb[1:]) / 2)[None, :]
    na = len(cosv)
    sino = torch.zeros(na, nu, nv, dtype=torch.float64, device=dev)
    back = torch.zeros(nx * ny, nv, dtype=torch.float64, device=dev) if y is not None else None
    for a, (c, s) in enumerate(zip(np.asarray(cosv, np.float64), np.asarray(sinv, np.float64))):
        c, s = float(c), float(s)
        w = abs(s) + abs(c)
        pc = (-s * xc + c * yc).reshape(-1)
        lo, hi = pc - 0.5 * d * w, pc + 0.5 * d * w
        j0 = torch.floor((lo - ub[0]) / du).long()
        for t in range(int(np.ceil(d * w / du)) + 2):
            j = j0 + t
            jc = j.clamp(0, nu - 1)
            ov = torch.minimum(hi, ub[jc + 1]) - torch.maximum(lo, ub[jc])
            ok = (j >= 0) & (j < nu) & (ov > 0)
            pix = torch.nonzero(ok).reshape(-1)
            wt = (ov[pix] / du / w)[:, None]
            sino[a].index_add_(0, jc[pix], wt * X[pix])
            if y is not None:
                back.index_add_(0, pix, wt * y[a].double()[jc[pix]])
    fp = d * sino
    if y is None:
        return fp, None
    return fp, d * (back @ Wv.T).reshape(nx, ny, nz)


def fp64_siddon_check(vol, views=(0, 90, 180, 270)):
    """Siddon FP/BP at the full grid and detector on four views against the
    float64 product of the same device trace."""
    import torch
    from vir_amd.sys_mat.siddon import SiddonProjector
    from vir_amd.workloads import cone256_geometry
    src, trg, grid = cone256_geometry(np.asarray(views))
    P = SiddonProjector(src, trg, grid["nPixels"], grid["dPixels"], structured=True)
    f32 = P.fp(vol)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    y = torch.rand(P.ray_shape, device="cuda", generator=gen)
    b32 = P.bp(y)
    off, rav, ln, _ = P.stored_csr()
    ray = torch.repeat_interleave(torch.arange(P.nrays, device="cuda"), off[1:] - off[:-1])
    keep = rav >= 0
    ray, idx, ln64 = ray[keep], rav[keep].long(), ln[keep].double()
    f64 = _coo_apply(ray, idx, ln64, vol.reshape(-1).double(), P.nrays)
    b64 = _coo_apply(idx, ray, ln64, y.reshape(-1).double(), P.nvox)
    return {"views": list(views), "rays": int(P.nrays), "fp": _err(f32, f64), "bp": _err(b32, b64)}


def cpu_baseline(src, trg, grid, target_s):
    """The reference's CPU path -- siddons() trace, CSR forward projection
    with a float64 accumulator (parallel over rays, as proj_numba's prange
    path) and CSR backprojection -- as the C oracle port on every host core
    this process may use (at most 16, the box's share), on a bounded
    contiguous sample of rays of the same workload."""
    root = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, root)
    from oracle import siddon as osd
    nP, dP = grid["nPixels"], grid["dPixels"]
    s = src.reshape(-1, 3)
    t = trg.reshape(-1, 3)
    rng = np.random.default_rng(0)
    vol = rng.uniform(0, 1, nP).astype(np.float32)
    nth = max(1, min(16, len(os.sched_getaffinity(0))))

    def run(n, start, threads):
        t0 = time.perf_counter()
        osd.fp_bp_threads(s[start:start + n], t[start:start + n], vol, nP, dP, nthreads=threads)
        return time.perf_counter() - t0

    def sample(threads, target):
        # grow the sample until it takes about `target` seconds (fixed set-up
        # costs make small samples look slow)
        n = 8192 * threads
        dt = run(n, start, threads)
        while dt < 0.6 * target and n < avail:
            n = int(min(avail, max(2 * n, n * target / max(dt, 1e-3))))
            dt = run(n, start, threads)
        return n, dt

    start = s.shape[0] // 4
    avail = s.shape[0] - start
    n, dt = sample(nth, target_s)
    n1, dt1 = sample(1, target_s / 2)
    return {"value": n / dt, "unit": "rays/s", "cores": nth, "kind": "port",
            "sample": f"{n} consecutive rays of the same scan (from view {start // (512 * 384)}), "
                      f"C oracle siddons() trace + float64 CSR FP + CSR BP on {nth} OpenMP "
                      f"threads, {dt:.1f} s",
            "single_thread": {"value": n1 / dt1, "unit": "rays/s", "cores": 1,
                              "sample": f"{n1} consecutive rays from the same start, {dt1:.1f} s"},
            "cpu_model": cpu_model()}


def events(n):
    import torch
    return [torch.cuda.Event(enable_timing=True) for _ in range(n)]


def sirt_fan_section(fp64=True):
    """512^2 ellipse phantom, 720 views over 2*pi, 888 equiangular bins
    (DSO 1000, DSD 1500 pixel units), Poisson noise I0 = 1e5 on the
    log-transformed sinogram, 50 graph-captured SIRT iterations with the
    distance-driven fan projector."""
    import torch
    from vir_amd.serve import fan_geometry
    from vir_amd.sirt import SIRT
    from vir_amd.sys_mat.dd import DDProjector
    from vir_amd.workloads import phantom
    g = fan_geometry(512, 720)
    P = DDProjector(g["img_shape"], g["ang"], g["nu"], "fan", DSO=g["DSO"], DSD=g["DSD"],
                    du=g["du"], detector="arc")
    x_true = phantom((512, 512), 10, seed=1003, supersample=4)
    p = P.fp(x_true)
    mu = 4.0 / float(p.max())
    gen = torch.Generator(device="cuda")
    gen.manual_seed(1003)
    I = torch.poisson(1e5 * torch.exp(-mu * p), generator=gen).clamp_(min=1.0)
    b = -torch.log(I / 1e5) / mu
    S = SIRT(P, b, graph=True)
    S.step()                       # capture + first iteration
    torch.cuda.synchronize()
    S = SIRT(P, b, graph=True)
    S.step()
    torch.cuda.synchronize()
    e0, e1 = events(2)
    e0.record()
    for _ in range(49):
        S.step()
    e1.record()
    torch.cuda.synchronize()
    ms_it = e0.elapsed_time(e1) / 49
    fp64 = fan_fp64_check(P, g, b, x_true) if fp64 else None
    torch.cuda.empty_cache()
    n_views, n_d, n_v = 720, g["nu"], 512 * 512
    b_fp = n_views * (4.0 * n_v + 4.0 * n_d)
    alg = 2 * b_fp + 16.0 * n_d * n_views + 16.0 * n_v
    rmse = float((S.x - x_true).norm() / x_true.norm())
    # FBP (Hann window) as the initial image: its time, and the SIRT
    # iterations from it that reach the error of 50 iterations from zero
    from vir_amd.fbp import FBP
    F = FBP(g["img_shape"], g["ang"], g["nu"], "fan", DSO=g["DSO"], DSD=g["DSD"], du=g["du"],
            detector="arc", filter_type="hann")
    x0 = F(b)
    torch.cuda.synchronize()
    f0, f1 = events(2)
    f0.record()
    x0 = F(b)
    f1.record()
    torch.cuda.synchronize()
    fbp_ms = f0.elapsed_time(f1)
    err0 = float((x0 - x_true).norm() / x_true.norm())
    W = SIRT(P, b, x0=x0, graph=True, W_sino=S.W_sino, W_img=S.W_img)
    n_warm = None
    for it in range(1, 51):
        W.step()
        if float((W.x - x_true).norm() / x_true.norm()) <= rmse:
            n_warm = it
            break
    return {"workload": "fan2d_dd_sirt_512", "views": n_views, "bins": n_d, "iterations": 50,
            "fp64_error": fp64,
            "ms_per_iteration": ms_it, "iterations_per_s": 1e3 / ms_it,
            "ms_50_iterations": 50 * ms_it,
            "alg_bytes_per_iteration": alg, "alg_GBps": alg / (ms_it * 1e-3) / 1e9,
            "rel_error_vs_phantom": rmse, "noise": "Poisson I0=1e5, mu=4/max(p)",
            "fbp_init": {"filter": "hann", "ms": fbp_ms, "rel_error_vs_phantom": err0,
                         "sirt_iterations_to_match_50_from_zero": n_warm}}


def fan_fp64_check(P, g, b, x_true, iters=2):
    """The fan projector pair and `iters` SIRT iterations (graph-replayed,
    as timed) at the full 512^2 / 720-view / 888-bin configuration against
    the float64 operator and recurrence."""
    import torch
    from vir_amd.sirt import SIRT
    R, C, V = dd_fan_arc_fp64_matrix(g["img_shape"], P.cos.cpu().numpy(), P.sin.cpu().numpy(),
                                     g["nu"], g["du"], g["DSO"], g["DSD"])
    nsino, npix = P.na * g["nu"], int(np.prod(g["img_shape"]))

    def A(x):
        return _coo_apply(R, C, V, x.reshape(-1).double(), nsino)

    def At(y):
        return _coo_apply(C, R, V, y.reshape(-1).double(), npix)

    gen = torch.Generator(device="cuda")
    gen.manual_seed(8)
    y = torch.rand(P.sino_shape, device="cuda", generator=gen)
    out = {"fp": _err(P.fp(x_true), A(x_true)), "bp": _err(P.bp(y), At(y))}
    b64 = b.reshape(-1).double()
    Ws = A(torch.ones(npix, device="cuda"))
    Wi = At(torch.ones(nsino, device="cuda"))
    ms, mi = Ws != 0, Wi != 0
    x = torch.zeros(npix, dtype=torch.float64, device="cuda")
    T = SIRT(P, b, graph=True)
    for _ in range(iters):
        r = A(x)
        r[ms] = (b64[ms] - r[ms]) / Ws[ms]
        u = At(r)
        u[mi] /= Wi[mi]
        x += u
        T.step()
    out["sirt_%d_iterations" % iters] = _err(T.x, x)
    out["reference"] = "float64 weights and recurrence (torch, on the device)"
    return out


def autograd_section(fp64=True):
    """Batch of 8 random-ellipsoid volumes 256x256x64, 3-D parallel beam,
    180 views over [0, pi), 384 x 64 detector at 1.0: 100 gradient steps on
    0.5*||A x - b||^2 + 1e-3*TV(x), A the separable-footprint projector
    inside torch.autograd."""
    import torch
    from vir_amd.autograd import least_squares, project, total_variation
    from vir_amd.sys_mat.pd import SFProjector
    from vir_amd.workloads import phantom_device
    nb, shape = 8, (256, 256, 64)
    ang = np.arange(180) * np.pi / 180
    P = SFProjector(shape, ang, 384, "par3d", nv=64, du=1.0, dv=1.0, batch=nb)
    truth = torch.stack([phantom_device(shape, 20, 1004 + k) for k in range(nb)])
    b = P.fp(truth)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(1004)
    xr = torch.randn(truth.shape, device="cuda", generator=gen)
    yr = torch.randn(b.shape, device="cuda", generator=gen)
    lhs = float((P.fp(xr).double() * yr.double()).sum())
    rhs = float((xr.double() * P.bp(yr).double()).sum())
    adj = abs(lhs - rhs) / abs(lhs)
    # FP64: one volume of the batch on 4 of the 180 views
    if fp64:
        sub = ang[::45]
        P1 = SFProjector(shape, sub, 384, "par3d", nv=64, du=1.0, dv=1.0)
        y1 = torch.rand(P1.sino_shape, device="cuda", generator=gen)
        f64, b64 = sf_par3d_fp64(truth[0], P1.cos.cpu().numpy(), P1.sin.cpu().numpy(), 384, 64, y=y1)
        fp64 = {"views": [int(round(float(a) * 180 / np.pi)) for a in sub],
                "fp": _err(P1.fp(truth[0]), f64), "bp": _err(P1.bp(y1), b64)}
        del P1, y1, f64, b64
    else:
        fp64 = None
    v = xr / xr.norm()
    for _ in range(8):             # power iteration for ||A||^2
        w = P.bp(P.fp(v))
        L = float(w.norm())
        v = w / L
    lam, eps = 1e-3, 1e-2
    lr = 1.0 / (L + lam * 12.0 / eps)

    def loss_fn(x):
        return least_squares(project(x, P), b) + lam * total_variation(x, eps=eps)

    x = torch.zeros_like(truth)
    l0 = float(loss_fn(x))
    e0, e1 = events(2)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(100):
        x.requires_grad_(True)
        loss = loss_fn(x)
        (gx,) = torch.autograd.grad(loss, x)
        x = (x.detach() - lr * gx)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 100
    l1 = float(loss_fn(x))
    b_fp = nb * 180 * (4.0 * (256 * 256 * 64) + 4.0 * (384 * 64))
    return {"workload": "par3d_sf_autograd_8x256x256x64", "steps": 100, "ms_per_step": ms,
            "steps_per_s": 1e3 / ms, "adjoint_rel_err": adj, "fp64_error": fp64,
            "loss_ratio_after_100": l1 / l0,
            "alg_bytes_fp": b_fp, "lr": lr}


def batched_section(P, vol, fp_ms, bp_ms, steps=3):
    """FP and BP of B volumes on the same rays (one trace per group): time
    per call and per volume against the single-volume calls, and item 0
    checked bit for bit against the single call."""
    import torch
    res = {}
    single_f = P.fp(vol)
    single_b = P.bp(single_f)
    for nb in (2, 4):
        vols = torch.stack([vol * (1.0 + 0.25 * j) for j in range(nb)])
        sinos = P.fp_batch(vols)
        backs = P.bp_batch(sinos)
        torch.cuda.synchronize()
        same = bool(torch.equal(sinos[0].view(torch.int32), single_f.view(torch.int32)) and
                    torch.equal(backs[0].view(torch.int32), single_b.view(torch.int32)))
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        tf, tb = [], []
        for _ in range(steps):
            e[0].record()
            P.fp_batch(vols, out=sinos)
            e[1].record()
            P.bp_batch(sinos, out=backs)
            e[2].record()
            torch.cuda.synchronize()
            tf.append(e[0].elapsed_time(e[1]))
            tb.append(e[1].elapsed_time(e[2]))
        f, b = float(np.mean(tf)), float(np.mean(tb))
        res[f"batch{nb}"] = {"fp_ms": f, "bp_ms": b, "fp_ms_per_volume": f / nb, "bp_ms_per_volume": b / nb,
                             "fp_speedup_per_volume": fp_ms * nb / f, "bp_speedup_per_volume": bp_ms * nb / b,
                             "item0_equals_single_call": same}
        del vols, sinos, backs
        torch.cuda.empty_cache()
    return res


def serving_section(n_serve, n_plain, n_small):
    """The serving trace (Poisson arrivals at 200/s, 60 % fan-beam jobs of
    128..1024^2 and 180..720 views, 40 % cone-beam jobs of 64..256^3 and
    120..360 views, 20 SIRT iterations each): the first n_serve requests
    replayed in real time under a 200 GB budget, with the projector cache
    (32 GB) and same-geometry grouping on; the first n_plain requests also
    run through the kernel-per-call path and through a fresh server (empty
    cache) with all arrivals at once, and so do the first n_small requests
    of at most 128^2 / 128^3."""
    import dataclasses
    from vir_amd.serve import Server, SirtBackend, plain_sirt, poisson_trace, summarize
    import torch
    trace = poisson_trace(10000, 200.0, seed=1005)
    warm = [dataclasses.replace(r, rid=i, arrival=0.0, iters=2) for i, r in
            enumerate([q for q in trace[:40] if q.n <= 128][:4])]
    Server(SirtBackend(n_streams=8, cache_bytes=0), 200 << 30, max_active=16).run(
        warm, keep_results=False)
    be = SirtBackend(n_streams=8, cache_bytes=32 << 30)
    sample = trace[:n_serve]
    srv = Server(be, 200 << 30, max_active=16, group=True, batch=True)
    res = srv.run(sample, keep_results=False,
                  progress=lambda r: log(f"[serve] request {r.rid} done in {r.latency:.2f}s")
                  if r.rid % 16 == 0 else None)
    out = summarize(res, srv.elapsed)
    out.update({"requests": n_serve, "arrival_rate_per_s": 200.0,
                "max_concurrent": srv.max_concurrent, "max_in_use_GB": srv.max_in_use / 1e9,
                "projector_cache": {"hits": be.cache_hits, "misses": be.cache_misses,
                                    "limit_GB": 32},
                "grouped_admissions": srv.grouped, "batched_admissions": srv.batched,
                "trace": "poisson_trace(10000, 200/s, seed 1005)[:%d]" % n_serve})
    del be, srv, res
    torch.cuda.empty_cache()
    sub = [dataclasses.replace(r, arrival=0.0) for r in trace[:n_plain]]
    t0 = time.perf_counter()
    done = 0
    for r in sub:
        plain_sirt(r)
        done += 1
        log(f"[serve] plain request {r.rid} ({r.kind} {r.n} x {r.views}) done")
        if time.perf_counter() - t0 > 90.0:
            break
    plain_s = time.perf_counter() - t0
    srv2 = Server(SirtBackend(n_streams=8, cache_bytes=0), 200 << 30, max_active=16)
    res2 = srv2.run(sub[:done], keep_results=False)
    served = summarize(res2, srv2.elapsed)
    out["per_call_path"] = {"requests": done, "jobs_per_s": done / plain_s,
                            "served_jobs_per_s_same_requests": served["jobs_per_s"],
                            "served_over_per_call": served["jobs_per_s"] / (done / plain_s)}
    del srv2, res2
    torch.cuda.empty_cache()
    # the same comparison on the trace's small jobs (fan 128^2 and cone
    # 64^3 / 128^3), whose per-call cost is host work rather than kernels
    small = [dataclasses.replace(r, arrival=0.0) for r in trace if r.n <= 128][:n_small]
    t0 = time.perf_counter()
    done = 0
    for r in small:
        plain_sirt(r)
        done += 1
        if time.perf_counter() - t0 > 60.0:
            break
    plain_s = time.perf_counter() - t0
    log(f"[serve] plain path: {done} small requests in {plain_s:.1f}s")
    srv3 = Server(SirtBackend(n_streams=8, cache_bytes=0), 200 << 30, max_active=16)
    res3 = srv3.run(small[:done], keep_results=False)
    served = summarize(res3, srv3.elapsed)
    out["per_call_path_small_jobs"] = {
        "requests": done, "sizes": sorted({f"{r.kind} {r.n}" for r in small[:done]}),
        "jobs_per_s": done / plain_s, "served_jobs_per_s_same_requests": served["jobs_per_s"],
        "served_over_per_call": served["jobs_per_s"] / (done / plain_s)}
    return out


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
                     "alg_bytes": dom_bytes, "l2_hit_rate": l2_hit_rate(dom_kernel)},
    }
    del sino, back
    if not args.no_extras:
        log("[bench] batched volumes on the same rays")
        out["batched"] = batched_section(P, vol, fp_ms, bp_ms)
        log(f"[bench] {out['batched']}")
    if not args.no_fp64:
        log("[bench] FP64 check on 4 views")
        out["fp64_error"] = fp64_siddon_check(vol)
        torch.cuda.empty_cache()
        log(f"[bench] {out['fp64_error']}")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(src, trg, grid, args.cpu_seconds)
    if not args.no_extras:
        del P, vol, src, trg
        torch.cuda.empty_cache()
        log("[bench] fan-beam SIRT section")
        out["sirt_fan"] = sirt_fan_section(fp64=not args.no_fp64)
        log(f"[bench] {out['sirt_fan']}")
        log("[bench] differentiable projector section")
        out["autograd"] = autograd_section(fp64=not args.no_fp64)
        log(f"[bench] {out['autograd']}")
        torch.cuda.empty_cache()
        if not args.no_serve:
            log("[bench] serving section")
            out["serving"] = serving_section(args.serve_requests, args.plain_requests, args.plain_small)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
