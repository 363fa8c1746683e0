"""Capture golden vectors for the Siddon system-matrix builder.

Runs ONLY in the build container, where the reference tree is mounted at
/root/reference.  It loads the reference's own numpy-only modules by file
path -- core/core_funcs.py (Grid3d, np_array), sys_mat/siddon.py (siddons)
and proj_geom.py (geom_circular) -- and records their input/output vectors
into tests/golden/siddon_golden.npz.  Nothing else pins siddons() bit for bit
(its numpy dtype promotions and np.isclose snapping decide indices at ties),
so its outputs are captured here as data; the reference code itself is not
copied or shipped.

Usage:  python tests/golden/make_siddon_golden.py [--ref /root/reference]
"""
import argparse
import hashlib
import importlib.util
import os
import sys
import time
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_reference(ref):
    def load(name, path):
        spec = importlib.util.spec_from_file_location(name, path)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        return mod

    core = load("vir_ref_core_funcs", os.path.join(ref, "core", "core_funcs.py"))
    # siddon.py does `import vir` and calls vir.Grid3d / vir.np_array: expose
    # the reference's own core functions under the package name it expects.
    vir = types.ModuleType("vir")
    vir.Grid3d = core.Grid3d
    vir.np_array = core.np_array
    sys.modules["vir"] = vir
    siddon = load("vir_ref_siddon", os.path.join(ref, "sys_mat", "siddon.py"))
    geom = load("vir_ref_proj_geom", os.path.join(ref, "proj_geom.py"))
    return siddon, geom


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def run_flat(siddons, src, trg, nPixels, dPixels, origin, epsilon=1e-5):
    try:
        ind, ln, off = siddons(src, trg, nPixels, dPixels, origin, flat=True,
                               epsilon=epsilon)
        return ind, ln, off, 0
    except ValueError:
        return None, None, None, 1


def run_unravel(siddons, src, trg, nPixels, dPixels, origin, epsilon=1e-5):
    """ravel=False object output flattened to (x,y,z,len,off,hit)."""
    rays = siddons(src, trg, nPixels, dPixels, origin, epsilon=epsilon)
    rays = rays.reshape(-1)
    xs, ys, zs, ls, off, hit = [], [], [], [], [0], []
    for r in rays:
        if r is None:
            hit.append(0)
            off.append(off[-1])
            continue
        hit.append(1)
        xs.append(np.asarray(r[0], np.int32))
        ys.append(np.asarray(r[1], np.int32))
        zs.append(np.asarray(r[2], np.int32))
        ls.append(np.asarray(r[3], np.float32))
        off.append(off[-1] + len(r[3]))
    cat = lambda v, dt: np.concatenate(v).astype(dt) if v else np.zeros(0, dt)
    return (cat(xs, np.int32), cat(ys, np.int32), cat(zs, np.int32),
            cat(ls, np.float32), np.array(off, np.int64), np.array(hit, np.uint8))


def catalog_rays():
    """Edge-case catalog of tests/Test_Siddons/siddons_test1.py:28-127,
    grouped by grid.  Each entry: (src, trg)."""
    c4 = [
        ([-2.1, .25, .25], [2.0, .25, .25]), ([-2.0, .25, .25], [2.0, .25, .25]),
        ([-1.9, .25, .25], [2.0, .25, .25]), ([-2.0, .25, .25], [1.9, .25, .25]),
        ([-2.0, .25, .25], [2.0, .25, .25]), ([-2.0, .25, .25], [2.1, .25, .25]),
        ([2.1, .25, .25], [-2.0, .25, .25]), ([2.0, .25, .25], [-2.0, .25, .25]),
        ([1.9, .25, .25], [-2.0, .25, .25]), ([2.0, .25, .25], [-1.9, .25, .25]),
        ([2.0, .25, .25], [-2.0, .25, .25]), ([2.0, .25, .25], [-2.1, .25, .25]),
        ([.25, 4.0, .25], [.25, 4.0, .25]), ([-4.0, .25, .25], [-4.0, .25, .25]),
        ([2.0, -2.0, 0.0], [-2.0, -2.0, 0.0]), ([2.0, -1.0, 0.0], [-2.0, -1.0, 0.0]),
        ([2.0, 0.0, 0.0], [-2.0, 0.0, 0.0]), ([2.0, 1.0, 0.0], [-2.0, 1.0, 0.0]),
        ([2.0, 2.0, 0.0], [-2.0, 2.0, 0.0]),
        ([-2.1, -2.1, .25], [2.1, 2.1, .25]), ([-2.0, -2.0, .25], [2.0, 2.0, .25]),
        ([-1.9, -1.9, .25], [1.9, 1.9, .25]),
        ([-4.0, 0.0, .25], [0.0, 4.0, .25]), ([0.0, 4.0, .25], [4.0, 0.0, .25]),
        ([4.0, 0.0, .25], [0.0, 4.0, .25]), ([4.0, 0.0, .25], [0.0, -4.0, .25]),
    ]
    c128 = [
        ([-64.1, 0.0, .25], [64.1, 0.0, .25]),
        ([-64.1, -0.0001, .25], [64.1, 0.0001, .25]),
        ([-64.1, -2, .25], [64.1, 0.0001, .25]),
    ]
    return [((4, 4, 1), 1.0, 0.0, c4), ((128, 128, 1), 1.0, 0.0, c128)]


def random_rays(rng, n, nPixels, dPixels, origin):
    """Mixed population: generic, axis-aligned, grid-snapped and interior."""
    nP = np.array(nPixels, float)
    dP = np.broadcast_to(np.array(dPixels, float), (3,))
    oP = np.broadcast_to(np.array(origin, float), (3,))
    half = nP * dP / 2
    src = np.empty((n, 3))
    trg = np.empty((n, 3))
    for i in range(n):
        kind = i % 6
        a = rng.uniform(-1.6, 1.6, 3) * half + oP
        b = rng.uniform(-1.6, 1.6, 3) * half + oP
        if kind == 1:  # axis-aligned in one or two components
            ax = rng.integers(0, 3)
            b[ax] = a[ax]
            if rng.random() < 0.5:
                ax2 = (ax + 1) % 3
                b[ax2] = a[ax2]
        elif kind == 2:  # endpoints snapped to plane coordinates
            a = (np.round(a / dP * 2) / 2) * dP
            b = (np.round(b / dP * 2) / 2) * dP
        elif kind == 3:  # one endpoint inside the grid
            a = rng.uniform(-0.9, 0.9, 3) * half + oP
        elif kind == 4:  # rays through grid corners / lines
            c = (np.round(rng.uniform(-1, 1, 3) * nP / 2) * dP) + oP
            dvec = rng.normal(size=3)
            a = c - dvec * half.max() * 2
            b = c + dvec * half.max() * 2
        src[i], trg[i] = a, b
    return src, trg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--out", default=os.path.join(HERE, "siddon_golden.npz"))
    args = ap.parse_args()
    siddon, geom = load_reference(args.ref)
    siddons = siddon.siddons
    rng = np.random.default_rng(20260101)
    out = {}
    t0 = time.time()

    # A: edge-case catalog, unravelled object output + flat output
    k = 0
    for nPixels, dPix, orig, rays in catalog_rays():
        for (s, t) in rays:
            s = np.array(s)
            t = np.array(t)
            key = f"cat{k:02d}"
            out[key + "_src"] = s
            out[key + "_trg"] = t
            out[key + "_grid"] = np.array(list(nPixels) + [dPix, dPix, dPix, orig, orig, orig])
            x, y, z, ln, off, hit = run_unravel(siddons, s, t, nPixels, dPix, orig)
            out[key + "_u"] = np.concatenate([x, y, z]).astype(np.int32)
            out[key + "_len"] = ln
            out[key + "_off"] = off
            out[key + "_hit"] = hit
            ind, fl, fo, raised = run_flat(siddons, s, t, nPixels, dPix, orig)
            out[key + "_raised"] = np.array(raised)
            k += 1
    out["n_cat"] = np.array(k)

    # B: random rays on small anisotropic / offset grids (flat + unravel)
    cases = [((5, 7, 3), (0.7, 1.3, 0.9), (0.3, -0.2, 0.1), 1.0e-5),
             ((6, 4, 1), 1.0, 0.0, 1.0e-5),
             ((8, 8, 8), 0.5, 0.0, 1.0e-5),
             ((9, 3, 5), (1.0, 0.25, 2.0), (-1.5, 0.75, 0.0), 1.0e-4),
             ((16, 16, 1), 2.0, (0.5, 0.5, 0.0), 1.0e-5)]
    for ci, (nP, dP, oP, epsv) in enumerate(cases):
        src, trg = random_rays(rng, 600, nP, dP, oP)
        key = f"rnd{ci}"
        out[key + "_src"] = src
        out[key + "_trg"] = trg
        out[key + "_grid"] = np.concatenate([np.array(nP, float).reshape(-1),
                                             np.broadcast_to(np.array(dP, float), (3,)),
                                             np.broadcast_to(np.array(oP, float), (3,))])
        out[key + "_eps"] = np.array(epsv)
        x, y, z, ln, off, hit = run_unravel(siddons, src, trg, nP, dP, oP, epsv)
        out[key + "_u"] = np.concatenate([x, y, z]).astype(np.int32)
        out[key + "_len"] = ln
        out[key + "_off"] = off
        out[key + "_hit"] = hit
    out["n_rnd"] = np.array(len(cases))

    # C: geom_circular outputs (par 2D, fan 2D, par 3D, cone 3D)
    views = np.arange(12) * np.pi / 12
    dets = (np.arange(9) - 4.0) * 1.5
    detz = (np.arange(5) - 2.0) * 0.75
    gcases = {
        "par2d": dict(geom="par"),
        "fan2d": dict(geom="fan", src_iso=40.0, det_iso=25.0),
        "par3d": dict(geom="par", DetsZ=detz),
        "cone3d": dict(geom="cone", src_iso=40.0, det_iso=25.0, DetsZ=detz),
    }
    for name, kw in gcases.items():
        s, t = geom.geom_circular(dets, views, **kw)
        out["geom_" + name + "_src"] = s
        out["geom_" + name + "_trg"] = t
    out["geom_views"] = views
    out["geom_dets"] = dets
    out["geom_detz"] = detz

    # D: structured geometries, full CSR (flat=True) stored for small ones
    def structured(name, nP, dP, s, t):
        ind, ln, off, raised = run_flat(siddons, s, t, nP, dP, 0.0)
        out[name + "_src"] = s.astype(np.float32)
        out[name + "_trg"] = t.astype(np.float32)
        out[name + "_grid"] = np.array(list(nP) + [dP] * 3 + [0.0] * 3, float)
        out[name + "_raised"] = np.array(raised)
        if not raised:
            out[name + "_ind"] = ind
            out[name + "_flen"] = ln
            out[name + "_foff"] = off

    v = np.arange(30) * np.pi / 30
    d = (np.arange(47) - 23.0) * 1.0
    s, t = geom.geom_circular(d, v, geom="par")
    structured("par2d_small", (32, 32, 1), 1.0, s, t)
    v = np.arange(24) * 2 * np.pi / 24
    d = (np.arange(40) - 19.5) * 1.0
    s, t = geom.geom_circular(d, v, geom="fan", src_iso=60.0, det_iso=40.0)
    structured("fan2d_small", (24, 24, 1), 1.0, s, t)
    v = np.arange(10) * 2 * np.pi / 10
    d = (np.arange(20) - 9.5) * 1.0
    dz = (np.arange(12) - 5.5) * 1.0
    s, t = geom.geom_circular(d, v, geom="cone", src_iso=30.0, det_iso=30.0, DetsZ=dz)
    structured("cone3d_small", (16, 16, 12), 0.5, s, t)

    # E: the 2D parallel 128^2 workload (180 views x 185 bins): hashes of the
    # whole CSR plus full entries for a sample of rays
    v = np.arange(180) * np.pi / 180
    d = (np.arange(185) - 92.0) * 1.0
    s, t = geom.geom_circular(d, v, geom="par")
    ind, ln, off, raised = run_flat(siddons, s, t, (128, 128, 1), 1.0, 0.0)
    assert not raised
    out["par128_views"] = v
    out["par128_dets"] = d
    out["par128_src_sha"] = np.array(sha(s.astype(np.float32)))
    out["par128_trg_sha"] = np.array(sha(t.astype(np.float32)))
    out["par128_ind_sha"] = np.array(sha(ind))
    out["par128_len_sha"] = np.array(sha(ln))
    out["par128_off_sha"] = np.array(sha(off))
    out["par128_nnz"] = np.array(ind.size)
    out["par128_len_sum64"] = np.array(ln.astype(np.float64).sum())
    pick = np.sort(rng.choice(off.size - 1, 300, replace=False))
    out["par128_pick"] = pick
    out["par128_pick_off"] = np.array([off[i] for i in pick] + [0])
    out["par128_pick_cnt"] = np.array([off[i + 1] - off[i] for i in pick])
    out["par128_pick_ind"] = np.concatenate([ind[off[i]:off[i + 1]] for i in pick])
    out["par128_pick_len"] = np.concatenate([ln[off[i]:off[i + 1]] for i in pick])

    # F: rays sampled from the 256^3 cone-beam workload (0.5 mm voxels,
    # src-iso 500, det-iso 500, 512x384 detector at 1 mm, 360 views / 2pi)
    vsel = np.array([0, 1, 45, 90, 91, 137, 180, 225, 270, 333])
    vall = np.arange(360) * 2 * np.pi / 360
    du = (np.arange(512) - 255.5) * 1.0
    dv = (np.arange(384) - 191.5) * 1.0
    s, t = geom.geom_circular(du, vall[vsel], geom="cone", src_iso=500.0,
                              det_iso=500.0, DetsZ=dv)
    iu = rng.integers(0, 512, 1500)
    iv = rng.integers(0, 384, 1500)
    ia = rng.integers(0, vsel.size, 1500)
    ss = s[ia, iu, iv]
    tt = t[ia, iu, iv]
    key = "cone256"
    out[key + "_src"] = ss
    out[key + "_trg"] = tt
    out[key + "_rayid"] = np.stack([vsel[ia], iu, iv], -1)
    out[key + "_grid"] = np.array([256, 256, 256, 0.5, 0.5, 0.5, 0, 0, 0], float)
    ind, ln, off, raised = run_flat(siddons, ss, tt, (256, 256, 256), 0.5, 0.0)
    assert not raised
    out[key + "_ind"] = ind
    out[key + "_flen"] = ln
    out[key + "_foff"] = off
    # geometry of the sampled views, to pin the f32 endpoints bit for bit
    out[key + "_view_src_sha"] = np.array(sha(s.astype(np.float32)))
    out[key + "_view_trg_sha"] = np.array(sha(t.astype(np.float32)))

    # G: a ray whose midpoint index leaves the grid: flat mode raises
    s = np.array([2.0, 2.0, -1.5])
    t = np.array([1.99999994012882, -1.9565179063616578, 0.14530362421344067])
    ind, fl, fo, raised = run_flat(siddons, s, t, (4, 4, 3), 1.0, 0.0)
    x, y, z, ln, off, hit = run_unravel(siddons, s, t, (4, 4, 3), 1.0, 0.0)
    out["oob_src"] = s
    out["oob_trg"] = t
    out["oob_grid"] = np.array([4, 4, 3, 1.0, 1.0, 1.0, 0, 0, 0], float)
    out["oob_raised"] = np.array(raised)
    out["oob_u"] = np.concatenate([x, y, z]).astype(np.int32)
    out["oob_len"] = ln
    out["oob_off"] = off
    out["oob_hit"] = hit

    np.savez_compressed(args.out, **out)
    print(f"wrote {args.out} ({os.path.getsize(args.out)/1e6:.2f} MB) in "
          f"{time.time()-t0:.1f}s")


if __name__ == "__main__":
    main()
