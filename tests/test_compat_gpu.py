"""compat=True modes of the drop-ins against the oracle flags that restate
the reference's quirks literally:

  * ray-driven AW / Joseph: rays parallel to an axis outside the image are
    clamped onto the border row (rd.py:66-73)  -> rd_ref(clamp_parallel=True)
  * SF fan forward: a zero pixel is skipped without advancing the o boundary
    of the next pixel's amplitude (pd.py:321-325)
                                          -> pd_ref.fan_fp_loop(stale_oy=True)
  * cone DD forward: the u buffer is re-zeroed only over each line's
    footprint and added whole (dd.py:317, 374)
                                          -> dd_ref.cone_fp(stale_buffer=True)

Bar: relative L2 error <= 1e-5, as for the default modes.  Every case also
checks that the quirk is exercised: the default mode differs from the
compat one by far more than the tolerance.  Images reach the border (that
is where the quirks act)."""
import numpy as np
import pytest

from oracle import dd_ref, pd_ref, rd_ref

pytestmark = pytest.mark.gpu
TOL = 1e-5


def rel(a, b):
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _axis_bin_offset(theta, DSD, du, nu, iu0):
    """su placing bin iu0's fan ray from a source at angle theta (float32)
    parallel to the x planes: v_x = -DSD c - u s = 0."""
    c, s = float(np.cos(np.float32(theta))), float(np.sin(np.float32(theta)))
    ustar = -DSD * c / s
    su = np.float32(ustar / du - (iu0 - (nu / 2.0 - 0.5)))
    u = float(np.float32(du)) * (iu0 - (nu / 2.0 - 0.5) + float(su))
    m = np.hypot(DSD, u)
    assert abs(-DSD * c - u * s) / m < 2e-7      # well inside |h_x| <= 1e-6
    return float(su)


def _rd_cases():
    a32 = np.float32
    par = dict(ang=np.array([0.0, np.pi / 2, 0.3, 1.1, 2.0, 2.9], a32), nu=41, du=0.9, su=0.3)
    th = 1.3
    fan = dict(ang=np.array([th, 0.4, 2.5, 4.0], a32), nu=41, du=1.1, DSO=40.0, DSD=70.0)
    fan["su"] = _axis_bin_offset(th, fan["DSD"], fan["du"], fan["nu"], 5)
    cone = dict(ang=np.array([th, 0.7, 3.3], a32), nu=23, du=1.6, DSO=35.0, DSD=60.0, nv=9,
                dv=1.2, sv=0.2)
    cone["su"] = _axis_bin_offset(th, cone["DSD"], cone["du"], cone["nu"], 2)
    return {"par": par, "fan": fan, "cone": cone}


def _rd_run(rd, geom, g, img, y, shape, d, joseph, compat):
    kw = dict(joseph=joseph, compat=compat)
    if geom == "par":
        fp = rd.aw_fp_par_2d(img, g["ang"], g["nu"], g["du"], g["su"], d, **kw)
        bp = rd.aw_bp_par_2d(y, g["ang"], shape, g["du"], g["su"], d, **kw)
    elif geom == "fan":
        fp = rd.aw_fp_fan_2d(img, g["ang"], g["nu"], g["DSO"], g["DSD"], g["du"], g["su"], d, **kw)
        bp = rd.aw_bp_fan_2d(y, g["ang"], shape, g["DSO"], g["DSD"], g["du"], g["su"], d, **kw)
    else:
        fp = rd.aw_fp_cone_3d(img, g["ang"], g["nu"], g["nv"], g["DSO"], g["DSD"], g["du"],
                              g["dv"], g["su"], g["sv"], d, **kw)
        bp = rd.aw_bp_cone_3d(y, g["ang"], shape, g["DSO"], g["DSD"], g["du"], g["dv"], g["su"],
                              g["sv"], d, **kw)
    return fp, bp


@pytest.mark.parametrize("geom", ["par", "fan", "cone"])
def test_aw_clamps_axis_parallel_rays_like_the_reference(geom):
    from vir_amd.sys_mat import rd
    g = _rd_cases()[geom]
    d = 0.9
    shape = (14, 12, 10) if geom == "cone" else (20, 17)
    rng = np.random.default_rng(11)
    img = rng.uniform(0.2, 1.0, shape).astype(np.float32)
    ref_fp = rd_ref.fp(img, geom, d_pix=d, clamp_parallel=True, **g)
    y = rng.uniform(0, 1, ref_fp.shape).astype(np.float32)
    kw = {k: v for k, v in g.items() if k not in ("nu", "nv", "ang")}
    ref_bp = rd_ref.bp(y, geom, g["ang"], shape, d_pix=d, clamp_parallel=True, **kw)
    fp, bp = _rd_run(rd, geom, g, img, y, shape, d, False, True)
    assert rel(fp, ref_fp) <= TOL
    assert rel(bp, ref_bp) <= TOL
    # the quirk is exercised: the default mode drops those rays
    fp0, bp0 = _rd_run(rd, geom, g, img, y, shape, d, False, False)
    assert rel(fp0, ref_fp) > 1e-3 and rel(bp0, ref_bp) > 1e-4
    assert rel(fp0, rd_ref.fp(img, geom, d_pix=d, **g)) <= TOL


def test_joseph_clamps_axis_parallel_rays_like_the_reference():
    """Views along the axes only: there every sample position is exact
    enough that no sample flips in or out of [t_entry, t_exit] between
    float32 and float64 (the image width is 37.2 sample steps)."""
    from vir_amd.sys_mat import rd
    g = dict(ang=np.array([0.0, np.pi / 2], np.float32), nu=41, du=0.9, su=0.3)
    d, shape = 0.93, (20, 17)
    rng = np.random.default_rng(12)
    img = rng.uniform(0.2, 1.0, shape).astype(np.float32)
    ref_fp = rd_ref.fp(img, "par", d_pix=d, joseph=True, clamp_parallel=True, **g)
    y = rng.uniform(0, 1, ref_fp.shape).astype(np.float32)
    ref_bp = rd_ref.bp(y, "par", g["ang"], shape, du=g["du"], su=g["su"], d_pix=d, joseph=True,
                       clamp_parallel=True)
    fp, bp = _rd_run(rd, "par", g, img, y, shape, d, True, True)
    assert rel(fp, ref_fp) <= TOL
    assert rel(bp, ref_bp) <= TOL
    fp0, bp0 = _rd_run(rd, "par", g, img, y, shape, d, True, False)
    assert rel(fp0, ref_fp) > 1e-3 and rel(bp0, ref_bp) > 1e-4


def test_sf_fan_forward_keeps_the_stale_boundary():
    from vir_amd.sys_mat import pd
    rng = np.random.default_rng(13)
    ang = np.linspace(0, 2 * np.pi, 9, endpoint=False)
    kw = dict(nu=36, DSO=40.0, DSD=60.0, du=0.9, su=0.1, d_pix=1.1)
    img = rng.uniform(0.1, 1.0, (16, 13)).astype(np.float32)
    img[:, :3] = 0.0          # zero runs at the start of every column
    img[4:7, 6:9] = 0.0       # and inside some
    img[::3, 10] = 0.0
    stale = pd_ref.fan_fp_loop(img, ang, **kw, stale_oy=True)
    fixed = pd_ref.fan_fp_loop(img, ang, **kw, stale_oy=False)
    got = pd.pd_fp_fan_2d(img, ang, kw["nu"], kw["DSO"], kw["DSD"], kw["du"], kw["su"],
                          kw["d_pix"], compat=True)
    assert rel(got, stale) <= TOL
    got0 = pd.pd_fp_fan_2d(img, ang, kw["nu"], kw["DSO"], kw["DSD"], kw["du"], kw["su"],
                           kw["d_pix"])
    assert rel(got0, fixed) <= TOL
    # the stale boundary moves the amplitude only (a few 1e-4 here)
    assert rel(fixed, stale) > 1e-4
    # nothing to skip: both modes agree
    dense = rng.uniform(0.1, 1.0, (16, 13)).astype(np.float32)
    a = pd.pd_fp_fan_2d(dense, ang, kw["nu"], kw["DSO"], kw["DSD"], kw["du"], kw["su"],
                        kw["d_pix"], compat=True)
    b = pd.pd_fp_fan_2d(dense, ang, kw["nu"], kw["DSO"], kw["DSD"], kw["du"], kw["su"],
                        kw["d_pix"])
    assert np.array_equal(a, b)


def test_cone_dd_forward_reuses_the_stale_buffer():
    """Cubic volume: the reference's cone sweeps unpack the transposed
    volume's shape as (nP, nz, no) while indexing it [pixel, line, slab]
    (dd.py:289, 302-330), so their loops stay in bounds only when the line
    and slab counts agree."""
    from vir_amd.sys_mat import dd
    rng = np.random.default_rng(14)
    shape = (10, 10, 10)
    img = rng.uniform(0.5, 1.0, shape).astype(np.float32)
    ang = (np.linspace(0, 2 * np.pi, 9, endpoint=False) + 0.1).astype(np.float32)
    DSO, DSD = 40.0, 70.0
    kw = dict(du=1.2, dv=1.1, su=0.1, d_pix=1.0)
    stale = dd_ref.cone_fp(img, ang, 30, 20, DSO, DSD, sv=0.0, stale_buffer=True, **kw)
    fixed = dd_ref.cone_fp(img, ang, 30, 20, DSO, DSD, sv=0.0, **kw)
    got = dd.dd_fp_cone_3d(img, ang, 30, 20, DSO, DSD, sv=0.0, compat=True, **kw)
    assert rel(got, stale) <= TOL
    assert rel(fixed, stale) > 1e-3
    got0 = dd.dd_fp_cone_3d(img, ang, 30, 20, DSO, DSD, sv=0.0, **kw)
    assert rel(got0, fixed) <= TOL
    # accumulate adds onto the output
    import torch
    P = dd.DDProjector(shape, ang, 30, "cone", nv=20, DSO=DSO, DSD=DSD, sv=0.0, compat=True, **kw)
    out = torch.ones(P.sino_shape, device="cuda")
    P.fp(torch.from_numpy(img).cuda(), out=out, accumulate=True)
    assert rel(out.cpu().numpy() - 1.0, stale) <= 1e-4
    # an object clear of the border: identical to the default
    small = np.zeros(shape, np.float32)
    small[3:7, 3:7, 3:7] = img[3:7, 3:7, 3:7]
    assert rel(dd.dd_fp_cone_3d(small, ang, 30, 20, DSO, DSD, sv=0.0, compat=True, **kw),
               dd.dd_fp_cone_3d(small, ang, 30, 20, DSO, DSD, sv=0.0, **kw)) <= TOL


def test_compat_is_refused_where_there_is_no_quirk():
    from vir_amd.sys_mat import dd, pd
    with pytest.raises(ValueError):
        pd.SFProjector((8, 8), [0.0], 10, "par", compat=True)
    with pytest.raises(ValueError):
        dd.DDProjector((8, 8), [0.0], 10, "fan", DSO=40.0, DSD=60.0, compat=True)
