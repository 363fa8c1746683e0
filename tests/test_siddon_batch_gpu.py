"""Batched Siddon projections (vir_siddon_*_batch through the C ABI): a batch
of volumes on one ray set is traced once per group, and every item must equal
the single-volume call on it bit for bit -- forward, backward (direct and
view-chunked plans, 2-D and 3-D sub-bricks, float32 and float64 walks),
VIR_ACCUMULATE, and the fused SIRT steps.  Batch sizes 1, 2, 3 and 5 cover
the FP groups of 4, 2, 1 and the BP groups of 2, 1.  One item's sinogram is
mostly zeros (both signs), so the rays the single call skips for that item
are traced for the others."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _projector(kind):
    import torch  # noqa: F401
    from vir_amd.sys_mat import siddon as sd
    from test_siddon_gpu import _geometry
    if kind == "cone256":
        from vir_amd.workloads import cone256_geometry
        s, t, grid = cone256_geometry(np.arange(0, 360, 45))
        return sd.SiddonProjector(s, t, grid["nPixels"], grid["dPixels"])
    s, t, nP, dP = _geometry(kind, np.random.default_rng(3))
    return sd.SiddonProjector(s, t, nP, dP)


def _items(rng, nb, n, zero_item=True):
    a = rng.uniform(-1, 1, (nb, n)).astype(np.float32)
    if zero_item and nb > 1:
        keep = rng.uniform(0, 1, n) < 0.1
        a[1] = np.where(keep, a[1], np.where(rng.uniform(0, 1, n) < 0.5, 0.0, -0.0)).astype(np.float32)
    return a


def _same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("kind", ["par2d", "fan2d", "cone3d", "par3d", "cone256"])
def test_batched_fp_bp_equal_single_calls(kind):
    import torch
    print("case", kind, flush=True)
    P = _projector(kind)
    rng = np.random.default_rng(21)
    for nb in (1, 2, 3, 5):
        vols = torch.from_numpy(_items(rng, nb, P.nvox, zero_item=False)).cuda()
        sinos = torch.from_numpy(_items(rng, nb, P.nrays)).cuda()
        fb = P.fp_batch(vols.reshape((nb,) + P.nPixels), check=True).cpu().numpy().reshape(nb, -1)
        bb = P.bp_batch(sinos.reshape((nb,) + tuple(P.ray_shape))).cpu().numpy().reshape(nb, -1)
        for j in range(nb):
            f1 = P.fp(vols[j]).cpu().numpy().reshape(-1)
            b1 = P.bp(sinos[j]).cpu().numpy().reshape(-1)
            assert _same(fb[j], f1), (kind, nb, j, "fp")
            assert _same(bb[j], b1), (kind, nb, j, "bp")
        # VIR_ACCUMULATE onto existing contents
        init = torch.from_numpy(_items(rng, nb, P.nvox, zero_item=False)).cuda()
        ba = P.bp_batch(sinos, out=init.clone(), accumulate=True).cpu().numpy().reshape(nb, -1)
        sinit = torch.from_numpy(_items(rng, nb, P.nrays, zero_item=False)).cuda()
        fa = P.fp_batch(vols, out=sinit.clone(), accumulate=True).cpu().numpy().reshape(nb, -1)
        for j in range(nb):
            b1 = P.bp(sinos[j], out=init[j].clone(), accumulate=True).cpu().numpy().reshape(-1)
            f1 = P.fp(vols[j], out=sinit[j].clone(), accumulate=True).cpu().numpy().reshape(-1)
            assert _same(ba[j], b1), (kind, nb, j, "bp accumulate")
            assert _same(fa[j], f1), (kind, nb, j, "fp accumulate")


@pytest.mark.parametrize("kind", ["fan2d", "cone3d", "par3d"])
def test_batched_fused_sirt_steps_equal_single_calls(kind):
    import torch
    print("case", kind, flush=True)
    P = _projector(kind)
    assert P.fused_sirt
    rng = np.random.default_rng(22)
    nb = 3
    x = torch.from_numpy(rng.uniform(0, 1, (nb, P.nvox)).astype(np.float32)).cuda()
    b = torch.from_numpy(rng.uniform(0, 5, (nb, P.nrays)).astype(np.float32)).cuda()
    w_s = rng.uniform(0.5, 2, P.nrays).astype(np.float32)
    w_s[::7] = 0.0
    w_i = rng.uniform(0.5, 2, P.nvox).astype(np.float32)
    w_i[::5] = 0.0
    w_s, w_i = torch.from_numpy(w_s).cuda(), torch.from_numpy(w_i).cuda()
    r = torch.empty(nb, P.nrays, dtype=torch.float32, device="cuda")
    P.fp_residual_batch(x, b, w_s, r)
    xb = x.clone()
    P.bp_update_batch(r, w_i, 0.37, xb)
    for j in range(nb):
        r1 = torch.empty(P.nrays, dtype=torch.float32, device="cuda")
        P.fp_residual(x[j].contiguous(), b[j].contiguous(), w_s, r1)
        assert _same(r[j].cpu().numpy(), r1.cpu().numpy()), (kind, j, "residual")
        x1 = x[j].clone()
        P.bp_update(r1, w_i, 0.37, x1)
        assert _same(xb[j].cpu().numpy(), x1.cpu().numpy()), (kind, j, "update")


def test_batch_argument_checks():
    import torch
    P = _projector("cone3d")
    with pytest.raises(ValueError):
        P.fp_batch(torch.zeros(P.nvox + 1, device="cuda"))
    with pytest.raises(ValueError):
        P.bp_batch(torch.zeros(2 * P.nrays - 1, device="cuda"))
    with pytest.raises(ValueError):
        P.fp_residual_batch(torch.zeros(2, P.nvox, device="cuda"), torch.zeros(P.nrays, device="cuda"),
                            torch.ones(P.nrays, device="cuda"), torch.empty(2, P.nrays, device="cuda"))
    with pytest.raises(ValueError):
        P.bp_update_batch(torch.zeros(2, P.nrays, device="cuda"), torch.ones(P.nvox, device="cuda"), 1.0,
                          torch.zeros(P.nvox, device="cuda"))
