"""vir_siddon_bp_rec_noalloc, through the C ABI, against vir_siddon_bp_rec
with a caller-provided workspace: the same bits on two calls in a row (the
second smaller), with and without VIR_ACCUMULATE, and a call that needs more
staging than the fixed array is refused with nothing launched."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _projector(n, nviews, ndet):
    from vir_amd.geometry import circular_projections, circular_rays
    from vir_amd.sys_mat.siddon import SiddonProjector
    views = np.arange(nviews) * 2 * np.pi / nviews
    dets = (np.arange(ndet) - (ndet - 1) / 2.0) * 2.0
    src, trg = circular_rays(dets, views, "cone", 4.0 * n, 4.0 * n, dets)
    mats = circular_projections(dets, views, "cone", 4.0 * n, 4.0 * n, dets)
    return SiddonProjector(src, trg, (n, n, n), 1.0, proj=mats)


def _bp_both(P, sino, vol0, flags):
    import torch
    from vir_amd import _lib
    from vir_amd.device import stream
    L = _lib.lib()
    nviews, nu, nv = P.layout
    fl = flags | P._ung
    ref = vol0.clone()
    _lib.check(L.vir_siddon_bp_rec(_lib.ptr(P.rec), nviews, nu, nv, _lib.ptr(P.proj), P.grid,
                                   _lib.ptr(sino), _lib.ptr(ref), fl, _lib.ptr(P._ws), stream()),
               "vir_siddon_bp_rec")
    got = vol0.clone()
    _lib.check(L.vir_siddon_bp_rec_noalloc(_lib.ptr(P.rec), nviews, nu, nv, _lib.ptr(P.proj), P.grid,
                                           _lib.ptr(sino), _lib.ptr(got), fl, stream()),
               "vir_siddon_bp_rec_noalloc")
    torch.cuda.synchronize()
    return ref, got


@pytest.mark.parametrize("accumulate", [False, True])
def test_noalloc_equals_workspace_version_twice(accumulate):
    import torch
    from vir_amd import _lib
    rng = np.random.default_rng(11)
    fl = _lib.VIR_ACCUMULATE if accumulate else 0
    for n, nviews, ndet in ((64, 40, 96), (32, 24, 48)):     # second call smaller
        print(f"noalloc: grid {n}^3, {nviews} views, {ndet}^2 detector, accumulate={accumulate}",
              flush=True)
        P = _projector(n, nviews, ndet)
        need = int(_lib.lib().vir_siddon_bp_workspace(P.grid, nviews))
        assert 0 < need <= (32 << 20)                       # a multi-chunk plan
        sino = torch.from_numpy(rng.uniform(0, 1, P.ray_shape).astype(np.float32)).cuda()
        vol0 = torch.from_numpy(rng.uniform(0, 1, P.nPixels).astype(np.float32)).cuda()
        ref, got = _bp_both(P, sino, vol0, fl)
        assert float(ref.abs().max()) > 0
        assert torch.equal(ref, got)


def test_noalloc_refuses_more_staging_than_it_has():
    import torch
    from vir_amd import _lib
    from vir_amd.device import stream
    # 256^3 with four views is a single-chunk plan; VIR_ACCUMULATE stages the
    # whole volume (64 MiB), more than the fixed 32 MiB
    P = _projector(256, 4, 32)
    assert int(_lib.lib().vir_siddon_bp_workspace(P.grid, 4)) == 0
    sino = torch.ones(P.ray_shape, dtype=torch.float32, device="cuda")
    vol = torch.full(P.nPixels, 3.0, dtype=torch.float32, device="cuda")
    nviews, nu, nv = P.layout
    rc = _lib.lib().vir_siddon_bp_rec_noalloc(_lib.ptr(P.rec), nviews, nu, nv, _lib.ptr(P.proj), P.grid,
                                              _lib.ptr(sino), _lib.ptr(vol),
                                              _lib.VIR_ACCUMULATE | P._ung, stream())
    assert rc == -1 and b"staging" in _lib.lib().vir_last_error()
    torch.cuda.synchronize()
    assert bool((vol == 3.0).all())
    # without VIR_ACCUMULATE the same call needs no staging and runs
    ref, got = _bp_both(P, sino, vol, 0)
    assert torch.equal(ref, got)
