"""Host-side argument checks: every bad input below is refused before any
kernel could read or write out of bounds (no GPU is needed to see it)."""
import numpy as np
import pytest
import torch

from oracle.projection import sd_b_proj_numpy
from vir_amd.autograd import least_squares, total_variation
from vir_amd.device import out_tensor
from vir_amd.proj_geom import check_projections, geom_circular, view_projections
from vir_amd.projection.proj import check_csr, last_of_repeats, sd_b_proj, sd_f_proj


def _csr():
    ind = np.array([0, 5, 7, 3, 3, 9], np.int32)
    ln = np.ones(6, np.float32)
    off = np.array([0, 3, 3, 6], np.int64)
    return ind, ln, off


def test_check_csr_accepts_a_valid_matrix():
    check_csr(*_csr(), 10)


@pytest.mark.parametrize("bad", [10, 11, -1])
def test_index_outside_the_volume_raises_index_error(bad):
    ind, ln, off = _csr()
    ind[4] = bad
    with pytest.raises(IndexError):
        check_csr(ind, ln, off, 10)


@pytest.mark.parametrize("off", [[1, 3, 3, 6], [0, 4, 3, 6], [0, 3, 3, 5], [0, 3, 3, 7]])
def test_inconsistent_offsets_raise_value_error(off):
    ind, ln, _ = _csr()
    with pytest.raises(ValueError):
        check_csr(ind, ln, np.array(off, np.int64), 10)


def test_lengths_must_match_indices():
    ind, ln, off = _csr()
    with pytest.raises(ValueError):
        check_csr(ind, ln[:5], off, 10)


def test_drop_ins_refuse_bad_matrices_before_launching():
    ind, ln, off = _csr()
    ind[0] = 1000
    with pytest.raises(IndexError):
        sd_f_proj(np.zeros((2, 5, 1), np.float32), (ind, ln, off))
    with pytest.raises(IndexError):
        sd_b_proj(np.zeros(3, np.float32), (ind, ln, off), (2, 5, 1))
    # a sinogram shorter than the ray count
    with pytest.raises(IndexError):
        sd_b_proj(np.zeros(2, np.float32), _csr(), (2, 5, 1))


def test_last_of_repeats():
    # ray 0 visits voxel 4 twice (entries 0 and 2); ray 1 visits 2 twice
    ind = np.array([4, 1, 4, 2, 2, 5], np.int32)
    ln = np.array([1, 2, 3, 4, 5, 6], np.float32)
    off = np.array([0, 3, 3, 6], np.int64)
    i2, l2, o2 = last_of_repeats(ind, ln, off)
    assert i2.tolist() == [1, 4, 2, 5]
    assert l2.tolist() == [2, 3, 5, 6]
    assert o2.tolist() == [0, 2, 2, 4]


def test_numpy_oracle_keeps_the_last_repeat():
    """The oracle is numpy's own fancy-index += (proj.py:134): a voxel
    repeated within a ray gets the last entry, across rays values add."""
    ind = np.array([4, 1, 4, 4], np.int32)
    ln = np.array([1, 2, 3, 10], np.float32)
    off = np.array([0, 3, 4], np.int64)
    out = sd_b_proj_numpy(np.array([1.0, 0.5], np.float32), ind, ln, off, (6,))
    assert out.tolist() == [0, 2, 0, 0, 3 + 5, 0]
    # and it equals the sequential backprojection of the reduced matrix
    i2, l2, o2 = last_of_repeats(ind, ln, off)
    ref = np.zeros(6, np.float32)
    for r, s in enumerate([1.0, 0.5]):
        for e in range(o2[r], o2[r + 1]):
            ref[i2[e]] += np.float32(l2[e] * np.float32(s))
    assert np.array_equal(ref, out)


def test_out_buffer_must_be_a_cuda_float32_tensor():
    with pytest.raises(TypeError):
        out_tensor(torch.zeros(4), (4,), "cuda")
    with pytest.raises(TypeError):
        out_tensor(np.zeros(4, np.float32), (4,), "cuda")


def test_loss_terms_refuse_host_tensors():
    with pytest.raises(TypeError):
        least_squares(torch.zeros(8), torch.zeros(8))
    with pytest.raises(TypeError):
        total_variation(torch.zeros(4, 4, 4))


def test_projection_matrices_must_match_the_rays():
    v = np.arange(5) * 2 * np.pi / 5
    du = np.arange(24) - 11.5
    dv = np.arange(12) - 5.5
    s, t = geom_circular(du, v, "cone", src_iso=60.0, det_iso=60.0, DetsZ=dv)
    s, t = s.astype(np.float32), t.astype(np.float32)
    M = view_projections(s, t)
    check_projections(s, t, M)
    swapped = M.reshape(-1, 3, 4)[:, [1, 0, 2]].reshape(-1, 12)
    with pytest.raises(ValueError):
        check_projections(s, t, swapped)
    pitch = M.copy()
    pitch[:, :4] *= 1.01      # a detector pitch off by 1 %
    with pytest.raises(ValueError):
        check_projections(s, t, pitch)
    with pytest.raises(ValueError):
        check_projections(s, t, np.roll(M, 1, axis=0))   # matrices of other views
