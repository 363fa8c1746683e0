"""geom_square (proj_geom.py:12-57) against golden vectors captured from the
reference (tests/golden/make_geom_square_golden.py): the host restatement
bit for bit in float64 (float32 inputs included), the C oracle's Siddon
trace of its rays bit for bit, and (GPU) the device generator equal to the
float32 cast and its trace equal to the reference's."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import siddon as osd
from vir_amd.proj_geom import geom_square

NCASES = 5
TRACE_CASE = 4


@pytest.fixture(scope="module")
def gs():
    return np.load(os.path.join(GOLDEN, "geom_square_golden.npz"))


def _inputs(gs, k):
    # distY was a Python scalar: keep it one, so numpy's promotion with
    # float32 angles is the reference's
    return gs[f"case{k}_binsx"], float(gs[f"case{k}_disty"]), gs[f"case{k}_angs"]


@pytest.mark.parametrize("k", range(NCASES))
def test_host_geom_square_is_the_reference(gs, k):
    s, t = geom_square(*_inputs(gs, k))
    assert s.dtype == np.float64 and s.shape == gs[f"case{k}_src"].shape
    assert np.array_equal(s.view(np.int64), gs[f"case{k}_src"].view(np.int64))
    assert np.array_equal(t.view(np.int64), gs[f"case{k}_trg"].view(np.int64))


def test_oracle_trace_of_square_rays(gs):
    s, t = gs[f"case{TRACE_CASE}_src"], gs[f"case{TRACE_CASE}_trg"]
    ind, ln, off = osd.siddons(s, t, (24, 24, 1), 1.0, 0.0, flat=True)
    assert np.array_equal(off, gs["trace_off"]) and np.array_equal(ind, gs["trace_ind"])
    assert np.array_equal(ln.view(np.int32), gs["trace_len"].view(np.int32))


@pytest.mark.gpu
@pytest.mark.parametrize("k", range(NCASES))
def test_device_square_rays(gs, k):
    from vir_amd.geometry import square_rays
    s, t = square_rays(*_inputs(gs, k))
    assert np.array_equal(s.cpu().numpy(), gs[f"case{k}_src"].astype(np.float32))
    assert np.array_equal(t.cpu().numpy().view(np.int32),
                          gs[f"case{k}_trg"].astype(np.float32).view(np.int32))


@pytest.mark.gpu
def test_gpu_trace_of_device_square_rays(gs):
    from vir_amd.geometry import square_rays
    from vir_amd.sys_mat.siddon import siddons
    s, t = square_rays(*_inputs(gs, TRACE_CASE))
    ind, ln, off = siddons(s, t, (24, 24, 1), 1.0, 0.0, flat=True)
    assert np.array_equal(off, gs["trace_off"]) and np.array_equal(ind, gs["trace_ind"])
    assert np.array_equal(ln.view(np.int32), gs["trace_len"].view(np.int32))
