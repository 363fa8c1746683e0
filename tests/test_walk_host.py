"""The Siddon walker's lean walks, checked on the CPU.

tests/host/walk_check.cpp compiles the device functions of
vir_amd/csrc/siddon_ray.h for the host (clang++, as hipcc's host pass would)
and compares, pair by pair, the walks the unguarded kernels take -- the
forward projector's three-phase walk and the backprojector's walk over a
sub-brick window, without the special values away from a ray's ends and in
voxel units on power-of-two grids -- with the general walk in world units:
cone rays of the 256^3 benchmark scan, the GPU tests' cone geometry, a
non-power-of-two grid, 4 mm and 0.25 mm voxels with an offset origin, and
sources near the grid.
With the margins of the lean walks dropped, or the reciprocal direction
left in world units in the voxel-unit walk (planted mistakes), the same check
must find mismatches."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.fixture(scope="module")
def walk_check(tmp_path_factory):
    cxx = CLANG if os.path.exists(CLANG) else shutil.which("clang++")
    if cxx is None:
        pytest.skip("clang++ not available")
    out = str(tmp_path_factory.mktemp("walk") / "walk_check")
    subprocess.run([cxx, "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                    "-I", os.path.join(HERE, "host", "include"),
                    "-I", os.path.join(HERE, "..", "vir_amd", "csrc"),
                    os.path.join(HERE, "host", "walk_check.cpp"), "-o", out], check=True)
    return out


def _write(path, s, t, n, d, o=(0.0, 0.0, 0.0)):
    s = np.asarray(s, np.float32).reshape(-1, 3)
    t = np.asarray(t, np.float32).reshape(-1, 3)
    with open(path, "wb") as f:
        np.array(list(n) + [len(s)], np.int32).tofile(f)
        np.array([d] * 3 + list(o), np.float32).tofile(f)
        np.array([1e-5], np.float64).tofile(f)
        s.tofile(f)
        t.tofile(f)


def _cases():
    from vir_amd.proj_geom import geom_circular
    from vir_amd.workloads import cone256_geometry
    rng = np.random.default_rng(17)
    out = {}
    s, t, _ = cone256_geometry(np.array([0, 51, 133, 290]))
    s, t = s.reshape(-1, 3), t.reshape(-1, 3)
    pick = rng.choice(len(s), 500, replace=False)
    out["cone256"] = (s[pick], t[pick], (256, 256, 256), 0.5, (0.0, 0.0, 0.0))
    v = np.arange(13) * 2 * np.pi / 13
    s, t = geom_circular((np.arange(36) - 17.5) * 1.0, v, "cone", src_iso=80.0, det_iso=80.0,
                         DetsZ=(np.arange(24) - 11.5) * 1.0)
    out["cone3d"] = (s, t, (40, 36, 20), 0.5, (0.0, 0.0, 0.0))
    out["cone3d_d07"] = (s, t, (40, 36, 20), 0.7, (0.0, 0.0, 0.0))
    v = np.arange(11) * 2 * np.pi / 11 + 0.1
    s, t = geom_circular((np.arange(24) - 11.5) * 5.3, v, "cone", src_iso=300.0, det_iso=260.0,
                         DetsZ=(np.arange(20) - 9.5) * 5.3)
    off = np.array([3.3, -2.7, 1.1])
    out["d4_offset"] = (s + off, t + off, (40, 40, 32), 4.0, tuple(off))
    out["d025_offset"] = ((s + off) / 16, (t + off) / 16, (40, 40, 32), 0.25, tuple(off / 16))
    v = np.arange(9) * 2 * np.pi / 9
    s, t = geom_circular((np.arange(30) - 14.5) * 1.3, v, "cone", src_iso=30.0, det_iso=40.0,
                         DetsZ=(np.arange(24) - 11.5) * 1.3)
    out["near_source"] = (s, t, (32, 32, 24), 1.0, (0.0, 0.0, 0.0))
    return out


def _run(exe, path, *extra):
    r = subprocess.run([exe, path, *extra], capture_output=True, text=True, timeout=300)
    counts = dict(zip(r.stdout.split()[0::2], map(int, r.stdout.split()[1::2])))
    return r.returncode, counts


@pytest.mark.parametrize("name", ["cone256", "cone3d", "cone3d_d07", "d4_offset", "d025_offset",
                                  "near_source"])
def test_lean_walks_visit_the_general_walks_pairs(walk_check, tmp_path, name):
    s, t, n, d, o = _cases()[name]
    path = str(tmp_path / f"{name}.bin")
    _write(path, s, t, n, d, o)
    rc, c = _run(walk_check, path)
    assert rc == 0 and c["fp_mismatch"] == 0 and c["bp_mismatch"] == 0, c
    # the lean walks carry most of the work, so the comparison covers them;
    # every power-of-two grid here is walked in voxel units
    assert c["fp_mid_pairs"] > 0.5 * c["fp_pairs"], c
    assert c["bp_mid_windows"] > 0.3 * c["bp_windows"], c
    assert c["unit_rays"] == (0 if name == "cone3d_d07" else c["rays"]), c


@pytest.mark.parametrize("mistake", ["loose", "unscaled_y"])
def test_planted_mistakes_are_caught(walk_check, tmp_path, mistake):
    s, t, n, d, o = _cases()["cone3d"]
    path = str(tmp_path / "cone3d.bin")
    _write(path, s, t, n, d, o)
    rc, c = _run(walk_check, path, mistake)
    assert rc == 1 and c["fp_mismatch"] > 0 and c["bp_mismatch"] > 0, c
