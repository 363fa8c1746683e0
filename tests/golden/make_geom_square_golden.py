"""Capture golden vectors for geom_square and a Siddon trace of its rays.

Runs ONLY in the build container, where the reference tree is mounted at
/root/reference.  It loads the reference's numpy-only proj_geom.py
(geom_square, proj_geom.py:12-57) and sys_mat/siddon.py by file path (see
make_siddon_golden.load_reference) and records inputs and outputs into
tests/golden/geom_square_golden.npz:

  case{k}_binsx, case{k}_disty, case{k}_angs   inputs (dtypes as given)
  case{k}_src, case{k}_trg                      float64 (nBins, nAngs, 4, 3)
  trace_ind, trace_len, trace_off               siddons(flat=True) of case 4's
                                                rays on a 24 x 24 x 1 grid

Usage:  python tests/golden/make_geom_square_golden.py [--ref /root/reference]
"""
import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_siddon_golden import load_reference  # noqa: E402

CASES = [
    (np.arange(8) - 3.5, 20.0, np.linspace(-0.6, 0.6, 5)),
    (np.linspace(-10, 10, 33).astype(np.float32), 17.3, np.linspace(0.1, 1.2, 7).astype(np.float32)),
    (np.arange(5), 3, np.array([0.0, np.pi / 4, np.pi / 2, -np.pi / 3])),
    (np.linspace(-63.5, 63.5, 128), 70.0, np.linspace(-np.pi / 4, np.pi / 4, 33)),
    (np.linspace(-11.5, 11.5, 24), 16.0, np.linspace(np.pi / 4, 3 * np.pi / 4, 9)),
]
TRACE_CASE = 4   # angles across the square (measured from the detector line)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    args = ap.parse_args()
    siddon, geom = load_reference(args.ref)
    out = {}
    for k, (bx, dy, ang) in enumerate(CASES):
        s, t = geom.geom_square(bx, dy, ang)
        out[f"case{k}_binsx"] = bx
        out[f"case{k}_disty"] = np.float64(dy)
        out[f"case{k}_angs"] = ang
        out[f"case{k}_src"] = s
        out[f"case{k}_trg"] = t
    s, t = out[f"case{TRACE_CASE}_src"], out[f"case{TRACE_CASE}_trg"]
    ind, ln, off = siddon.siddons(s, t, (24, 24, 1), 1.0, 0.0, flat=True)
    out["trace_ind"], out["trace_len"], out["trace_off"] = ind, ln, off
    path = os.path.join(HERE, "geom_square_golden.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(CASES)} cases, trace nnz={ind.size}")


if __name__ == "__main__":
    main()
