"""ORACLE: the float32 accumulation loops of the reference's Siddon list
reductions -- test infrastructure only.

  list2array_loop   sys_mat/siddon.py:96-163 (per entry array.flat[i] += len,
                    float32, entries in list order)
  rays_ave_loop     sys_mat/siddon.py:432-487 (np.unique, then float32
                    += in entry order, then / nRays)

Pinned in tests/test_siddon_lists.py against scipy's float64 column sums and
by merging identical rays.
"""
import numpy as np


def list2array_loop(ind, ln, nPixels):
    out = np.zeros(nPixels, dtype=np.float32)
    flat = out.reshape(-1)
    for i in range(len(ind)):
        flat[ind[i]] += np.float32(ln[i])
    return out


def rays_ave_loop(ind, ln, nrays):
    ind = np.asarray(ind)
    uniq, inv = np.unique(ind, return_inverse=True)
    acc = np.zeros(uniq.size, dtype=np.float32)
    for i, k in enumerate(inv):
        acc[k] += np.float32(ln[i])
    acc /= np.float32(nrays)
    return uniq, acc
