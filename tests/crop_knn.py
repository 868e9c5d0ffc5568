"""Brute-force kNN over a crop of a large cloud, for checking single queries at full size."""
import numpy as np

import octree_amd as A
import pyoracle as O


def crop_brute_knn(sx, tree, q, k, kth_d2):
    """Exact k nearest of centre q by brute force over the points of the cube that holds every point
    within the k-th squared distance kth_d2 (found with the oracle's cube search), in (d^2, index)
    order; d^2 as distPointsSquared (linear_octree.hpp:844-851)."""
    half = float(np.sqrt(kth_d2)) * (1.0 + 1e-9) + 1e-9
    lists, _ = O.radius(tree, sx, A.K_CUBE, A.A_PRUNE, np.asarray(q, np.float64)[None, :], half)
    cand = np.asarray(lists[0], np.int64)
    p = sx[cand]
    dx, dy, dz = q[0] - p[:, 0], q[1] - p[:, 1], q[2] - p[:, 2]
    d2 = (dx * dx + dy * dy) + dz * dz
    o = np.lexsort((cand, d2))[:k]
    return cand[o], d2[o]
