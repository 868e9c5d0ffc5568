"""Clouds for the batch tests: uneven sizes (log-uniform, plus tiny ones), uniform boxes and
LiDAR-style tiles with their own bounding boxes far from the origin, Morton and Hilbert, different
leaf sizes."""
import numpy as np

import octree_amd as A


def make_clouds(seed=21, count=14):
    rng = np.random.default_rng(seed)
    sizes = [1, 7, 130] + [int(v) for v in np.exp(rng.uniform(np.log(1e3), np.log(2.5e5), count - 3))]
    clouds = []
    for i, n in enumerate(sizes):
        off = rng.uniform(-1e5, 1e5, 3)
        if i % 2:
            xyz = A.uniform_box(n, seed=100 + i, size=(60.0, 60.0, 15.0), offset=tuple(off))
        else:
            ext = max(20.0, 200.0 * np.sqrt(n / 2.5e5))
            xyz = A.lidar_tile(n, seed=100 + i, extent=ext, offset=(off[0], off[1]))
        enc = A.ENC_HILBERT if i % 3 else A.ENC_MORTON
        ml = (32, 128, 16)[i % 3]
        clouds.append((xyz, enc, ml))
    return clouds


def clustered_duplicates(clusters=200, per=17, seed=5):
    """`clusters` groups of `per` identical points: with maxPointsLeaf < per every group splits down
    to the depth limit (21 levels), many more internal nodes than points / maxPointsLeaf."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 1000.0, (clusters, 3))
    return np.repeat(c, per, axis=0)
