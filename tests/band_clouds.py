"""Adversarial clouds for the FP32 screening of the searches (constructed on the CPU, no GPU).

The FP32 + recheck searches decide a point or node in FP32 when it lies outside a proven rounding
band, and re-evaluate it in FP64 otherwise. These clouds put points and node boxes exactly on, or a
few units in the last place around, the thresholds the reference's FP64 expressions decide on:

- boundary_cloud: points whose FP64 squared distance to a query is exactly r^2 (sphere `<` excludes
  them, circle `<=` includes them; kernel_3d.hpp:70, kernel_2d.hpp:62), the representable points on
  either side of that boundary along random directions, and points at r^2 (1 +- 2^-e) for e across
  the FP32 band's width, at LAS-like coordinate offsets;
- magnitude_cloud: points and radii near the magnitudes where the leaf-recentred FP32 terms overflow
  or approach the overflow guard (amax + |b|^2 + R near 1e30 and beyond FLT_MAX);
- node_boundary_cloud: a lattice in a power-of-two box (exact node centres and radii) with query
  centres placed so that the node judgements of boxIntersect / boxOverlap hit their equalities
  (dx == r + rl for OUTSIDE, sum of (|d| + rl)^2 == R for INSIDE, sum of (|d| - rl)^2 == R for the
  overlap test; kernel_3d.hpp:181-223, 277-302, kernel_2d.hpp:107-144, 195-218);
- lattice_cloud: a unit lattice with kNN queries at lattice and half-lattice points, where many
  neighbours tie at the k-th distance (the (d^2, index) order decides).
"""
import numpy as np


def d2_sphere(p, q):
    """isInside's FP64 expression for the sphere (kernel_3d.hpp:127-138): d1*d1 + d2*d2 + d3*d3."""
    d1, d2, d3 = p[..., 0] - q[0], p[..., 1] - q[1], p[..., 2] - q[2]
    return (d1 * d1 + d2 * d2) + d3 * d3


def d2_circle(p, q):
    d1, d2 = p[..., 0] - q[0], p[..., 1] - q[1]
    return d1 * d1 + d2 * d2


# exact offsets with |v|^2 == r^2 in FP64 for r in {2.5, 3, 5}
EXACT = {2.5: [(2.5, 0.0, 0.0), (1.5, 2.0, 0.0)], 3.0: [(3.0, 0.0, 0.0), (1.0, 2.0, 2.0)], 5.0: [(5.0, 0.0, 0.0), (3.0, 4.0, 0.0)]}


def _perms_signs(v, rng, count):
    out = []
    for _ in range(count):
        p = rng.permutation(3)
        s = rng.choice([-1.0, 1.0], 3)
        out.append(np.array(v)[p] * s)
    return out


def _boundary_pair(q, u, r, d2f):
    """The representable points q + t u on both sides of d2f == r^2 (largest t with d2 < R, smallest
    with d2 >= R), found by bisection on t and a nextafter walk of the result."""
    R = r * r
    lo, hi = 0.0, 2.0 * r
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if mid == lo or mid == hi:
            break
        if d2f(q + mid * u, q) < R:
            lo = mid
        else:
            hi = mid
    return q + lo * u, q + hi * u


def boundary_cloud(r, seed=1, nq=48, per=12, offset=(651000.0, 6861000.0, 120.0), background=40_000, extent=60.0):
    """Returns (points, query centres). Queries sit on a 2^-10 m grid (so q +- exact offsets are exact)."""
    rng = np.random.default_rng(seed)
    off = np.asarray(offset)
    Q = off + rng.integers(0, int(extent * 1024), (nq, 3)) / 1024.0
    pts = [off + rng.uniform(0, extent, (background, 3))]
    R = r * r
    for q in Q:
        extra = []
        for v in EXACT[r]:
            for w in _perms_signs(v, rng, 4):
                p = q + w
                extra.append(p)
                for ax in range(3):  # one ulp either way on each axis
                    for dirn in (-np.inf, np.inf):
                        p2 = p.copy()
                        p2[ax] = np.nextafter(p2[ax], dirn)
                        extra.append(p2)
        for _ in range(per):
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            extra.extend(_boundary_pair(q, u, r, d2_sphere))
            u2 = np.array([u[0], u[1], 0.0])
            u2 /= np.linalg.norm(u2)
            a, b = _boundary_pair(q, u2, r, d2_circle)
            dz = rng.uniform(-3.0, 3.0)
            extra.extend([a + [0, 0, dz], b + [0, 0, dz]])
        for e in range(10, 31):  # across the FP32 band (2^-18 relative) and well inside / outside it
            for sgn in (-1.0, 1.0):
                u = rng.normal(size=3)
                u /= np.linalg.norm(u)
                extra.append(q + r * np.sqrt(1.0 + sgn * 2.0 ** -e) * u)
        pts.append(np.array(extra))
    return np.vstack(pts), Q


def magnitude_cloud(mag, seed=2, n=400):
    """Points spread over +-mag around a centre at mag (leaf-recentred FP32 terms ~mag^2), queries at
    some of the points and between them, radius mag (FP32 exponent range: mag ~ 1e15 puts
    amax + |b|^2 + R near 1e30, mag ~ 2e19 beyond FLT_MAX)."""
    rng = np.random.default_rng(seed)
    c = np.array([mag, -mag, 0.5 * mag])
    pts = c + rng.uniform(-mag, mag, (n, 3))
    q = np.vstack([pts[:20], c + rng.uniform(-mag, mag, (12, 3))])
    # exact axis offsets at distance mag from the first queries
    extra = [q[i] + s * np.eye(3)[ax] * mag for i in range(4) for ax in range(3) for s in (-1.0, 1.0)]
    return np.vstack([pts, extra]), q


def node_boundary_cloud(seed=3, n=20_000):
    """A cloud on the integer lattice of [0, 1024]^3 with both corners present: Box(center, radii) is
    exactly [0, 1024]^3, halfLengths = 2^-12 and level radii 2^(9 - L), so node centres and radii are
    exact and cell faces fall on lattice planes."""
    rng = np.random.default_rng(seed)
    pts = rng.integers(0, 1025, (n, 3)).astype(np.float64)
    pts = np.vstack([pts, [[0.0, 0.0, 0.0], [1024.0, 1024.0, 1024.0]]])
    return pts


def node_boundary_queries(centers, rl, seed=4, count=300):
    """Query centres and radii on the node-judgement thresholds of the given nodes (centre c, radius rl
    on every axis, a power of two), with (2, 3, 6) * rl, |.|^2 = 49 rl^2, all exact:
      kind 0  OUTSIDE equality |q - c| == r + rl on one axis (boxIntersect's `>`; the box kernels'
              faces q -+ r == c +- rl), r in {1, 2.5} * rl
      kind 1  INSIDE equality sum (|d| + rl)^2 == R: |d| = (1, 2, 5) rl, r = 7 rl
      kind 2  overlap equality sum (|d| - rl)^2 == R: |d| = (3, 4, 7) rl, r = 7 rl
      kind 3  two axes inside the node's slab (|d| < rl), one at r + rl exactly."""
    rng = np.random.default_rng(seed)
    qs, rs = [], []
    for _ in range(count):
        i = int(rng.integers(0, len(centers)))
        c, h = centers[i], float(rl[i])
        kind = int(rng.integers(0, 4))
        s = rng.choice([-1.0, 1.0], 3)
        perm = rng.permutation(3)
        if kind == 0:
            r = float(rng.choice([1.0, 2.5])) * h
            d = np.zeros(3)
            d[perm[0]] = h + r
            d[perm[1]] = float(rng.integers(0, 3)) * h
        elif kind == 1:
            d = np.array([1.0, 2.0, 5.0])[perm] * h
            r = 7.0 * h
        elif kind == 2:
            d = np.array([3.0, 4.0, 7.0])[perm] * h
            r = 7.0 * h
        else:
            r = 2.5 * h
            d = np.array([h + r, 0.5 * h, 0.0])[perm]
        qs.append(c + s * d)
        rs.append(r)
    return np.array(qs), np.array(rs)


def lattice_cloud(side=40, offset=(651000.0, 6861000.0, 100.0)):
    """Unit lattice side^3 at a LAS-like offset; kNN queries at lattice points (6 neighbours at 1, 12 at
    sqrt 2, 8 at sqrt 3, ...) and cube centres (8 neighbours at sqrt 0.75), deep inside and at the border."""
    g = np.arange(side, dtype=np.float64)
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) + np.asarray(offset)
    rng = np.random.default_rng(5)
    lat = np.asarray(offset) + rng.integers(0, side, (120, 3)).astype(np.float64)
    half = np.asarray(offset) + rng.integers(0, side - 1, (120, 3)) + 0.5
    corner = np.asarray(offset) + np.array([[0.0, 0.0, 0.0], [side - 1.0, side - 1.0, side - 1.0], [-0.5, 3.0, 7.0]])
    return pts, np.vstack([lat, half, corner])
