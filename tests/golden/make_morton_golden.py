#!/usr/bin/env python3
"""Writes tests/golden/morton_ref.npz: Morton keys computed by the reference's own libmorton
(inc/encoding/libmorton/morton.h, compiled from the reference sources by oracle/Makefile into
oracle/_ref/libref_morton.so with the reference's Release flags).

Inputs: edge triples (0, 2^21 - 1, every single bit) plus 4096 seeded random triples < 2^21 for the
3-D keys (xyz, keys); edge pairs (0, 2^31 - 1, every single bit) plus 2048 seeded random pairs < 2^31
for the 2-D keys of MortonEncoder2D (xy2, keys2).
Run after `make -C oracle` in a tree where the reference is mounted; the .npz is data only."""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SO = os.path.join(ROOT, "oracle", "_ref", "libref_morton.so")


def triples():
    m = (1 << 21) - 1
    edge = [(0, 0, 0), (m, m, m), (m, 0, 0), (0, m, 0), (0, 0, m), (1000, 2000, 3000)]
    for b in range(21):
        edge += [(1 << b, 0, 0), (0, 1 << b, 0), (0, 0, 1 << b), (1 << b, 1 << b, 1 << b)]
    rng = np.random.default_rng(20240601)
    rnd = rng.integers(0, 1 << 21, size=(4096, 3), dtype=np.uint64)
    return np.vstack([np.array(edge, np.uint64), rnd])


def pairs():
    m = (1 << 31) - 1
    edge = [(0, 0), (m, m), (m, 0), (0, m), (5, 9)]
    for b in range(31):
        edge += [(1 << b, 0), (0, 1 << b), (1 << b, 1 << b)]
    rng = np.random.default_rng(20240602)
    return np.vstack([np.array(edge, np.uint64), rng.integers(0, 1 << 31, size=(2048, 2), dtype=np.uint64)])


def main():
    if not os.path.exists(SO):
        sys.exit(f"{SO} missing: build it with make -C oracle where the reference is mounted")
    L = C.CDLL(SO)
    L.ref_morton3d_encode.restype = C.c_uint64
    L.ref_morton3d_encode.argtypes = [C.c_uint64] * 3
    t = triples()
    L.ref_morton2d_encode.restype = C.c_uint64
    L.ref_morton2d_encode.argtypes = [C.c_uint64] * 2
    keys = np.array([L.ref_morton3d_encode(int(x), int(y), int(z)) for x, y, z in t], np.uint64)
    p = pairs()
    keys2 = np.array([L.ref_morton2d_encode(int(x), int(y)) for x, y in p], np.uint64)
    np.savez_compressed(os.path.join(HERE, "morton_ref.npz"), xyz=t, keys=keys, xy2=p, keys2=keys2)
    print(f"wrote {len(keys)} reference 3-D and {len(keys2)} 2-D Morton keys")


if __name__ == "__main__":
    main()
