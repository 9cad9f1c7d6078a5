"""The reflected (real-ground) tensor restatement, oracle/ground.py, pinned on
the CPU against the reference's own vie_halfspace_cuboids (compiled from
vie_halfspace.c into oracle/_ref), against the vacuum substrate (no
reflection), and against image theory in the perfect-conductor limit
(the reflected field of a source above a very dense substrate is the field of
its mirror image with currents diag(-1, -1, +1), which the reference's PEC
tensor vie_cuboid_pec_n adds to the free-space one)."""
import math

import numpy as np
import pytest

from oracle import ground as G

K = 2 * math.pi * 200e6 / 299792458.0


def test_table_equals_the_reference_pair_by_pair(ref):
    shape, h, z0, eps = (6, 5, 4), 0.05, 0.075, 13 - 0.45j
    T = G.table(shape, h, z0, K, eps)
    worst = 0.0
    for dx, dy, it, js in [(0, 0, 0, 0), (1, 0, 0, 0), (2, 3, 1, 2), (5, 4, 3, 3), (3, 1, 0, 3), (0, 4, 2, 1),
                           (-2, 1, 1, 0), (4, -3, 2, 2), (-1, -2, 0, 1)]:
        n, order = ref.halfspace_cuboids([dx * h, dy * h, z0 + it * h], [h] * 3, [0, 0, z0 + js * h], [h] * 3, K, eps)
        mine = G.tensor(T, dx, dy, it + js)
        err = float(np.max(np.abs(mine - n)) / np.max(np.abs(n)))
        worst = max(worst, err)
    assert worst < 1e-8, worst


def test_fresnel_matches_the_reference(ref):
    for eps in (13 - 0.45j, 4.0, 80 - 30j):
        for u in (0.0, 0.3, 0.99, 1.0, 1.7, 3.6, 25.0):
            rs, rp = G.fresnel(eps, u)
            rrs, rrp = ref.halfspace_fresnel(eps, u)
            assert abs(rs - rrs) < 1e-14 and abs(rp - rrp) < 1e-14


def test_vacuum_substrate_reflects_nothing():
    T = G.table((3, 3, 2), 0.05, 0.1, K, 1.0)
    assert np.all(T == 0)


def test_dense_substrate_is_the_mirror_image(ref):
    h, z0 = 0.05, 0.125
    T = G.table((3, 2, 2), h, z0, K, 1e10 - 1e10j, order=64)
    for dx, dy, it, js in [(0, 0, 0, 0), (1, 1, 1, 0), (2, 0, 1, 1)]:
        tc, sc = [dx * h, dy * h, z0 + it * h], [0, 0, z0 + js * h]
        image = ref.cuboid_pec_n(tc, [h] * 3, sc, [h] * 3, K) - ref.cuboid_n(tc, [h] * 3, sc, [h] * 3, K)
        mine = G.tensor(T, dx, dy, it + js)
        assert np.max(np.abs(mine - image)) / np.max(np.abs(image)) < 1e-3
