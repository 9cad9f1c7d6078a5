"""Pins of the mesh-check oracle (oracle/mesh_check.c: validate_box_
relationships, dielectric_mesh.c:75-123, and evaluate_clearance,
dielectric_mesh.c:125-316) against the reference's own dielectric_mesh_build
(oracle/_ref/libmeshref.so, built from dielectric_mesh.c and shared.c), the
reference tests' overlap / touching / clearance cases
(tests/dielectric_mesh.py:163-230), and brute-force segment-box distances."""
import math

import numpy as np
import pytest

from oracle import mesh_check as M

needs_ref = pytest.mark.skipif(not M.ref_available(), reason="oracle/_ref/libmeshref.so not built here")


def random_boxes(rng, n, overlap_every=0):
    """Boxes on a jittered lattice (disjoint), a few enlarged to overlap."""
    g = int(math.ceil(n ** (1 / 3)))
    idx = np.array([(i % g, (i // g) % g, i // (g * g)) for i in range(n)], float)
    lo = idx * 0.1 + rng.uniform(0, 0.03, (n, 3))
    hi = lo + rng.uniform(0.02, 0.07, (n, 3))
    if overlap_every:
        for k in range(overlap_every, n, overlap_every):
            hi[k] += 0.09  # reaches into the next lattice cell
    return lo, hi


@needs_ref
def test_reference_overlap_case_and_touching_boxes():
    # tests/dielectric_mesh.py:163-169 and :171-184
    ok, rec, val, line, *_ = M.reference_mesh([[0, 0, 0], [0.5, 0, 0]], [[1, 1, 1], [1.5, 1, 1]])
    assert not ok and rec == "BOX" and val == "1,2" and line == 2
    assert M.box_overlap([[0, 0, 0], [0.5, 0, 0]], [[1, 1, 1], [1.5, 1, 1]]) == (0, 1)
    lo, hi = [[0, 0, 0], [0.1, 0, 0]], [[0.1, 0.1, 0.1], [0.2, 0.1, 0.1]]
    assert M.reference_mesh(lo, hi, frequency_mhz=10)[0]
    assert M.box_overlap(lo, hi) is None


@needs_ref
@pytest.mark.parametrize("seed", range(6))
def test_first_overlapping_pair_matches_reference(seed):
    rng = np.random.default_rng(seed)
    lo, hi = random_boxes(rng, 60, overlap_every=(0, 7, 13, 29, 59, 0)[seed])
    ok, rec, val, line, *_ = M.reference_mesh(lo, hi)
    pair = M.box_overlap(lo, hi)
    if pair is None:
        assert ok, (rec, val)
    else:
        # the reference reports the region ids (box index + 1 here) and the line of the right box
        assert not ok and rec == "BOX" and val == f"{pair[0] + 1},{pair[1] + 1}" and line == pair[1] + 1


def test_overlap_is_interior_only():
    """Shared faces, edges and corners are no overlap; any positive interior is."""
    lo, hi = [[0, 0, 0]], [[1, 1, 1]]
    for shift in ([1, 0, 0], [1, 1, 0], [1, 1, 1], [0, -1, 0]):
        s = np.array(shift, float)
        assert M.box_overlap(lo + [s], hi + [1 + s]) is None
    eps = 1e-9
    assert M.box_overlap(lo + [[1 - eps, 0, 0]], hi + [[2, 1, 1]]) == (0, 1)
    # the loop order: the first left box wins, then the first right box
    lo3, hi3 = [[0, 0, 0], [2, 0, 0], [0.5, 0, 0], [1.5, 0, 0]], [[1, 1, 1], [3, 1, 1], [0.9, 1, 1], [2.5, 1, 1]]
    assert M.box_overlap(lo3, hi3) == (0, 2)


def _format(r):
    kind = "patch" if r["worst_is_patch"] else "wire"
    return f"box={r['worst_box']} {kind}={r['worst_conductor']} measured={r['worst_gap']:.9g} " \
           f"required={r['worst_required']:.9g}"


@needs_ref
def test_reference_clearance_cases():
    box = ([[0, 0, 0]], [[0.1, 0.1, 0.1]])
    cl = M.reference_mesh(*box, frequency_mhz=10)[7]
    # a wire through the box (tests/dielectric_mesh.py:208-215)
    w = M.wire([0.05, 0.05, -0.1], [0.05, 0.05, 0.2], 0.001)
    ok, rec, val, *_ = M.reference_mesh(*box, frequency_mhz=10, segments=w)
    r = M.clearance(*box, cl, segments=w)
    assert not ok and rec == "CLEARANCE" and val == _format(r) and r["worst_margin"] < -1e-12
    # the default far wire (tests/dielectric_mesh.py:30): clear, same gaps
    far = M.wire([100, 0, 0], [100, 0, 0.1], 0.001)
    ok, rec, val, line, gap, req, has, _ = M.reference_mesh(*box, frequency_mhz=10, segments=far)
    r = M.clearance(*box, cl, segments=far)
    assert ok and has == 1 and gap == r["minimum_gap"] and req == r["required_gap"]
    # a patch next to the box (tests/dielectric_mesh.py:218-230)
    patches = (np.array([[0.05, 0.05, 0.15]]), np.array([0.01]))
    ok, rec, val, *_ = M.reference_mesh(*box, frequency_mhz=10, segments=far, patches=patches)
    r = M.clearance(*box, cl, segments=far, patches=patches)
    assert not ok and rec == "CLEARANCE" and val == _format(r) and r["worst_is_patch"] == 1


@needs_ref
@pytest.mark.parametrize("seed", range(5))
def test_clearance_matches_reference(seed):
    rng = np.random.default_rng(100 + seed)
    lo, hi = random_boxes(rng, 27)
    cl = M.reference_mesh(lo, hi, frequency_mhz=300.0)[7]
    far = 1.5 if seed % 2 else 0.0  # odd seeds clear the boxes: the smallest gap is compared
    segs = []
    for _ in range(4):
        a = rng.uniform(-0.3, 0.6, 3) + far
        b = a + rng.normal(0, 0.3 if not far else 0.05, 3)
        segs.append(M.wire(a, b, rng.uniform(1e-4, 5e-3), segments=int(rng.integers(1, 6))))
    seg = [np.concatenate([s[k] for s in segs]) for k in range(8)]
    pts = rng.uniform(-0.5, 0.8, (3, 3)) + far
    area = rng.uniform(1e-4, 4e-3, 3)
    ok, rec, val, line, gap, req, has, _ = M.reference_mesh(lo, hi, frequency_mhz=300.0, segments=seg,
                                                           patches=(pts, area))
    r = M.clearance(lo, hi, cl, segments=seg, patches=(pts, area))
    assert ok == bool(seed % 2)
    if ok:
        assert r["worst_margin"] >= -1e-12 and gap == r["minimum_gap"] and req == r["required_gap"]
    else:
        assert rec == "CLEARANCE" and val == _format(r)


@pytest.mark.parametrize("seed", range(20))
def test_segment_box_distance_is_the_minimum_over_the_segment(seed):
    rng = np.random.default_rng(seed)
    c, h = rng.normal(0, 1, 3), rng.uniform(0.1, 1.0, 3)
    s = rng.normal(0, 1.5, 3)
    e = s + rng.normal(0, 1.5, 3) if seed % 5 else s.copy()  # one in five is a point
    d2 = M.segment_box_d2(s, e, c, h)
    t = np.linspace(0, 1, 200001)[:, None]
    p = s + t * (e - s)
    q = np.clip(p, c - h / 2, c + h / 2)
    brute = np.min(np.sum((p - q) ** 2, axis=1))
    assert d2 <= brute + 1e-12 and brute - d2 <= 1e-7 * max(1.0, brute)
