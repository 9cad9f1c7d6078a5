"""GPU parity of the mesher's relationship checks (vie_hip_box_overlap,
vie_hip_conductor_clearance) against the oracle's restatement of
validate_box_relationships (dielectric_mesh.c:75-123) and evaluate_clearance
(dielectric_mesh.c:125-316), through the C-ABI: the first overlapping pair
exactly, and the clearance figures bit for bit, across pair tiles, at the
per-voxel 64^3 model's full size."""
import numpy as np
import pytest

from nec2c_amd import vie
from oracle import mesh_check as M

pytestmark = pytest.mark.gpu


def lattice(n, pitch=0.5 / 64):
    """Per-voxel boxes of an n^3 block (touching faces only: no overlap)."""
    i = np.indices((n, n, n)).reshape(3, -1).T.astype(float)
    lo = i * pitch
    return lo, lo + pitch


@pytest.mark.parametrize("count", [0, 1, 2, 255, 256, 257, 700])
@pytest.mark.parametrize("seed", [0, 1])
def test_first_pair_matches_oracle(count, seed):
    rng = np.random.default_rng(seed + count)
    lo = rng.uniform(0, 4, (count, 3))
    hi = lo + rng.uniform(0.05, 0.4, (count, 3))
    got = vie.box_overlap(lo, hi)
    want = M.box_overlap(lo, hi)
    print(count, seed, got, want, flush=True)
    assert got == want


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_non_finite_corners_take_the_exact_scan(bad):
    """A non-finite corner goes through the same single scan kernel, whose
    fmax / fmin treat NaN as the reference's C library does (the other
    operand), so the first overlapping pair is the reference's."""
    lo, hi = lattice(8, 1.0)  # touching only: no pair overlaps until a corner goes bad
    for corner, k, a in ((lo, 300, 0), (hi, 77, 1), (lo, 450, 2)):
        c = corner.copy()
        c[k, a] = bad
        args = (c, hi) if corner is lo else (lo, c)
        want = M.box_overlap(*args)
        if np.isnan(bad):  # fmax / fmin take the other box's coordinate: new overlaps appear
            assert want is not None
        assert vie.box_overlap(*args) == want


def test_no_overlap_in_a_voxel_lattice_and_planted_pairs():
    lo, hi = lattice(24)  # 13824 boxes, 9.6e7 pairs
    assert vie.box_overlap(lo, hi) is None
    for k in (0, 255, 256, 5000, len(lo) - 1):
        # a duplicate of box k appended at the end overlaps box k only
        lo2, hi2 = np.vstack([lo, lo[k]]), np.vstack([hi, hi[k]])
        assert vie.box_overlap(lo2, hi2) == (k, len(lo))
    # two plants: the earlier left box wins, whatever its right partner
    lo2, hi2 = np.vstack([lo, lo[9000], lo[300]]), np.vstack([hi, hi[9000], hi[300]])
    assert vie.box_overlap(lo2, hi2) == (300, len(lo) + 1)


def _conductors(rng, near):
    segs = []
    for _ in range(5):
        a = rng.uniform(-0.1, 0.6, 3) + (0 if near else 1.0)
        b = a + rng.normal(0, 0.2, 3)
        segs.append(M.wire(a, b, rng.uniform(1e-4, 3e-3), segments=int(rng.integers(1, 8))))
    seg = [np.concatenate([s[k] for s in segs]) for k in range(8)]
    pts = rng.uniform(-0.2, 0.7, (4, 3)) + (0 if near else 1.0)
    return seg, (pts, rng.uniform(1e-5, 1e-3, 4))


@pytest.mark.parametrize("near", [True, False])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_clearance_matches_oracle_bitwise(near, seed):
    rng = np.random.default_rng(10 + seed)
    lo, hi = lattice(20, 0.5 / 20)
    lo, hi = lo + rng.uniform(0, 1e-3, lo.shape), hi - rng.uniform(0, 1e-3, hi.shape)
    cl = np.max(hi - lo, axis=1)
    cl[::97] = 0.0  # boxes without cells are skipped
    seg, pat = _conductors(rng, near)
    got = vie.conductor_clearance(lo, hi, cl, segments=seg, patches=pat)
    want = M.clearance(lo, hi, cl, segments=seg, patches=pat)
    print(near, seed, got, flush=True)
    for k, v in want.items():
        assert got[k] == v, (k, got[k], v)
    assert got["has_conductors"] == 1 and got["insufficient"] == (want["worst_margin"] < -1e-12)
    assert got["insufficient"] == near


def test_clearance_edge_cases():
    lo, hi = lattice(4)
    cl = np.max(hi - lo, axis=1)
    seg = M.wire([0, 0, 1], [0, 0, 2], 1e-3)
    # no conductors, or no box with cells: nothing evaluated (mesh->has_conductors = 0)
    for args in (dict(segments=None), dict(segments=seg, cell=np.zeros_like(cl))):
        r = vie.conductor_clearance(lo, hi, args.get("cell", cl), segments=args["segments"])
        assert r["has_conductors"] == 0 and r["worst_box"] == 0 and r["minimum_gap"] == np.finfo(float).max
    # patches only
    pat = (np.array([[0.5, 0.5, 0.5], [-1.0, 0.0, 0.0]]), np.array([1e-4, 2e-4]))
    got = vie.conductor_clearance(lo, hi, cl, patches=pat)
    want = M.clearance(lo, hi, cl, patches=pat)
    assert all(got[k] == v for k, v in want.items())
    with pytest.raises(ValueError):
        vie.conductor_clearance(lo, hi, cl[:-1], segments=seg)


def test_per_voxel_64cube_dipole():
    """The 64^3 per-voxel model (262144 BOX records) with the 21-segment
    dipole of the coupled benchmark: no overlap, clearance as the oracle."""
    lo, hi = lattice(64)
    assert vie.box_overlap(lo, hi) is None
    cl = np.full(len(lo), 0.5 / 64)
    seg = M.wire([-0.1, 0.25, 0.25 - 0.24], [-0.1, 0.25, 0.25 + 0.24], 1e-3, segments=21)
    got = vie.conductor_clearance(lo, hi, cl, segments=seg)
    want = M.clearance(lo, hi, cl, segments=seg)
    assert all(got[k] == v for k, v in want.items())
    # the kept device buffers are freed; the next call allocates again
    assert vie.lib().vie_hip_release_workspaces() == 1
    assert vie.box_overlap(lo[:300], hi[:300]) is None
