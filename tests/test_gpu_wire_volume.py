"""GPU parity of the conductor-to-volume coupling (vie_hip_conductor_volume_columns)
against the oracle's C restatement of NEC's classic near field averaged over
the cells (oracle/wire_volume.c: fields.c, calculations.c,
dielectric_wire.c:316-420), through the C-ABI.  Geometries follow the
reference's revision-1 examples (the 11-segment dipole of
examples/dielectric-r1/free-space.nec, the cells on a PEC plane of
pec-materials.nec), plus patches, cells inside a thick wire's radius, the
lumped-element range and the 64^3 lattice of the sweep (sampled cells)."""
import math

import numpy as np
import pytest

from nec2c_amd import vie, workloads as W
from oracle import wire_volume as V

pytestmark = pytest.mark.gpu


def segments(c: V.Conductors):
    return (c.x, c.y, c.z, c.si, c.bi, c.cab, c.sab, c.salp)


def patches(c: V.Conductors):
    return None if c.m == 0 else (c.patch_points, c.patch_area, c.t1, c.t2)


def nec_like_sources(rng, n, m=0, columns=None, dense=True):
    """One column per segment with NEC's sparsity for a straight wire (the
    basis of unknown j spans segments j-1, j, j+1), a patch column per patch,
    and one dense column."""
    stride = 3 * n + 3 * m
    cols = []
    for j in range(n):
        v = np.zeros(stride, complex)
        for s in (j - 1, j, j + 1):
            if 0 <= s < n:
                v[3 * s:3 * s + 3] = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        cols.append(v)
    for p in range(m):
        v = np.zeros(stride, complex)
        v[3 * n + 3 * p:3 * n + 3 * p + 3] = rng.standard_normal(3)
        cols.append(v)
    if dense:
        cols.append(rng.standard_normal(stride) + 1j * rng.standard_normal(stride))
    return np.array(cols)


def block(origin, shape, pitch):
    return W.cube_cells(shape, pitch, origin=origin)


def rowwise_rel(a, b):
    """Largest error relative to each row's largest entry; a row that is zero
    in the oracle must be exactly zero."""
    scale = np.abs(b).max(axis=1)
    zero = scale == 0
    assert (a[zero] == 0).all()
    return float(np.max(np.abs(a[~zero] - b[~zero]).max(axis=1) / scale[~zero]))


def check(k0, wlam, centers, sizes, chi, cond, src, rkh=1.0, ground=0, tol=1e-12):
    got = vie.conductor_volume_columns(k0, wlam, centers, sizes, chi, segments(cond), src, patches(cond), rkh=rkh,
                                       ground=ground)
    ref = V.conductor_volume(k0, wlam, centers, sizes, chi, cond, src, rkh=rkh, ground=ground, threads=8)
    err = rowwise_rel(got, ref)
    print("row-relative error", err, flush=True)
    assert got.shape == ref.shape and np.isfinite(got).all()
    assert err < tol
    return got


@pytest.mark.parametrize("ground", [0, 1])
def test_example_dipole_and_cube(ground):
    """The 11-segment 0.48 m dipole of the examples at 300 MHz and a 4x4x4
    block of 1 cm cells 0.1 m from it (free space), or on z = 0 (PEC)."""
    f = 300.0
    wlam, k0 = 299.8 / f, 2 * math.pi * f * 1e6 / V.C0
    dip = V.straight_wire([0.0, 0.0, 0.36 if ground == 0 else 0.05], [0.0, 0.0, 0.84 if ground == 0 else 0.53], 11,
                          0.001)
    cond = V.scaled(dip, wlam)
    z0 = 0.58 if ground == 0 else 0.0
    centers = block((0.10, -0.02, z0), (4, 4, 4), 0.01)
    sizes = np.full_like(centers, 0.01)
    chi = np.linspace(1.0, 3.0, len(centers)) - 0.05j
    rng = np.random.default_rng(ground)
    check(k0, wlam, centers, sizes, chi, cond, nec_like_sources(rng, 11), ground=ground)


@pytest.mark.parametrize("ground", [0, 1])
def test_wire_and_patches(ground):
    f = 250.0
    wlam, k0 = 299.8 / f, 2 * math.pi * f * 1e6 / V.C0
    w = V.straight_wire([0.05, 0.0, 0.1], [0.1, 0.05, 0.5], 7, 0.002)
    pts = np.array([[0.3, 0.1, 0.2], [0.25, -0.1, 0.35], [0.2, 0.2, 0.15]])
    t1 = np.array([[1.0, 0, 0], [0, 1.0, 0], [0.6, 0.8, 0]])
    t2 = np.array([[0, 0.6, 0.8], [0, 0, 1.0], [0, 0, 1.0]])
    cond = V.scaled(V.Conductors(w.x, w.y, w.z, w.si, w.bi, w.cab, w.sab, w.salp, pts, np.full(3, 0.004), t1, t2),
                    wlam)
    centers = block((0.15, -0.05, 0.05), (3, 4, 3), 0.02)
    sizes = np.full_like(centers, 0.02)
    sizes[::2, 2] = 0.015  # irregular cells
    rng = np.random.default_rng(7 + ground)
    check(k0, wlam, centers, sizes, 2.0 - 0.3j, cond, nec_like_sources(rng, 7, 3), ground=ground)


def test_points_inside_a_thick_wire_and_lumped_range():
    """A 1 cm-radius wire through a column of cells: Gauss points within 0.9
    of its radius switch NEC's kernel radius; and a short rkh puts far
    segments on the lumped-element formula."""
    f = 300.0
    wlam, k0 = 299.8 / f, 2 * math.pi * f * 1e6 / V.C0
    w = V.straight_wire([0.12, 0.0, 0.3], [0.12, 0.0, 0.9], 12, 0.01)
    cond = V.scaled(w, wlam)
    centers = block((0.11, -0.01, 0.55), (1, 1, 4), 0.02)  # cells around the wire axis
    centers = np.concatenate([centers, block((0.5, 0.3, 0.2), (2, 2, 2), 0.03)])
    sizes = np.concatenate([np.full((4, 3), 0.02), np.full((8, 3), 0.03)])
    rng = np.random.default_rng(11)
    src = nec_like_sources(rng, 12)
    check(k0, wlam, centers, sizes, 1.5, cond, src)
    check(k0, wlam, centers, sizes, 1.5, cond, src, rkh=0.3)


def test_sweep_lattice_sampled():
    """The 64^3 lattice of the sweep (edge 0.5 m) with the 21-segment dipole
    0.1 m from a face at 250 MHz: every column on all cells, a seeded sample
    of cells checked against the oracle."""
    f = 250.0
    wlam, k0 = 299.8 / f, 2 * math.pi * f * 1e6 / V.C0
    n, pitch = 64, 0.5 / 64
    centers = W.cube_cells(n, pitch)
    sizes = np.full_like(centers, pitch)
    dip = V.straight_wire([-0.1, 0.25, 0.01], [-0.1, 0.25, 0.49], 21, 0.001)
    cond = V.scaled(dip, wlam)
    rng = np.random.default_rng(3)
    src = nec_like_sources(rng, 21, dense=False)
    chi = np.full(len(centers), 3.0 - 0.1j)
    got = vie.conductor_volume_columns(k0, wlam, centers, sizes, chi, segments(cond), src)
    assert got.shape == (21, 3 * len(centers)) and np.isfinite(got).all()
    pick = np.concatenate([rng.choice(len(centers), 40, replace=False), [0, len(centers) - 1]])
    ref = V.conductor_volume(k0, wlam, centers[pick], sizes[pick], chi[pick], cond, src, threads=8)
    cols = (3 * pick[:, None] + np.arange(3)[None, :]).reshape(-1)
    err = rowwise_rel(got[:, cols], ref)
    print("64^3 sampled error", err, flush=True)
    assert err < 1e-12


def test_zero_columns_and_skipped_segments():
    """A column with no current yields zeros; segments that radiate in no
    column are never evaluated; rejections name the problem."""
    f = 300.0
    wlam, k0 = 299.8 / f, 2 * math.pi * f * 1e6 / V.C0
    cond = V.scaled(V.straight_wire([0.0, 0.0, 0.36], [0.0, 0.0, 0.84], 11, 0.001), wlam)
    centers = block((0.10, -0.02, 0.58), (2, 2, 2), 0.02)
    sizes = np.full_like(centers, 0.02)
    src = np.zeros((3, 33), complex)
    src[1, 15:18] = [1.0, 0.5j, -0.25]  # segment 5 only
    got = check(k0, wlam, centers, sizes, 2.0, cond, src)
    assert (got[0] == 0).all() and (got[2] == 0).all()
    bad = list(segments(cond))
    bad[3] = bad[3].copy()
    bad[3][4] = 0.0
    with pytest.raises(RuntimeError, match="segment length"):
        vie.conductor_volume_columns(k0, wlam, centers, sizes, 2.0, bad, src)
    with pytest.raises(RuntimeError, match="finite ground"):
        vie.conductor_volume_columns(k0, wlam, centers, sizes, 2.0, segments(cond), src, ground=2)


def test_workspaces_released_and_reallocated():
    """vie_hip_release_workspaces frees the kept buffers; the next call
    allocates again and gives the same result."""
    f = 300.0
    wlam, k0 = 299.8 / f, 2 * math.pi * f * 1e6 / V.C0
    cond = V.scaled(V.straight_wire([0.0, 0.0, 0.36], [0.0, 0.0, 0.84], 11, 0.001), wlam)
    centers = block((0.10, -0.02, 0.58), (2, 2, 2), 0.02)
    sizes = np.full_like(centers, 0.02)
    src = nec_like_sources(np.random.default_rng(1), 11)
    a = vie.conductor_volume_columns(k0, wlam, centers, sizes, 2.0, segments(cond), src)
    assert vie.lib().vie_hip_release_workspaces() == 1
    b = vie.conductor_volume_columns(k0, wlam, centers, sizes, 2.0, segments(cond), src)
    assert np.array_equal(a, b)
