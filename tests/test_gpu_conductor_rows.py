"""GPU parity of the volume-to-conductor rows (vie_hip_volume_conductor_rows)
against the oracle's restatement of dielectric_wire.c:70-214 and :1164-1262,
through the C-ABI: irregular cells in free space and over a PEC plane, wires
and patches in one call, the 64^3 lattice with the sweep's dipole (sampled
entries), and the reference's failure report for an evaluation that meets
a non-finite match point."""
import math

import numpy as np
import pytest

from nec2c_amd import vie, workloads as W
from oracle import conductor_rows as R

pytestmark = pytest.mark.gpu


def unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def geometry(rng, cells, wires, patches):
    c = rng.uniform(0.0, 0.2, (cells, 3)) + np.array([0.0, 0.0, 0.05])
    s = rng.uniform(0.005, 0.02, (cells, 3))
    wp = rng.uniform(-0.3, -0.05, (wires, 3))
    wt = unit(rng.standard_normal((wires, 3)))
    pp = rng.uniform(0.3, 0.5, (patches, 3))
    t1 = unit(rng.standard_normal((patches, 3)))
    t2 = unit(np.cross(t1, rng.standard_normal((patches, 3))))
    ps = rng.uniform(1e-3, 1e-2, patches)
    return c, s, wp, wt, pp, t1, t2, ps


def rowwise_rel(a, b):
    return float(np.max(np.abs(a - b).max(axis=1) / np.abs(b).max(axis=1)))


@pytest.mark.parametrize("ground", [0, 1])
def test_rows_match_oracle(ground):
    rng = np.random.default_rng(20 + ground)
    c, s, wp, wt, pp, t1, t2, ps = geometry(rng, 37, 5, 3)
    k0 = W.k0_of(300e6)
    got = vie.volume_conductor_rows(k0, c, s, wp, wt, pp, t1, t2, ps, ground=ground)
    ref = R.conductor_rows(k0, c, s, wp, wt, pp, t1, t2, ps, ground=ground)
    assert got.shape == (5 + 2 * 3, 3 * 37)
    err = rowwise_rel(got, ref)
    print("ground", ground, "row-relative error", err, flush=True)
    assert err < 1e-12


@pytest.mark.parametrize("freq", [2e9, 12e9])
def test_rows_match_oracle_electrically_large_cells(freq):
    """Cells of 0.2-0.7 (2 GHz) and beyond 1 (12 GHz) in k |size| / 2: the
    kernel's wider phase series and its per-node sincos."""
    rng = np.random.default_rng(int(freq // 1e9))
    c, s, wp, wt, pp, t1, t2, ps = geometry(rng, 29, 3, 2)
    k0 = W.k0_of(freq)
    for ground in (0, 1):
        got = vie.volume_conductor_rows(k0, c, s, wp, wt, pp, t1, t2, ps, ground=ground)
        ref = R.conductor_rows(k0, c, s, wp, wt, pp, t1, t2, ps, ground=ground)
        err = rowwise_rel(got, ref)
        print(freq, ground, "row-relative error", err, flush=True)
        assert err < 1e-12


def test_wires_only_and_patches_only():
    rng = np.random.default_rng(30)
    c, s, wp, wt, pp, t1, t2, ps = geometry(rng, 300, 4, 2)
    k0 = W.k0_of(150e6)
    gw = vie.volume_conductor_rows(k0, c, s, wp, wt)
    gp = vie.volume_conductor_rows(k0, c, s, patch_points=pp, patch_t1=t1, patch_t2=t2, patch_scale=ps)
    both = vie.volume_conductor_rows(k0, c, s, wp, wt, pp, t1, t2, ps)
    # one thread per pair: the rows of a combined call are the separate calls' rows
    assert np.array_equal(both[:4], gw) and np.array_equal(both[4:], gp)
    assert rowwise_rel(gw, R.conductor_rows(k0, c, s, wp, wt)) < 1e-12


def test_sweep_lattice_sampled():
    """The 64^3 lattice of the sweep (edge 0.5 m) with the 21-segment dipole
    0.1 m from a face: every row, a seeded sample of cells checked against
    the oracle, the rest checked for finiteness."""
    n, pitch = 64, 0.5 / 64
    centers = W.cube_cells(n, pitch)
    sizes = np.full_like(centers, pitch)
    seg = 0.48 / 21
    wp = np.array([[-0.1, 0.25, 0.25 - 0.24 + (i + 0.5) * seg] for i in range(21)])
    wt = np.tile([0.0, 0.0, 1.0], (21, 1))
    k0 = W.k0_of(250e6)
    got = vie.volume_conductor_rows(k0, centers, sizes, wp, wt)
    assert got.shape == (21, 3 * len(centers)) and np.isfinite(got).all()
    pick = np.random.default_rng(5).choice(len(centers), 60, replace=False)
    pick = np.concatenate([pick, [0, len(centers) - 1]])
    ref = R.conductor_rows(k0, centers[pick], sizes[pick], wp, wt)
    cols = (3 * pick[:, None] + np.arange(3)[None, :]).reshape(-1)
    err = float(np.max(np.abs(got[:, cols] - ref)) / np.abs(ref).max())
    print("64^3 sampled error", err, flush=True)
    assert err < 1e-12


def test_failure_names_the_first_failing_pair():
    rng = np.random.default_rng(40)
    c, s, wp, wt, pp, t1, t2, ps = geometry(rng, 10, 3, 2)
    k0 = W.k0_of(300e6)
    bad = wp.copy()
    bad[1, 0] = math.nan
    with pytest.raises(RuntimeError, match=r"wire 2, volume cell 1: volume-to-wire interaction failed"):
        vie.volume_conductor_rows(k0, c, s, bad, wt, pp, t1, t2, ps)
    cb = c.copy()
    cb[6, 2] = math.nan  # cell 7 fails for every conductor; the first is wire 1
    with pytest.raises(RuntimeError, match=r"wire 1, volume cell 7"):
        vie.volume_conductor_rows(k0, cb, s, wp, wt, pp, t1, t2, ps)
    bp = pp.copy()
    bp[1, 1] = math.nan
    with pytest.raises(RuntimeError, match=r"patch 2, volume cell 1: volume-to-patch interaction failed"):
        vie.volume_conductor_rows(k0, c, s, wp, wt, bp, t1, t2, ps)
    # the library recovers: a good call after the failures
    ok = vie.volume_conductor_rows(k0, c, s, wp, wt, pp, t1, t2, ps)
    assert rowwise_rel(ok, R.conductor_rows(k0, c, s, wp, wt, pp, t1, t2, ps)) < 1e-12


def test_leading_dimension_leaves_other_columns():
    """Rows land in a wider system row (ld > 3 cells) without touching the
    columns beyond the volume block."""
    rng = np.random.default_rng(50)
    c, s, wp, wt, pp, t1, t2, ps = geometry(rng, 9, 2, 1)
    k0 = W.k0_of(200e6)
    ld = 3 * 9 + 5
    out = np.full((4, ld), 7.0 + 7.0j)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (c, s, wp, wt, pp, t1, t2, ps)]
    ok = vie.lib().vie_hip_volume_conductor_rows(k0, 9, *[a.ctypes.data for a in arrs[:2]], 2,
                                                 *[a.ctypes.data for a in arrs[2:4]], 1,
                                                 *[a.ctypes.data for a in arrs[4:]], 0, out.ctypes.data, ld)
    assert ok, vie.last_error()
    ref = R.conductor_rows(k0, c, s, wp, wt, pp, t1, t2, ps)
    assert rowwise_rel(out[:, :27], ref) < 1e-12
    assert (out[:, 27:] == 7.0 + 7.0j).all()
