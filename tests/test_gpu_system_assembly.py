"""GPU: the conductor rows (B) and columns (D) assembled straight into a
system's device storage (dielectric_fft_assemble_conductor_rows / _columns)
give the same operator as the same blocks computed on the device, copied to
the host and written through dielectric_fft_entry (the reference's path,
dielectric_wire.c:1164-1262 and :1132-1136); reading an entry afterwards
brings the device blocks back into the host storage unchanged."""
import ctypes
import math

import numpy as np
import pytest

from nec2c_amd import vie, workloads as W

pytestmark = pytest.mark.gpu


def nec_like_sources(rng, n):
    src = np.zeros((n, 3 * n), complex)
    for j in range(n):
        for s in (j - 1, j, j + 1):
            if 0 <= s < n:
                src[j, 3 * s:3 * s + 3] = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    return src


def case(pec):
    n, pitch = 10, 0.5 / 64
    f = 300.0
    wlam, k0 = 299.8 / f, W.k0_of(f * 1e6)
    lift = 4 * pitch if pec else 0.0  # the PEC gap is a multiple of half the pitch
    centers = W.cube_cells(n, pitch) + np.array([0.0, 0.0, lift])
    sizes = np.full_like(centers, pitch)
    eps = np.full(len(centers), 3.0 - 0.2j)
    segs = 11
    start, end = np.array([-0.1, 0.04, 0.01 + lift]), np.array([-0.1, 0.04, 0.49 + lift])
    seg = W.nec_straight_wire(start, end, segs, 0.001, wlam)
    points = np.stack([seg[0], seg[1], seg[2]], -1) * wlam
    tangents = np.stack([seg[5], seg[6], seg[7]], -1)
    rng = np.random.default_rng(4 if pec else 3)
    src = nec_like_sources(rng, segs)
    C = W.Dipole(center=(-0.1, 0.04, 0.25 + lift), length=0.48, segments=segs).impedance_matrix(k0)
    return dict(n=n, pitch=pitch, wlam=wlam, k0=k0, centers=centers, sizes=sizes, eps=eps, seg=seg, points=points,
                tangents=tangents, src=src, C=C, segs=segs)


def build(cs, pec):
    s = vie.System(cs["centers"], [vie.MaterialSpec(1, 3.0)], cs["segs"], cs["k0"], pec=pec, pitch=cs["pitch"])
    s.set_permittivities(cs["eps"])
    return s


@pytest.mark.parametrize("pec", [False, True])
def test_device_assembly_matches_host_blocks(pec):
    cs = case(pec)
    g = 1 if pec else 0
    chi = cs["eps"] - 1.0
    rows = vie.volume_conductor_rows(cs["k0"], cs["centers"], cs["sizes"], cs["points"], cs["tangents"], ground=g)
    cols = vie.conductor_volume_columns(cs["k0"], cs["wlam"], cs["centers"], cs["sizes"], chi, cs["seg"], cs["src"],
                                        ground=g)
    host = build(cs, pec)
    host.set_blocks(C=cs["C"], B=rows, D=cols.T)
    dev = build(cs, pec)
    dev.set_blocks(C=cs["C"])
    dev.assemble_rows(cs["k0"], cs["centers"], cs["sizes"], cs["points"], cs["tangents"], ground=g)
    dev.assemble_columns(cs["k0"], cs["wlam"], cs["centers"], cs["sizes"], chi, cs["seg"], cs["src"], ground=g)
    rng = np.random.default_rng(9)
    x = rng.standard_normal(host.order) + 1j * rng.standard_normal(host.order)
    for prec in (vie.FP64, vie.MIXED):
        a, b = host.apply_ex(x, precision=prec), dev.apply_ex(x, precision=prec)
        print("pec", pec, "precision", prec, "max diff", np.abs(a - b).max(), flush=True)
        assert np.array_equal(a, b)
    # an entry read brings B and D back into the host storage, bit for bit
    L = vie.lib()
    c = cs["segs"]
    for r, q in ((0, 0), (5, 100), (10, 3 * len(cs["centers"]) - 1)):
        p = L.dielectric_fft_entry(dev.s, r, c + q)
        assert complex(*(ctypes.c_double * 2).from_address(p)) == rows[r, q]
        p = L.dielectric_fft_entry(dev.s, c + q, r)
        assert complex(*(ctypes.c_double * 2).from_address(p)) == cols[r, q]
    assert np.array_equal(host.apply_ex(x), dev.apply_ex(x))


def test_device_assembly_rejects_a_mismatch():
    cs = case(False)
    s = build(cs, False)
    with pytest.raises(RuntimeError, match="conductor count"):
        s.assemble_rows(cs["k0"], cs["centers"], cs["sizes"], cs["points"][:5], cs["tangents"][:5])
    with pytest.raises(ValueError, match="cell count"):
        s.assemble_columns(cs["k0"], cs["wlam"], cs["centers"][:10], cs["sizes"][:10], 2.0, cs["seg"], cs["src"])
    with pytest.raises(ValueError, match="one column per conductor"):
        s.assemble_columns(cs["k0"], cs["wlam"], cs["centers"], cs["sizes"], 2.0, cs["seg"], cs["src"][:4])
