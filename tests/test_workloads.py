"""The synthetic benchmark inputs (nec2c_amd/workloads.py)."""
import numpy as np

from nec2c_amd import workloads as W


def test_random_permittivity_range_and_passivity():
    eps = W.random_permittivity((16, 16, 16), seed=2)
    assert eps.shape == (16, 16, 16)
    assert eps.real.min() >= 1.5 and eps.real.max() <= 10.0
    assert np.all(eps.imag <= 0) and np.all(-eps.imag / eps.real <= 0.05 + 1e-12)
    assert np.array_equal(eps, W.random_permittivity((16, 16, 16), seed=2))


def test_dipole_impedance_is_physical():
    d = W.Dipole(center=(0, 0, 0))
    for f in (250e6, 300e6, 350e6):
        k = W.k0_of(f)
        Z = d.impedance_matrix(k)
        assert np.allclose(Z, Z.T, rtol=1e-10, atol=1e-9)          # reciprocity
        zin = 1 / np.linalg.solve(Z, d.rhs())[d.feed_index()]
        assert 20 < zin.real < 400                                  # radiating near resonance
    zr = 1 / np.linalg.solve(d.impedance_matrix(W.k0_of(300e6)), d.rhs())[d.feed_index()]
    assert 60 < zr.real < 110 and abs(zr.imag) < 80


def test_coupling_blocks_reciprocal_structure():
    n, h = 3, 0.05
    cent = W.cube_cells(n, h, origin=(0.0, 0.0, 0.0))
    eps = np.full(n ** 3, 4.0 - 0.1j)
    d = W.Dipole(center=(-0.1, 0.075, 0.075))
    k = W.k0_of(300e6)
    B, D = W.coupling_blocks(d, cent, h, eps, k)
    assert B.shape == (21, 3 * n ** 3) and D.shape == (3 * n ** 3, 21)
    # both blocks come from the same symmetric dyadic: D_iz ~ (eps-1) G_iz, B_zj ~ G_zj
    omega = k * W.C0
    ratio = D[2::3, :].T / ((eps - 1)[None, :] * B[:, 2::3])
    expected = -(1j * omega * W.MU0 / k ** 2) / h ** 3
    assert np.allclose(ratio, expected, rtol=1e-10)
