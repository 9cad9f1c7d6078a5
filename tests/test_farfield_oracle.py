"""Pins of the far-field oracle (oracle/farfield.py, a restatement of
dielectric_solver.c:47-217 and dielectric_ground.c:133-144) against what the
mathematics fixes: the Hertzian-dipole far field of one cell, the cell factor
against brute-force quadrature, the radiated power of a small dipole, the
translation phase law, image theory over a PEC plane, and a statement-by-
statement loop over cells."""
import math

import numpy as np
import pytest

from oracle import farfield as F


def rand_dirs(rng, n):
    d = rng.standard_normal((n, 3))
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def test_single_cell_is_the_dipole_far_field():
    rng = np.random.default_rng(1)
    k0 = 2 * math.pi * 300e6 / F.C0
    c = np.array([[0.03, -0.02, 0.05]])
    s = np.array([[0.01, 0.02, 0.015]])
    J = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    d = rand_dirs(rng, 7)
    a = F.radiation_amplitudes(c, s, J, k0, d)
    V = float(np.prod(s))
    for q in range(len(d)):
        factor = np.prod([math.sin(0.5 * k0 * d[q, i] * s[0, i]) / (0.5 * k0 * d[q, i] * s[0, i]) for i in range(3)])
        phase = np.exp(1j * k0 * d[q] @ c[0])
        # E = i k eta0 / (4 pi) (p x r) x r with p = J V (here c0 mu0 = eta0)
        expect = 1j * k0 * F.C0 * F.MU_0 / (4 * math.pi) * V * factor * phase * np.cross(np.cross(J, d[q]), d[q])
        assert np.allclose(a[q], expect, rtol=1e-13, atol=0)
        assert abs(a[q] @ d[q]) < 1e-12 * np.linalg.norm(a[q])  # transverse


def test_cell_factor_against_quadrature():
    k0 = 40.0
    c = np.array([[0.1, 0.2, -0.05]])
    s = np.array([[0.03, 0.05, 0.02]])
    d = rand_dirs(np.random.default_rng(2), 5)
    x, w = np.polynomial.legendre.leggauss(24)
    pts = [c[0, a] + 0.5 * s[0, a] * x for a in range(3)]
    got = F.cell_phase_average(c, s, d, k0, 1.0)[:, 0]
    for q in range(len(d)):
        ph = np.exp(1j * k0 * (d[q, 0] * pts[0][:, None, None] + d[q, 1] * pts[1][None, :, None]
                                + d[q, 2] * pts[2][None, None, :]))
        avg = np.einsum("i,j,k,ijk->", w, w, w, ph) / 8.0
        assert abs(got[q] - avg) < 1e-13


def test_small_dipole_power():
    k0 = 2 * math.pi * 100e6 / F.C0
    h = 1e-4
    J = np.array([0.0, 0.0, 1.0 + 0.5j])
    P = F.scattered_power(np.zeros((1, 3)), np.full((1, 3), h), J, k0)
    p = abs(J[2]) * h ** 3
    # |a|^2 = (k eta0 p / 4 pi)^2 sin^2(theta); integral of sin^2 over the sphere is 8 pi / 3
    expect = (k0 * F.C0 * F.MU_0 * p / (4 * math.pi)) ** 2 * (8 * math.pi / 3) / (2 * F.ETA0)
    assert abs(P - expect) < 2e-5 * expect


def test_translation_phase():
    rng = np.random.default_rng(3)
    k0 = 15.0
    c = rng.uniform(-0.2, 0.2, (20, 3))
    s = rng.uniform(0.01, 0.03, (20, 3))
    J = rng.standard_normal(60) + 1j * rng.standard_normal(60)
    d = rand_dirs(rng, 9)
    t = np.array([0.3, -0.1, 0.25])
    a0 = F.radiation_amplitudes(c, s, J, k0, d)
    a1 = F.radiation_amplitudes(c + t, s, J, k0, d)
    assert np.allclose(a1, a0 * np.exp(1j * k0 * d @ t)[:, None], rtol=1e-12, atol=1e-14 * np.abs(a0).max())


@pytest.mark.parametrize("comp,sign", [(2, 1.0), (0, -1.0), (1, -1.0)])
def test_pec_ground_is_image_theory(comp, sign):
    """Over a PEC plane the reference's reflection equals the image current
    (horizontal components reversed, vertical kept) at the mirrored cell."""
    rng = np.random.default_rng(4 + comp)
    k0 = 12.0
    c = np.array([[0.05, -0.03, 0.2]])
    s = np.array([[0.02, 0.02, 0.02]])
    J = np.zeros(3, complex)
    J[comp] = 1.0 - 0.3j
    d = rand_dirs(rng, 11)
    d[:, 2] = np.abs(d[:, 2])  # the upper hemisphere
    a = F.radiation_amplitudes(c, s, J, k0, d, ground=1)
    Jimg = sign * J
    ci = c * np.array([1, 1, -1])
    expect = F.radiation_amplitudes(c, s, J, k0, d) + F.radiation_amplitudes(ci, s, Jimg, k0, d)
    assert np.allclose(a, expect, rtol=1e-12, atol=1e-14 * np.abs(expect).max())


@pytest.mark.parametrize("ground", [0, 1])
def test_vectorised_matches_loop(ground):
    rng = np.random.default_rng(5)
    k0 = 20.0
    c = rng.uniform(0.01, 0.3, (13, 3))
    s = rng.uniform(1e-6, 0.04, (13, 3))
    J = rng.standard_normal(39) + 1j * rng.standard_normal(39)
    d = rand_dirs(rng, 4)
    a = F.radiation_amplitudes(c, s, J, k0, d, ground=ground)
    for q in range(len(d)):
        ref = F.radiation_amplitude_loop(c, s, J, k0, d[q], ground=ground)
        assert np.allclose(a[q], ref, rtol=1e-12, atol=0)


def test_golden_directions():
    d = F.golden_directions(974)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0)
    assert abs(d[:, 2].mean()) < 1e-12 and d[:, 2].max() < 1
    h = F.golden_directions(974, ground=1)
    assert (h[:, 2] > 0).all()
