"""Pins of the conductor-to-volume oracle (oracle/wire_volume.c: NEC's classic
near field of frozen basis coefficients, fields.c / calculations.c, averaged
over volume cells as dielectric_wire.c:316-420 does) against what the
mathematics fixes: the field of sine, cosine and constant current filaments
with their end charges by brute-force quadrature, the Romberg integral
against a high-order rule, the lumped current element far away, the patch
field as an elementary current, image theory over a PEC plane, the thin-wire
radius inside a wire, and the 6^3 cell average with the -i omega eps0 chi
scaling."""
import math

import numpy as np
import pytest

from oracle import wire_volume as V

K = 2 * math.pi


def filament(s, z, rh, kind, n=400):
    """(E_z, E_rho) / C of a current filament on [-s/2, s/2] of the axis with
    its end charges: k^2 int I G dz' + grad int I'_total G dz', G = e^{-ikR}/R."""
    x, w = np.polynomial.legendre.leggauss(n)
    zs, ws = 0.5 * s * x, 0.5 * s * w
    if kind == "sine":
        I, Ip, ends = np.sin(K * zs), K * np.cos(K * zs), (math.sin(-K * s / 2), math.sin(K * s / 2))
    elif kind == "cosine":
        I, Ip, ends = np.cos(K * zs), -K * np.sin(K * zs), (math.cos(K * s / 2), math.cos(K * s / 2))
    else:
        I, Ip, ends = np.ones_like(zs), 0 * zs, (1.0, 1.0)

    def grad(zp):
        R = np.sqrt((z - zp) ** 2 + rh * rh)
        f = -(1 + 1j * K * R) * np.exp(-1j * K * R) / R ** 3
        return f * (z - zp), f * rh

    R = np.sqrt((z - zs) ** 2 + rh * rh)
    gz, gr = grad(zs)
    ez = K * K * np.sum(ws * I * np.exp(-1j * K * R) / R) + np.sum(ws * Ip * gz)
    er = np.sum(ws * Ip * gr)
    (a1, b1), (a2, b2) = grad(np.array(-s / 2)), grad(np.array(s / 2))
    ez += ends[0] * a1 - ends[1] * a2
    er += ends[0] * b1 - ends[1] * b2
    return ez, er


@pytest.mark.parametrize("s,z,rh", [(0.05, 0.3, 0.2), (0.04, -0.1, 0.15), (0.1, 0.02, 0.3)])
def test_thin_wire_kernel_is_the_filament_field(s, z, rh):
    """eksc = -CONST1 (the filament field) for all six components; the
    closed forms to NEC's rounded 2 pi, the constant current's axial term to
    the Romberg tolerance."""
    e = V.eksc(s, z, rh)
    c = -V.NEC_CONST1
    for kind, iz, ir, tol in (("sine", 0, 1, 1e-8), ("cosine", 2, 3, 1e-8), ("constant", 4, 5, 2e-6)):
        ez, er = filament(s, z, rh, kind)
        assert abs(e[iz] - c * ez) < tol * abs(c * ez), kind
        assert abs(e[ir] - c * er) < 1e-8 * abs(c * er), kind


def test_romberg_integral():
    """NEC's Romberg rule accepts a step when successive estimates differ by
    less than 1e-4 relative: within that of the exact integral, and, for an
    integrand smooth over the whole interval, Simpson's rule on one step."""
    x, w = np.polynomial.legendre.leggauss(300)
    f = lambda t, zpk, rkb2: (np.cos(np.sqrt(rkb2 + (t - zpk) ** 2)) + 1j * np.sin(np.sqrt(rkb2 + (t - zpk) ** 2))) \
        / np.sqrt(rkb2 + (t - zpk) ** 2)
    for zpk, rkb2, half in ((1.9, 0.8, 0.2), (-0.3, 0.05, 0.15), (0.0, 1.7, 0.4), (0.1, 0.002, 0.3)):
        exact = np.sum(half * w * f(half * x, zpk, rkb2))
        assert abs(V.intx(zpk, rkb2, -half, half) - exact) < 1e-4 * abs(exact)
    zpk, rkb2, half = 0.5, 900.0, 0.05
    simpson = half / 3 * (f(-half, zpk, rkb2) + 4 * f(0.0, zpk, rkb2) + f(half, zpk, rkb2))
    assert abs(V.intx(zpk, rkb2, -half, half) - simpson) < 1e-15 * abs(simpson)


def one_segment(direction=(0.0, 0.6, 0.8), s=0.01, center=(0.0, 0.0, 0.0), radius=1e-4):
    d = np.asarray(direction, float)
    return V.Conductors(np.array([center[0]]), np.array([center[1]]), np.array([center[2]]), np.array([s]),
                        np.array([radius]), d[:1], d[1:2], d[2:])


def test_lumped_element_beyond_rkh():
    c = one_segment()
    p = np.array([1.3, -0.4, 0.9])  # 1.63 wavelengths away
    for kind in range(3):
        src = np.zeros(3, complex)
        src[kind] = 1.0
        lumped = V.near_field(c, src, p, 1.0, rkh=1.0)
        exact = V.near_field(c, src, p, 1.0, rkh=100.0)
        if kind == 1:
            assert np.abs(lumped).max() == 0.0  # the odd (sine) current has no lumped moment
        else:
            # a short element: (s / r)^2 ~ 4e-5
            assert np.abs(lumped - exact).max() < 1e-4 * np.abs(exact).max()


def dyadic(r):
    R = np.linalg.norm(r)
    g = np.exp(-1j * V.NEC_TP * R) / R
    k = V.NEC_TP
    return g * (k * k - 1j * k / R - 1 / R ** 2) * np.eye(3) + g * (-k * k + 3j * k / R + 3 / R ** 2) / R ** 2 * np.outer(r, r)


def patch_only(center, area, t1, t2):
    z = np.zeros(0)
    return V.Conductors(z, z, z, z, z, z, z, z, np.array([center]), np.array([area]), np.array([t1]), np.array([t2]))


def test_patch_field_is_an_elementary_current():
    c0 = np.array([0.1, 0.2, 0.3])
    pc = patch_only(c0, 0.002, [1.0, 0, 0], [0, 0.6, 0.8])
    for cur in (np.array([1.0, 0, 0]), np.array([0, 0.6, 0.8]), np.array([0.3, -0.2, 0.5j])):
        for p in (np.array([0.5, -0.3, 0.9]), np.array([-0.2, 0.7, 0.1])):
            e = V.near_field(pc, cur.astype(complex), p, 1.0)
            # only the current's components along t1 and t2 radiate
            tang = (cur @ [1.0, 0, 0]) * np.array([1.0, 0, 0]) + (cur @ [0, 0.6, 0.8]) * np.array([0, 0.6, 0.8])
            expect = -V.NEC_CONST1 * 0.002 * (dyadic(p - c0) @ tang) * (4.771341188 / 4.771341189)
            assert np.allclose(e, expect, rtol=1e-12, atol=0)


def mirrored(c: V.Conductors):
    """Segments and patches with their images over z = 0 appended."""
    cat = np.concatenate
    return V.Conductors(cat([c.x, c.x]), cat([c.y, c.y]), cat([c.z, -c.z]), cat([c.si, c.si]), cat([c.bi, c.bi]),
                        cat([c.cab, c.cab]), cat([c.sab, c.sab]), cat([c.salp, -c.salp]),
                        cat([c.patch_points, c.patch_points * [1, 1, -1]]), cat([c.patch_area, c.patch_area]),
                        cat([c.t1, c.t1 * [1, 1, -1]]), cat([c.t2, c.t2 * [1, 1, -1]]))


def test_pec_is_image_theory():
    """Over a PEC plane: the free-space field of the conductors plus their
    images (z and the vertical direction mirrored, currents reversed)."""
    w = V.straight_wire([0.0, 0.0, 0.1], [0.05, 0.02, 0.45], 5, 1e-3)
    p = patch_only([0.3, 0.1, 0.2], 0.003, [1.0, 0, 0], [0, 0.6, 0.8])
    c = V.Conductors(w.x, w.y, w.z, w.si, w.bi, w.cab, w.sab, w.salp, p.patch_points, p.patch_area, p.t1, p.t2)
    rng = np.random.default_rng(2)
    src = rng.standard_normal(3 * 5 + 3) + 1j * rng.standard_normal(3 * 5 + 3)
    # the image patch's current: its components along the mirrored tangents reversed
    both = np.concatenate([src[:15], -src[:15], src[15:], -src[15:] * [1, 1, -1]])
    for point in ([0.4, -0.2, 0.3], [0.2, 0.3, 0.05], [1.5, 0.2, 0.7]):
        g = V.near_field(c, src, point, 1.0, ground=1)
        f = V.near_field(mirrored(c), both, point, 1.0, ground=0)
        assert np.allclose(g, f, rtol=1e-12, atol=1e-14 * np.abs(f).max())


def test_thin_wire_radius_inside_a_wire():
    """A point on the axis of one segment sees every segment through the
    reduced kernel with that segment's radius: sqrt(rho^2 + a^2)."""
    a = 0.004
    w = V.straight_wire([0.0, 0.0, 0.0], [0.0, 0.0, 0.2], 4, a)
    src = np.zeros(12, complex)
    src[3 * 3:3 * 3 + 3] = [1.0, 0.5j, -0.2]  # segment 3 radiates
    p = np.array([0.0, 0.0, 0.03])  # on segment 0's axis
    e = V.near_field(w, src, p, 1.0)
    k = V.eksc(w.si[3], p[2] - w.z[3], a)
    expect_z = k[4] * 1.0 + k[0] * 0.5j + k[2] * -0.2
    assert abs(e[2] - expect_z) < 1e-13 * abs(expect_z) and abs(e[0]) == 0 and abs(e[1]) == 0
    # off the wire (beyond 0.9 a from every axis) the kernel radius is zero
    q = np.array([0.02, 0.0, 0.03])
    k = V.eksc(w.si[3], q[2] - w.z[3], 0.02)
    assert abs(V.near_field(w, src, q, 1.0)[2] - (k[4] + 0.5j * k[0] - 0.2 * k[2])) < 1e-13 * abs(k[4])


def test_cell_average_and_column_scaling():
    wlam = 299.8 / 300.0
    k0 = 2 * math.pi * 300e6 / V.C0
    w = V.scaled(V.straight_wire([0.0, 0.0, 0.36], [0.0, 0.0, 0.84], 11, 0.001), wlam)
    centers = np.array([[0.11, 0.0, 0.59], [0.13, -0.01, 0.61]])
    sizes = np.array([[0.02, 0.02, 0.02], [0.02, 0.01, 0.02]])
    chi = np.array([2.0 - 0.1j, 1.5])
    rng = np.random.default_rng(3)
    sources = rng.standard_normal((2, 33)) + 1j * rng.standard_normal((2, 33))
    sources[0, :15] = 0.0  # a column with zero segments (the reference skips them)
    got = V.conductor_volume(k0, wlam, centers, sizes, chi, w, sources, threads=2)
    x, wt = np.polynomial.legendre.leggauss(6)
    omega = k0 * V.C0
    for col in range(2):
        for cell in range(2):
            acc = np.zeros(3, complex)
            for i in range(6):
                for j in range(6):
                    for l in range(6):
                        p = centers[cell] + 0.5 * sizes[cell] * np.array([x[i], x[j], x[l]])
                        acc += wt[i] * wt[j] * wt[l] / 8 * V.near_field(w, sources[col], p, wlam)
            expect = -1j * omega * V.EPSILON_0 * chi[cell] * acc
            assert np.allclose(got[col, 3 * cell:3 * cell + 3], expect, rtol=1e-12, atol=1e-13 * np.abs(expect).max())
