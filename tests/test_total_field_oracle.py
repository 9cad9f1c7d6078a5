"""Pins of the coupled far-field oracle (oracle/total_field.py: ffld / fflds,
radiation.c:36-373; dw_far_field and dw_radiated_power,
dielectric_wire.c:768-840) against what the mathematics fixes: NEC's
segment sums equal 2 pi times the radiation integral of the current
A + B sin(ks) + C cos(ks) along the segment (brute-force quadrature), a patch
radiates as a short segment of the same moment, image theory over a PEC
plane, the radiated power of a short dipole, and the interference of the
wire and volume fields."""
import math

import numpy as np
import pytest

from oracle import farfield as F
from oracle import total_field as T


def random_segment(rng):
    si = rng.uniform(0.01, 0.2)
    t = rng.standard_normal(3)
    t /= np.linalg.norm(t)
    c = rng.uniform(-0.5, 0.5, 3)
    return si, t, c


@pytest.mark.parametrize("trial", range(4))
def test_segment_sum_is_the_radiation_integral(trial):
    rng = np.random.default_rng(trial)
    si, t, c = random_segment(rng)
    cur = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    theta, phi = rng.uniform(0.05, 3.1), rng.uniform(0, 2 * math.pi)
    phx, phy, roz, thx, thy, thz, rox, roy = T._unit_vectors(theta, phi)
    rhat = np.array([rox, roy, roz])
    got = T._segment_sum(([c[0]], [c[1]], [c[2]], [si], [t[0]], [t[1]], [t[2]]), cur, rox, roy, roz)
    x, w = np.polynomial.legendre.leggauss(80)
    s, ws = 0.5 * si * x, 0.5 * si * w
    current = cur[0] + cur[1] * np.sin(T.NEC_TP * s) + cur[2] * np.cos(T.NEC_TP * s)
    expect = T.NEC_TP * np.sum(ws * current * np.exp(1j * T.NEC_TP * (rhat @ c + (rhat @ t) * s))) * t
    assert np.allclose(got, expect, rtol=1e-12, atol=1e-14 * np.abs(expect).max())


def test_broadside_direction_uses_the_series():
    """omega -> 0 (the direction normal to the segment): NEC's small-argument
    forms agree with the radiation integral too."""
    seg = ([0.0], [0.0], [0.0], [0.05], [0.0], [0.0], [1.0])
    cur = np.array([1.0 + 0.5j, -0.3j, 0.7])
    got = T._segment_sum(seg, cur, 1.0, 0.0, 0.0)  # r = x, t = z
    x, w = np.polynomial.legendre.leggauss(40)
    s, ws = 0.025 * x, 0.025 * w
    current = cur[0] + cur[1] * np.sin(T.NEC_TP * s) + cur[2] * np.cos(T.NEC_TP * s)
    assert abs(got[2] - T.NEC_TP * np.sum(ws * current)) < 1e-9 * abs(got[2])


def test_patch_radiates_as_a_short_segment():
    v, phase = np.array([0.3, -0.5, 0.8]), 1.0 - 0.4j
    J = v * phase
    area = 1e-4
    p = np.array([0.1, -0.2, 0.3])
    length = 1e-5
    t = v / np.linalg.norm(v)
    A = np.linalg.norm(v) * phase * area / length  # the same moment: A l t = J area
    for theta, phi in ((0.4, 1.0), (2.0, 4.0), (1.3, 0.2)):
        ep = T.conductor_far_field(theta, phi, patches=([p], [area]), pcur=J)
        seg = ([p[0]], [p[1]], [p[2]], [length], [t[0]], [t[1]], [t[2]])
        es = T.conductor_far_field(theta, phi, seg, np.array([A, 0.0, 0.0]))
        assert np.allclose(ep, es, rtol=1e-6, atol=0)


def test_pec_is_image_theory():
    rng = np.random.default_rng(5)
    n = 4
    si = rng.uniform(0.02, 0.05, n)
    t = rng.standard_normal((n, 3))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    c = rng.uniform(0.05, 0.4, (n, 3))
    cur = rng.standard_normal(3 * n) + 1j * rng.standard_normal(3 * n)
    pts = rng.uniform(0.05, 0.3, (2, 3))
    pcur = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    seg = (c[:, 0], c[:, 1], c[:, 2], si, t[:, 0], t[:, 1], t[:, 2])
    mseg = (np.r_[c[:, 0], c[:, 0]], np.r_[c[:, 1], c[:, 1]], np.r_[c[:, 2], -c[:, 2]], np.r_[si, si],
            np.r_[t[:, 0], t[:, 0]], np.r_[t[:, 1], t[:, 1]], np.r_[t[:, 2], -t[:, 2]])
    mcur = np.r_[cur, -cur]
    mpts = np.r_[pts, pts * [1, 1, -1]]
    mpcur = np.r_[pcur, -(pcur.reshape(2, 3) * [1, 1, -1]).reshape(-1)]
    for theta, phi in ((0.3, 0.5), (1.2, 2.5), (1.5, 5.0)):
        g = T.conductor_far_field(theta, phi, seg, cur, (pts, [1e-3, 2e-3]), pcur, ground=1)
        f = T.conductor_far_field(theta, phi, mseg, mcur, (mpts, [1e-3, 2e-3, 1e-3, 2e-3]), mpcur, ground=0)
        assert np.allclose(g, f, rtol=1e-12, atol=1e-14 * np.abs(f).max())


def test_short_dipole_power():
    """P = eta k^2 (I l)^2 / (12 pi) for a short constant current I (A), through
    the 3896-point rule (NEC's eta; NEC holds lengths in wavelengths and
    currents in amperes per wavelength, so the coefficient is I / wlam)."""
    wlam = 299.8 / 100.0
    l = 1e-3  # wavelengths
    I = 1.0 + 0.5j
    seg = ([0.0], [0.0], [0.0], [l], [0.0], [0.0], [1.0])
    P = T.radiated_power(3896, wlam, 2 * math.pi / wlam, seg=seg, cur=np.array([I / wlam, 0, 0]))
    k = 2 * math.pi / wlam
    expect = T.NEC_ETA * k * k * (abs(I) * l * wlam) ** 2 / (12 * math.pi)
    assert abs(P - expect) < 1e-5 * expect


def test_wire_and_volume_fields_interfere():
    """The total field adds the complex fields: the power of wire + volume
    differs from the sum of their powers by the cross term, and each part
    alone matches its own formula."""
    wlam = 299.8 / 300.0
    k0 = 2 * math.pi * 300e6 / F.C0
    seg = ([0.0], [0.0], [0.1 / wlam], [0.03 / wlam], [0.0], [0.0], [1.0])
    cur = np.array([1.0, 0.2j, 0.1])
    centers = np.array([[0.05, 0.0, 0.1]])
    sizes = np.array([[0.01, 0.01, 0.01]])
    J = np.array([0.0, 0.0, 2e4 + 1e4j])
    both = dict(seg=seg, cur=cur, centers=centers, sizes=sizes, current=J)
    for theta, phi in ((0.7, 0.2), (2.2, 1.9)):
        et, ep = T.total_far_field(theta, phi, wlam, k0, **both)
        wt, wp = T.conductor_far_field(theta, phi, seg, cur)
        d = np.array([[math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)]])
        a = F.radiation_amplitudes(centers, sizes, J, k0, d)[0]
        t = np.array([math.cos(theta) * math.cos(phi), math.cos(theta) * math.sin(phi), -math.sin(theta)])
        assert abs(et - (wt * wlam + a @ t)) < 1e-12 * abs(et)
    p_both = T.radiated_power(974, wlam, k0, **both)
    p_wire = T.radiated_power(974, wlam, k0, seg=seg, cur=cur)
    p_vol = T.radiated_power(974, wlam, k0, centers=centers, sizes=sizes, current=J)
    assert abs(p_both - p_wire - p_vol) > 1e-3 * p_both
    # the volume part alone is the scattered power with NEC's eta in place of the exact one
    assert abs(p_vol - F.scattered_power(centers, sizes, J, k0) * F.ETA0 / T.NEC_ETA) < 1e-12 * p_vol
