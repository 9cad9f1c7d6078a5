"""GPU parity of the coupled far field (vie_hip_total_far_field,
vie_hip_total_radiated_power) against the oracle's restatement of ffld /
fflds (radiation.c:36-373) and dw_far_field / dw_radiated_power
(dielectric_wire.c:768-840), through the C-ABI: wires, patches and a volume
current together, free space and over a PEC plane."""
import math

import numpy as np
import pytest

from nec2c_amd import vie, workloads as W
from oracle import total_field as T

pytestmark = pytest.mark.gpu


def scene(rng, ground):
    f = 300.0
    wlam, k0 = 299.8 / f, 2 * math.pi * f * 1e6 / 299792458.0
    lift = 0.05 if ground else 0.0
    seg = W.nec_straight_wire([0.0, 0.0, 0.36 + lift], [0.0, 0.0, 0.84 + lift], 11, 0.001, wlam)
    seg7 = (seg[0], seg[1], seg[2], seg[3], seg[5], seg[6], seg[7])  # no radius in the far field
    scur = (rng.standard_normal(33) + 1j * rng.standard_normal(33)) / wlam
    pts = (rng.uniform(0.1, 0.3, (3, 3)) + [0, 0, lift]) / wlam
    area = np.array([4e-4, 3e-4, 5e-4]) / wlam ** 2
    pcur = (rng.standard_normal(9) + 1j * rng.standard_normal(9)) / wlam
    centers = W.cube_cells((3, 3, 3), 0.02, origin=(0.10, -0.03, 0.58 if not ground else 0.02))
    sizes = np.full_like(centers, 0.02)
    J = 1e3 * (rng.standard_normal(3 * len(centers)) + 1j * rng.standard_normal(3 * len(centers)))
    return wlam, k0, seg7, scur, (pts, area), pcur, centers, sizes, J


@pytest.mark.parametrize("ground", [0, 1])
def test_far_field_matches_oracle(ground):
    rng = np.random.default_rng(40 + ground)
    wlam, k0, seg, scur, patches, pcur, centers, sizes, J = scene(rng, ground)
    th = list(rng.uniform(0.05, math.pi / 2 if ground else math.pi - 0.05, 25)) + [math.pi / 2, 0.3]
    ph = list(rng.uniform(0, 2 * math.pi, 25)) + [0.0, 1.0]
    et, ep = vie.total_far_field(k0, wlam, th, ph, seg, scur, patches, pcur, centers, sizes, J, ground=ground)
    ref = [T.total_far_field(a, b, wlam, k0, seg=seg, cur=scur, patches=patches, pcur=pcur, centers=centers,
                             sizes=sizes, current=J, ground=ground) for a, b in zip(th, ph)]
    rt = np.array([r[0] for r in ref])
    rp = np.array([r[1] for r in ref])
    scale = max(np.abs(rt).max(), np.abs(rp).max())
    err = max(np.abs(et - rt).max(), np.abs(ep - rp).max()) / scale
    print("ground", ground, "relative error", err, flush=True)
    assert err < 1e-12


@pytest.mark.parametrize("ground", [0, 1])
def test_radiated_power_matches_oracle(ground):
    rng = np.random.default_rng(50 + ground)
    wlam, k0, seg, scur, patches, pcur, centers, sizes, J = scene(rng, ground)
    for count in (974, 3896):
        p = vie.total_radiated_power(k0, wlam, count, seg, scur, patches, pcur, centers, sizes, J, ground=ground)
        pr = T.radiated_power(count, wlam, k0, seg=seg, cur=scur, patches=patches, pcur=pcur, centers=centers,
                              sizes=sizes, current=J, ground=ground)
        print("ground", ground, count, p, pr, flush=True)
        assert abs(p - pr) < 1e-12 * pr


def test_wires_only_and_volume_only():
    rng = np.random.default_rng(60)
    wlam, k0, seg, scur, patches, pcur, centers, sizes, J = scene(rng, 0)
    th, ph = [0.4, 1.1, 2.9], [0.2, 3.0, 5.5]
    et, ep = vie.total_far_field(k0, wlam, th, ph, segments=seg, segment_current=scur)
    for k in range(3):
        rt, rp = T.total_far_field(th[k], ph[k], wlam, k0, seg=seg, cur=scur)
        assert abs(et[k] - rt) < 1e-12 * abs(rt) and abs(ep[k] - rp) < 1e-12 * max(abs(rp), abs(rt))
    et, ep = vie.total_far_field(k0, wlam, th, ph, centers=centers, sizes=sizes, volume_current=J)
    for k in range(3):
        rt, rp = T.total_far_field(th[k], ph[k], wlam, k0, centers=centers, sizes=sizes, current=J)
        assert abs(et[k] - rt) < 1e-12 * abs(rt) and abs(ep[k] - rp) < 1e-12 * max(abs(rp), abs(rt))
