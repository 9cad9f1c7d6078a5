"""GPU parity of the scattered far field (vie_hip_radiation_amplitudes,
vie_hip_scattered_power) against the oracle's restatement of
dielectric_solver.c:124-217 and dielectric_ground.c:133-144, through the C-ABI."""
import math

import numpy as np
import pytest

from nec2c_amd import vie, workloads as W
from oracle import farfield as F

pytestmark = pytest.mark.gpu


def dirs(rng, n, upper=False):
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    if upper:
        d[:, 2] = np.abs(d[:, 2])
    return d


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("ground", [0, 1, 2])
def test_amplitudes_match_oracle_irregular_cells(ground):
    rng = np.random.default_rng(10 + ground)
    n, k0 = 1000, 2 * math.pi * 300e6 / F.C0
    c = rng.uniform(0.0, 0.4, (n, 3))
    s = rng.uniform(0.005, 0.03, (n, 3))
    J = rng.standard_normal(3 * n) + 1j * rng.standard_normal(3 * n)
    d = dirs(rng, 37, upper=ground > 0)
    rs = rp = None
    if ground == 2:
        rs = -0.6 + 0.2j + 0.1 * rng.standard_normal(len(d))
        rp = 0.5 - 0.3j + 0.1 * rng.standard_normal(len(d))
    a = vie.radiation_amplitudes(c, s, J, k0, d, ground=ground, rs=rs, rp=rp)
    ref = F.radiation_amplitudes(c, s, J, k0, d, ground=ground, rs=rs, rp=rp)
    print("ground", ground, "rel", rel(a, ref), flush=True)
    assert rel(a, ref) < 1e-12


def test_lattice_and_power_match_oracle():
    """A uniform lattice (the kernel's per-direction cell factor) carrying a
    random current, and the 974-direction scattered power in free space and
    over a PEC plane."""
    n, pitch = 12, 0.5 / 64
    centers = W.cube_cells(n, pitch) + np.array([0.0, 0.0, 0.05])
    sizes = np.full_like(centers, pitch)
    rng = np.random.default_rng(7)
    J = rng.standard_normal(3 * len(centers)) + 1j * rng.standard_normal(3 * len(centers))
    k0 = W.k0_of(350e6)
    d = dirs(rng, 101)
    assert rel(vie.radiation_amplitudes(centers, sizes, J, k0, d), F.radiation_amplitudes(centers, sizes, J, k0, d)) < 1e-12
    for ground in (0, 1):
        p = vie.scattered_power(centers, sizes, J, k0, ground=ground)
        pr = F.scattered_power(centers, sizes, J, k0, ground=ground)
        print("power", ground, p, pr, flush=True)
        assert abs(p - pr) <= 1e-12 * pr


def test_edge_cases():
    k0 = 10.0
    d = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0]])
    # no cells: zero field
    a = vie.radiation_amplitudes(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, complex), k0, d)
    assert a.shape == (2, 3) and not a.any()
    # a direction straight up over the ground (the reflection's horizontal frame is degenerate)
    c, s = np.array([[0.0, 0.0, 0.1]]), np.array([[0.01, 0.01, 0.01]])
    J = np.array([1.0, 0.5j, -0.2])
    up = np.array([[0.0, 0.0, 1.0]])
    assert rel(vie.radiation_amplitudes(c, s, J, k0, up, ground=1), F.radiation_amplitudes(c, s, J, k0, up, ground=1)) < 1e-13
    # more directions than one tile and a ragged last tile
    rng = np.random.default_rng(3)
    d2 = dirs(rng, 4099)
    c2 = rng.uniform(0, 0.1, (3, 3))
    J2 = rng.standard_normal(9) + 0j
    assert rel(vie.radiation_amplitudes(c2, np.full((3, 3), 0.01), J2, k0, d2),
               F.radiation_amplitudes(c2, np.full((3, 3), 0.01), J2, k0, d2)) < 1e-12
