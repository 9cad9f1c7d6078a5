"""The real-ground operator on the device (dielectric_fft_prepare_ground):
the reflected-tensor table against the oracle (oracle/ground.py, pinned to
the reference's vie_halfspace_cuboids) and the reference itself, the volume
operator against a dense product (the reference's own free-space FFT
operator plus the reflected tensor of every pair), FP32 against FP64,
vacuum ground against free space, every x-pass variant alike, and solves
against a dense solve of the same matrix."""
import math

import numpy as np
import pytest

from nec2c_amd import vie, workloads as W
from oracle import ground as G

pytestmark = pytest.mark.gpu

K = 2 * math.pi * 200e6 / 299792458.0
EPS_G = 13 - 0.45j


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def lattice(shape, holes=True, seed=4):
    xyz = np.stack(np.meshgrid(*[np.arange(n) for n in shape], indexing="ij"), -1).reshape(-1, 3)
    if holes:
        xyz = xyz[np.arange(len(xyz)) % 7 != 3]
    eps = W.random_permittivity(shape, seed).reshape(-1)[np.ravel_multi_index(xyz.T, shape)]
    return xyz, eps


def ground_system(xyz, eps, h, z0, k, conductors=0, eps_g=EPS_G):
    centers = h * (xyz + 0.5)
    centers[:, 2] += z0 - 0.5 * h
    s = vie.System(centers, [vie.MaterialSpec(1, 1.5)], conductors, k, pitch=h, ground=eps_g)
    s.set_permittivities(eps)
    return s


def dense_reference(ref, shape, xyz, eps, h, z0, k, eps_g=EPS_G):
    """Free-space part from the reference's own FFT operator, reflected part
    from the oracle table; returns apply(x) of the whole volume operator."""
    plan = ref.plan(shape, xyz, h, k, eps[0])
    plan.set_permittivities(eps)
    NR = G.reflected_matrix(G.table(shape, h, z0, k, eps_g), xyz)
    chi = np.repeat(eps - 1, 3)
    return (lambda x: plan.apply(x) - chi * (NR @ x)), plan, NR


def test_device_table_matches_the_oracle_and_the_reference(ref):
    for shape, h, z0 in [((6, 5, 4), 0.05, 0.075), ((20, 16, 10), 0.02, 0.05)]:
        T = vie.ground_table(shape, h, z0, K, EPS_G)
        To = G.table(shape, h, z0, K, EPS_G)
        err = rel(T, To)
        print(f"table {shape}: device vs oracle {err:.2e}", flush=True)
        assert err < 1e-12
    T = vie.ground_table((6, 5, 4), 0.05, 0.075, K, EPS_G)
    for dx, dy, it, js in [(0, 0, 0, 0), (3, 1, 0, 3), (5, 4, 3, 3)]:
        n, _ = ref.halfspace_cuboids([dx * 0.05, dy * 0.05, 0.075 + it * 0.05], [0.05] * 3,
                                     [0, 0, 0.075 + js * 0.05], [0.05] * 3, K, EPS_G)
        assert rel(G.tensor(T, dx, dy, it + js), n) < 1e-8


@pytest.mark.parametrize("shape,h,z0", [((6, 5, 4), 0.05, 0.1), ((9, 7, 5), 0.04, 0.06)])
def test_ground_operator_matches_the_dense_product(ref, shape, h, z0):
    xyz, eps = lattice(shape)
    s = ground_system(xyz, eps, h, z0, K)
    apply_ref, plan, NR = dense_reference(ref, shape, xyz, eps, h, z0, K)
    x = np.random.default_rng(1).standard_normal(s.order) + 1j * np.random.default_rng(2).standard_normal(s.order)
    yr = apply_ref(x)
    y64 = s.apply_ex(x, vie.FP64)
    y32 = s.apply_ex(x, vie.MIXED)
    e64, e32 = rel(y64, yr), rel(y32, yr)
    # the reflected part is a real share of the product (not a negligible term)
    share = np.max(np.abs(np.repeat(eps - 1, 3) * (NR @ x))) / np.max(np.abs(yr))
    print(f"{shape}: FP64 {e64:.2e}, FP32 {e32:.2e}, reflected share {share:.2e}", flush=True)
    assert share > 1e-3
    assert e64 < 1e-10 and e32 < 2e-6
    s.close()


def test_reflected_part_of_the_product_pair_by_pair_against_the_reference(ref):
    """On a small lattice the reflected matrix is assembled from the
    reference's own vie_halfspace_cuboids for every pair of cells."""
    shape, h, z0 = (4, 3, 3), 0.05, 0.1
    xyz, eps = lattice(shape, holes=False)
    n = len(xyz)
    NR = np.zeros((3 * n, 3 * n), complex)
    cache = {}
    for a in range(n):
        for b in range(n):
            key = (xyz[a][0] - xyz[b][0], xyz[a][1] - xyz[b][1], xyz[a][2], xyz[b][2])
            if key not in cache:
                tc = [key[0] * h, key[1] * h, z0 + key[2] * h]
                cache[key], _ = ref.halfspace_cuboids(tc, [h] * 3, [0, 0, z0 + key[3] * h], [h] * 3, K, EPS_G)
            NR[3 * a:3 * a + 3, 3 * b:3 * b + 3] = cache[key]
    plan = ref.plan(shape, xyz, h, K, eps[0])
    plan.set_permittivities(eps)
    s = ground_system(xyz, eps, h, z0, K)
    x = np.random.default_rng(3).standard_normal(s.order) + 0.3j
    yr = plan.apply(x) - np.repeat(eps - 1, 3) * (NR @ x)
    err = rel(s.apply_ex(x, vie.FP64), yr)
    print(f"pair-by-pair reference: {err:.2e}", flush=True)
    assert err < 2e-7     # the reference's own quadrature tolerance (vie_halfspace.c:13)
    s.close()


def test_vacuum_ground_is_free_space():
    shape, h, z0 = (6, 5, 4), 0.05, 0.1
    xyz, eps = lattice(shape)
    sg = ground_system(xyz, eps, h, z0, K, eps_g=1.0)
    centers = h * (xyz + 0.5)
    centers[:, 2] += z0 - 0.5 * h
    sf = vie.System(centers, [vie.MaterialSpec(1, 1.5)], 0, K, pitch=h)
    sf.set_permittivities(eps)
    x = np.random.default_rng(5).standard_normal(sg.order) + 0j
    assert rel(sg.apply_ex(x, vie.FP64), sf.apply_ex(x, vie.FP64)) < 1e-15
    sg.close()
    sf.close()


@pytest.mark.parametrize("shape", [(20, 12, 16), (72, 6, 8)])
def test_every_x_pass_variant_gives_the_same_ground_product(shape):
    h, z0 = 0.02, 0.03
    xyz, eps = lattice(shape)
    s = ground_system(xyz, eps, h, z0, K)
    x = np.random.default_rng(6).standard_normal(s.order) + 0.5j
    base = {}
    ok = [v for v in range(16) if s.set_variants([0, 0, v, 0, 0], [0, 0, v, 0, 0])]
    assert any(v & 4 == 0 for v in ok) and any(v & 4 for v in ok)   # whole ky rows and paired halves
    assert any(v & 8 for v in ok)                                    # the XCD-grouped tile order
    # the register-resident kernel exists in FP32 only
    r16 = [v for v in range(16, 32) if s.set_variants([0, 0, ok[0], 0, 0], [0, 0, v, 0, 0])]
    if shape[0] > 64:
        assert r16
    for v64, v32 in [(v, v) for v in ok] + [(ok[0], v) for v in r16]:
        assert s.set_variants([0, 0, v64, 0, 0], [0, 0, v32, 0, 0])
        for prec, tol in ((vie.FP64, 1e-14), (vie.MIXED, 1e-6)):
            y = s.apply_ex(x, prec)
            if prec not in base:
                base[prec] = y
            assert rel(y, base[prec]) <= tol, (v64, v32, prec)
    print("ground x-pass variants", ok, "FP32 register-resident", r16, flush=True)
    s.close()


def test_ground_solves_match_a_dense_solve(ref):
    shape, h, z0 = (6, 5, 4), 0.05, 0.1
    xyz, eps = lattice(shape)
    s = ground_system(xyz, eps, h, z0, K)
    apply_ref, _, _ = dense_reference(ref, shape, xyz, eps, h, z0, K)
    n = s.order
    A = np.column_stack([apply_ref(np.eye(n, 1, -j).ravel().astype(complex)) for j in range(n)])
    b = np.random.default_rng(7).standard_normal(n) + 1j * np.random.default_rng(8).standard_normal(n)
    xd = np.linalg.solve(A, b)
    for prec in (vie.FP64, vie.MIXED):
        ok, x, rep = s.solve(b, precision=prec, restart=60, maximum_iterations=600, relative_tolerance=1e-10)
        r = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        print(f"ground solve precision {prec}: {rep.iterations} it, dense residual {r:.2e}, "
              f"error {rel(x, xd):.2e}", flush=True)
        assert ok and r < 1.05e-10 and rel(x, xd) < 1e-7
    s.close()
