"""Pins of the volume-to-conductor oracle (oracle/conductor_rows.py, a
restatement of dielectric_wire.c:70-214 and :1164-1262) against what the
mathematics fixes: numpy's Gauss-Legendre rule, finite differences of the
scalar Green function for both dyadics, the point-source limit, image theory
over a PEC plane, and a loop over the row assembly."""
import math

import numpy as np
import pytest

from oracle import conductor_rows as R


def test_gauss_rule_is_gauss_legendre():
    for order in (6, 8, 10):
        x, w = R.gauss(order)
        xr, wr = np.polynomial.legendre.leggauss(order)
        assert np.allclose(x, xr, atol=2e-15) and np.allclose(w, wr, atol=2e-15)


def green(p, k0):
    r = np.linalg.norm(p)
    return np.exp(-1j * k0 * r) / (4 * math.pi * r)


def fd_hessian(p, k0, h=1e-4):
    H = np.zeros((3, 3), complex)
    e = np.eye(3) * h
    for a in range(3):
        for b in range(3):
            H[a, b] = (green(p + e[a] + e[b], k0) - green(p + e[a] - e[b], k0) - green(p - e[a] + e[b], k0)
                       + green(p - e[a] - e[b], k0)) / (4 * h * h)
    return H


def test_e_dyadic_is_k2_plus_grad_grad_of_g():
    """A tiny cell far from the point: N ~ V (k^2 I + grad grad) g(r), with the
    Hessian from finite differences (a flipped sign or a dropped term fails)."""
    k0 = 7.0
    obs = np.array([0.9, -0.4, 0.7])
    c, s = np.array([0.05, 0.02, -0.01]), np.array([1e-3, 2e-3, 1.5e-3])
    N = R.point_cuboid_n(obs, c, s, k0)
    d = obs - c
    expect = np.prod(s) * (k0 ** 2 * green(d, k0) * np.eye(3) + fd_hessian(d, k0))
    # the cell's finite size enters at (s/r)^2 ~ 3e-6, the differences at ~1e-8
    assert np.allclose(N, expect, rtol=2e-5, atol=0)
    assert np.allclose(N, N.T, rtol=1e-13, atol=0)


def test_h_dyadic_is_grad_g_cross():
    k0 = 5.0
    obs = np.array([-0.6, 0.8, 0.3])
    c, s = np.array([0.0, 0.1, 0.0]), np.array([2e-3, 1e-3, 1e-3])
    H = R.point_cuboid_h(obs, c, s, k0)
    d = obs - c
    h = 1e-6
    # column q: (gradient of g with respect to the observation point) x e_q
    grad = np.array([(green(d + h * e, k0) - green(d - h * e, k0)) / (2 * h) for e in np.eye(3)])
    expect = np.prod(s) * np.stack([np.cross(grad, e) for e in np.eye(3)], axis=1)
    assert np.allclose(H, expect, rtol=2e-5, atol=0)


def test_point_source_limit_converges():
    k0 = 10.0
    obs = np.array([0.5, 0.5, 0.5])
    errs = []
    for size in (0.04, 0.02, 0.01):
        s = np.full(3, size)
        N = R.point_cuboid_n(obs, np.zeros(3), s, k0)
        P = size ** 3 * (k0 ** 2 * green(obs, k0) * np.eye(3) + fd_hessian(obs, k0, 1e-3))
        errs.append(np.abs(N - P).max() / np.abs(P).max())
    # second order in the cell size
    assert errs[0] > errs[1] > errs[2] and errs[2] < 1e-3
    assert 3.5 < errs[0] / errs[1] < 4.5 and 3.5 < errs[1] / errs[2] < 4.5


@pytest.mark.parametrize("magnetic", [False, True])
def test_pec_is_image_theory(magnetic):
    """Over a PEC plane: the field of a current equals free space plus the image
    current (horizontal components reversed, vertical kept) at the mirrored cell."""
    k0 = 6.0
    obs = np.array([0.3, -0.2, 0.4])
    c, s = np.array([0.1, 0.05, 0.15]), np.array([0.02, 0.03, 0.02])
    f = R.point_cuboid_h if magnetic else R.point_cuboid_n
    got = R.point_background(obs, c, s, k0, magnetic, ground=1)
    ci = c * np.array([1, 1, -1])
    for q in range(3):
        J = np.eye(3)[q]
        Jimg = J * np.array([-1.0, -1.0, 1.0])
        expect = f(obs, c, s, k0) @ J + f(obs, ci, s, k0) @ Jimg
        assert np.allclose(got @ J, expect, rtol=1e-13, atol=0)


MU_0 = 1.0 / (R.EPSILON_0 * R.C0 ** 2)  # consistent with k0 = omega / c0


def _gl(n, a, b):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (b - a) * x + 0.5 * (b + a), 0.5 * (b - a) * w


def _mixed_potential_field(obs, center, size, q, k0, n=14):
    """E at obs of the uniform current density e_q (A/m^2) in a cuboid, in the
    mixed-potential form E = -i w A - grad(phi) under exp(+i w t): the vector
    potential A = mu0 int g J dV over the volume, and the scalar potential of
    the surface charge sigma = (J . n) / (i w) on the two faces normal to q
    (the current has no divergence inside).  Independent of the (k^2 + grad
    grad) dyadic and of the 1/(i w eps0) the row assembly applies to it."""
    omega = k0 * R.C0
    axes = [_gl(n, center[a] - size[a] / 2, center[a] + size[a] / 2) for a in range(3)]
    X, Y, Z = np.meshgrid(axes[0][0], axes[1][0], axes[2][0], indexing="ij")
    W = np.einsum("i,j,k->ijk", axes[0][1], axes[1][1], axes[2][1])
    Rv = np.stack([obs[0] - X, obs[1] - Y, obs[2] - Z], -1)
    r = np.linalg.norm(Rv, axis=-1)
    A = MU_0 * np.sum(W * np.exp(-1j * k0 * r) / (4 * math.pi * r))
    E = np.zeros(3, complex)
    E[q] = -1j * omega * A
    others = [a for a in range(3) if a != q]
    (u, wu), (v, wv) = axes[others[0]][0:2], axes[others[1]][0:2]
    U, V = np.meshgrid(u, v, indexing="ij")
    WF = np.outer(wu, wv)
    for side in (+1, -1):
        pts = np.zeros(U.shape + (3,))
        pts[..., q] = center[q] + side * size[q] / 2
        pts[..., others[0]], pts[..., others[1]] = U, V
        Rf = obs - pts
        rf = np.linalg.norm(Rf, axis=-1)
        # grad_obs of exp(-ikR)/(4 pi R)
        dg = (-(1 + 1j * k0 * rf) * np.exp(-1j * k0 * rf) / (4 * math.pi * rf ** 3))[..., None] * Rf
        sigma = side / (1j * omega)
        E -= (sigma / R.EPSILON_0) * np.sum(WF[..., None] * dg, axis=(0, 1))
    return E


def _biot_savart_field(obs, center, size, q, k0, n=14):
    """H at obs of the uniform current density e_q in a cuboid:
    int (1 + ikR) exp(-ikR) / (4 pi R^2) (J x R_hat) dV, R = obs - r'."""
    axes = [_gl(n, center[a] - size[a] / 2, center[a] + size[a] / 2) for a in range(3)]
    X, Y, Z = np.meshgrid(axes[0][0], axes[1][0], axes[2][0], indexing="ij")
    W = np.einsum("i,j,k->ijk", axes[0][1], axes[1][1], axes[2][1])
    Rv = np.stack([obs[0] - X, obs[1] - Y, obs[2] - Z], -1)
    r = np.linalg.norm(Rv, axis=-1)
    f = W * (1 + 1j * k0 * r) * np.exp(-1j * k0 * r) / (4 * math.pi * r ** 3)
    J = np.eye(3)[q]
    return np.sum(f[..., None] * np.cross(J, Rv), axis=(0, 1, 2))


def test_wire_rows_are_the_tested_field_of_the_cell_current():
    """A wire row entry is t . E(r_w) in V/m of the cell's uniform current
    density e_q (the volume unknown is a current density: the volume rows
    scale the wire field by -i w eps0 chi).  E comes from the mixed
    potentials above, so a wrong 1/(i w eps0) (sign, eps0 vs mu0, omega
    power) or a dropped charge term fails."""
    rng = np.random.default_rng(3)
    k0 = 2 * math.pi * 300e6 / R.C0
    centers = rng.uniform(0.1, 0.3, (3, 3))
    sizes = rng.uniform(0.01, 0.03, (3, 3))
    wp = rng.uniform(-0.25, -0.05, (2, 3))
    wt = rng.standard_normal((2, 3))
    wt /= np.linalg.norm(wt, axis=1, keepdims=True)
    rows = R.conductor_rows(k0, centers, sizes, wp, wt, ground=0)
    for w in range(2):
        for cc in range(3):
            for q in range(3):
                want = wt[w] @ _mixed_potential_field(wp[w], centers[cc], sizes[cc], q, k0)
                got = rows[w, 3 * cc + q]
                assert abs(got - want) <= 1e-8 * abs(rows[w]).max(), (w, cc, q, got, want)


def test_patch_rows_follow_nec_patch_equations():
    """NEC's patch equations (matrix.c:426-434, cmss) are psalp t1 . (n x H)
    for the first and -psalp t2 . (n x H) for the second (n = t1 x t2: the
    self terms -1/2 and +1/2 carry those signs).  H here is the Biot-Savart
    field of the cell's uniform current, and n x H is formed explicitly, so a
    swapped row order, a sign or a wrong H fails."""
    rng = np.random.default_rng(5)
    k0 = 2 * math.pi * 250e6 / R.C0
    centers = rng.uniform(0.1, 0.3, (2, 3))
    sizes = rng.uniform(0.01, 0.03, (2, 3))
    pp = rng.uniform(-0.25, -0.05, (2, 3))
    t1 = rng.standard_normal((2, 3))
    t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
    t2 = np.cross(rng.standard_normal((2, 3)), t1)
    t2 /= np.linalg.norm(t2, axis=1, keepdims=True)
    scale = np.array([0.7, 1.3])
    rows = R.conductor_rows(k0, centers, sizes, patch_points=pp, patch_t1=t1, patch_t2=t2, patch_scale=scale,
                            ground=0)
    for p in range(2):
        n = np.cross(t1[p], t2[p])
        for cc in range(2):
            for q in range(3):
                H = _biot_savart_field(pp[p], centers[cc], sizes[cc], q, k0)
                nxh = np.cross(n, H)
                first, second = scale[p] * (t1[p] @ nxh), -scale[p] * (t2[p] @ nxh)
                tol = 1e-8 * abs(rows[2 * p:2 * p + 2]).max()
                assert abs(rows[2 * p, 3 * cc + q] - first) <= tol
                assert abs(rows[2 * p + 1, 3 * cc + q] - second) <= tol


@pytest.mark.parametrize("ground", [0, 1])
def test_c_restatement_matches_numpy(ground):
    """oracle/conductor_rows.c (the CPU baseline) entry for entry against the
    numpy restatement, threaded, and its first-failure index."""
    rng = np.random.default_rng(8 + ground)
    k0 = 2 * math.pi * 250e6 / R.C0
    centers = rng.uniform(0.0, 0.2, (6, 3)) + np.array([0, 0, 0.05])
    sizes = rng.uniform(0.01, 0.02, (6, 3))
    wp = rng.uniform(-0.3, -0.05, (3, 3))
    wt = rng.standard_normal((3, 3))
    wt /= np.linalg.norm(wt, axis=1, keepdims=True)
    pp = rng.uniform(0.3, 0.5, (2, 3))
    t1 = np.array([[1.0, 0, 0], [0, 0.6, 0.8]])
    t2 = np.array([[0, 1.0, 0], [1.0, 0, 0]])
    ps = [0.3, 0.9]
    ref = R.conductor_rows(k0, centers, sizes, wp, wt, pp, t1, t2, ps, ground=ground)
    got, failed = R.conductor_rows_c(k0, centers, sizes, wp, wt, pp, t1, t2, ps, ground=ground, threads=3)
    assert failed == -1
    scale = np.abs(ref).max(axis=1, keepdims=True)
    assert np.max(np.abs(got - ref) / scale) < 1e-13
    bad = pp.copy()
    bad[1, 2] = math.nan
    _, failed = R.conductor_rows_c(k0, centers, sizes, wp, wt, bad, t1, t2, ps, ground=ground, threads=3)
    assert failed == (3 + 1) * 6 + 0
