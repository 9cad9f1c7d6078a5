"""Pins the oracle's FFT operator, hybrid system and GMRES restatement
(oracle/oracle.py) against brute force, closed forms, the committed golden
products of the reference build, and oracle/_ref itself."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_operator_small.json")
HOLES = ((5, 3, 2), [0, 0, 0, 4, 2, 1, 1, 0, 0, 0, 1, 0, 0, 0, 1, 1, 1, 1])
FAR = ((9, 5, 3), [0, 0, 0, 8, 4, 2, 4, 1, 0])


def cvec(a):
    return np.array([complex(*v) for v in a])


def dense_operator(shape, xyz, h, k, er):
    """Matrix of -(eps-1) N_cd + eps delta from pairwise cuboid integrals
    (the reference's brute-force check, tests/vie_fft_test.c:19-25)."""
    xyz = np.asarray(xyz).reshape(-1, 3)
    n = 3 * len(xyz)
    A = np.zeros((n, n), complex)
    for c, pc in enumerate(xyz):
        for d, pd in enumerate(xyz):
            N = O.cuboid_n(h * pc, [h] * 3, h * pd, [h] * 3, k, rtol=5e-7, atol=5e-10, far_order=16)
            A[3 * c:3 * c + 3, 3 * d:3 * d + 3] = -(er - 1) * N + (er * np.eye(3) if c == d else 0)
    return A


@pytest.mark.parametrize("geom,h,k,er", [(HOLES, .25, .8, 3.0), (HOLES, .25, 0.0, 3.0),
                                          (FAR, .2, 4.0, 3.0), (HOLES, .25, .8, 3 - .4j),
                                          (((1, 1, 1), [0, 0, 0]), .25, .2, 3 - .4j)])
def test_operator_equals_dense_brute_force(geom, h, k, er):
    shape, xyz = geom
    vo = O.VolumeOperator(shape, xyz, h, k, er)
    A = dense_operator(shape, xyz, h, k, er)
    n = vo.order
    for i in range(n):
        e = np.zeros(n, complex)
        e[i] = 1 + 2j
        y = vo.apply(e)
        assert np.max(np.abs(y - A @ e) / np.maximum(1, np.abs(A @ e))) < 2e-8


def test_fft_convolution_is_circular_convolution():
    # brute-force circular convolution of the padded tensor with the source
    shape, xyz = HOLES
    vo = O.VolumeOperator(shape, xyz, .25, .8, 3.0)
    P = vo.spec.padded
    T = vo.spec.spatial
    rng = np.random.default_rng(2)
    x = rng.standard_normal(vo.order) + 1j * rng.standard_normal(vo.order)
    src = np.zeros((3,) + P, complex)
    xv = x.reshape(-1, 3)
    for c, p in enumerate(vo.xyz):
        src[:, p[0], p[1], p[2]] = xv[c]
    dst = np.zeros_like(src)
    for i in range(3):
        for j in range(3):
            t = T[O.tensor_index(i, j)]
            for c, p in enumerate(vo.xyz):
                for q, r in enumerate(vo.xyz):
                    d = tuple((p[a] - r[a]) % P[a] for a in range(3))
                    dst[i, p[0], p[1], p[2]] += t[d] * src[j, r[0], r[1], r[2]]
    direct = np.array([[dst[i, p[0], p[1], p[2]] for i in range(3)] for p in vo.xyz])
    assert np.max(np.abs(vo.convolve(x) - direct)) < 1e-13 * np.max(np.abs(direct))


def test_vacuum_is_identity():
    shape = (4, 3, 3)
    xyz = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, 3)
    vo = O.VolumeOperator(shape, xyz, .25, .7, 1.0)
    x = np.arange(vo.order) * (1 + .5j)
    assert np.max(np.abs(vo.apply(x) - x)) < 1e-14 * np.max(np.abs(x))


def test_spectrum_symmetry_even_odd():
    # the padded tensor is even (diagonal) / odd (off-diagonal) per axis
    s = O.KernelSpectrum((4, 3, 5), .1, 2.0)
    T = s.spatial
    P = s.padded
    neg = [np.r_[0, np.arange(p - 1, 0, -1)] for p in P]
    for c, (i, j) in enumerate(O.TENSOR_PAIRS):
        for ax in range(3):
            flipped = np.take(T[c], neg[ax], axis=ax)
            sign = -1 if (ax == i) != (ax == j) else 1
            assert np.max(np.abs(flipped - sign * T[c])) == 0


def test_golden_reference_products():
    with open(GOLDEN) as f:
        cases = json.load(f)["cases"]
    for c in cases:
        er = complex(*c["epsilon"])
        vo = O.VolumeOperator(c["shape"], c["xyz"], c["pitch"], c["k0"], er)
        x, y = cvec(c["x"]), cvec(c["y"])
        assert np.max(np.abs(vo.apply(x) - y)) < 1e-12 * np.max(np.abs(y)), c["name"]
        assert np.max(np.abs(vo.precondition(x) - cvec(c["precondition"]))) < 1e-14 * np.max(np.abs(x))
        sol = np.zeros_like(y)
        ok, rep = O.gmres(vo.apply, vo.precondition, y, sol, relative_tolerance=1e-11, maximum_iterations=150)
        assert ok and rep.iterations == c["iterations"] and rep.restarts == c["restarts"], c["name"]
        assert np.max(np.abs(sol - cvec(c["solve"]))) < 1e-9


def test_gmres_restatement_matches_reference_and_numpy(ref):
    rng = np.random.default_rng(4)
    n = 40
    A = np.diag(3 + 0.2j + rng.standard_normal(n)) + 0.3 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    b = A @ x
    okr, xr, repr_ = ref.dense_gmres(A, b, restart=7, rtol=1e-12)
    xo = np.zeros(n, complex)
    ok, rep = O.gmres(lambda v: A @ v, lambda v: v / np.diag(A), b, xo, restart=7, relative_tolerance=1e-12)
    assert okr and ok
    assert (rep.iterations, rep.restarts, rep.operator_applications) == \
        (repr_.iterations, repr_.restarts, repr_.operator_applications)
    assert np.max(np.abs(xo - xr)) < 1e-12
    assert np.max(np.abs(xo - np.linalg.solve(A, b))) < 1e-10


def test_hybrid_pec_matches_reference(ref):
    mats = [O.Material(1, 3, .2), O.Material(2, 2, .4)]
    for gap, pec in ((0.0, False), (0.0, True), (0.125, True)):
        cells = [O.Cell((.25 * i, -.25 * i, .125 + gap + .25 * i), (.25, .25, .25), i + 1) for i in range(2)]
        hs = O.HybridSystem(cells, mats, 0, .5, pec=pec)
        rs = ref.system(cells, mats, 0, .5, pec=pec)
        x = np.sin(1 + np.arange(6)) + 1j * np.cos(.3 * np.arange(6))
        assert np.max(np.abs(hs.apply(x) - rs.apply(x))) < 1e-13
        assert np.max(np.abs(hs.precondition(x) - rs.precondition(x))) < 1e-14


def test_hybrid_pec_dense_agreement(ref):
    # PEC image operator vs the reference's dense PEC cuboid kernel
    # (tests/dielectric_fft_test.c:8-46, tolerance 2e-7)
    mats = [O.Material(1, 3, .2), O.Material(2, 2, .4)]
    cells = [O.Cell((.25 * i, -.25 * i, .125 + .25 * i), (.25, .25, .25), i + 1) for i in range(2)]
    hs = O.HybridSystem(cells, mats, 0, .5, pec=True)
    x = np.sin(1 + np.arange(6)) + 1j * np.cos(.3 * np.arange(6))
    exp = np.zeros(6, complex)
    for i in range(2):
        e = mats[i].epsilon(.5)
        exp[3 * i:3 * i + 3] = e * x[3 * i:3 * i + 3]
        for j in range(2):
            N = ref.cuboid_pec_n(cells[i].center, cells[i].size, cells[j].center, cells[j].size, .5)
            exp[3 * i:3 * i + 3] -= (e - 1) * N @ x[3 * j:3 * j + 3]
    assert np.max(np.abs(hs.apply(x) - exp)) < 2e-7


def test_hybrid_conductor_blocks_and_rejections():
    mats = [O.Material(1, 3.0), O.Material(2, 2.0)]
    cells = [O.Cell((0, 0, 0), (.25,) * 3, 1), O.Cell((.25, 0, 0), (.25,) * 3, 1)]
    hs = O.HybridSystem(cells, mats, 2, .5)
    assert hs.order == 8
    with pytest.raises(O.PrepareError):
        O.HybridSystem([O.Cell((0, 0, 0), (.25,) * 3, 1), O.Cell((.3, 0, 0), (.25,) * 3, 1)], mats, 0, .5)
    with pytest.raises(O.PrepareError):
        O.HybridSystem([O.Cell((0, 0, 0), (.25,) * 3, 1), O.Cell((.25, 0, 0), (.25, .25, .125), 1)], mats, 0, .5)
    with pytest.raises(O.PrepareError):
        O.HybridSystem([O.Cell((0, 0, 0), (.25,) * 3, 7)], mats, 0, .5)
