"""Pins the oracle's kernel quadrature (oracle/vie_oracle.c): the reference's
independent golden tensors, closed forms and invariances of
tests/vie_kernel_test.c, and entry-for-entry agreement with the reference's
own vie_kernel.c / vie_fft.c build (oracle/_ref)."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "vie_reference_tensors.json")


def load_cases():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", load_cases(), ids=lambda c: c["name"])
def test_golden_tensors(case):
    # tolerance of tests/vie_kernel_test.c:226
    N = O.cuboid_n(case["test_center"], case["test_size"], case["source_center"], case["source_size"], case["k0"])
    assert N is not None
    exp = np.array([complex(*v) for v in case["tensor"]]).reshape(3, 3)
    assert np.all(np.abs(N - exp) <= 3e-10 + 3e-8 * np.abs(exp))


def test_static_cube_self_is_two_thirds():
    # depolarisation of a cube: N = diag(2/3) (tests/vie_kernel_test.c:58)
    N = O.cuboid_n([0, 0, 0], [1, 1, 1], [0, 0, 0], [1, 1, 1], 0.0)
    assert np.max(np.abs(N - np.diag([2 / 3] * 3))) < 2e-9
    assert O.relation([0, 0, 0], [1, 1, 1], [0, 0, 0], [1, 1, 1]) == 0


def test_static_rectangular_self_trace():
    N = O.cuboid_n([.2, -.4, .1], [.2, .5, .9], [.2, -.4, .1], [.2, .5, .9], 0.0)
    assert abs(np.trace(N) - 2.0) < 3e-9
    assert np.max(np.abs(N - np.diag(np.diag(N)))) < 2e-10


def test_translation_and_reciprocity():
    a, asz = np.array([0.1, -0.2, 0.3]), np.array([0.3, 0.2, 0.1])
    b, bsz = np.array([0.65, 0.15, -0.05]), np.array([0.2, 0.25, 0.15])
    ab = O.cuboid_n(a, asz, b, bsz, 2.3)
    ba = O.cuboid_n(b, bsz, a, asz, 2.3)
    shift = np.array([17.0, 14.0, 11.0])
    moved = O.cuboid_n(a + shift, asz, b + shift, bsz, 2.3)
    assert np.max(np.abs(ab - moved)) < 3e-12
    va, vb = np.prod(asz), np.prod(bsz)
    assert np.max(np.abs(va * ab - vb * ba.T)) < 2e-11


def test_axis_permutation():
    perm = [2, 0, 1]
    a, asz = np.array([0.0, 0.1, 0.2]), np.array([0.2, 0.3, 0.4])
    b, bsz = np.array([0.7, -0.2, 0.4]), np.array([0.1, 0.25, 0.35])
    orig = O.cuboid_n(a, asz, b, bsz, 1.7)
    pn = O.cuboid_n(a[perm], asz[perm], b[perm], bsz[perm], 1.7)
    assert np.max(np.abs(pn - orig[np.ix_(perm, perm)])) < 2e-11


def test_static_limit_is_continuous():
    a, sz = [0.0, 0.0, 0.0], [0.2, 0.3, 0.4]
    b = [0.2, 0.0, 0.0]
    assert O.relation(a, sz, b, sz) == 1
    z = O.cuboid_n(a, sz, b, sz, 0.0)
    s = O.cuboid_n(a, sz, b, sz, 1e-8)
    assert np.max(np.abs(z - s)) < 2e-9


def test_far_rule_agrees_with_face_pair_quadrature():
    # two independent quadratures of the same far-pair integral
    h, k = 0.1, 3.0
    for v in ([4, 0, 0], [5, 3, 1], [3, 3, 3]):
        far = O.far_kernel(v, h, k)
        assert far is not None
        direct = O.cuboid_n(np.array(v) * h, [h] * 3, [0, 0, 0], [h] * 3, k, rtol=1e-12, atol=1e-14,
                            far_order=48, max_order=48)
        assert np.max(np.abs(far - direct)) < 1e-8 * np.max(np.abs(direct)) + 1e-12


def test_near_offsets_are_not_far():
    # gap^2 <= 4 goes to the certified cuboid integral (vie_fft.c:323-324)
    assert O.far_kernel([3, 1, 1], 0.1, 1.0) is None
    assert O.far_kernel([2, 2, 2], 0.1, 1.0) is None
    assert O.far_kernel([4, 0, 0], 0.1, 1.0) is not None


def test_matches_reference_build_exactly(ref):
    pairs = [([0, 0, 0], [.25] * 3, [0, 0, 0], [.25] * 3, 0.8),
             ([.25, 0, 0], [.25] * 3, [0, 0, 0], [.25] * 3, 0.8),
             ([.25, .25, 0], [.25] * 3, [0, 0, 0], [.25] * 3, 0.8),
             ([.5, .25, .25], [.25] * 3, [0, 0, 0], [.25] * 3, 2.0),
             ([.1, -.2, .3], [.3, .2, .1], [.48, .03, .12], [.2, .25, .15], 3.2)]
    for tc, ts, sc, ss, k in pairs:
        for kw in (dict(), dict(rtol=5e-7, atol=5e-10, far_order=16)):
            a = O.cuboid_n(tc, ts, sc, ss, k, **kw)
            b = ref.cuboid_n(tc, ts, sc, ss, k, **kw)
            assert np.max(np.abs(a - b)) <= 1e-15 * max(1.0, np.max(np.abs(b)))


def test_offset_table_matches_reference_spectrum(ref):
    # the whole spectrum pipeline: oracle operator vs the reference operator
    shape = (6, 4, 3)
    xyz = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, 3)
    vo = O.VolumeOperator(shape, xyz, 0.05, 3.0, 4 - 0.5j)
    rp = ref.plan(shape, xyz, 0.05, 3.0, 4 - 0.5j)
    x = np.random.default_rng(0).standard_normal(vo.order) + 1j
    assert np.max(np.abs(vo.apply(x) - rp.apply(x))) < 1e-13 * np.max(np.abs(rp.apply(x)))
