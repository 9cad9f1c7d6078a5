"""Device solves at a size where every reduction runs grid-stride with more
blocks than keys (48^3 cells + 21-segment dipole, 331,797 unknowns): the
FP64 and mixed solutions must meet 1e-8 on the reference's own operator
(oracle/_ref) and agree on the feed impedance."""
import numpy as np
import pytest

from nec2c_amd import vie, workloads as W

pytestmark = pytest.mark.gpu


def test_coupled_48cube_fp64_and_mixed_meet_reference_residual(ref):
    n, pitch, f = 48, 0.5 / 64, 200e6
    k0 = W.k0_of(f)
    eps = W.random_permittivity((n, n, n), seed=7).reshape(-1)
    centers = W.cube_cells(n, pitch)
    dip = W.Dipole(center=(-0.1, 0.1875, 0.1875))
    s = vie.System(centers, [vie.MaterialSpec(1, 1.5)], dip.segments, k0, pitch=pitch)
    s.set_permittivities(eps)
    C = dip.impedance_matrix(k0)
    B, D = W.coupling_blocks(dip, centers, pitch, eps, k0)
    s.set_blocks(C=C, B=B, D=D)
    b = np.zeros(s.order, complex)
    b[:21] = dip.rhs()
    opts = dict(restart=120, maximum_iterations=2000, relative_tolerance=1e-8)
    ok64, x64, r64 = s.solve(b, precision=vie.FP64, **opts)
    ok32, x32, r32 = s.solve(b, precision=vie.MIXED, **opts)
    print("fp64", r64.as_dict(), "mixed", r32.as_dict(), flush=True)
    assert ok64 and ok32
    xyz = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    rp = ref.plan((n, n, n), xyz, pitch, k0, eps[0])
    rp.set_permittivities(eps)
    rc = rp.coupled(C, B, D)
    for x in (x64, x32):
        res = np.linalg.norm(b - rc.apply(x)) / np.linalg.norm(b)
        assert res <= 1.05e-8, res
    fi = dip.feed_index()
    assert abs(1 / x32[fi] - 1 / x64[fi]) / abs(1 / x64[fi]) < 1e-6
