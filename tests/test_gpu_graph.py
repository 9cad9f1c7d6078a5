"""Recorded Arnoldi steps (VIE_HIP_GRAPH=1): the matvec step of a device solve
is recorded once as a hipGraph and replayed for every Arnoldi step of every
cycle, with the cycle starts, restarts and true-residual checks launched
directly between the replays.  The replayed solve must equal the one with
direct launches bit for bit (the same kernels in the same order), in FP64
and in mixed precision, across several restart cycles, and a second solve
on the same workspace must replay the recorded step again with the same
result."""
import numpy as np
import pytest

from nec2c_amd import vie, workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def coupled():
    shape, pitch, f = (14, 10, 12), 0.02, 300e6
    k0 = W.k0_of(f)
    centers = W.cube_cells(shape, pitch)
    eps = W.random_permittivity(shape, seed=5).reshape(-1)
    dipole = W.Dipole(center=(-0.05, 0.1, 0.12), length=0.3, segments=11)
    s = vie.System(centers, [vie.MaterialSpec(1, 4.0)], dipole.segments, k0, pitch=pitch)
    s.set_permittivities(eps)
    B, D = W.coupling_blocks(dipole, centers, pitch, eps, k0)
    s.set_blocks(C=dipole.impedance_matrix(k0), B=B, D=D)
    b = np.zeros(s.order, dtype=np.complex128)
    b[: dipole.segments] = dipole.rhs()
    yield s, b
    s.close()


@pytest.mark.parametrize("precision", [vie.FP64, vie.MIXED])
def test_recorded_steps_equal_direct_launches(coupled, precision, monkeypatch):
    s, b = coupled
    # restart 12: many cycles, so the recorded step is replayed across restarts
    opts = dict(restart=12, maximum_iterations=2000, relative_tolerance=1e-8)
    monkeypatch.delenv("VIE_HIP_GRAPH", raising=False)
    ok0, x0, rep0 = s.solve(b, precision=precision, **opts)
    monkeypatch.setenv("VIE_HIP_GRAPH", "1")
    ok1, x1, rep1 = s.solve(b, precision=precision, **opts)
    ok2, x2, rep2 = s.solve(b, precision=precision, **opts)
    monkeypatch.delenv("VIE_HIP_GRAPH")
    print(f"precision {precision}: direct {rep0.iterations} it ({rep0.restarts} restarts), "
          f"recorded {rep1.iterations} / {rep2.iterations} it, residual {rep1.relative_residual:.2e}", flush=True)
    assert ok0 and ok1 and ok2
    assert rep0.restarts >= 3
    assert rep1.iterations == rep0.iterations and rep2.iterations == rep0.iterations
    assert np.array_equal(x1, x0) and np.array_equal(x2, x0)
